"""GPU parity of many-deflector ray shooting and images (LensModelBulk)."""

import numpy as np
import numpy.testing as npt
import pytest

from jaxtronomy_amd import workloads as W
from jaxtronomy_amd.Data.imaging_data import ImageData
from jaxtronomy_amd.Data.psf import PSF
from jaxtronomy_amd.ImSim.image_model import ImageModel
from jaxtronomy_amd.LensModel.lens_model_bulk import LensModelBulk
from jaxtronomy_amd.LightModel.light_model import LightModel
from oracle import lensim_oracle as O

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def test_single_plane_bulk_nfw_pair():
    """test/test_LensModel/test_single_plane_bulk.py:28-40: two NFW, NULL-padded to 10."""
    lens_model_list = ["NFW", "NFW"]
    kw = {"Rs": 1.3, "alpha_Rs": 2.18, "center_x": 0.1, "center_y": -2.1}
    lm = LensModelBulk(lens_model_list)
    x = np.tile(np.linspace(-5, 5, 20), 20)
    y = np.repeat(np.linspace(-5, 5, 20), 20)
    rs = lm.prepare_ray_shooting_kwargs(lens_model_list, [kw, kw], num_deflectors=10)
    bx, by = lm.ray_shooting(x, y, rs)
    wx, wy = O.ray_shooting(x, y, lens_model_list, [kw, kw])
    npt.assert_allclose(_np(bx), wx, atol=1e-12, rtol=1e-12)
    npt.assert_allclose(_np(by), wy, atol=1e-12, rtol=1e-12)
    with pytest.raises(ValueError):
        lm.prepare_ray_shooting_kwargs(lens_model_list + ["NFW"], [kw, kw])
    with pytest.raises(ValueError):
        lm.prepare_ray_shooting_kwargs(lens_model_list, [kw, kw], num_deflectors=1)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("mixed", 2e-6)])
def test_lens_model_bulk_mixed_types(precision, tol):
    """test/test_LensModel/test_lens_model_bulk.py (single plane): EPL, TNFW, TNFW, SHEAR."""
    lens_model_list = ["EPL", "TNFW", "TNFW", "SHEAR"]
    kwargs = [{"theta_E": 1.5, "gamma": 1.7, "e1": 0.1, "e2": 0.2},
              {"Rs": 1.3, "alpha_Rs": 2.18, "r_trunc": 1.5, "center_x": 0.1, "center_y": -2.1},
              {"Rs": 1.4, "alpha_Rs": 3.18, "r_trunc": 1.5, "center_x": -0.11, "center_y": 1.1},
              {"gamma1": 1.1, "gamma2": 0.3}]
    lm = LensModelBulk(list(set(lens_model_list)), precision=precision)
    x = np.tile(np.linspace(-5, 5, 20), 20)
    y = np.repeat(np.linspace(-5, 5, 20), 20)
    wx, wy = O.ray_shooting(x, y, lens_model_list, kwargs)
    for nd in (None, 6):
        rs = lm.prepare_ray_shooting_kwargs(lens_model_list, kwargs, num_deflectors=nd)
        bx, by = lm.ray_shooting(x, y, rs)
        scale = max(np.max(np.abs(x - wx)), np.max(np.abs(y - wy)))
        assert np.max(np.abs(_np(bx) - wx)) / scale < tol
        assert np.max(np.abs(_np(by) - wy)) / scale < tol


def _subhalos(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        Rs = rng.uniform(0.02, 0.15)
        out.append(dict(Rs=Rs, alpha_Rs=rng.uniform(0.002, 0.02), r_trunc=5 * Rs,
                        center_x=rng.uniform(-3, 3), center_y=rng.uniform(-3, 3)))
    return out


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("mixed", 2e-6)])
def test_many_tnfw_uniform_type(precision, tol):
    subs = _subhalos(300, 3)
    names = ["TNFW"] * len(subs)
    lm = LensModelBulk(["TNFW"], precision=precision)
    rng = np.random.default_rng(4)
    x = rng.uniform(-3, 3, 2000)
    y = rng.uniform(-3, 3, 2000)
    x[0], y[0] = subs[0]["center_x"], subs[0]["center_y"]  # a ray on a subhalo centre
    bx, by = lm.ray_shooting(x, y, lm.prepare_ray_shooting_kwargs(names, subs))
    wx, wy = O.ray_shooting(x, y, names, subs)
    scale = max(np.max(np.abs(x - wx)), np.max(np.abs(y - wy)))
    assert max(np.max(np.abs(_np(bx) - wx)), np.max(np.abs(_np(by) - wy))) / scale < tol


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("mixed", 1e-5)])
def test_substructure_image(precision, tol):
    """EPL + SHEAR main deflector and 200 TNFW subhalos, Sersic source, 60x60 s=2,
    Moffat PSF: bulk image against the oracle."""
    cfg = dict(W.EPL180, numPix=60, supersampling=2, deltaPix=0.05)
    g = W.grid_dict(cfg)
    kl, ks, _ = W.epl180_draw(7)
    subs = _subhalos(200, 9)
    names = ["EPL", "SHEAR"] + ["TNFW"] * len(subs)
    lm = LensModelBulk(["EPL", "SHEAR", "TNFW"], precision=precision)
    kernel = W.moffat_kernel(21, 2.0, 3.0)
    im = ImageModel(ImageData(**W.data_kwargs(cfg)), PSF("PIXEL", kernel_point_source=kernel), lm,
                    LightModel(["SERSIC_ELLIPSE"]), kwargs_numerics={"supersampling_factor": 2},
                    precision=precision)
    got = _np(im.image(lm.prepare_ray_shooting_kwargs(names, kl + subs, num_deflectors=256), ks))
    want = O.image(g, names, kl + subs, ["SERSIC_ELLIPSE"], ks, [], [], kernel=kernel)
    assert np.max(np.abs(got - want)) / np.max(np.abs(want)) < tol
