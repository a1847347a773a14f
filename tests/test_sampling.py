"""Param: parameter vectors <-> packed parameter blocks (no device work)."""

import numpy as np
import pytest

from jaxtronomy_amd import workloads as W
from jaxtronomy_amd.Data.imaging_data import ImageData
from jaxtronomy_amd.Data.psf import PSF
from jaxtronomy_amd.ImSim.image_model import ImageModel
from jaxtronomy_amd.LensModel.lens_model import LensModel
from jaxtronomy_amd.LightModel.light_model import LightModel
from jaxtronomy_amd.Sampling.likelihood import Param, prepare_logL_func


def _im():
    cfg = dict(W.EPL180, numPix=20, supersampling=1)
    return ImageModel(ImageData(**W.data_kwargs(cfg, np.zeros((20, 20)))), PSF("NONE"), LensModel(["EPL", "SHEAR"]),
                      LightModel(["SERSIC_ELLIPSE"]), LightModel(["SERSIC_ELLIPSE"]))


def test_param_ordering_fixed_and_round_trip():
    im = _im()
    p = Param(im, kwargs_fixed_lens=[{"center_x": 0.0, "center_y": 0.0}, {"ra_0": 0.0, "dec_0": 0.0}],
              kwargs_fixed_lens_light=[{"center_x": 0.0, "center_y": 0.0}],
              kwargs_lower_lens=[{"theta_E": 0.1, "gamma": 1.5}, {}], kwargs_upper_lens=[{"theta_E": 3.0}, {}])
    n, names = p.num_param()
    # EPL 4 free, SHEAR 2, source Sersic 7, lens light 5: the 18 free parameters of the gradient workload
    assert n == 18
    assert names[:6] == ["theta_E_lens0", "gamma_lens0", "e1_lens0", "e2_lens0", "gamma1_lens1", "gamma2_lens1"]
    kl, ks, kll = W.epl180_draw(0)
    args = p.kwargs2args(kl, ks, kll, image_model=im)
    packed = p.args2params(np.stack([args, args * 1.01]))
    want, _ = im.pack(kl, ks, kll)
    np.testing.assert_array_equal(packed[0], want[0])
    kw = p.params2kwargs(packed[0], im)
    assert kw["kwargs_lens"][0]["theta_E"] == kl[0]["theta_E"] and kw["kwargs_lens_light"][0]["center_x"] == 0.0
    lo, hi = p.param_limits()
    assert lo[0] == 0.1 and hi[0] == 3.0 and lo[1] == 1.5 and np.isinf(hi[1])
    with pytest.raises(ValueError):
        p.args2params(np.zeros((2, 5)))


def test_param_linear_solver_drops_amplitudes():
    im = _im()
    p = Param(im, linear_solver=True)
    n, names = p.num_param()
    assert n == 6 + 4 + 6 + 6 and not any(nm.startswith("amp_") for nm in names)
    f = prepare_logL_func("gpu", lambda a: np.sum(a, axis=1))
    np.testing.assert_array_equal(f(np.ones((3, 2))), [2.0, 2.0, 2.0])
    with pytest.raises(ValueError):
        prepare_logL_func("tpu", None)
