"""Gradients of the log-likelihood with a LensModelBulk deflector list (many
subhalos): the FP64 form against central differences of the C oracle's logL,
against central differences of its own FP64 logL on a list of several types,
in the kwargs form of ray_shooting_kwargs["all_kwargs"], and the mixed form
against the FP64 form at a substructure-sized scene."""

import numpy as np
import pytest
import torch

from jaxtronomy_amd import backend
from jaxtronomy_amd import workloads as W
from jaxtronomy_amd.Data.imaging_data import ImageData
from jaxtronomy_amd.Data.psf import PSF
from jaxtronomy_amd.ImSim.image_model import ImageModel
from jaxtronomy_amd.LensModel.lens_model_bulk import LensModelBulk
from jaxtronomy_amd.LightModel.light_model import LightModel

pytestmark = pytest.mark.gpu

# lensim_profile codes of the C oracle's type list
EPL, NFW, SIE, TNFW, SHEAR, SERSIC_ELLIPSE = 1, 4, 2, 5, 6, 33


def _ring_subhalos(n, seed, kinds=("TNFW",)):
    """n subhalos on the Einstein-ring annulus (0.7" < r < 1.7"), where the
    arc's flux makes their gradients large; kinds cycle through the list."""
    rng = np.random.default_rng(seed)
    names, out = [], []
    for q in range(n):
        kind = kinds[q % len(kinds)]
        r, phi = rng.uniform(0.7, 1.7), rng.uniform(0, 2 * np.pi)
        cx, cy = r * np.cos(phi), r * np.sin(phi)
        if kind == "SIE":
            kw = dict(theta_E=rng.uniform(0.02, 0.08), e1=rng.normal(0, 0.1), e2=rng.normal(0, 0.1),
                      center_x=cx, center_y=cy)
        else:
            Rs = rng.uniform(0.05, 0.2)
            kw = dict(Rs=Rs, alpha_Rs=rng.uniform(0.01, 0.05), center_x=cx, center_y=cy)
            if kind == "TNFW":
                kw["r_trunc"] = 5 * Rs
        names.append(kind)
        out.append(kw)
    return names, out


def _model(cfg, kernel, unique, precision, data=None):
    return ImageModel(ImageData(**W.data_kwargs(cfg, data)), PSF("PIXEL", kernel_point_source=kernel),
                      LensModelBulk(list(unique), precision=precision), LightModel(["SERSIC_ELLIPSE"]),
                      kwargs_numerics={"supersampling_factor": cfg["supersampling"]}, precision=precision)


def _scene(cfg, kernel, kinds, n_sub, seed, num_deflectors=None):
    """(names, truth kwargs, source kwargs, noisy data) of EPL + SHEAR + subhalos,
    the data from the FP64 image of the truth."""
    kl, ks, _ = W.epl180_draw(seed)
    kl = [dict(kl[0], center_x=0.0, center_y=0.0), kl[1]]
    sn, subs = _ring_subhalos(n_sub, seed + 100, kinds)
    names = ["EPL", "SHEAR"] + sn
    probe = _model(cfg, kernel, ("EPL", "SHEAR") + tuple(kinds), "fp64")
    rs = probe.LensModel.prepare_ray_shooting_kwargs(names, kl + subs, num_deflectors=num_deflectors)
    clean = probe.image(rs, ks).cpu().numpy()
    data = W.noisy(clean, cfg["exposure_time"], cfg["background_rms"], np.random.default_rng(seed + 7))
    return names, kl + subs, ks, data


def _perturb(kwargs, seed, frac=0.03):
    rng = np.random.default_rng(seed)
    out = []
    for kw in kwargs:
        d = dict(kw)
        for k in d:
            if k in ("ra_0", "dec_0"):
                continue
            d[k] = d[k] + rng.normal(0, frac) * max(abs(d[k]), 0.05)
        out.append(d)
    return out


def test_bulk_tnfw_grad_vs_c_oracle_fd():
    """EPL + SHEAR inline and 50 TNFW in the bulk loop, 40x40 at 0.08" s=2, 11x11
    Moffat: every parameter's FP64 gradient against central differences of the
    C oracle's logL of the same packed parameters."""
    from oracle import c_oracle

    cfg = dict(W.EPL180, numPix=40, deltaPix=0.08, supersampling=2)
    kernel = W.moffat_kernel(11, 1.5, 3.0)
    names, truth, ks, data = _scene(cfg, kernel, ("TNFW",), 50, 11)
    im = _model(cfg, kernel, ("EPL", "SHEAR", "TNFW"), "fp64", data)
    point = _perturb(truth, 12)
    rs = im.LensModel.prepare_ray_shooting_kwargs(names, point)
    params, n, m, keep = im.pack_bulk(rs, ks)
    assert m.n_lens == 2 and m.n_bulk == 50 and m.bulk_type == TNFW
    logl, grad = im.likelihood_and_grad_batch(backend.to_device_f64(params), n, model=m)
    del keep
    g = grad.cpu().numpy()[0]
    types = [EPL, SHEAR] + [TNFW] * 50 + [SERSIC_ELLIPSE]
    grid = W.grid_dict(cfg)

    def ll(p):
        return c_oracle.loglike(52, 1, 0, types, p, grid, data, kernel=kernel, exposure_time=cfg["exposure_time"],
                                background_rms=cfg["background_rms"])[0]

    base = ll(params)
    assert abs(float(logl[0]) - base) < 1e-10 * abs(base)
    slots = [(0, i) for i in range(6)] + [(1, 0), (1, 1)] + [(j, i) for j in range(2, 52) for i in range(5)] + \
            [(52, i) for i in range(7)]
    fd = np.empty(len(slots))
    for q, (j, i) in enumerate(slots):
        h = 1e-6 * max(abs(params[0, j, i]), 0.05)
        pp, pm = params.copy(), params.copy()
        pp[0, j, i] += h
        pm[0, j, i] -= h
        fd[q] = (ll(pp) - ll(pm)) / (2 * h)
    got = np.array([g[j, i] for (j, i) in slots])
    # each subhalo parameter kind against the largest of its kind (far subhalos
    # have tiny gradients whose differences sit at the roundoff of logL / h)
    kind = [("lens" if j < 2 else "src" if j == 52 else "sub%d" % i) for (j, i) in slots]
    for kd in set(kind):
        sel = np.array([k == kd for k in kind])
        scale = np.max(np.abs(fd[sel]))
        err = np.max(np.abs(got[sel] - fd[sel])) / scale
        assert err < 2e-6, (kd, err)
    big = np.abs(fd) > 1e-2 * np.max(np.abs(fd))
    assert np.max(np.abs(got[big] - fd[big]) / np.abs(fd[big])) < 1e-4
    # the subhalo gradients matter: the largest is a sizeable fraction of the main deflector's
    assert np.max(np.abs(got[8:258])) > 1e-3 * np.max(np.abs(got[:8]))


def _fd_batch(im, params, m, slots, rel=1e-6):
    """Central differences of the GPU FP64 logL, all perturbations in one batch."""
    B = params.shape[0]
    assert B == 1
    reps = np.repeat(params, 2 * len(slots), axis=0)
    hs = []
    for q, (j, i) in enumerate(slots):
        h = rel * max(abs(params[0, j, i]), 0.05)
        reps[2 * q, j, i] += h
        reps[2 * q + 1, j, i] -= h
        hs.append(h)
    ll = im.likelihood_batch(backend.to_device_f64(reps), reps.shape[0], model=m).cpu().numpy()
    return (ll[0::2] - ll[1::2]) / (2 * np.array(hs))


def test_bulk_mixed_types_grad_vs_fd_and_kwargs_form():
    """EPL inline and 105 deflectors of three types (35 each of TNFW, NFW, SIE) in
    the bulk loop with a per-deflector type (more minority deflectors than the
    inline slots hold), NULL padding to 128: the FP64 gradient against central
    differences of the FP64 logL, and the kwargs form of the gradient."""
    cfg = dict(W.EPL180, numPix=48, deltaPix=0.07, supersampling=2)
    kernel = W.moffat_kernel(11, 1.5, 3.0)
    kinds = ("TNFW", "NFW", "SIE")
    n_sub = 105
    names, truth, ks, data = _scene(cfg, kernel, kinds, n_sub, 21, num_deflectors=128)
    im = _model(cfg, kernel, ("EPL", "SHEAR") + kinds, "fp64", data)
    # a list without the SHEAR (the NULL padding holds the rest): every non-EPL deflector is bulk
    names = ["EPL"] + names[2:]
    point = [_perturb(truth, 22)[0]] + _perturb(truth[2:], 23)
    rs = im.LensModel.prepare_ray_shooting_kwargs(names, point, num_deflectors=128)
    params, n, m, keep = im.pack_bulk(rs, ks)
    assert m.n_lens == 1 and m.n_bulk == n_sub and m.bulk_type == 0
    logl, grad = im.likelihood_and_grad_batch(backend.to_device_f64(params), n, model=m)
    g = grad.cpu().numpy()[0]
    npar = {"TNFW": 5, "NFW": 4, "SIE": 5}
    slots = [(0, i) for i in range(6)] + [(1 + d, i) for d in range(n_sub) for i in range(npar[names[1 + d]])] + \
            [(1 + n_sub, i) for i in range(7)]
    fd = _fd_batch(im, params, m, slots)
    del keep
    got = np.array([g[j, i] for (j, i) in slots])
    scale = np.max(np.abs(fd[6:-7]))
    assert np.max(np.abs(got[6:-7] - fd[6:-7])) / scale < 1e-5
    assert np.max(np.abs(got[:6] - fd[:6]) / np.abs(fd[:6])) < 1e-5
    big = np.abs(fd) > 1e-2 * np.max(np.abs(fd))
    assert np.max(np.abs(got[big] - fd[big]) / np.abs(fd[big])) < 1e-4
    # slots beyond each deflector's kwargs stay zero
    assert np.all(g[1:1 + n_sub, 6:] == 0)
    # kwargs form: {kwarg: [n_deflectors]} in the dict's deflector order, zero on padding
    logl2, glens, gsrc, glight = im.likelihood_gradient(rs, ks)
    assert float(logl2) == float(logl[0]) and glight == []
    assert set(glens) == set(rs["all_kwargs"])
    for kw in glens:
        assert glens[kw].shape == (128,) and np.all(glens[kw][1 + n_sub:] == 0)
    np.testing.assert_array_equal(glens["theta_E"][0], g[0, 0])
    for d in range(n_sub):
        row = g[1 + d]
        if names[1 + d] == "SIE":
            np.testing.assert_array_equal([glens[k][1 + d] for k in ("theta_E", "e1", "e2", "center_x", "center_y")],
                                          row[:5])
            assert glens["Rs"][1 + d] == 0 and glens["alpha_Rs"][1 + d] == 0
        else:
            want = row[:npar[names[1 + d]]]
            keys = ["Rs", "alpha_Rs"] + (["r_trunc"] if names[1 + d] == "TNFW" else []) + ["center_x", "center_y"]
            np.testing.assert_array_equal([glens[k][1 + d] for k in keys], want)
            assert glens["theta_E"][1 + d] == 0
    np.testing.assert_array_equal([gsrc[0][k] for k in ("amp", "R_sersic", "n_sersic", "e1", "e2", "center_x",
                                                         "center_y")], g[1 + n_sub, :7])


@pytest.mark.parametrize("n_sub", [200, 300])
def test_bulk_grad_mixed_vs_fp64(n_sub):
    """120x120 at 0.04" s=2, 21x21 Moffat, EPL + SHEAR + n_sub TNFW (300: two
    groups of 256 deflectors per block), batch of 8 points: the mixed form's
    gradient within 1e-4 of the FP64 form's, per parameter kind relative to
    that kind's gradient scale, and logL within 1e-6."""
    cfg = dict(W.EPL180, numPix=120, supersampling=2)
    kernel = W.moffat_kernel(21, cfg["psf_fwhm"] / cfg["deltaPix"], cfg["psf_beta"])
    names, truth, ks, data = _scene(cfg, kernel, ("TNFW",), n_sub, 31)
    im64 = _model(cfg, kernel, ("EPL", "SHEAR", "TNFW"), "fp64", data)
    im32 = _model(cfg, kernel, ("EPL", "SHEAR", "TNFW"), "mixed", data)
    pts = [_perturb(truth, 40 + b) for b in range(8)]
    lm = im64.LensModel
    rs = lm.stack_ray_shooting_kwargs([lm.prepare_ray_shooting_kwargs(names, p) for p in pts])
    src = W.stack_kwargs([ks] * 8)
    params, n, m, keep = im64.pack_bulk(rs, src)
    assert n == 8 and m.n_bulk == n_sub
    pt = backend.to_device_f64(params)
    l64, g64 = im64.likelihood_and_grad_batch(pt, n, model=m)
    l32, g32 = im32.likelihood_and_grad_batch(pt, n, model=m)
    del keep
    g64, g32 = g64.cpu().numpy(), g32.cpu().numpy()
    assert np.max(np.abs(((l32 - l64) / l64).cpu().numpy())) < 1e-6
    kinds = {"EPL": (0, range(6)), "SHEAR": (1, range(2)), "SRC": (2 + n_sub, range(7))}
    for name, (j, rng_i) in kinds.items():
        for i in rng_i:
            scale = np.max(np.abs(g64[:, j, i]))
            assert np.max(np.abs(g32[:, j, i] - g64[:, j, i])) / scale < 1e-4, (name, i)
    for i in range(5):
        a, b = g32[:, 2:2 + n_sub, i], g64[:, 2:2 + n_sub, i]
        assert np.max(np.abs(a - b)) / np.max(np.abs(b)) < 1e-4, ("TNFW", i)
