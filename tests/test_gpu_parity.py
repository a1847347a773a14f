"""GPU parity: the HIP path (through the C ABI) against the FP64 oracle.

Bars: the plain FP64 form within 1e-10 relative (image error measured as
max|delta| / max|reference|); the mixed form within 1e-5 for images and 1e-6
relative for the log-likelihood; deflections in the mixed form within 2e-6 of
the deflection scale.
"""

import numpy as np
import numpy.testing as npt
import pytest
import torch

from jaxtronomy_amd import workloads as W
from jaxtronomy_amd.Data.imaging_data import ImageData
from jaxtronomy_amd.Data.psf import PSF
from jaxtronomy_amd.ImSim.image_model import ImageModel
from jaxtronomy_amd.LensModel.lens_model import LensModel
from jaxtronomy_amd.LightModel.light_model import LightModel
from oracle import lensim_oracle as O

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10
MIXED_IMG_TOL = 1e-5
MIXED_LOGL_TOL = 1e-6


def _np(t):
    return t.detach().cpu().numpy() if torch.is_tensor(t) else np.asarray(t)


def img_err(got, want):
    return float(np.max(np.abs(got - want)) / np.max(np.abs(want)))


def _points(seed=0, n=500, r=3.0):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-r, r, n)
    y = rng.uniform(-r, r, n)
    # exact centres, axes and the origin are edge cases of every profile
    x[:4] = [0.0, 0.0, 1e-12, 0.3]
    y[:4] = [0.0, 0.5, 0.0, 0.0]
    return x, y


LENS_CASES = [
    ("EPL", dict(theta_E=1.2, gamma=1.9, e1=0.1, e2=-0.05, center_x=0.1, center_y=-0.2)),
    ("EPL", dict(theta_E=1.0, gamma=2.2, e1=0.3, e2=0.2, center_x=0.0, center_y=0.0)),
    ("EPL", dict(theta_E=1.0, gamma=2.0, e1=0.0, e2=0.0, center_x=0.0, center_y=0.0)),  # f == 0
    ("EPL", dict(theta_E=0.9, gamma=1.7, e1=-0.55, e2=0.45, center_x=0.0, center_y=0.0)),  # nmax 200-500
    ("SIE", dict(theta_E=1.3, e1=0.12, e2=0.07, center_x=0.0, center_y=0.0)),
    ("SIE", dict(theta_E=1.3, e1=0.0, e2=0.0, center_x=0.05, center_y=0.0)),  # q clamp
    ("NIE", dict(theta_E=1.1, e1=-0.2, e2=0.1, s_scale=0.2, center_x=0.0, center_y=0.1)),
    ("NFW", dict(Rs=0.8, alpha_Rs=0.35, center_x=0.0, center_y=0.0)),
    ("NFW", dict(Rs=0.5, alpha_Rs=0.35, center_x=0.5, center_y=0.0)),  # X == 1 at (0, 0)
    ("TNFW", dict(Rs=0.4, alpha_Rs=0.2, r_trunc=2.0, center_x=0.1, center_y=0.1)),
    ("SHEAR", dict(gamma1=0.03, gamma2=-0.02, ra_0=0.1, dec_0=0.0)),
    ("CONVERGENCE", dict(kappa=-0.03, ra_0=0.1, dec_0=-0.2)),
]


def defl_err(bx, by, x, y, ax, ay):
    """max |beta - beta_ref| / max |alpha_ref|: deflection error on the scale of
    the deflection field (pointwise relative errors are meaningless where the
    reference formula itself cancels, e.g. NFW's g(X) at its X = 1e-6 clamp)."""
    scale = max(np.max(np.abs(ax)), np.max(np.abs(ay)), 1e-300)
    return max(np.max(np.abs(_np(bx) - (x - ax))), np.max(np.abs(_np(by) - (y - ay)))) / scale


@pytest.mark.parametrize("name,kw", LENS_CASES)
def test_ray_shooting_fp64(name, kw):
    x, y = _points()
    lm = LensModel([name])
    bx, by = lm.ray_shooting(x, y, [kw])
    func, names = O.LENS_FUNCS[name]
    ax, ay = func(x, y, *[kw.get(n, 0.0) for n in names])
    assert defl_err(bx, by, x, y, ax, ay) < FP64_TOL
    # away from the ill-conditioned centre every ray matches pointwise
    far = np.hypot(x - kw.get("center_x", kw.get("ra_0", 0.0)), y - kw.get("center_y", kw.get("dec_0", 0.0))) > 1e-3
    npt.assert_allclose(_np(bx)[far], (x - ax)[far], rtol=FP64_TOL, atol=1e-13)
    npt.assert_allclose(_np(by)[far], (y - ay)[far], rtol=FP64_TOL, atol=1e-13)


@pytest.mark.parametrize("idx", [4, 5, 6, 9])
def test_fp64_trig_deflections_radial_sweep(idx):
    """SIE / NIE and TNFW in FP64, whose arctangents, area tangents and
    logarithms are the library's own series: deflections on a radial sweep
    (1e-4 to 1e2 around the centre, all directions) within 1e-13 of the
    oracle's numpy field, on the scale of the field, beyond 1e-3 of the centre
    and, for the TNFW, beyond its scale radius: inside, its g(X) = O(X^2) is a
    difference of O(ln X) terms in the reference formula itself, which the
    oracle's numpy forms and the library's forms both carry, to 5e-11 (inside
    the 1e-10 bar that the whole sweep is held to)."""
    name, kw = LENS_CASES[idx]
    cx, cy = kw.get("center_x", 0.0), kw.get("center_y", 0.0)
    r = np.logspace(-4, 2, 3001)
    phi = np.linspace(0, 6 * np.pi, r.size)
    x, y = cx + r * np.cos(phi), cy + r * np.sin(phi)
    bx, by = LensModel([name]).ray_shooting(x, y, [kw])
    func, names = O.LENS_FUNCS[name]
    ax, ay = func(x, y, *[kw.get(n, 0.0) for n in names])
    assert defl_err(bx, by, x, y, ax, ay) < FP64_TOL
    far = r > (kw["Rs"] if name == "TNFW" else 1e-3)
    assert defl_err(_np(bx)[far], _np(by)[far], x[far], y[far], ax[far], ay[far]) < 1e-13


@pytest.mark.parametrize("name,kw", LENS_CASES)
def test_ray_shooting_mixed(name, kw):
    x, y = _points()
    lm = LensModel([name], precision="mixed")
    bx, by = lm.ray_shooting(x, y, [kw])
    func, names = O.LENS_FUNCS[name]
    ax, ay = func(x, y, *[kw.get(n, 0.0) for n in names])
    assert defl_err(bx, by, x, y, ax, ay) < 2e-6


def test_epl_q_sweep_batched():
    """Scenario of test/test_LensModel/test_Profiles/test_epl.py:190-206: 200 axis
    ratios in [0.03, 1] x 101 points, as one batch of 200 parameter sets."""
    y = np.linspace(-10, 10, 101)
    x = np.ones_like(y)
    qs = np.linspace(0.03, 1, 200)[::-1]
    c = (1 - qs) / (1 + qs)
    kw = dict(theta_E=1.7, gamma=1.9, e1=c, e2=np.zeros_like(c), center_x=0.0, center_y=0.0)
    bx, by = LensModel(["EPL"]).ray_shooting(x, y, [kw])
    assert bx.shape == (200, 101)
    for i in range(200):
        ax, ay = O.epl_derivatives(x, y, 1.7, 1.9, c[i], 0.0)
        npt.assert_allclose(_np(bx[i]), x - ax, rtol=FP64_TOL, atol=1e-12)
        npt.assert_allclose(_np(by[i]), y - ay, rtol=FP64_TOL, atol=1e-12)


@pytest.mark.parametrize("gamma", [1.5, 1.9, 2.3])
def test_epl_q_sweep_mixed(gamma):
    """The mixed form over the same axis-ratio sweep: its float correction
    delta = 2F1 - 1 switches from the z-series (f < 0.44) to the w-series and
    to the double long series (nmax >= 100) along the sweep; the deflection
    stays within 2e-6 of the deflection scale throughout, including points
    on the axes and next to the centre (|Z| = 1e-12 takes the direction of Z
    unclamped)."""
    y = np.concatenate([np.linspace(-10, 10, 101), [1e-12, 0.0]])
    x = np.concatenate([np.ones(101), [0.0, 1e-12]])
    qs = np.linspace(0.03, 1, 200)[::-1]
    c = (1 - qs) / (1 + qs)
    kw = dict(theta_E=1.7, gamma=gamma, e1=c, e2=np.zeros_like(c), center_x=0.0, center_y=0.0)
    bx, by = LensModel(["EPL"], precision="mixed").ray_shooting(x, y, [kw])
    for i in range(200):
        ax, ay = O.epl_derivatives(x, y, 1.7, gamma, c[i], 0.0)
        assert defl_err(bx[i], by[i], x, y, ax, ay) < 2e-6, (i, c[i])


def test_ray_shooting_profile_sum_and_k():
    x, y = _points(1)
    names = ["EPL", "SHEAR", "NFW"]
    kws = [LENS_CASES[0][1], LENS_CASES[10][1], LENS_CASES[7][1]]
    lm = LensModel(names)
    bx, by = lm.ray_shooting(x, y, kws)
    wx, wy = O.ray_shooting(x, y, names, kws)
    npt.assert_allclose(_np(bx), wx, rtol=FP64_TOL, atol=1e-13)
    npt.assert_allclose(_np(by), wy, rtol=FP64_TOL, atol=1e-13)
    bx1, _ = lm.ray_shooting(x, y, kws, k=1)
    wx1, _ = O.ray_shooting(x, y, ["SHEAR"], [kws[1]])
    npt.assert_allclose(_np(bx1), wx1, rtol=FP64_TOL, atol=1e-13)


LIGHT_CASES = [
    ("SERSIC_ELLIPSE", dict(amp=23.0, R_sersic=0.6, n_sersic=7, e1=0.05, e2=0.03, center_x=0.0, center_y=0.0)),
    ("SERSIC_ELLIPSE", dict(amp=3.0, R_sersic=0.1, n_sersic=1.2, e1=-0.3, e2=0.2, center_x=0.1, center_y=0.0)),
    ("SERSIC", dict(amp=31.0, R_sersic=0.1, n_sersic=2, center_x=0.0, center_y=0.0)),
    ("GAUSSIAN", dict(amp=2.0, sigma=0.3, center_x=0.2, center_y=-0.1)),
]


@pytest.mark.parametrize("precision,tol", [("fp64", FP64_TOL), ("mixed", 2e-5)])
@pytest.mark.parametrize("name,kw", LIGHT_CASES)
def test_surface_brightness(name, kw, precision, tol):
    x, y = _points(2)
    f = LightModel([name], precision=precision).surface_brightness(x, y, [kw])
    func, names = O.LIGHT_FUNCS[name]
    want = func(x, y, *[kw.get(n, 0.0) for n in names])
    assert img_err(_np(f), want) < tol


@pytest.mark.parametrize("n_sersic", [0.5, 1.0, 2.5, 4.0, 8.0])
def test_fp64_sersic_radial_sweep_pointwise(n_sersic):
    """The FP64 Sersic (its powers and exponentials by the library's own
    base-2 logarithm and exponential) point by point against the oracle's
    numpy exp / log over radii from 1e-6 to 2e3 R_sersic (past the 1000 R_sersic
    window): relative 1e-13 wherever the profile is above 1e-30 of its peak
    (further out the exponent's own rounding, |b_n (1 - frac^(1/n))| ulp, grows
    to the bound: 1e-11 down to 1e-250 of the peak)."""
    rs = 0.3
    r = rs * np.logspace(-6, np.log10(2000.0), 4001)
    phi = np.linspace(0, 2 * np.pi, r.size)
    x, y = r * np.cos(phi), r * np.sin(phi)
    kw = dict(amp=5.0, R_sersic=rs, n_sersic=n_sersic, e1=0.1, e2=-0.05, center_x=0.0, center_y=0.0)
    got = _np(LightModel(["SERSIC_ELLIPSE"]).surface_brightness(x, y, [kw]))
    func, names = O.LIGHT_FUNCS["SERSIC_ELLIPSE"]
    want = func(x, y, *[kw.get(n, 0.0) for n in names])
    rel = np.abs(got - want) / np.where(want > 0, want, 1.0)
    big, tail = want > 1e-30 * want.max(), want > 1e-250 * want.max()
    assert np.max(rel[big]) < 1e-13
    assert np.max(rel[tail]) < 1e-11
    assert np.all(got[~tail] < 1e-240 * want.max())


# ------------------------------------------------------------------ images --

def _small_model(psf_type, precision, draw=0):
    cfg = W.SMALL
    data = ImageData(**W.data_kwargs(cfg))
    if psf_type == "GAUSSIAN":
        psf = PSF("GAUSSIAN", fwhm=cfg["psf_fwhm"], truncation=cfg["psf_truncation"])
    else:
        psf = PSF("PIXEL", kernel_point_source=O.gaussian_kernel(cfg["psf_fwhm"], cfg["deltaPix"], 6))
    im = ImageModel(data, psf, LensModel(["SIE", "SHEAR"]), LightModel(["SERSIC_ELLIPSE"]),
                    kwargs_numerics={"supersampling_factor": cfg["supersampling"]}, precision=precision)
    return im, cfg


@pytest.mark.parametrize("psf_type", ["GAUSSIAN", "PIXEL"])
def test_small_config_images_fp64(psf_type):
    """20 seeded draws of the 60x60 SIE+SHEAR+Sersic scene, against the oracle."""
    im, cfg = _small_model(psf_type, "fp64")
    assert im.kernel.shape == (11, 11)
    g = W.grid_dict(cfg)
    for d in range(20):
        kl, ks = W.small_draw(d)
        got = _np(im.image(kl, ks))
        want = O.image(g, ["SIE", "SHEAR"], kl, ["SERSIC_ELLIPSE"], ks, [], [], psf_type=psf_type,
                       kernel=im.kernel)
        assert img_err(got, want) < FP64_TOL, d


def test_small_config_batch_equals_single_and_mixed():
    im, cfg = _small_model("GAUSSIAN", "fp64")
    imm, _ = _small_model("GAUSSIAN", "mixed")
    draws = [W.small_draw(d) for d in range(16)]
    kl = W.stack_kwargs([d[0] for d in draws])
    ks = W.stack_kwargs([d[1] for d in draws])
    batch = _np(im.image(kl, ks))
    batch_m = _np(imm.image(kl, ks))
    assert batch.shape == (16, 60, 60)
    for i in (0, 7, 15):
        single = _np(im.image(*draws[i]))
        npt.assert_array_equal(batch[i], single)
        assert img_err(batch_m[i], single) < MIXED_IMG_TOL


@pytest.mark.parametrize("lens", [["SIE", "SHEAR"], ["NFW", "SIE"]])
def test_small_config_mixed_loglike(lens):
    """The mixed form's logL within 1e-6 of the FP64 form's on SIE scenes with
    compact sources (draw 0 has R_sersic 0.15", n 3.7: there one float ulp of a
    systematic deflection error moves the logL by ~6e-7), each checked against
    the oracle's logL for the FP64 form."""
    cfg = W.SMALL
    g = W.grid_dict(cfg)
    gk = O.gaussian_kernel(cfg["psf_fwhm"], cfg["deltaPix"], cfg["psf_truncation"])
    for d in range(6):
        kl, ks = W.small_draw(d)
        if lens[0] == "NFW":
            kl = [dict(Rs=0.8, alpha_Rs=0.3, center_x=0.05, center_y=-0.02), kl[0]]
        want_img = O.image(g, lens, kl, ["SERSIC_ELLIPSE"], ks, [], [], psf_type="GAUSSIAN", kernel=gk)
        data = W.noisy(want_img, cfg["exposure_time"], cfg["background_rms"], np.random.default_rng(d))
        want = O.log_likelihood(want_img, data, background_rms=cfg["background_rms"],
                                exposure_map=cfg["exposure_time"])
        ll = {}
        for precision in ("fp64", "mixed"):
            im = ImageModel(ImageData(**W.data_kwargs(cfg, data)),
                            PSF("GAUSSIAN", fwhm=cfg["psf_fwhm"], truncation=cfg["psf_truncation"]),
                            LensModel(lens), LightModel(["SERSIC_ELLIPSE"]), precision=precision)
            ll[precision] = float(im.likelihood_data_given_model(kl, ks))
        assert abs(ll["fp64"] - want) < FP64_TOL * abs(want), d
        assert abs(ll["mixed"] - ll["fp64"]) < MIXED_LOGL_TOL * abs(ll["fp64"]), (d, ll)


def _epl180_model(precision):
    cfg = W.EPL180
    data = ImageData(**W.data_kwargs(cfg))
    psf = PSF("PIXEL", kernel_point_source=W.moffat_kernel(cfg["psf_size"], cfg["psf_fwhm"] / cfg["deltaPix"],
                                                           cfg["psf_beta"]))
    im = ImageModel(data, psf, LensModel(["EPL", "SHEAR"]), LightModel(["SERSIC_ELLIPSE"]),
                    LightModel(["SERSIC_ELLIPSE"]), kwargs_numerics={"supersampling_factor": cfg["supersampling"]},
                    precision=precision)
    return im, cfg


def test_epl180_images_fp64_and_mixed():
    im, cfg = _epl180_model("fp64")
    imm, _ = _epl180_model("mixed")
    g = W.grid_dict(cfg)
    for d in range(3):
        kl, ks, kll = W.epl180_draw(d)
        want = O.image(g, ["EPL", "SHEAR"], kl, ["SERSIC_ELLIPSE"], ks, ["SERSIC_ELLIPSE"], kll, kernel=im.kernel)
        got = _np(im.image(kl, ks, kll))
        assert img_err(got, want) < FP64_TOL, d
        got_m = _np(imm.image(kl, ks, kll))
        assert img_err(got_m, want) < MIXED_IMG_TOL, d


def test_source_lens_light_split_and_unconvolved():
    im, cfg = _epl180_model("fp64")
    g = W.grid_dict(cfg)
    kl, ks, kll = W.epl180_draw(5)
    src = _np(im.source_surface_brightness(ks, kl))
    want = O.image(g, ["EPL", "SHEAR"], kl, ["SERSIC_ELLIPSE"], ks, [], [], kernel=im.kernel)
    assert img_err(src, want) < FP64_TOL
    ll = _np(im.lens_surface_brightness(kll, unconvolved=True))
    want = O.image(g, [], [], [], [], ["SERSIC_ELLIPSE"], kll, unconvolved=True)
    assert img_err(ll, want) < FP64_TOL
    de = _np(im.source_surface_brightness(ks, kl, de_lensed=True))
    want = O.image(g, [], [], ["SERSIC_ELLIPSE"], ks, [], [], kernel=im.kernel)
    assert img_err(de, want) < FP64_TOL


def test_reference_image_model_scene():
    """Scenario of test/test_ImSim/test_image_model.py:28-110 without point sources:
    100x100 at 0.05", SIE+SHEAR, SERSIC lens light, SERSIC_ELLIPSE source,
    17x17 PIXEL kernel, supersampling 3."""
    from jaxtronomy_amd.Util import param_util
    from jaxtronomy_amd.Util.simulation_util import data_configure_simple

    kwargs_data = data_configure_simple(100, 0.05, 100, 0.05, inverse=True)
    kernel = np.zeros((17, 17))
    kernel[5:-5, 5:-5] = 1
    kernel[7:-7, 7:-7] = 3
    kernel[9, 9] = 7
    e1, e2 = param_util.phi_q2_ellipticity(0.2, 0.8)
    kwargs_lens = [dict(theta_E=1.0, center_x=0, center_y=0, e1=e1, e2=e2), dict(gamma1=0.01, gamma2=0.01)]
    e1s, e2s = param_util.phi_q2_ellipticity(0.2, 0.9)
    kwargs_source = [dict(amp=23.0, R_sersic=0.6, n_sersic=7, center_x=0, center_y=0, e1=e1s, e2=e2s)]
    kwargs_lens_light = [dict(amp=31.0, R_sersic=0.1, n_sersic=2, center_x=0, center_y=0)]
    im = ImageModel(ImageData(**kwargs_data), PSF("PIXEL", kernel_point_source=kernel), LensModel(["SIE", "SHEAR"]),
                    LightModel(["SERSIC_ELLIPSE"]), LightModel(["SERSIC"]),
                    kwargs_numerics={"supersampling_factor": 3})
    got = _np(im.image(kwargs_lens, kwargs_source, kwargs_lens_light))
    g = dict(nx=100, ny=100, s=3, ra_at_xy_0=kwargs_data["ra_at_xy_0"], dec_at_xy_0=kwargs_data["dec_at_xy_0"],
             transform_pix2angle=kwargs_data["transform_pix2angle"])
    want = O.image(g, ["SIE", "SHEAR"], kwargs_lens, ["SERSIC_ELLIPSE"], kwargs_source, ["SERSIC"],
                   kwargs_lens_light, kernel=kernel / kernel.sum())
    assert img_err(got, want) < FP64_TOL
    # likelihood of the model against noisy data drawn from it (image_model.py:162-221)
    rng = np.random.default_rng(7)
    data = W.noisy(want, 100.0, 0.05, rng)
    kwargs_data["image_data"] = data
    im2 = ImageModel(ImageData(**kwargs_data), PSF("PIXEL", kernel_point_source=kernel), LensModel(["SIE", "SHEAR"]),
                     LightModel(["SERSIC_ELLIPSE"]), LightModel(["SERSIC"]),
                     kwargs_numerics={"supersampling_factor": 3})
    logl = float(im2.likelihood_data_given_model(kwargs_lens, kwargs_source, kwargs_lens_light))
    want_ll = O.log_likelihood(want, data, background_rms=0.05, exposure_map=100.0)
    npt.assert_allclose(logl, want_ll, rtol=FP64_TOL)


# -------------------------------------------------------------- likelihood --

@pytest.mark.parametrize("precision,tol", [("fp64", FP64_TOL), ("mixed", MIXED_LOGL_TOL)])
def test_epl180_loglike_batch(precision, tol):
    cfg = W.EPL180
    g = W.grid_dict(cfg)
    kernel = W.moffat_kernel(cfg["psf_size"], cfg["psf_fwhm"] / cfg["deltaPix"], cfg["psf_beta"])
    truth = W.epl180_draw(0)
    model_truth = O.image(g, ["EPL", "SHEAR"], truth[0], ["SERSIC_ELLIPSE"], truth[1], ["SERSIC_ELLIPSE"],
                          truth[2], kernel=kernel)
    rng = np.random.default_rng(11)
    data = W.noisy(model_truth, cfg["exposure_time"], cfg["background_rms"], rng)
    mask = np.ones_like(data)
    mask[:10, :] = 0  # a masked strip
    im = ImageModel(ImageData(**W.data_kwargs(cfg, data)), PSF("PIXEL", kernel_point_source=kernel),
                    LensModel(["EPL", "SHEAR"]), LightModel(["SERSIC_ELLIPSE"]), LightModel(["SERSIC_ELLIPSE"]),
                    kwargs_numerics={"supersampling_factor": 3}, likelihood_mask=mask, precision=precision)
    draws = [truth] + [W.epl180_draw(d) for d in range(1, 4)]
    kl = W.stack_kwargs([d[0] for d in draws])
    ks = W.stack_kwargs([d[1] for d in draws])
    kll = W.stack_kwargs([d[2] for d in draws])
    logl = _np(im.likelihood_data_given_model(kl, ks, kll))
    for i, d in enumerate(draws):
        m = O.image(g, ["EPL", "SHEAR"], d[0], ["SERSIC_ELLIPSE"], d[1], ["SERSIC_ELLIPSE"], d[2], kernel=kernel)
        want = O.log_likelihood(m, data, mask, background_rms=cfg["background_rms"],
                                exposure_map=cfg["exposure_time"])
        assert abs(logl[i] - want) / abs(want) < tol, (i, logl[i], want)


@pytest.mark.parametrize("kw", [
    dict(amp=39.85, R_sersic=0.4769, n_sersic=2.408, e1=-0.0082, e2=0.0267, center_x=-0.0284, center_y=0.0510),
    dict(amp=60.47, R_sersic=1.2019, n_sersic=4.5, e1=-0.0205, e2=0.0947, center_x=0.0, center_y=0.0),
    dict(amp=3.0, R_sersic=0.12, n_sersic=1.0, e1=0.2, e2=-0.1, center_x=0.3, center_y=-0.2)])
def test_mixed_sersic_unbiased(kw):
    """The float Sersic's error against the FP64 oracle has no one-signed part:
    the mean signed relative error in each radial band stays below 4e-8 (a
    float b_n, amplitude or ellipse matrix, or v_log_f32 of a large |log2|,
    each leaves 1e-7 to 3e-7 of bias), and the rms below 4e-7."""
    n = 6000
    rng = np.random.default_rng(7)
    r = np.geomspace(0.003, 6.0, n) * kw["R_sersic"]
    phi = rng.uniform(0, 2 * np.pi, n)
    x, y = kw["center_x"] + r * np.cos(phi), kw["center_y"] + r * np.sin(phi)
    got = _np(LightModel(["SERSIC_ELLIPSE"], precision="mixed").surface_brightness(x, y, [kw]))
    func, names = O.LIGHT_FUNCS["SERSIC_ELLIPSE"]
    want = func(x, y, *[kw.get(nm, 0.0) for nm in names])
    rel = (got - want) / want
    q = r / kw["R_sersic"]
    for lo, hi in ((0.003, 0.1), (0.1, 0.5), (0.5, 1.0), (1.0, 2.0), (2.0, 6.0)):
        sel = (q >= lo) & (q < hi)
        assert abs(rel[sel].mean()) < 4e-8, (lo, hi, rel[sel].mean())
    assert np.sqrt(np.mean(rel ** 2)) < 4e-7


def test_mixed_loglike_perturbed_points():
    """Log-likelihoods of poor fits amplify any systematic model error: at 64
    points around the truth (120x120 s=2, 21x21 PSF, 18 free parameters
    perturbed), the mixed form stays within 1e-6 of the FP64 form, which itself
    matches the oracle."""
    cfg = dict(W.EPL180, numPix=120, supersampling=2)
    g = W.grid_dict(cfg)
    kernel = W.moffat_kernel(21, cfg["psf_fwhm"] / cfg["deltaPix"], cfg["psf_beta"])
    truth = W.epl180_draw(0)
    m = O.image(g, ["EPL", "SHEAR"], truth[0], ["SERSIC_ELLIPSE"], truth[1], ["SERSIC_ELLIPSE"], truth[2],
                kernel=kernel)
    data = W.noisy(m, cfg["exposure_time"], cfg["background_rms"], np.random.default_rng(3))

    def model(precision):
        return ImageModel(ImageData(**W.data_kwargs(cfg, data)), PSF("PIXEL", kernel_point_source=kernel),
                          LensModel(["EPL", "SHEAR"]), LightModel(["SERSIC_ELLIPSE"]), LightModel(["SERSIC_ELLIPSE"]),
                          kwargs_numerics={"supersampling_factor": 2}, precision=precision)

    im64, im32 = model("fp64"), model("mixed")
    base, _ = im64.pack(*truth)
    rng = np.random.default_rng(12)
    p = np.repeat(base, 64, axis=0)
    for k, i in [(0, 0), (0, 1), (0, 2), (0, 3), (0, 4), (0, 5), (1, 0), (1, 1), (2, 0), (2, 1), (2, 2), (2, 3),
                 (2, 4), (3, 0), (3, 1), (3, 2), (3, 3), (3, 4)]:
        p[:, k, i] += rng.normal(0, 0.02, 64) * max(abs(base[0, k, i]), 0.1)
    params = torch.as_tensor(p, device="cuda")
    l64 = _np(im64.likelihood_batch(params))
    l32 = _np(im32.likelihood_batch(params))
    assert np.max(np.abs(l32 - l64) / np.abs(l64)) < MIXED_LOGL_TOL
    for j in range(2):  # the packed rows back as kwargs (param_names order), through the oracle
        kw = [{nm: float(p[j, k, i]) for i, nm in enumerate(names)} for k, names in enumerate(im64._names)]
        mj = O.image(g, ["EPL", "SHEAR"], kw[:2], ["SERSIC_ELLIPSE"], kw[2:3], ["SERSIC_ELLIPSE"], kw[3:4],
                     kernel=kernel)
        want = O.log_likelihood(mj, data, background_rms=cfg["background_rms"], exposure_map=cfg["exposure_time"])
        assert abs(l64[j] - want) / abs(want) < FP64_TOL


def test_noise_map_likelihood_and_gaussian_psf():
    im, cfg = _small_model("GAUSSIAN", "fp64")
    g = W.grid_dict(cfg)
    kl, ks = W.small_draw(3)
    m = O.image(g, ["SIE", "SHEAR"], kl, ["SERSIC_ELLIPSE"], ks, [], [], psf_type="GAUSSIAN", kernel=im.kernel)
    rng = np.random.default_rng(5)
    noise_map = np.full(m.shape, 0.02) + 0.01 * rng.uniform(size=m.shape)
    data = m + rng.normal(0, 1, m.shape) * noise_map
    kw = W.data_kwargs(cfg, data)
    kw["noise_map"] = noise_map
    im2 = ImageModel(ImageData(**kw), PSF("GAUSSIAN", fwhm=0.1, truncation=6), LensModel(["SIE", "SHEAR"]),
                     LightModel(["SERSIC_ELLIPSE"]))
    logl = float(im2.likelihood_data_given_model(kl, ks))
    npt.assert_allclose(logl, O.log_likelihood(m, data, noise_map=noise_map), rtol=FP64_TOL)


# ------------------------------------------------------------ convolution --

@pytest.mark.parametrize("shape", [(3, 3), (5, 11), (11, 11), (17, 17), (21, 9), (41, 41), (61, 61), (15, 75),
                                   (101, 101)])
@pytest.mark.parametrize("precision,tol", [("fp64", FP64_TOL), ("mixed", 2e-6)])
def test_pixel_psf_shapes_asymmetric(shape, precision, tol):
    """Asymmetric (orientation-sensitive) and non-square kernels, both forms;
    the mixed form runs them on the matrix cores."""
    rng = np.random.default_rng(shape[0] * 100 + shape[1])
    kernel = rng.uniform(0.1, 1.0, size=shape)
    kernel[0, :] *= 3.0  # break the symmetry
    cfg = dict(W.SMALL, numPix=70, supersampling=2)
    g = W.grid_dict(cfg)
    im = ImageModel(ImageData(**W.data_kwargs(cfg)), PSF("PIXEL", kernel_point_source=kernel),
                    LensModel(["SIE", "SHEAR"]), LightModel(["SERSIC_ELLIPSE"]), LightModel(["GAUSSIAN"]),
                    kwargs_numerics={"supersampling_factor": 2}, precision=precision)
    kl, ks = W.small_draw(4)
    kll = [dict(amp=30.0, sigma=0.4, center_x=0.1, center_y=0.0)]
    got = _np(im.image(kl, ks, kll))
    want = O.image(g, ["SIE", "SHEAR"], kl, ["SERSIC_ELLIPSE"], ks, ["GAUSSIAN"], kll, kernel=kernel / kernel.sum())
    assert img_err(got, want) < tol


@pytest.mark.parametrize("n,batch", [(70, 2), (180, 64), (256, 64)])
def test_mfma_tile_widths(n, batch):
    """The matrix-core convolution picks its tile width (32, 48 or 64 columns)
    from the image size and batch so that blocks spread evenly over the CUs;
    these sizes take each width.  Every image of the mixed form stays within
    the mixed image bar of the FP64 form (itself pinned to the oracle)."""
    cfg = dict(W.EPL180, numPix=n, supersampling=1)
    kernel = W.moffat_kernel(41, cfg["psf_fwhm"] / cfg["deltaPix"], cfg["psf_beta"])
    draws = [W.epl180_draw(d) for d in range(batch)]
    kl = W.stack_kwargs([d[0] for d in draws])
    ks = W.stack_kwargs([d[1] for d in draws])
    kll = W.stack_kwargs([d[2] for d in draws])
    out = {}
    for precision in ("fp64", "mixed"):
        im = ImageModel(ImageData(**W.data_kwargs(cfg)), PSF("PIXEL", kernel_point_source=kernel),
                        LensModel(["EPL", "SHEAR"]), LightModel(["SERSIC_ELLIPSE"]), LightModel(["SERSIC_ELLIPSE"]),
                        precision=precision)
        out[precision] = _np(im.image(kl, ks, kll)).astype(np.float64)
    for i in range(batch):
        assert img_err(out["mixed"][i], out["fp64"][i]) < MIXED_IMG_TOL, i


def test_mixed_gaussian_psf_edge_padding_and_signed_kernel():
    cfg = W.SMALL
    g = W.grid_dict(cfg)
    kl, ks = W.small_draw(6)
    im = ImageModel(ImageData(**W.data_kwargs(cfg)), PSF("GAUSSIAN", fwhm=0.15, truncation=5),
                    LensModel(["SIE", "SHEAR"]), LightModel(["SERSIC_ELLIPSE"]), precision="mixed")
    want = O.image(g, ["SIE", "SHEAR"], kl, ["SERSIC_ELLIPSE"], ks, [], [], psf_type="GAUSSIAN", kernel=im.kernel)
    assert img_err(_np(im.image(kl, ks)), want) < 2e-6
    # a kernel with negative lobes, without normalisation
    k = np.outer([1.0, -0.5, 2.0, -0.5, 1.0], [0.5, 1.0, -1.0, 1.0, 0.25])
    im2 = ImageModel(ImageData(**W.data_kwargs(cfg)),
                     PSF("PIXEL", kernel_point_source=k, kernel_point_source_normalisation=False),
                     LensModel(["SIE", "SHEAR"]), LightModel(["SERSIC_ELLIPSE"]), precision="mixed")
    want = O.image(g, ["SIE", "SHEAR"], kl, ["SERSIC_ELLIPSE"], ks, [], [], kernel=k)
    # cancelling lobes amplify the float rounding of the unconvolved image: the
    # stated mixed-form bar applies
    assert img_err(_np(im2.image(kl, ks)), want) < MIXED_IMG_TOL


@pytest.mark.parametrize("variant", ["pixel_split", "pixel_full", "pixel_regular", "gaussian"])
@pytest.mark.parametrize("precision,tol,ltol", [("fp64", FP64_TOL, FP64_TOL), ("mixed", MIXED_IMG_TOL,
                                                                                  MIXED_LOGL_TOL)])
def test_supersampled_convolution(variant, precision, tol, ltol):
    """supersampling_convolution=True (ImSim/Numerics/numerics.py:103-116): the
    rays convolved with the supersampled kernel before the s x s mean, plus the
    split kernel's low-resolution part; image and logL against the oracle."""
    s = 3
    cfg = dict(W.EPL180, numPix=40, supersampling=s, deltaPix=0.08)
    g = W.grid_dict(cfg)
    kl, ks, kll = W.epl180_draw(5)
    numerics = {"supersampling_factor": s, "supersampling_convolution": True}
    if variant == "gaussian":
        psf = PSF("GAUSSIAN", fwhm=0.2, truncation=3)
        kernel_hr = O.gaussian_kernel(0.2, cfg["deltaPix"], 3, s, True)
        kernel_low = None
        psf_type = "GAUSSIAN"
    elif variant == "pixel_regular":
        # a kernel at the regular resolution, interpolated onto the supersampled grid (subgrid_kernel)
        psf = PSF("PIXEL", kernel_point_source=W.moffat_kernel(9, 2.0, 3.0))
        psf_type = "PIXEL"
        numerics["supersampling_kernel_size"] = None
        kernel_low, kernel_hr = None, psf.kernel_point_source_supersampled(s)
        assert kernel_hr.shape == (27, 27)
    else:
        k_super = W.moffat_kernel(7 * s, 2.0 * s, 3.0)
        psf = PSF("PIXEL", kernel_point_source=k_super, point_source_supersampling_factor=s)
        psf_type = "PIXEL"
        if variant == "pixel_split":
            numerics["supersampling_kernel_size"] = 3
            from jaxtronomy_amd.Data.psf import split_kernel
            kernel_low, kernel_hr = split_kernel(k_super / k_super.sum(), 3, s)
        else:
            numerics["supersampling_kernel_size"] = None
            kernel_low, kernel_hr = None, k_super / k_super.sum()
    want = O.image(g, ["EPL", "SHEAR"], kl, ["SERSIC_ELLIPSE"], ks, ["SERSIC_ELLIPSE"], kll, psf_type=psf_type,
                   kernel=kernel_hr, supersampling_convolution=True, kernel_low=kernel_low)
    data = W.noisy(want, cfg["exposure_time"], cfg["background_rms"], np.random.default_rng(11))
    im = ImageModel(ImageData(**W.data_kwargs(cfg, data)), psf, LensModel(["EPL", "SHEAR"]),
                    LightModel(["SERSIC_ELLIPSE"]), LightModel(["SERSIC_ELLIPSE"]), kwargs_numerics=numerics,
                    precision=precision)
    npt.assert_allclose(im.kernel, kernel_hr, rtol=1e-14, atol=1e-17)
    got = _np(im.image(kl, ks, kll)).astype(np.float64)
    assert img_err(got, want) < tol
    ll = float(im.likelihood_data_given_model(kl, ks, kll))
    ll_want = O.log_likelihood(want, data, background_rms=cfg["background_rms"], exposure_map=cfg["exposure_time"])
    assert abs(ll - ll_want) < ltol * abs(ll_want)
    # differs from the regular-grid convolution of the same PSF
    if variant != "gaussian":
        plain = O.image(g, ["EPL", "SHEAR"], kl, ["SERSIC_ELLIPSE"], ks, ["SERSIC_ELLIPSE"], kll,
                        kernel=psf.kernel_point_source)
        assert img_err(plain, want) > 1e-4


@pytest.mark.parametrize("psf_type", ["PIXEL", "GAUSSIAN"])
@pytest.mark.parametrize("precision,tol,ltol", [("fp64", FP64_TOL, FP64_TOL), ("mixed", MIXED_IMG_TOL,
                                                                                  MIXED_LOGL_TOL)])
def test_flux_evaluate_indexes_sub_frame(psf_type, precision, tol, ltol):
    """flux_evaluate_indexes: rays only in the mask's pixels, the model of the
    square sub-frame around them convolved there, zero elsewhere; image and logL
    (over the whole frame) against the oracle, with x_min != y_min."""
    cfg = dict(W.EPL180, numPix=40, supersampling=2, deltaPix=0.1)
    g = W.grid_dict(cfg)
    kl, ks, kll = W.epl180_draw(8)
    mask = np.zeros((40, 40), bool)
    yy, xx = np.mgrid[0:40, 0:40]
    mask[(yy - 17) ** 2 + (xx - 21) ** 2 < 13 ** 2] = True  # a disc: square bounding box 4..30 x 8..34
    if psf_type == "PIXEL":
        kernel = W.moffat_kernel(9, 2.0, 3.0)
        psf = PSF("PIXEL", kernel_point_source=kernel)
    else:
        psf = PSF("GAUSSIAN", fwhm=0.25, truncation=3)
        kernel = O.gaussian_kernel(0.25, cfg["deltaPix"], 3)
    want = O.image(g, ["EPL", "SHEAR"], kl, ["SERSIC_ELLIPSE"], ks, ["SERSIC_ELLIPSE"], kll, psf_type=psf_type,
                   kernel=kernel, flux_evaluate_indexes=mask)
    data = W.noisy(O.image(g, ["EPL", "SHEAR"], kl, ["SERSIC_ELLIPSE"], ks, ["SERSIC_ELLIPSE"], kll,
                           psf_type=psf_type, kernel=kernel), cfg["exposure_time"], cfg["background_rms"],
                   np.random.default_rng(4))
    im = ImageModel(ImageData(**W.data_kwargs(cfg, data)), psf, LensModel(["EPL", "SHEAR"]),
                    LightModel(["SERSIC_ELLIPSE"]), LightModel(["SERSIC_ELLIPSE"]),
                    kwargs_numerics={"supersampling_factor": 2, "flux_evaluate_indexes": mask}, precision=precision)
    assert im._frame == (5, 9, 25)  # rows 5..29, columns 9..33
    got = _np(im.image(kl, ks, kll)).astype(np.float64)
    assert got.shape == (40, 40)
    assert img_err(got, want) < tol
    assert np.all(got[:4] == 0)
    ll = float(im.likelihood_data_given_model(kl, ks, kll))
    ll_want = O.log_likelihood(want, data, background_rms=cfg["background_rms"], exposure_map=cfg["exposure_time"])
    assert abs(ll - ll_want) < ltol * abs(ll_want)


_PAIR_KWARGS = {
    "NFW": dict(Rs=0.8, alpha_Rs=0.6, center_x=0.1, center_y=-0.05),
    "SIE": dict(theta_E=1.1, e1=0.05, e2=-0.03, center_x=0.0, center_y=0.02),
    "NIE": dict(theta_E=1.0, e1=-0.04, e2=0.06, s_scale=0.05, center_x=0.01, center_y=0.0),
    "EPL": dict(theta_E=1.2, gamma=2.1, e1=0.08, e2=-0.02, center_x=0.0, center_y=0.0),
    "SHEAR": dict(gamma1=0.03, gamma2=-0.02),
}


@pytest.mark.gpu
@pytest.mark.parametrize("lenses", [("NFW", "SIE"), ("SIE", "NFW"), ("SIE",), ("EPL",), ("NIE", "SHEAR")])
@pytest.mark.parametrize("lens_light", [False, True])
@pytest.mark.parametrize("precision,tol", [("fp64", FP64_TOL), ("mixed", MIXED_IMG_TOL)])
def test_fast_render_deflector_lists(lenses, lens_light, precision, tol):
    """The deflector lists the compile-time render serves (one or two
    deflectors: EPL or SIE / NIE with or without SHEAR, NFW with SIE in either
    order), with and without lens light, against the oracle's image."""
    cfg = W.SMALL
    kernel = O.gaussian_kernel(cfg["psf_fwhm"], cfg["deltaPix"], 6)
    kl = [dict(_PAIR_KWARGS[n]) for n in lenses]
    ks = [dict(amp=20.0, R_sersic=0.3, n_sersic=2.5, e1=0.1, e2=-0.05, center_x=0.05, center_y=0.1)]
    kll = [dict(amp=8.0, R_sersic=0.8, n_sersic=3.5, e1=-0.05, e2=0.02, center_x=0.0, center_y=0.0)]
    im = ImageModel(ImageData(**W.data_kwargs(cfg)), PSF("PIXEL", kernel_point_source=kernel), LensModel(list(lenses)),
                    LightModel(["SERSIC_ELLIPSE"]), LightModel(["SERSIC_ELLIPSE"] if lens_light else []),
                    kwargs_numerics={"supersampling_factor": 2}, precision=precision)
    got = _np(im.image(kl, ks, kll if lens_light else None))
    g = dict(W.grid_dict(cfg), s=2)
    want = O.image(g, list(lenses), kl, ["SERSIC_ELLIPSE"], ks, ["SERSIC_ELLIPSE"] if lens_light else [],
                   kll if lens_light else [], kernel=kernel)
    assert img_err(got, want) < tol, (lenses, lens_light, precision, img_err(got, want))
