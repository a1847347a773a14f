"""Pins the FP64 oracle against independent mathematics (no GPU).

Each test checks the restatement against something that does not share its
formula: scipy/mpmath special functions, closed forms, invariants (divergence
of the deflection equals twice the convergence, curl-free deflection),
quadrature of the 3-D density, and brute-force or library convolution.
"""

import mpmath
import numpy as np
import numpy.testing as npt
import pytest
import scipy.integrate as integrate
import scipy.ndimage as ndimage
import scipy.signal as signal
import scipy.special as special

from oracle import lensim_oracle as O


# ---------------------------------------------------------------- hyp2f1 ---

def test_hyp2f1_lopez_temme_vs_scipy():
    # scenario of test/test_Util/test_hyp2f1_util.py:90-111
    a, b, c = 0.3, 0.7, 2.2
    z = np.array([0.3, 0.5, -0.6, 0.2, -3.2]) + 1j * np.array([0.5, 2.1, 0.1, -1.6, 1.1])
    npt.assert_allclose(O.hyp2f1_lopez_temme_8(a, b, c, z, nmax=75), special.hyp2f1(a, b, c, z),
                        rtol=1e-8, atol=1e-8)
    z = np.array([0.99 + 0.001j])
    npt.assert_allclose(O.hyp2f1_lopez_temme_8(a, b, c, z, nmax=200), special.hyp2f1(a, b, c, z),
                        rtol=1e-5, atol=1e-5)
    z = -10.99 + 0.001j
    npt.assert_allclose(O.hyp2f1_lopez_temme_8(a, b, c, z, nmax=75), special.hyp2f1(a, b, c, z),
                        rtol=1e-8, atol=1e-8)


@pytest.mark.parametrize("f", [0.05, 0.2, 0.5, 0.65, 0.8, 0.9])
@pytest.mark.parametrize("t", [0.7, 1.0, 1.3])
def test_hyp2f1_epl_branch_vs_mpmath(f, t):
    """The EPL call 2F1(1, t/2; 2-t/2; z), |z| = f, with the reference's nmax rule,
    agrees with a 50-digit evaluation."""
    mpmath.mp.dps = 50
    phis = np.linspace(-np.pi, np.pi, 7)
    z = f * np.exp(1j * phis)
    got = O.hyp2f1_lopez_temme_8(1.0, t / 2, 2 - t / 2, z, nmax=O.epl_nmax(f))
    want = np.array([complex(mpmath.hyp2f1(1, t / 2, 2 - t / 2, mpmath.mpc(zz.real, zz.imag))) for zz in z])
    npt.assert_allclose(got, want, rtol=2e-11, atol=0)


def test_epl_nmax_thresholds():
    assert O.epl_nmax(0.0) == 1
    assert [O.epl_nmax(f) for f in (0.1, 0.3, 0.5, 0.7, 0.8, 0.9, 0.95)] == [10, 20, 35, 50, 100, 200, 500]
    assert O.epl_nmax(0.12) == 20 and O.epl_nmax(0.4) == 35 and O.epl_nmax(0.92) == 500


# ----------------------------------------------------------- deflectors ---

def _div_curl(func, x, y, h=1e-5):
    fx_px, fy_px = func(x + h, y)
    fx_mx, fy_mx = func(x - h, y)
    fx_py, fy_py = func(x, y + h)
    fx_my, fy_my = func(x, y - h)
    div = (fx_px - fx_mx) / (2 * h) + (fy_py - fy_my) / (2 * h)
    curl = (fy_px - fy_mx) / (2 * h) - (fx_py - fx_my) / (2 * h)
    return div, curl


def _test_points(seed=0, n=40, rmin=0.2, rmax=3.0):
    rng = np.random.default_rng(seed)
    r = rng.uniform(rmin, rmax, n)
    ph = rng.uniform(-np.pi, np.pi, n)
    return r * np.cos(ph) + 0.05, r * np.sin(ph) - 0.03


def test_epl_round_closed_form():
    x, y = _test_points()
    theta_E, gamma = 1.3, 2.3
    ax, ay = O.epl_derivatives(x, y, theta_E, gamma, 0.0, 0.0, 0.05, -0.03)
    R = np.hypot(x - 0.05, y + 0.03)
    t = gamma - 1
    npt.assert_allclose(np.hypot(ax, ay), theta_E ** t * R ** (1 - t), rtol=1e-13)
    npt.assert_allclose(ax * (y + 0.03) - ay * (x - 0.05), 0, atol=1e-13)


@pytest.mark.parametrize("gamma,e1,e2", [(1.8, 0.05, -0.08), (2.0, 0.2, 0.1), (2.2, -0.3, 0.25),
                                          (1.9, 0.45, -0.3), (2.1, -0.02, 0.01)])
def test_epl_divergence_is_twice_kappa(gamma, e1, e2):
    theta_E, cx, cy = 1.2, 0.05, -0.03
    x, y = _test_points(1)
    func = lambda xx, yy: O.epl_derivatives(xx, yy, theta_E, gamma, e1, e2, cx, cy)
    div, curl = _div_curl(func, x, y)
    phi, q = O.ellipticity2phi_q(e1, e2)
    b, t = theta_E * np.sqrt(q), gamma - 1
    xr, yr = O.rotate(x - cx, y - cy, phi)
    kappa = (2 - t) / 2 * (b / np.hypot(q * xr, yr)) ** t
    npt.assert_allclose(div, 2 * kappa, rtol=2e-6)
    npt.assert_allclose(curl, 0, atol=2e-6)


@pytest.mark.parametrize("e1,e2", [(0.1, -0.05), (-0.25, 0.2)])
def test_epl_isothermal_equals_sie(e1, e2):
    x, y = _test_points(2)
    ax, ay = O.epl_derivatives(x, y, 1.1, 2.0, e1, e2, 0.1, 0.2)
    sx, sy = O.sie_derivatives(x, y, 1.1, e1, e2, 0.1, 0.2)
    npt.assert_allclose(ax, sx, rtol=1e-8, atol=1e-9)
    npt.assert_allclose(ay, sy, rtol=1e-8, atol=1e-9)


@pytest.mark.parametrize("e1,e2,s", [(0.1, 0.2, 1e-10), (-0.3, 0.05, 0.1), (0.02, -0.01, 0.3)])
def test_nie_divergence_is_twice_kappa(e1, e2, s):
    theta_E, cx, cy = 1.4, -0.1, 0.07
    x, y = _test_points(3)
    func = lambda xx, yy: O.nie_derivatives(xx, yy, theta_E, e1, e2, s, cx, cy)
    div, curl = _div_curl(func, x, y)
    b, ss, q, phi = O.nie_param_conv(theta_E, e1, e2, s)
    xr, yr = O.rotate(x - cx, y - cy, phi)
    kappa = b / 2.0 * (q ** 2 * (ss ** 2 + xr ** 2) + yr ** 2) ** (-0.5)
    npt.assert_allclose(div, 2 * kappa, rtol=1e-5)
    npt.assert_allclose(curl, 0, atol=1e-5)


def test_sie_round_limit_is_sis():
    x, y = _test_points(4)
    ax, ay = O.sie_derivatives(x, y, 1.3, 0.0, 0.0, 0.2, -0.1)
    sx, sy = O.sis_derivatives(x, y, 1.3, 0.2, -0.1)
    npt.assert_allclose(ax, sx, rtol=1e-7)
    npt.assert_allclose(ay, sy, rtol=1e-7)


def _m2d_from_density(rho, R):
    """Mass in a cylinder of radius R from a spherical density: M3d(R) plus the
    part of the outer shells inside the cylinder."""
    inner = integrate.quad(lambda r: 4 * np.pi * r ** 2 * rho(r), 0, R, limit=400, epsabs=0, epsrel=1e-12)[0]
    outer = integrate.quad(lambda r: 4 * np.pi * r ** 2 * rho(r) * (1 - np.sqrt(1 - (R / r) ** 2)),
                           R, np.inf, limit=400, epsabs=0, epsrel=1e-12)[0]
    return inner + outer


@pytest.mark.parametrize("R", [0.05, 0.3, 0.7, 1.0, 1.6, 4.0])
def test_nfw_deflection_from_projected_mass(R):
    Rs, alpha_Rs = 0.8, 0.35
    rho0 = O.nfw_alpha2rho0(alpha_Rs, Rs)
    rho = lambda r: rho0 / (r / Rs * (1 + r / Rs) ** 2)
    m2d = _m2d_from_density(rho, R)
    ax, ay = O.nfw_derivatives(np.array([R * 0.6]), np.array([R * 0.8]), Rs, alpha_Rs)
    npt.assert_allclose(np.hypot(ax, ay), m2d / (np.pi * R), rtol=1e-7)


def test_nfw_alpha_at_rs_and_exact_branch():
    Rs, alpha_Rs = 0.8, 0.35
    ax, ay = O.nfw_derivatives(np.array([Rs, Rs * (1 + 1e-6)]), np.array([0.0, 0.0]), Rs, alpha_Rs)
    npt.assert_allclose(ax, alpha_Rs, rtol=1e-5)  # X==1 branch and its neighbour
    assert ax[0] == pytest.approx(alpha_Rs, rel=1e-14)


@pytest.mark.parametrize("R", [0.05, 0.3, 1.0, 2.5, 7.0])
def test_tnfw_deflection_from_projected_mass(R):
    Rs, alpha_Rs, r_trunc = 0.6, 0.2, 3.0
    rho0 = O.nfw_alpha2rho0(alpha_Rs, Rs)
    rho = lambda r: r_trunc ** 2 / (r_trunc ** 2 + r ** 2) * rho0 / (r / Rs * (1 + r / Rs) ** 2)
    m2d = _m2d_from_density(rho, R)
    ax, ay = O.tnfw_derivatives(np.array([R * 0.8]), np.array([-R * 0.6]), Rs, alpha_Rs, r_trunc)
    npt.assert_allclose(np.hypot(ax, ay), m2d / (np.pi * R), rtol=1e-7)


def test_tnfw_large_truncation_is_nfw():
    x, y = _test_points(5)
    a = O.tnfw_derivatives(x, y, 0.5, 0.3, 1e5, 0.05, -0.03)
    b = O.nfw_derivatives(x, y, 0.5, 0.3, 0.05, -0.03)
    npt.assert_allclose(a, b, rtol=1e-4)


def test_shear_linear():
    x, y = np.array([1.0, -2.0, 0.5]), np.array([0.3, 1.0, -4.0])
    ax, ay = O.shear_derivatives(x, y, 0.03, -0.02, 0.1, 0.2)
    M = np.array([[0.03, -0.02], [-0.02, -0.03]])
    v = M @ np.vstack([x - 0.1, y - 0.2])
    npt.assert_allclose(ax, v[0], rtol=1e-15)
    npt.assert_allclose(ay, v[1], rtol=1e-15)


def test_convergence_mass_sheet():
    """A sheet of constant convergence: |alpha(R)| = M_2d(R) / (pi R) with M_2d
    integrated by quadrature (kappa pi R^2), radial from (ra_0, dec_0), and
    div alpha = 2 kappa by central differences."""
    from scipy.integrate import quad

    kappa, ra0, dec0 = -0.013, 0.4, -0.7
    rng = np.random.default_rng(3)
    x, y = rng.uniform(-3, 3, 50), rng.uniform(-3, 3, 50)
    ax, ay = O.convergence_derivatives(x, y, kappa, ra0, dec0)
    for xi, yi, axi, ayi in zip(x[:10], y[:10], ax[:10], ay[:10]):
        R = np.hypot(xi - ra0, yi - dec0)
        m2d = quad(lambda r: 2 * np.pi * r * kappa, 0, R)[0]
        npt.assert_allclose(np.hypot(axi, ayi), abs(m2d) / (np.pi * R), rtol=1e-12)
        npt.assert_allclose(axi * (yi - dec0) - ayi * (xi - ra0), 0.0, atol=1e-15)  # radial
        assert np.sign(axi * (xi - ra0) + ayi * (yi - dec0)) == np.sign(kappa)
    h = 1e-5
    div = ((O.convergence_derivatives(x + h, y, kappa, ra0, dec0)[0] -
            O.convergence_derivatives(x - h, y, kappa, ra0, dec0)[0]) +
           (O.convergence_derivatives(x, y + h, kappa, ra0, dec0)[1] -
            O.convergence_derivatives(x, y - h, kappa, ra0, dec0)[1])) / (2 * h)
    npt.assert_allclose(div, 2 * kappa, rtol=1e-8)


# ---------------------------------------------------------------- light ---

def _sersic_total_flux(amp, R_sersic, n):
    """Closed form of sersic_utils.py:197-215 written with scipy's gamma."""
    bn = 1.9992 * n - 0.3271
    return amp * R_sersic ** 2 * 2 * np.pi * n * np.exp(bn) / bn ** (2 * n) * special.gamma(2 * n)


@pytest.mark.parametrize("n", [1.0, 2.5, 4.0])
def test_sersic_round_total_flux(n):
    amp, Rs = 7.0, 0.3
    f = integrate.quad(lambda r: 2 * np.pi * r * O.sersic(np.array([r]), np.array([0.0]), amp, Rs, n)[0],
                       0, 1000 * Rs, limit=1000, points=[1e-4, Rs, 10 * Rs], epsrel=1e-10)[0]
    # flux inside the 1000 R_sersic window: total x P(2n, b_n 1000^(1/n))
    bn = 1.9992 * n - 0.3271
    tot = _sersic_total_flux(amp, Rs, n)
    inside = tot * special.gammainc(2 * n, bn * 1000.0 ** (1 / n))
    # inside R < 1e-4 the profile is flat at its value there (smoothing core)
    core_true = tot * special.gammainc(2 * n, bn * (1e-4 / Rs) ** (1 / n))
    core_flat = np.pi * 1e-8 * amp * np.exp(-bn * ((1e-4 / Rs) ** (1 / n) - 1))
    npt.assert_allclose(f, inside - core_true + core_flat, rtol=1e-8)


def test_sersic_ellipse_area_preserving_and_axes():
    amp, Rs, n, e1, e2 = 3.0, 0.4, 1.0, 0.15, -0.1
    g = np.linspace(-6, 6, 2401)
    xx, yy = np.meshgrid(g, g)
    f = O.sersic_elliptic(xx, yy, amp, Rs, n, e1, e2).sum() * (g[1] - g[0]) ** 2
    npt.assert_allclose(f, _sersic_total_flux(amp, Rs, n), rtol=1e-4)
    phi, q = O.ellipticity2phi_q(e1, e2)
    d = np.array([0.1, 0.3, 0.9])
    major = O.sersic_elliptic(d * np.cos(phi), d * np.sin(phi), amp, Rs, n, e1, e2)
    minor = O.sersic_elliptic(-q * d * np.sin(phi), q * d * np.cos(phi), amp, Rs, n, e1, e2)
    npt.assert_allclose(major, minor, rtol=1e-12)
    # the amplitude is the surface brightness at the product-average half-light radius
    on_ring = O.sersic_elliptic(np.array([Rs / np.sqrt(q) * np.cos(phi)]), np.array([Rs / np.sqrt(q) * np.sin(phi)]),
                                amp, Rs, n, e1, e2)
    npt.assert_allclose(on_ring, amp, rtol=1e-12)


def test_sersic_floors_and_window():
    # R below the smoothing scale is floored; beyond 1000 R_sersic the profile is zero
    a = O.sersic_elliptic(np.array([0.0, 5e-5]), np.array([0.0, 0.0]), 2.0, 0.2, 3.0, 0.0, 0.0)
    assert a[0] == a[1]
    z = O.sersic_elliptic(np.array([0.2 * 1000.5]), np.array([0.0]), 2.0, 0.2, 1.0, 0.0, 0.0)
    assert z[0] == 0.0


# --------------------------------------------------------------- numerics ---

def test_subgrid_means_equal_pixel_centres():
    M = np.array([[-0.05, 0.0], [0.0, 0.05]])
    for s in (1, 2, 3, 5):
        x, y = O.subgrid_coordinates(6, 6, M, 0.125, -0.125, s)
        xb = O.re_size(x.reshape(6 * s, 6 * s), s)
        yb = O.re_size(y.reshape(6 * s, 6 * s), s)
        ra, dec = O.coordinate_grid(6, 6, M, 0.125, -0.125)
        npt.assert_allclose(xb, ra, atol=1e-15)
        npt.assert_allclose(yb, dec, atol=1e-15)
        # x is the fastest index and the sub-pixel spacing is delta/s
        npt.assert_allclose(np.diff(x.reshape(6 * s, 6 * s), axis=1), -0.05 / s, rtol=1e-12)
        npt.assert_allclose(np.diff(y.reshape(6 * s, 6 * s), axis=0), 0.05 / s, rtol=1e-12)


def test_re_size_block_mean():
    rng = np.random.default_rng(0)
    a = rng.normal(size=(9, 12))
    r = O.re_size(a, 3)
    want = np.array([[a[3 * i:3 * i + 3, 3 * j:3 * j + 3].mean() for j in range(4)] for i in range(3)])
    npt.assert_allclose(r, want, rtol=1e-14)


def test_convolve_same_matches_fft_and_ndimage():
    rng = np.random.default_rng(1)
    img = rng.normal(size=(23, 23))
    ker = rng.uniform(size=(7, 5))  # asymmetric on purpose: orientation matters
    npt.assert_allclose(O.convolve_same(img, ker), signal.fftconvolve(img, ker, mode="same"), atol=1e-12)
    k2 = rng.uniform(size=(5, 5))
    npt.assert_allclose(O.convolve_same(img, k2, mode="edge"), ndimage.convolve(img, k2, mode="nearest"),
                        atol=1e-12)


def test_gaussian_convolution_matches_gaussian_filter():
    rng = np.random.default_rng(2)
    img = rng.uniform(size=(30, 30))
    fwhm, pix, trunc = 0.1, 0.05, 4.0
    k = O.gaussian_kernel(fwhm, pix, trunc)
    sigma_pix = fwhm / (2 * np.sqrt(2 * np.log(2))) / pix
    want = ndimage.gaussian_filter(img, sigma_pix, mode="nearest", truncate=trunc)
    npt.assert_allclose(O.gaussian_convolve(img, k), want, atol=1e-12)
    assert k.shape == (2 * round(sigma_pix * trunc) + 1,) * 2


def test_moffat_kernel_fwhm_and_norm():
    k = O.moffat_kernel(41, 2.0, 3.0)
    npt.assert_allclose(k.sum(), 1.0, rtol=1e-14)
    a = 2.0 / (2 * np.sqrt(2 ** (1 / 3) - 1))
    prof = lambda r: (1 + r ** 2 / a ** 2) ** -3.0
    npt.assert_allclose(prof(1.0) / prof(0.0), 0.5, rtol=1e-12)


def test_log_likelihood_hand_computed():
    model = np.array([[1.0, 2.0, -1.0], [0.0, 4.0, 0.5], [3.0, 1.0, 2.0]])
    data = np.array([[1.5, 2.0, 0.0], [0.2, 3.0, 0.5], [2.0, 1.5, 1.0]])
    mask = np.array([[1, 1, 1], [1, 0, 1], [1, 1, 1]])
    bkg, t = 0.5, 2.0
    total = 0.0
    for i in range(3):
        for j in range(3):
            var = max(model[i, j], 0.0) / t + bkg ** 2
            total += mask[i, j] * (model[i, j] - data[i, j]) ** 2 / var
    got = O.log_likelihood(model, data, mask, background_rms=bkg, exposure_map=t)
    npt.assert_allclose(got, -total / 2, rtol=1e-15)
    nm = np.full((3, 3), 0.3)
    got2 = O.log_likelihood(model, data, mask, noise_map=nm)
    npt.assert_allclose(got2, -np.sum(mask * (model - data) ** 2 / 0.09) / 2, rtol=1e-14)


# ----------------------------------------------------------------- image ---

def _grid(n, d, s):
    M = np.array([[-d, 0.0], [0.0, d]])
    return dict(nx=n, ny=n, transform_pix2angle=M, ra_at_xy_0=(n - 1) / 2 * d, dec_at_xy_0=-(n - 1) / 2 * d, s=s)


def test_image_flux_conservation_unlensed():
    g = _grid(80, 0.05, 3)
    kw = dict(amp=5.0, R_sersic=0.25, n_sersic=1.5, e1=0.1, e2=-0.05, center_x=0.02, center_y=-0.01)
    ker = O.moffat_kernel(11, 2.0, 3.0)
    img = O.image(g, [], [], ["SERSIC_ELLIPSE"], [kw], [], [], kernel=ker)
    npt.assert_allclose(img.sum(), _sersic_total_flux(5.0, 0.25, 1.5), rtol=2e-3)
    # unlensed source and lens light of the same profile are the same image
    img2 = O.image(g, [], [], [], [], ["SERSIC_ELLIPSE"], [kw], kernel=ker)
    npt.assert_allclose(img, img2, rtol=1e-12, atol=1e-15)


def test_image_point_mass_like_ring():
    """A round SIS lensing a round source centred behind it makes an image whose
    brightness depends only on the radius (ring symmetry of the composition)."""
    g = _grid(61, 0.05, 1)
    img = O.image(g, ["SIS"], [dict(theta_E=1.0)], ["SERSIC"], [dict(amp=1.0, R_sersic=0.1, n_sersic=1.0)],
                  [], [], psf_type="NONE")
    npt.assert_allclose(img, img.T, rtol=1e-12, atol=1e-15)
    npt.assert_allclose(img, img[::-1, ::-1], rtol=1e-12, atol=1e-15)
    ring = img[30, 30 + 20]  # r = 1.0 arcsec = theta_E: brightest
    assert ring == img.max()


def test_sub_frame_semantics():
    """NumericsSubFrame restated: outside the rectangle zero; deep inside equal to
    the full-frame model; unconvolved masked pixels zero; all-True equals no mask;
    the rectangle's grid starts at map_pix2coord(x_min, y_min) while it is placed
    at [x_min:, y_min:] (the reference's index order)."""
    from jaxtronomy_amd import workloads as W

    cfg = dict(W.EPL180, numPix=24, supersampling=2, deltaPix=0.12)
    g = W.grid_dict(cfg)
    kl, ks, kll = W.epl180_draw(3)
    kern = W.moffat_kernel(5, 1.5, 3.0)
    args = (g, ["EPL", "SHEAR"], kl, ["SERSIC_ELLIPSE"], ks, ["SERSIC_ELLIPSE"], kll)
    full = O.image(*args, kernel=kern)
    np.testing.assert_array_equal(O.image(*args, kernel=kern, flux_evaluate_indexes=np.ones((24, 24), bool)), full)
    mask = np.zeros((24, 24), bool)
    mask[4:20, 4:20] = True  # x_min = y_min: no transposition
    mask[10, 11] = False
    sub = O.image(*args, kernel=kern, flux_evaluate_indexes=mask)
    assert np.all(sub[:4] == 0) and np.all(sub[:, 20:] == 0)
    np.testing.assert_allclose(sub[6:8, 6:18], full[6:8, 6:18], rtol=1e-13)  # 2+ pixels from edge and hole
    un = O.image(*args, unconvolved=True, flux_evaluate_indexes=mask)
    assert un[10, 11] == 0 and un[10, 12] != 0
    # x_min != y_min: local pixel (i, j) is full-grid pixel (col x_min + j, row y_min + i),
    # written at [x_min + i, y_min + j]
    mask2 = np.zeros((24, 24), bool)
    mask2[2:12, 7:17] = True
    un2 = O.image(*args, unconvolved=True, flux_evaluate_indexes=mask2)
    un_full = O.image(*args, unconvolved=True)
    np.testing.assert_allclose(un2[2:12, 7:17], un_full[7:17, 2:12], rtol=1e-13)


def test_linear_solve_recovers_amplitudes():
    """get_param_WLS restated: noiseless data of known amplitudes are recovered
    exactly; the solution equals an independent weighted lstsq; the
    marginalisation constant is -1/2 log det M."""
    from jaxtronomy_amd import workloads as W

    cfg = dict(W.EPL180, numPix=30, supersampling=1, deltaPix=0.1)
    g = W.grid_dict(cfg)
    kl, ks, kll = W.epl180_draw(2)
    kern = W.moffat_kernel(7, 1.5, 3.0)
    clean = O.image(g, ["EPL", "SHEAR"], kl, ["SERSIC_ELLIPSE"], ks, ["SERSIC_ELLIPSE"], kll, kernel=kern)
    mask = np.ones((30, 30))
    mask[:3] = 0
    args = (g, ["EPL", "SHEAR"], kl, ["SERSIC_ELLIPSE"], ks, ["SERSIC_ELLIPSE"], kll)
    logl, B, model, M_inv = O.linear_solve(*args, clean, mask=mask, background_rms=0.01, exposure_map=1000.0,
                                           kernel=kern, source_marg=True)
    npt.assert_allclose(B, [ks[0]["amp"], kll[0]["amp"]], rtol=1e-10)
    npt.assert_allclose(model[3:], clean[3:], rtol=1e-9, atol=1e-12)
    assert np.all(model[:3] == 0)
    # noisy data: the WLS solution equals lstsq of the whitened system
    noisy = W.noisy(clean, 1000.0, 0.01, np.random.default_rng(0))
    logl, B, model, M_inv = O.linear_solve(*args, noisy, mask=mask, background_rms=0.01, exposure_map=1000.0,
                                           kernel=kern, source_marg=True)
    cols = [O.image(g, ["EPL", "SHEAR"], kl, ["SERSIC_ELLIPSE"], [dict(ks[0], amp=1.0)], [], [], kernel=kern),
            O.image(g, ["EPL", "SHEAR"], kl, [], [], ["SERSIC_ELLIPSE"], [dict(kll[0], amp=1.0)], kernel=kern)]
    sel = mask.ravel() > 0
    var = np.maximum(noisy, 0) / 1000.0 + 1e-4
    sw = 1 / np.sqrt(var.ravel()[sel])
    Aw = np.array([c.ravel()[sel] * sw for c in cols]).T
    B_ls = np.linalg.lstsq(Aw, noisy.ravel()[sel] * sw, rcond=None)[0]
    npt.assert_allclose(B, B_ls, rtol=1e-8)
    M = Aw.T @ Aw
    chi = O.log_likelihood(model, noisy, mask=mask, background_rms=0.01, exposure_map=1000.0)
    npt.assert_allclose(logl - chi, -0.5 * np.linalg.slogdet(M)[1], rtol=1e-8)
    npt.assert_allclose(M_inv, np.linalg.inv(M), rtol=1e-8)
