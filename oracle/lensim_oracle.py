"""FP64 CPU oracle for the strong-lensing image-simulation hot path.

TEST INFRASTRUCTURE ONLY.  This module is the checker: it is imported by
``tests/``, by ``__graft_entry__.smoke()`` and by the ``cpu_baseline`` leg of
``bench.py`` and by nothing else.  The product path (``jaxtronomy_amd``) never
imports it and fails loudly when its HIP library is missing.

It restates, in plain numpy float64 / complex128, the algorithm of JAXtronomy
0.1.3 for ray shooting -> Sersic light -> supersample pooling -> PSF
convolution -> chi^2 log-likelihood.  Every function cites the reference file
and lines it follows (paths relative to the reference's ``jaxtronomy/``
package).  The lenstronomy helpers that JAXtronomy imports but does not vendor
(coordinate grid, sub-grid, PSF normalisation) are restated from their public
behaviour and are pinned by invariants in ``tests/test_oracle.py``.

Pinning (how each function is checked independently of this file):
  hyp2f1_lopez_temme_8   scipy.special.hyp2f1 and mpmath.hyp2f1 (50 digits)
  epl_derivatives        q=1 closed form; div(alpha)=2 kappa by central FD;
                         gamma=2 equals the singular isothermal ellipse (NIE s->0)
  nie_derivatives        div(alpha)=2 kappa_NIE by FD; q->1 limit = SIS
  nfw/tnfw_derivatives   |alpha| = M_2d(R)/(pi R) with M_2d integrated from the
                         3-D density by scipy quadrature; alpha(Rs)=alpha_Rs
  shear_derivatives      exact linear form
  convergence_derivatives  |alpha(R)| = M_2d(R)/(pi R) of a uniform sheet
                         (quadrature), radial; div(alpha) = 2 kappa by FD
  sersic_*               closed-form flux integral vs fine-grid quadrature
  coordinate/sub grid    sub-pixel block means equal pixel centres, spacing
  re_size                exact block mean
  convolve_same          scipy.signal.fftconvolve(..., 'same')
  gaussian_convolve      scipy.ndimage.gaussian_filter(mode='nearest')
  log_likelihood         hand-computed 3x3 case
  gaussian_prior_logl    scipy.stats.norm.logpdf plus its normalisation
  likelihood_module_logl hand-built cases of each branch (bounds, NaN, +-inf)
  beta_double_source_plane  1 when the two sources coincide; the ratio of
                         deflections of an SIS whose Einstein radius is set by
                         the distances (scipy quadrature)
"""

import numpy as np

__all__ = [
    "ellipticity2phi_q", "phi_q2_ellipticity", "rotate",
    "hyp2f1_lopez_temme_8", "epl_nmax", "epl_major_axis_derivatives",
    "epl_derivatives", "nie_derivatives", "sie_derivatives", "sis_derivatives",
    "nfw_derivatives", "tnfw_derivatives", "shear_derivatives", "convergence_derivatives",
    "LENS_FUNCS", "LIGHT_FUNCS",
    "sersic_elliptic", "sersic", "gaussian_light", "b_n",
    "alpha", "ray_shooting", "surface_brightness",
    "coordinate_grid", "make_subgrid", "subgrid_coordinates", "re_size",
    "convolve_same", "gaussian_kernel", "gaussian_convolve",
    "moffat_kernel", "image", "log_likelihood", "c_d_model",
    "gaussian_prior_logl", "likelihood_module_logl", "beta_double_source_plane",
]

# --------------------------------------------------------------------------
# Util/param_util.py, Util/util.py
# --------------------------------------------------------------------------


def ellipticity2phi_q(e1, e2):
    """Util/param_util.py:88-100 (c clipped at 0.9999)."""
    phi = np.arctan2(e2, e1) / 2
    c = np.sqrt(e1 ** 2 + e2 ** 2)
    c = np.minimum(c, 0.9999)
    q = (1 - c) / (1 + c)
    return phi, q


def phi_q2_ellipticity(phi, q):
    """Util/param_util.py:74-84."""
    e1 = (1.0 - q) / (1.0 + q) * np.cos(2 * phi)
    e2 = (1.0 - q) / (1.0 + q) * np.sin(2 * phi)
    return e1, e2


def rotate(x, y, angle):
    """Util/util.py:103-114 (counter-clockwise rotation of the coordinates)."""
    c, s = np.cos(angle), np.sin(angle)
    return x * c + y * s, -x * s + y * c


# --------------------------------------------------------------------------
# Gauss hypergeometric series (Util/hyp2f1_util.py:136-176)
# --------------------------------------------------------------------------


def hyp2f1_lopez_temme_8(a, b, c, z, nmax=75):
    """Lopez & Temme (2013) eq. 8 series, Util/hyp2f1_util.py:136-176.

    Sum over n = 0..nmax-1 of (a)_n/n! * (z/(z-2))^n * F_n with F_n = 2F1(-n, b;
    c; 2) from Gauss' contiguous recurrence (DLMF 15.5.11), times (1-z/2)^-a.
    """
    z = np.asarray(z, dtype=complex)
    w = z / (z - 2.0)
    sum_prefactor = a * w
    prev_prev_F = 1.0
    prev_F = 1.0 - 2.0 * b / c
    total = 1 + sum_prefactor * prev_F
    for n in range(2, nmax):
        sum_prefactor = sum_prefactor * ((a + n - 1.0) / n * w)
        new_F = ((n - 1.0) * prev_prev_F - (2.0 * b - c) * prev_F) / (c + n - 1.0)
        total = total + sum_prefactor * new_F
        prev_prev_F = prev_F
        prev_F = new_F
    return (1.0 - z / 2.0) ** (-a) * total


def epl_nmax(f):
    """Number of series terms chosen from f=(1-q)/(1+q), LensModel/Profiles/epl.py:311-317.

    Returns 1 for f == 0 (the reference returns exactly one there).
    """
    if f == 0:
        return 1
    if f < 0.12:
        return 10
    if f < 0.4:
        return 20
    if f < 0.6:
        return 35
    if f < 0.73:
        return 50
    if f < 0.86:
        return 100
    if f < 0.92:
        return 200
    return 500


def epl_major_axis_derivatives(x, y, b, t, q):
    """LensModel/Profiles/epl.py:364-394 (Tessore & Metcalf 2015 eqs. 22-23)."""
    x = np.asarray(x, dtype=float)
    y = np.asarray(y, dtype=float)
    Z = q * x + y * 1j
    R = np.abs(Z)
    R = np.maximum(R, 0.000000001)
    f = (1.0 - q) / (1.0 + q)
    with np.errstate(all="ignore"):
        if f == 0:
            hyp = np.ones_like(Z)
        else:
            zz = -f * Z / np.conj(Z)
            hyp = hyp2f1_lopez_temme_8(1.0, t / 2.0, 2.0 - t / 2.0, zz, nmax=epl_nmax(f))
        R_omega = Z * hyp
        alpha = 2 / (1 + q) * (b / R) ** t * R_omega
        ar = np.nan_to_num(alpha.real, posinf=10 ** 10, neginf=-(10 ** 10))
        ai = np.nan_to_num(alpha.imag, posinf=10 ** 10, neginf=-(10 ** 10))
    return ar, ai


def epl_derivatives(x, y, theta_E, gamma, e1, e2, center_x=0.0, center_y=0.0):
    """EPL.derivatives, LensModel/Profiles/epl.py:178-200 with _param_conv :109-124."""
    t = gamma - 1
    phi_G, q = ellipticity2phi_q(e1, e2)
    b = theta_E * np.sqrt(q)
    x_, y_ = rotate(np.asarray(x, float) - center_x, np.asarray(y, float) - center_y, phi_G)
    fx_, fy_ = epl_major_axis_derivatives(x_, y_, b, t, q)
    return rotate(fx_, fy_, -phi_G)


# --------------------------------------------------------------------------
# NIE / SIE (LensModel/Profiles/nie.py, sie.py)
# --------------------------------------------------------------------------


def nie_param_conv(theta_E, e1, e2, s_scale):
    """NIE._param_conv, LensModel/Profiles/nie.py:182-200, :232-246."""
    phi_G, q = ellipticity2phi_q(e1, e2)
    theta_E_conv = theta_E / (np.sqrt((1.0 + q ** 2) / (2.0 * q)))
    b = theta_E_conv * np.sqrt((1 + q ** 2) / 2)
    s = s_scale / np.sqrt(q)
    return b, s, q, phi_G


def nie_major_axis_derivatives(x, y, b, s, q):
    """NIEMajorAxis.derivatives, LensModel/Profiles/nie.py:280-295 (Keeton & Kochanek 1998)."""
    q = 0.99999999 if q >= 1 else q
    psi = np.sqrt(q ** 2 * (s ** 2 + x ** 2) + y ** 2)
    sq = np.sqrt(1.0 - q ** 2)
    f_x = b / sq * np.arctan(sq * x / (psi + s))
    f_y = b / sq * np.arctanh(sq * y / (psi + q ** 2 * s))
    return f_x, f_y


def nie_derivatives(x, y, theta_E, e1, e2, s_scale, center_x=0.0, center_y=0.0):
    """NIE.derivatives, LensModel/Profiles/nie.py:86-107."""
    b, s, q, phi_G = nie_param_conv(theta_E, e1, e2, s_scale)
    x_, y_ = rotate(np.asarray(x, float) - center_x, np.asarray(y, float) - center_y, phi_G)
    fx_, fy_ = nie_major_axis_derivatives(x_, y_, b, s, q)
    return rotate(fx_, fy_, -phi_G)


def sie_derivatives(x, y, theta_E, e1, e2, center_x=0.0, center_y=0.0):
    """SIE.derivatives, LensModel/Profiles/sie.py:104-124 (NIE with s_scale=1e-10, :56)."""
    return nie_derivatives(x, y, theta_E, e1, e2, 0.0000000001, center_x, center_y)


def sis_derivatives(x, y, theta_E, center_x=0.0, center_y=0.0):
    """Singular isothermal sphere alpha = theta_E (x, y)/r (closed form, used as a
    q->1 pin for the NIE and EPL restatements)."""
    dx = np.asarray(x, float) - center_x
    dy = np.asarray(y, float) - center_y
    r = np.sqrt(dx ** 2 + dy ** 2)
    return theta_E * dx / r, theta_E * dy / r


# --------------------------------------------------------------------------
# NFW / TNFW (LensModel/Profiles/nfw.py, tnfw.py)
# --------------------------------------------------------------------------


def nfw_alpha2rho0(alpha_Rs, Rs):
    """NFW.alpha2rho0, LensModel/Profiles/nfw.py:384-395."""
    return alpha_Rs / (4.0 * Rs ** 2 * (1.0 + np.log(1.0 / 2.0)))


def nfw_g(X):
    """NFW.g, LensModel/Profiles/nfw.py:345-365 (exact X == 1 branch kept)."""
    c = 0.000001
    X = np.where(X < c, c, X)
    with np.errstate(all="ignore"):
        lt = np.log(X / 2.0) + 1 / np.sqrt(1 - X ** 2) * np.arccosh(1.0 / X)
        gt = np.log(X / 2) + 1 / np.sqrt(X ** 2 - 1) * np.arccos(np.minimum(1.0 / X, 1.0))
    return np.where(X < 1, lt, np.where(X == 1, 1 + np.log(1.0 / 2.0), gt))


def nfw_derivatives(x, y, Rs, alpha_Rs, center_x=0.0, center_y=0.0):
    """NFW.derivatives, LensModel/Profiles/nfw.py:86-105 and nfw_alpha :257-279."""
    rho0 = nfw_alpha2rho0(alpha_Rs, Rs)
    Rs = 0.0000001 if Rs < 0.0000001 else Rs
    x_ = np.asarray(x, float) - center_x
    y_ = np.asarray(y, float) - center_y
    R = np.sqrt(x_ ** 2 + y_ ** 2)
    R = np.maximum(R, 0.00000001)
    X = R / Rs
    gx = nfw_g(X)
    a = 4 * rho0 * Rs * gx / X ** 2
    return a * x_, a * y_


def tnfw_F(x):
    """TNFW.F, LensModel/Profiles/tnfw.py:288-304."""
    x = np.maximum(x, 0.001)
    with np.errstate(all="ignore"):
        lt = (1 - x ** 2) ** -0.5 * np.arctanh((1 - x ** 2) ** 0.5)
        gt = (x ** 2 - 1) ** -0.5 * np.arctan((x ** 2 - 1) ** 0.5)
    result = np.where(x < 1, lt, x)
    return np.where(x > 1, gt, result)


def tnfw_L(x, tau):
    """TNFW._L, LensModel/Profiles/tnfw.py:276-286."""
    x = np.maximum(x, 0.001)
    return np.log(x * (tau + np.sqrt(tau ** 2 + x ** 2)) ** -1)


def tnfw_g(x, tau):
    """TNFW._g, LensModel/Profiles/tnfw.py:329-349 (Baltz, Marshall & Oguri 2009)."""
    x = np.maximum(x, 0.001)
    return (
        tau ** 2 / (tau ** 2 + 1) ** 2
        * (
            (tau ** 2 + 1 + 2 * (x ** 2 - 1)) * tnfw_F(x)
            + tau * np.pi
            + (tau ** 2 - 1) * np.log(tau)
            + np.sqrt(tau ** 2 + x ** 2) * (-np.pi + tnfw_L(x, tau) * (tau ** 2 - 1) / tau)
        )
    )


def tnfw_derivatives(x, y, Rs, alpha_Rs, r_trunc, center_x=0.0, center_y=0.0):
    """TNFW.derivatives, LensModel/Profiles/tnfw.py:67-87 and tnfw_alpha :226-248."""
    rho0 = nfw_alpha2rho0(alpha_Rs, Rs)
    x_ = np.asarray(x, float) - center_x
    y_ = np.asarray(y, float) - center_y
    R = np.sqrt(x_ ** 2 + y_ ** 2)
    R = np.maximum(R, 0.001 * Rs)
    xx = R / Rs
    xx = np.maximum(xx, 0.001)
    tau = r_trunc / Rs
    gx = tnfw_g(xx, tau)
    a = 4 * rho0 * Rs * gx / xx ** 2
    return a * x_, a * y_


def shear_derivatives(x, y, gamma1, gamma2, ra_0=0.0, dec_0=0.0):
    """Shear.derivatives, LensModel/Profiles/shear.py:37-53."""
    x_ = np.asarray(x, float) - ra_0
    y_ = np.asarray(y, float) - dec_0
    return gamma1 * x_ + gamma2 * y_, gamma2 * x_ - gamma1 * y_


def convergence_derivatives(x, y, kappa, ra_0=0.0, dec_0=0.0):
    """Convergence.derivatives (a mass sheet), LensModel/Profiles/convergence.py:34-46."""
    return kappa * (np.asarray(x, float) - ra_0), kappa * (np.asarray(y, float) - dec_0)


# name -> (derivatives function, ordered param names as in each profile's
# ``param_names``: epl.py:50, sie.py:40, nie.py, nfw.py:48, tnfw.py:27, shear.py:16,
# convergence.py:15)
LENS_FUNCS = {
    "EPL": (epl_derivatives, ["theta_E", "gamma", "e1", "e2", "center_x", "center_y"]),
    "SIE": (sie_derivatives, ["theta_E", "e1", "e2", "center_x", "center_y"]),
    "NIE": (nie_derivatives, ["theta_E", "e1", "e2", "s_scale", "center_x", "center_y"]),
    "SIS": (sis_derivatives, ["theta_E", "center_x", "center_y"]),
    "NFW": (nfw_derivatives, ["Rs", "alpha_Rs", "center_x", "center_y"]),
    "TNFW": (tnfw_derivatives, ["Rs", "alpha_Rs", "r_trunc", "center_x", "center_y"]),
    "SHEAR": (shear_derivatives, ["gamma1", "gamma2", "ra_0", "dec_0"]),
    "CONVERGENCE": (convergence_derivatives, ["kappa", "ra_0", "dec_0"]),
}


def _call(func, names, x, y, kw):
    return func(x, y, *[kw.get(n, 0.0) for n in names])


def alpha(x, y, lens_model_list, kwargs_lens, alpha_scaling=1.0):
    """SinglePlane.alpha, LensModel/single_plane.py:112-139."""
    fx = np.zeros_like(np.asarray(x, float))
    fy = np.zeros_like(fx)
    for name, kw in zip(lens_model_list, kwargs_lens):
        func, names = LENS_FUNCS[name]
        ax, ay = _call(func, names, x, y, kw)
        fx = fx + ax
        fy = fy + ay
    return fx * alpha_scaling, fy * alpha_scaling


def ray_shooting(x, y, lens_model_list, kwargs_lens, alpha_scaling=1.0):
    """SinglePlane.ray_shooting, LensModel/single_plane.py:38-56."""
    x = np.asarray(x, float)
    y = np.asarray(y, float)
    dx, dy = alpha(x, y, lens_model_list, kwargs_lens, alpha_scaling)
    return x - dx, y - dy


# --------------------------------------------------------------------------
# Light profiles (LightModel/Profiles/sersic_ellipse.py, sersic.py, gaussian.py;
# LensModel/Profiles/sersic_utils.py)
# --------------------------------------------------------------------------


def b_n(n):
    """SersicUtil.b_n, LensModel/Profiles/sersic_utils.py:67-80."""
    return np.maximum(1.9992 * n - 0.3271, 0.00001)


def _r_sersic(R, R_sersic, n_sersic, max_R_frac=1000.0):
    """SersicUtil._r_sersic, LensModel/Profiles/sersic_utils.py:239-267."""
    R_ = np.maximum(0.0001, R)
    R_sersic_ = np.maximum(0.0001, R_sersic)
    bn = b_n(n_sersic)
    R_frac = R_ / R_sersic_
    with np.errstate(all="ignore"):
        exponent = -bn * (R_frac ** (1.0 / n_sersic) - 1.0)
        result = np.where(R_frac <= max_R_frac, np.exp(exponent), 0)
    return np.nan_to_num(result)


def sersic_elliptic(x, y, amp, R_sersic, n_sersic, e1, e2, center_x=0.0, center_y=0.0):
    """SersicElliptic.function, LightModel/Profiles/sersic_ellipse.py:62-93, with the
    product-average distance of sersic_utils.py:83-110 / param_util.py:103-124."""
    R_sersic = np.maximum(0, R_sersic)
    x_s = np.asarray(x, float) - center_x
    y_s = np.asarray(y, float) - center_y
    norm = np.maximum(np.sqrt(np.abs(1 - e1 ** 2 - e2 ** 2)), 0.000001)
    x_ = ((1 - e1) * x_s - e2 * y_s) / norm
    y_ = (-e2 * x_s + (1 + e1) * y_s) / norm
    R = np.sqrt(x_ ** 2 + y_ ** 2)
    return amp * _r_sersic(R, R_sersic, n_sersic)


def sersic(x, y, amp, R_sersic, n_sersic, center_x=0.0, center_y=0.0):
    """Sersic.function, LightModel/Profiles/sersic.py:53-73 (round: e1=e2=0)."""
    x_s = np.asarray(x, float) - center_x
    y_s = np.asarray(y, float) - center_y
    R = np.sqrt(x_s ** 2 + y_s ** 2)
    return amp * _r_sersic(R, R_sersic, n_sersic)


def gaussian_light(x, y, amp, sigma, center_x=0.0, center_y=0.0):
    """Gaussian.function, LightModel/Profiles/gaussian.py:27-42."""
    c = amp / (2 * np.pi * sigma ** 2)
    r2 = (np.asarray(x, float) - center_x) ** 2 / sigma ** 2 + (np.asarray(y, float) - center_y) ** 2 / sigma ** 2
    return c * np.exp(-r2 / 2.0)


LIGHT_FUNCS = {
    "SERSIC_ELLIPSE": (sersic_elliptic, ["amp", "R_sersic", "n_sersic", "e1", "e2", "center_x", "center_y"]),
    "SERSIC": (sersic, ["amp", "R_sersic", "n_sersic", "center_x", "center_y"]),
    "GAUSSIAN": (gaussian_light, ["amp", "sigma", "center_x", "center_y"]),
}


def surface_brightness(x, y, light_model_list, kwargs_list):
    """LightModelBase.surface_brightness, LightModel/light_model_base.py:230-252."""
    flux = np.zeros_like(np.asarray(x, float))
    for name, kw in zip(light_model_list, kwargs_list):
        func, names = LIGHT_FUNCS[name]
        flux = flux + _call(func, names, x, y, kw)
    return flux


# --------------------------------------------------------------------------
# Grid, pooling, convolution (ImSim/Numerics/*.py, Util/image_util.py)
# --------------------------------------------------------------------------


def coordinate_grid(nx, ny, transform_pix2angle, ra_at_xy_0, dec_at_xy_0):
    """Pixel-centre coordinates, x (column) fastest: lenstronomy
    Coordinates1D.coordinate_grid as used at ImSim/Numerics/grid.py:48.
    Returns 2-D (ny, nx) arrays."""
    M = np.asarray(transform_pix2angle, float)
    i, j = np.meshgrid(np.arange(nx), np.arange(ny))
    ra = i * M[0, 0] + j * M[0, 1] + ra_at_xy_0
    dec = i * M[1, 0] + j * M[1, 1] + dec_at_xy_0
    return ra.astype(float), dec.astype(float)


def make_subgrid(ra2d, dec2d, s):
    """lenstronomy make_subgrid as used at ImSim/Numerics/grid.py:60-62: each pixel
    splits into s x s sub-pixels at offsets (-1/2 + 1/(2s) + k/s) pixel steps.
    Returns flat arrays on the (ny*s, nx*s) grid, x fastest."""
    ny, nx = ra2d.shape
    if s == 1:
        return ra2d.ravel().copy(), dec2d.ravel().copy()
    d_ra_x = ra2d[0, 1] - ra2d[0, 0] if nx > 1 else 0.0
    d_ra_y = ra2d[1, 0] - ra2d[0, 0] if ny > 1 else 0.0
    d_dec_x = dec2d[0, 1] - dec2d[0, 0] if nx > 1 else 0.0
    d_dec_y = dec2d[1, 0] - dec2d[0, 0] if ny > 1 else 0.0
    ra_new = np.zeros((ny * s, nx * s))
    dec_new = np.zeros((ny * s, nx * s))
    for i in range(s):
        for j in range(s):
            oj = -1 / 2.0 + 1 / (2.0 * s) + j / float(s)
            oi = -1 / 2.0 + 1 / (2.0 * s) + i / float(s)
            ra_new[i::s, j::s] = ra2d + d_ra_x * oj + d_ra_y * oi
            dec_new[i::s, j::s] = dec2d + d_dec_x * oj + d_dec_y * oi
    return ra_new.ravel(), dec_new.ravel()


def subgrid_coordinates(nx, ny, transform_pix2angle, ra_at_xy_0, dec_at_xy_0, s):
    """RegularGrid.coordinates_evaluate, ImSim/Numerics/grid.py:22-72."""
    ra2d, dec2d = coordinate_grid(nx, ny, transform_pix2angle, ra_at_xy_0, dec_at_xy_0)
    return make_subgrid(ra2d, dec2d, s)


def re_size(image, factor=1):
    """image_util.re_size, Util/image_util.py:79-98 (block mean)."""
    if factor == 1:
        return image
    ny, nx = image.shape
    return image.reshape(ny // factor, factor, nx // factor, factor).mean(3).mean(1)


def convolve_same(image, kernel, mode="zero"):
    """2-D convolution with 'same' output, exact FP64 direct sum.

    mode='zero' is scipy/jax ``fftconvolve(image, kernel, 'same')``
    (ImSim/Numerics/convolution.py:74-86); mode='edge' first pads the image
    by replicating its edge, as GaussianConvolution does
    (convolution.py:217-236).  Works on (..., ny, nx) stacks.
    """
    image = np.asarray(image, float)
    kernel = np.asarray(kernel, float)
    ky, kx = kernel.shape
    cy, cx = (ky - 1) // 2, (kx - 1) // 2
    pad = ((0, 0),) * (image.ndim - 2) + ((cy, cy), (cx, cx))
    padded = np.pad(image, pad, mode="edge" if mode == "edge" else "constant")
    ny, nx = image.shape[-2:]
    out = np.zeros_like(image)
    for a in range(ky):
        for b in range(kx):
            # out[y, x] += img[y + cy - a, x + cx - b] * k[a, b]
            oy = 2 * cy - a
            ox = 2 * cx - b
            out += kernel[a, b] * padded[..., oy:oy + ny, ox:ox + nx]
    return out


def gaussian_kernel(fwhm, pixel_scale, truncation, supersampling_factor=1, supersampling_convolution=False):
    """GaussianConvolution kernel, ImSim/Numerics/convolution.py:177-206, :254-273
    (sigma from fwhm2sigma, Util/util.py:62-69; Gaussian.function on an integer grid,
    normalised to unit sum)."""
    sigma = fwhm / (2 * np.sqrt(2 * np.log(2)))
    sigma_scaled = sigma / pixel_scale
    if supersampling_convolution:
        sigma_scaled *= supersampling_factor
    r = max(round(sigma_scaled * truncation), 1)
    n = 2 * r + 1
    g = np.arange(n) - (n - 1) / 2.0
    xx, yy = np.meshgrid(g, g)
    k = gaussian_light(xx, yy, 1.0, sigma_scaled)
    return k / np.sum(k)


def gaussian_convolve(image, kernel):
    """GaussianConvolution.convolution2d, ImSim/Numerics/convolution.py:217-236."""
    return convolve_same(image, kernel, mode="edge")


def moffat_kernel(size, fwhm_pix, beta):
    """Unit-sum Moffat PSF sampled at pixel centres, I ~ (1 + (r/a)^2)^-beta with
    a = fwhm / (2 sqrt(2^(1/beta) - 1)).  (Workload definition, not a reference
    function.)"""
    a = fwhm_pix / (2.0 * np.sqrt(2.0 ** (1.0 / beta) - 1.0))
    g = np.arange(size) - (size - 1) / 2.0
    xx, yy = np.meshgrid(g, g)
    k = (1.0 + (xx ** 2 + yy ** 2) / a ** 2) ** (-beta)
    return k / np.sum(k)


def image(grid, lens_model_list, kwargs_lens, source_list, kwargs_source,
          lens_light_list, kwargs_lens_light, psf_type="PIXEL", kernel=None,
          flux_scaling=1.0, alpha_scaling=1.0, unconvolved=False,
          source_add=True, lens_light_add=True, supersampling_convolution=False, kernel_low=None,
          flux_evaluate_indexes=None, _eval=None):
    """ImageModel.image, ImSim/image_model.py:431-492.

    grid: dict(nx, ny, transform_pix2angle, ra_at_xy_0, dec_at_xy_0, s).
    Source light: rays -> source profiles -> x flux_scaling (:313-368) ->
    re_size_convolve (numerics.py:158-177); lens light: profiles on the grid ->
    re_size_convolve -> x flux_scaling (:370-393).
    supersampling_convolution (s > 1): ``kernel`` is sampled on the supersampled
    grid and convolves the ray fluxes before the s x s mean; ``kernel_low`` (a
    split kernel's low-resolution part, or None) convolves the mean of the
    unconvolved rays and is added (SubgridKernelConvolution.re_size_convolve,
    ImSim/Numerics/convolution.py:154-166; GaussianConvolution.re_size_convolve,
    :238-252).
    flux_evaluate_indexes (bool [ny, nx] or None): NumericsSubFrame
    (ImSim/Numerics/numerics_subframe.py:61-191) -- the model of the square
    rectangle [x_min:x_max+1, y_min:y_max+1] around the True pixels, on a grid
    starting at pixel (x_min, y_min) (map_pix2coord), with the rays of False
    pixels not shot (RegularGrid, grid.py:47-58, :126-150), convolved inside
    the rectangle and zero elsewhere.
    """
    if flux_evaluate_indexes is not None:
        mask = np.asarray(flux_evaluate_indexes, bool)
        a, b = np.where(mask)
        x_min, x_max, y_min, y_max = a.min(), a.max(), b.min(), b.max()
        n = x_max - x_min + 1
        assert y_max - y_min + 1 == n, "square sub-frames only"
        M = np.asarray(grid["transform_pix2angle"], float)
        sub = dict(grid, nx=n, ny=n, ra_at_xy_0=grid["ra_at_xy_0"] + M[0, 0] * x_min + M[0, 1] * y_min,
                   dec_at_xy_0=grid["dec_at_xy_0"] + M[1, 0] * x_min + M[1, 1] * y_min)
        sub_img = image(sub, lens_model_list, kwargs_lens, source_list, kwargs_source, lens_light_list,
                        kwargs_lens_light, psf_type, kernel, flux_scaling, alpha_scaling, unconvolved, source_add,
                        lens_light_add, supersampling_convolution, kernel_low, None,
                        _eval=mask[x_min:x_max + 1, y_min:y_max + 1])
        full = np.zeros((grid["ny"], grid["nx"]))
        full[x_min:x_max + 1, y_min:y_max + 1] = sub_img
        return full
    nx, ny, s = grid["nx"], grid["ny"], grid["s"]
    M = np.asarray(grid["transform_pix2angle"], float)
    pixel_width = np.sqrt(abs(np.linalg.det(M)))
    x, y = subgrid_coordinates(nx, ny, M, grid["ra_at_xy_0"], grid["dec_at_xy_0"], s)
    mode = "edge" if psf_type == "GAUSSIAN" else "zero"

    if _eval is not None:  # rays of pixels outside the mask carry no flux
        keep = np.repeat(np.repeat(np.asarray(_eval, bool), s, axis=0), s, axis=1).ravel()
    else:
        keep = None

    def re_size_convolve(flux):
        if keep is not None:
            flux = np.where(keep, flux, 0.0)
        hr = flux.reshape(ny * s, nx * s)
        if supersampling_convolution and s > 1 and not unconvolved and psf_type != "NONE":
            img = re_size(convolve_same(hr, kernel, mode=mode), s)
            if kernel_low is not None:
                img = img + convolve_same(re_size(hr, s), kernel_low, mode="zero")
            return img * pixel_width ** 2
        img = re_size(hr, s)
        if not unconvolved and psf_type != "NONE":
            img = convolve_same(img, kernel, mode=mode)
        return img * pixel_width ** 2

    model = np.zeros((ny, nx))
    if source_add and len(source_list) > 0:
        bx, by = ray_shooting(x, y, lens_model_list, kwargs_lens, alpha_scaling)
        flux = surface_brightness(bx, by, source_list, kwargs_source) * flux_scaling
        model = model + re_size_convolve(flux)
    if lens_light_add and len(lens_light_list) > 0:
        flux = surface_brightness(x, y, lens_light_list, kwargs_lens_light)
        model = model + re_size_convolve(flux) * flux_scaling
    return model


def c_d_model(model, background_rms=None, exposure_map=None, noise_map=None):
    """ImageNoise.C_D_model / covariance_matrix, Data/image_noise.py:100-137, with the
    exposure floor of :52-54."""
    if noise_map is not None:
        return np.asarray(noise_map, float) ** 2
    exposure_map = np.where(np.asarray(exposure_map, float) <= 10 ** (-10), 10 ** (-10), exposure_map)
    d_pos = np.where(model >= 0, model, 0)
    return d_pos / exposure_map + background_rms ** 2


def log_likelihood(model, data, mask=None, background_rms=None, exposure_map=None,
                   noise_map=None, additional_error_map=0.0):
    """ImageData._log_likelihood, Data/imaging_data.py:159-182 (diagonal method)."""
    model = np.asarray(model, float)
    if mask is None:
        mask = np.ones_like(model)
    c_d = c_d_model(model, background_rms, exposure_map, noise_map)
    chi2 = (model - data) ** 2
    chi2 = chi2 / (c_d + np.abs(additional_error_map))
    chi2 = chi2 * mask
    return -np.sum(chi2, axis=(-2, -1)) / 2


# ---------------------------------------------------------------------------
# Linear amplitude solve (ImSim/image_linear_solve.py, ImSim/de_lens.py)

EPSILON = np.finfo(float).eps


def linear_solve(grid, lens_model_list, kwargs_lens, source_list, kwargs_source, lens_light_list, kwargs_lens_light,
                 data, mask=None, background_rms=None, exposure_map=None, noise_map=None, psf_type="PIXEL",
                 kernel=None, flux_scaling=1.0, source_marg=False, linear_prior=None, **image_kwargs):
    """ImageLinearFit.image_linear_solve + likelihood_data_given_model
    (ImSim/image_linear_solve.py:98-181, :229-340).

    Response of each light profile at amp = 1 (linear_response_matrix
    :362-455, LinearBasis.functions_split LightModel/linear_basis.py:25-92),
    weights from the data-based variance (ImageModel.error_response
    image_model.py:577-596 -> ImageData.C_D), get_param_WLS (de_lens.py:10-36),
    model zero outside the mask (array_masked2image), logL with the model-based
    variance, plus the marginalisation constant (de_lens.py:40-77).
    Returns (logL, amplitudes, model image, M_inv or None).
    """
    data = np.asarray(data, float)
    if mask is None:
        mask = np.ones_like(data)
    m1 = np.asarray(mask, bool).ravel()
    cols = []
    for i in range(len(source_list)):
        kw = dict(kwargs_source[i], amp=1.0)
        cols.append(image(grid, lens_model_list, kwargs_lens, [source_list[i]], [kw], [], [], psf_type=psf_type,
                          kernel=kernel, flux_scaling=flux_scaling, **image_kwargs))
    for i in range(len(lens_light_list)):
        kw = dict(kwargs_lens_light[i], amp=1.0)
        cols.append(image(grid, lens_model_list, kwargs_lens, [], [], [lens_light_list[i]], [kw], psf_type=psf_type,
                          kernel=kernel, flux_scaling=flux_scaling, **image_kwargs))
    A = np.nan_to_num(np.array([c.ravel()[m1] for c in cols]))  # (n_lin, n_masked)
    if noise_map is not None:
        c_d = np.asarray(noise_map, float) ** 2
    else:
        c_d = np.where(data >= 0, data, 0) / exposure_map + background_rms ** 2
    w = 1.0 / c_d.ravel()[m1]
    d = data.ravel()[m1]
    M = A.dot((w * A).T)
    R = A.dot(w * d)
    stable = np.linalg.cond(M) < 5 / EPSILON
    if stable:
        B = np.linalg.solve(M, R)
        if not np.all(np.isfinite(B)):
            B = np.zeros(len(R))
        M_inv = np.nan_to_num(np.linalg.inv(M), nan=0, posinf=0, neginf=0)
    else:
        B = np.zeros(len(R))
        M_inv = np.zeros_like(M)
    flat = np.zeros(data.size)
    flat[m1] = A.T.dot(B)
    model = flat.reshape(data.shape)
    logL = log_likelihood(model, data, mask=mask, background_rms=background_rms, exposure_map=exposure_map,
                          noise_map=noise_map)
    if source_marg:
        if linear_prior is None:
            sign, log_det = np.linalg.slogdet(M_inv)
            logL += -1e15 if sign == 0 else sign * log_det / 2
        else:
            v = np.linalg.eigvals(M_inv).real
            v_abs = np.minimum(np.abs(v), linear_prior ** 2)
            with np.errstate(divide="ignore", invalid="ignore"):
                log_det = np.sum(np.log(v_abs)) * np.prod(np.sign(v))
            m = len(v)
            logL += -1e15 if np.isnan(log_det) else (log_det / 2 + m / 2.0 * np.log(np.pi / 2.0)
                                                      - m * np.log(linear_prior))
    return logL, B, model, M_inv


def multi_plane_ray_shooting(x, y, lens_model_list, kwargs_lens, T_ij_list, T_z_list, reduced2physical_factor):
    """MultiPlaneBulk.ray_shooting, LensModel/MultiPlane/multi_plane_bulk.py:152-212:
    the scan over redshift-sorted planes (the last entries of the distance lists
    are the source plane and any NULL fillers)."""
    ax = np.asarray(x, float).copy()
    ay = np.asarray(y, float).copy()
    px = np.zeros_like(ax)
    py = np.zeros_like(ay)
    n = len(lens_model_list)
    for p in range(len(T_ij_list)):
        px = px + ax * T_ij_list[p]
        py = py + ay * T_ij_list[p]
        if p < n:
            func, names = LENS_FUNCS[lens_model_list[p]]
            fx, fy = _call(func, names, px / T_z_list[p], py / T_z_list[p], kwargs_lens[p])
            ax = ax - fx * reduced2physical_factor[p]
            ay = ay - fy * reduced2physical_factor[p]
    return px / T_z_list[-1], py / T_z_list[-1]


# ---------------------------------------------------------------------------
# Sampling/likelihood.py: the sampler-facing log-likelihood of one vector

def gaussian_prior_logl(kwargs_list, prior_list):
    """lenstronomy PriorLikelihood's Gaussian prior of one kwargs list (as constructed at
    Sampling/likelihood.py:153-172 and evaluated at :353-356): for each
    [index, param_name, mean, sigma], -(kwargs_list[index][param_name] - mean)^2 / sigma^2 / 2."""
    logl = 0.0
    for index, name, mean, sigma in prior_list or []:
        logl -= (kwargs_list[index][name] - mean) ** 2 / sigma ** 2 / 2.0
    return logl


def likelihood_module_logl(args, lower, upper, image_logl, prior_logl=0.0, custom_logl=0.0, check_bounds=True):
    """LikelihoodModule.logL / log_likelihood, Sampling/likelihood.py:290-371, for the
    imaging likelihood of one parameter vector: a bound hit gives -1e18
    (:312-318); otherwise the custom addition, the prior and the image
    likelihood (NaN -> -1e15, image_likelihood.py:93) are summed in that order
    and the sum's NaN becomes 1e-18, +-inf the largest finite value (:370)."""
    args = np.asarray(args, float)
    if check_bounds and bool(np.any((args < lower) | (args > upper))):
        return -1e18
    logl = 0.0
    logl += custom_logl
    logl += prior_logl
    logl += float(np.nan_to_num(image_logl, nan=-1e15))
    return float(np.nan_to_num(logl, nan=1e-18))


def beta_double_source_plane(d_xy, z_lens, z_source_1, z_source_2):
    """Deflection ratio for a source at z_source_1 when the kwargs are normalised
    for z_source_2 (lenstronomy LensCosmo.beta_double_source_plane, called at
    LensModel/lens_model.py:180-189): D_ds1 / D_s1 * D_s2 / D_ds2 with the
    angular diameter distance function d_xy(z_observer, z_source)."""
    return d_xy(z_lens, z_source_1) / d_xy(0.0, z_source_1) * d_xy(0.0, z_source_2) / d_xy(z_lens, z_source_2)
