"""Seeded synthetic workloads (numpy only, no device work).

  small  SIE + SHEAR deflector, SERSIC_ELLIPSE source, 60x60 at 0.05"/pix,
         supersampling 1, 11x11 Gaussian PSF (FWHM 0.1", truncation 6),
         noise sigma 0.01 + Poisson at exposure 1000.
  epl180 EPL + SHEAR deflector, SERSIC_ELLIPSE lens light and source,
         180x180 at 0.04"/pix, supersampling 3, 41x41 Moffat PSF (beta 3,
         FWHM 0.08", unit sum).
Parameter draws use numpy.random.default_rng(20261202 + 1000 * config + draw).
"""

import numpy as np

from .Util.simulation_util import data_configure_simple

SEED0 = 20261202


def seed(config, draw):
    return SEED0 + 1000 * config + draw


def moffat_kernel(size, fwhm_pix, beta):
    """Unit-sum Moffat PSF I ~ (1 + r^2/a^2)^-beta, a = fwhm / (2 sqrt(2^(1/beta) - 1)),
    sampled at pixel centres of a size x size stamp."""
    a = fwhm_pix / (2.0 * np.sqrt(2.0 ** (1.0 / beta) - 1.0))
    g = np.arange(size) - (size - 1) / 2.0
    xx, yy = np.meshgrid(g, g)
    k = (1.0 + (xx ** 2 + yy ** 2) / a ** 2) ** (-beta)
    return k / np.sum(k)


def small_draw(draw):
    """One seeded small-config scene: (kwargs_lens, kwargs_source)."""
    rng = np.random.default_rng(seed(1, draw))
    kwargs_lens = [
        dict(theta_E=rng.uniform(0.8, 1.6), e1=rng.normal(0, 0.1), e2=rng.normal(0, 0.1), center_x=0.0,
             center_y=0.0),
        dict(gamma1=rng.normal(0, 0.03), gamma2=rng.normal(0, 0.03), ra_0=0.0, dec_0=0.0),
    ]
    kwargs_source = [
        dict(amp=rng.uniform(1, 100), R_sersic=rng.uniform(0.1, 0.5), n_sersic=rng.uniform(1, 4),
             e1=rng.normal(0, 0.1), e2=rng.normal(0, 0.1), center_x=rng.uniform(-0.1, 0.1),
             center_y=rng.uniform(-0.1, 0.1)),
    ]
    return kwargs_lens, kwargs_source


def epl180_draw(draw):
    """One seeded 180x180 EPL scene: (kwargs_lens, kwargs_source, kwargs_lens_light)."""
    rng = np.random.default_rng(seed(2, draw))
    kwargs_lens = [
        dict(theta_E=rng.uniform(0.8, 1.6), gamma=rng.uniform(1.8, 2.2), e1=rng.normal(0, 0.1),
             e2=rng.normal(0, 0.1), center_x=0.0, center_y=0.0),
        dict(gamma1=rng.normal(0, 0.03), gamma2=rng.normal(0, 0.03), ra_0=0.0, dec_0=0.0),
    ]
    kwargs_source = [
        dict(amp=rng.uniform(1, 100), R_sersic=rng.uniform(0.1, 0.5), n_sersic=rng.uniform(1, 4),
             e1=rng.normal(0, 0.1), e2=rng.normal(0, 0.1), center_x=rng.uniform(-0.1, 0.1),
             center_y=rng.uniform(-0.1, 0.1)),
    ]
    kwargs_lens_light = [
        dict(amp=rng.uniform(1, 100), R_sersic=rng.uniform(0.5, 1.5), n_sersic=rng.uniform(2, 5),
             e1=rng.normal(0, 0.1), e2=rng.normal(0, 0.1), center_x=0.0, center_y=0.0),
    ]
    return kwargs_lens, kwargs_source, kwargs_lens_light


# Serving mix: grids, supersampling factors, deflector sets and PSF sizes
# that requests are drawn from.
SERVE_MIX = dict(grids=(64, 100, 180, 240), supersampling=(1, 3, 5),
                 deflectors=(("SIE", "SHEAR"), ("EPL", "SHEAR"), ("NFW", "SIE")), psf_sizes=(11, 21, 31, 41),
                 deltaPix=0.04, psf_fwhm=0.08, psf_beta=3.0, exposure_time=1000.0, background_rms=0.01)


def mix_buckets(n, seed_=0):
    """n (grid, s, deflectors, psf size) combinations of the serving mix: the
    first 12 take every (grid, s) pair once, the deflector sets in turn (four
    each) and seeded PSF sizes."""
    rng = np.random.default_rng(seed(4, seed_))
    m = SERVE_MIX
    ng, ns, nd = len(m["grids"]), len(m["supersampling"]), len(m["deflectors"])
    out = []
    for i in range(n):
        g = m["grids"][i % ng]
        s = m["supersampling"][(i // ng) % ns]
        d = m["deflectors"][(i + i // ng) % nd]
        out.append((g, s, d, int(rng.choice(m["psf_sizes"]))))
    return out


def mix_draw(deflectors, draw):
    """One seeded scene for a deflector set of the serving mix: (kwargs_lens,
    kwargs_source, kwargs_lens_light) with one SERSIC_ELLIPSE source and lens
    light."""
    rng = np.random.default_rng(seed(4, 1000 + draw))
    kl = []
    for t in deflectors:
        if t in ("SIE", "EPL"):
            k = dict(theta_E=rng.uniform(0.8, 1.6), e1=rng.normal(0, 0.1), e2=rng.normal(0, 0.1),
                     center_x=rng.normal(0, 0.02), center_y=rng.normal(0, 0.02))
            if t == "EPL":
                k["gamma"] = rng.uniform(1.8, 2.2)
            kl.append(k)
        elif t == "SHEAR":
            kl.append(dict(gamma1=rng.normal(0, 0.03), gamma2=rng.normal(0, 0.03), ra_0=0.0, dec_0=0.0))
        elif t == "NFW":
            kl.append(dict(Rs=rng.uniform(2.0, 6.0), alpha_Rs=rng.uniform(0.3, 1.0), center_x=rng.normal(0, 0.05),
                           center_y=rng.normal(0, 0.05)))
        else:
            raise ValueError(t)
    ks = [dict(amp=rng.uniform(1, 100), R_sersic=rng.uniform(0.1, 0.5), n_sersic=rng.uniform(1, 4),
               e1=rng.normal(0, 0.1), e2=rng.normal(0, 0.1), center_x=rng.uniform(-0.1, 0.1),
               center_y=rng.uniform(-0.1, 0.1))]
    kll = [dict(amp=rng.uniform(1, 100), R_sersic=rng.uniform(0.5, 1.5), n_sersic=rng.uniform(2, 5),
                e1=rng.normal(0, 0.1), e2=rng.normal(0, 0.1), center_x=0.0, center_y=0.0)]
    return kl, ks, kll


def stack_kwargs(kwargs_lists):
    """List (over draws) of kwargs lists -> one kwargs list with array values."""
    out = []
    for slot in zip(*kwargs_lists):
        out.append({k: np.array([d[k] for d in slot]) for k in slot[0]})
    return out


SMALL = dict(numPix=60, deltaPix=0.05, supersampling=1, psf_fwhm=0.1, psf_truncation=6, exposure_time=1000.0,
             background_rms=0.01)
EPL180 = dict(numPix=180, deltaPix=0.04, supersampling=3, psf_size=41, psf_fwhm=0.08, psf_beta=3.0,
              exposure_time=1000.0, background_rms=0.01)


def data_kwargs(cfg, image=None):
    kw = data_configure_simple(cfg["numPix"], cfg["deltaPix"], cfg["exposure_time"], cfg["background_rms"],
                               inverse=True)
    if image is not None:
        kw["image_data"] = np.asarray(image, dtype=np.float64)
    return kw


def grid_dict(cfg):
    """Grid description (oracle layout) of a config."""
    kw = data_kwargs(cfg)
    return dict(nx=cfg["numPix"], ny=cfg["numPix"], s=cfg["supersampling"], ra_at_xy_0=kw["ra_at_xy_0"],
                dec_at_xy_0=kw["dec_at_xy_0"], transform_pix2angle=kw["transform_pix2angle"])


def noisy(image, exposure_time, background_rms, rng):
    """Poisson (counts at exposure_time) plus Gaussian background noise."""
    img = np.asarray(image, dtype=np.float64)
    counts = rng.poisson(np.maximum(img, 0) * exposure_time)
    return counts / exposure_time + rng.normal(0, background_rms, img.shape)


# ---------------------------------------------------------------------------
# Substructure workload: EPL + SHEAR main deflector plus n_sub TNFW subhalos,
# SERSIC_ELLIPSE source, 300x300 at 0.03"/pix, supersampling 3, 41x41 Moffat.
SUBSTRUCTURE = dict(numPix=300, deltaPix=0.03, supersampling=3, psf_size=41, psf_fwhm=0.08, psf_beta=3.0,
                    exposure_time=1000.0, background_rms=0.01, n_sub=1000, z_lens=0.5, z_source=1.5)

# Flat LambdaCDM (H0 = 70 km/s/Mpc, Omega_m = 0.3) and physical constants
_H0, _OM = 70.0, 0.3
_C_KMS = 299792.458
_RHO_CRIT0 = 2.77536627e11 * (_H0 / 100.0) ** 2  # M_sun / Mpc^3
_C2_4PIG = 1.6624541e18  # c^2 / (4 pi G) in M_sun / Mpc
_ARCSEC = np.pi / 180.0 / 3600.0


def _comoving(z1, z2, n=4096):
    """Comoving distance between redshifts z1 < z2 (Mpc), Simpson's rule."""
    z = np.linspace(z1, z2, n + 1)
    f = 1.0 / np.sqrt(_OM * (1 + z) ** 3 + 1 - _OM)
    h = (z2 - z1) / n
    return _C_KMS / _H0 * h / 3.0 * (f[0] + f[-1] + 4 * f[1:-1:2].sum() + 2 * f[2:-1:2].sum())


def nfw_physical_to_angle(m200, c, z_lens, z_source):
    """(Rs, alpha_Rs) in arcsec of an NFW halo of mass M200 (M_sun) and
    concentration c: r200 from 200 x the critical density at z_lens,
    rho0 = 200/3 rho_c c^3 / (ln(1+c) - c/(1+c)), and
    alpha_Rs = 4 rho0 Rs^2 (1 + ln 1/2) / Sigma_crit, in angles."""
    dd = _comoving(0.0, z_lens) / (1 + z_lens)
    ds = _comoving(0.0, z_source) / (1 + z_source)
    dds = _comoving(z_lens, z_source) / (1 + z_source)
    sigma_crit = _C2_4PIG * ds / (dd * dds)
    rho_c = _RHO_CRIT0 * (_OM * (1 + z_lens) ** 3 + 1 - _OM)
    m200 = np.asarray(m200, dtype=np.float64)
    c = np.asarray(c, dtype=np.float64)
    r200 = (3.0 * m200 / (4.0 * np.pi * 200.0 * rho_c)) ** (1.0 / 3.0)
    rs = r200 / c
    rho0 = 200.0 / 3.0 * rho_c * c ** 3 / (np.log(1 + c) - c / (1 + c))
    alpha_rs = rho0 * 4.0 * rs ** 2 * (1.0 + np.log(0.5)) / sigma_crit
    return rs / dd / _ARCSEC, alpha_rs / dd / _ARCSEC


def concentration(m200):
    """Fixed concentration-mass relation c = 15 (M / 1e8 M_sun)^-0.1."""
    return 15.0 * (np.asarray(m200, dtype=np.float64) / 1e8) ** -0.1


def substructure_draw(draw, n_sub=None, cfg=None):
    """One seeded substructure scene: (kwargs_lens main [EPL, SHEAR], subhalo
    kwargs list [TNFW] * n_sub, kwargs_source).  log10 M ~ U[6, 9], positions
    U[-3, 3]" squared, r_trunc = 5 Rs."""
    cfg = SUBSTRUCTURE if cfg is None else cfg
    n_sub = cfg["n_sub"] if n_sub is None else n_sub
    rng = np.random.default_rng(seed(5, draw))
    kwargs_lens = [
        dict(theta_E=rng.uniform(0.8, 1.6), gamma=rng.uniform(1.8, 2.2), e1=rng.normal(0, 0.1),
             e2=rng.normal(0, 0.1), center_x=0.0, center_y=0.0),
        dict(gamma1=rng.normal(0, 0.03), gamma2=rng.normal(0, 0.03), ra_0=0.0, dec_0=0.0),
    ]
    kwargs_source = [
        dict(amp=rng.uniform(1, 100), R_sersic=rng.uniform(0.1, 0.5), n_sersic=rng.uniform(1, 4),
             e1=rng.normal(0, 0.1), e2=rng.normal(0, 0.1), center_x=rng.uniform(-0.1, 0.1),
             center_y=rng.uniform(-0.1, 0.1)),
    ]
    m = 10.0 ** rng.uniform(6.0, 9.0, n_sub)
    cx = rng.uniform(-3.0, 3.0, n_sub)
    cy = rng.uniform(-3.0, 3.0, n_sub)
    rs, alpha_rs = nfw_physical_to_angle(m, concentration(m), cfg["z_lens"], cfg["z_source"])
    subs = [dict(Rs=float(rs[i]), alpha_Rs=float(alpha_rs[i]), r_trunc=float(5.0 * rs[i]),
                 center_x=float(cx[i]), center_y=float(cy[i])) for i in range(n_sub)]
    return kwargs_lens, subs, kwargs_source


# ---------------------------------------------------------------------------
# Line-of-sight halo workload: the shape of the multi-plane ray-tracing
# benchmark that the reference publishes for a pyHalo CDM realisation
# (notebooks/ray_shooting_through_dark_matter_substructure_with_GPU.ipynb:
# z_lens 0.5, z_source 2.5, cone opening 5.5", 5130 TNFW halos and 126
# CONVERGENCE sheets, 5500 deflector slots, a 400 x 400 grid over +-10").
# pyHalo is not available here: the halos are drawn to the same counts, with
# M ~ M^-1.9 on [1e6, 1e10] M_sun, one redshift plane per sheet, field halos
# per plane in proportion to the cone's comoving volume and a quarter of the
# halos as subhalos on the main lens plane, positions uniform in the cone's
# disc, r_trunc = 5 Rs, and each plane's sheet the negative of the mean
# convergence its halos add over the disc (kappa = -sum alpha_Rs-mass / area).
LOS = dict(z_lens=0.5, z_source=2.5, n_halo=5130, n_planes=126, cone_arcsec=5.5, num_deflectors=5500,
           num_pix=400, half_width=10.0)


def los_halo_draw(draw, cfg=None):
    """(lens_model_list, kwargs_lens, lens_redshift_list) of one line-of-sight
    realisation: n_halo TNFW and n_planes CONVERGENCE deflectors."""
    cfg = LOS if cfg is None else cfg
    rng = np.random.default_rng(seed(7, draw))
    zl, zs, n_planes = cfg["z_lens"], cfg["z_source"], cfg["n_planes"]
    z_planes = (np.arange(n_planes) + 0.5) * (zs / n_planes)
    main = int(np.argmin(np.abs(z_planes - zl)))
    z_planes[main] = zl
    dc = np.array([_comoving(0.0, z, n=256) for z in z_planes])
    vol = dc ** 2 * np.gradient(dc)  # comoving volume of a fixed-angle cone slice
    n_sub = cfg["n_halo"] // 4
    counts = np.floor((cfg["n_halo"] - n_sub) * vol / vol.sum()).astype(int)
    counts[np.argsort(-vol)[:cfg["n_halo"] - n_sub - counts.sum()]] += 1
    counts[main] += n_sub
    u = rng.uniform(size=cfg["n_halo"])
    a, lo, hi = -0.9, 1e6 ** -0.9, 1e10 ** -0.9  # M^-1.9: the CDF of M^(1 - 1.9)
    masses = (lo + u * (hi - lo)) ** (1.0 / a)
    radius = cfg["cone_arcsec"] / 2.0
    names, kwargs, z_list = [], [], []
    q = 0
    for p, zp in enumerate(z_planes):
        m = masses[q:q + counts[p]]
        q += counts[p]
        rs, alpha_rs = nfw_physical_to_angle(m, concentration(m), zp, zs)
        r = radius * np.sqrt(rng.uniform(size=len(m)))
        phi = rng.uniform(0, 2 * np.pi, len(m))
        for i in range(len(m)):
            names.append("TNFW")
            kwargs.append(dict(Rs=float(rs[i]), alpha_Rs=float(alpha_rs[i]), r_trunc=float(5.0 * rs[i]),
                               center_x=float(r[i] * np.cos(phi[i])), center_y=float(r[i] * np.sin(phi[i]))))
            z_list.append(float(zp))
        # the sheet: minus the halos' projected mass (~ alpha_Rs Rs each, in angle units) over the disc
        kappa = -float(np.sum(alpha_rs * rs)) / (np.pi * radius ** 2) if len(m) else 0.0
        names.append("CONVERGENCE")
        kwargs.append(dict(kappa=kappa, ra_0=0.0, dec_0=0.0))
        z_list.append(float(zp))
    return names, kwargs, z_list


def los_grid(cfg=None):
    """The benchmark's ray coordinates: a num_pix x num_pix grid over
    +-half_width arcsec (x fastest), as numpy float64 arrays."""
    cfg = LOS if cfg is None else cfg
    v = np.linspace(-cfg["half_width"], cfg["half_width"], cfg["num_pix"])
    return np.tile(v, cfg["num_pix"]), np.repeat(v, cfg["num_pix"])
