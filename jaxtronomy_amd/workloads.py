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
