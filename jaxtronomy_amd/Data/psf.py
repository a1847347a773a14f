"""PSF: the point spread function description used by ImageModel.

Follows the public behaviour of lenstronomy's ``PSF`` class, which the
reference imports (Util/class_creator.py:7, ImSim/image_model.py:64): PIXEL
kernels must be odd-sized and are normalised to unit sum; GAUSSIAN carries a
FWHM and a truncation (in units of sigma, default 5).
"""

import numpy as np


class PSF(object):
    def __init__(self, psf_type="NONE", fwhm=None, truncation=5, pixel_size=None, kernel_point_source=None,
                 psf_variance_map=None, point_source_supersampling_factor=1,
                 kernel_point_source_normalisation=True, kernel_point_source_init=None):
        self.psf_type = psf_type
        self._pixel_size = pixel_size
        self.psf_variance_map_bool = psf_variance_map is not None
        self._psf_variance_map = psf_variance_map
        self._point_source_supersampling_factor = point_source_supersampling_factor
        if psf_type == "GAUSSIAN":
            if fwhm is None:
                raise ValueError("fwhm must be set for GAUSSIAN psf type!")
            self._fwhm = float(fwhm)
            self._truncation = truncation
            self._kernel = None
        elif psf_type == "PIXEL":
            if kernel_point_source is None:
                raise ValueError("kernel_point_source needs to be specified for PIXEL PSF type!")
            k = np.asarray(kernel_point_source, dtype=np.float64)
            if k.ndim != 2 or k.shape[0] % 2 == 0 or k.shape[1] % 2 == 0:
                raise ValueError("kernel needs to have odd axis number, not %s" % (k.shape,))
            if kernel_point_source_normalisation:
                k = k / np.sum(k)
            self._kernel = k
            self._fwhm = None
            self._truncation = truncation
        elif psf_type == "NONE":
            self._kernel = None
            self._fwhm = None
            self._truncation = truncation
        else:
            raise ValueError("psf_type %s not supported!" % psf_type)

    @property
    def kernel_point_source(self):
        return self._kernel

    @property
    def fwhm(self):
        return self._fwhm

    @property
    def truncation(self):
        return self._truncation

    @property
    def psf_variance_map(self):
        return self._psf_variance_map

    def set_pixel_size(self, deltaPix):
        self._pixel_size = deltaPix


def cut_psf(psf_data, psf_size):
    """Centre crop of an odd kernel to psf_size x psf_size (lenstronomy kernel_util.cut_psf
    with normalisation=False, as used at ImSim/Numerics/numerics.py:195-213)."""
    n = psf_data.shape[0]
    if psf_size > n:
        return psf_data
    c = (n - 1) // 2
    r = (psf_size - 1) // 2
    return psf_data[c - r:c + r + 1, c - r:c + r + 1]


def gaussian_pixel_kernel(fwhm, pixel_scale, truncation, supersampling_factor=1, supersampling_convolution=False):
    """Kernel of GaussianConvolution (ImSim/Numerics/convolution.py:177-206, :254-273):
    sigma_pix = fwhm / (2 sqrt(2 ln 2)) / pixel_scale, radius max(round(sigma_pix *
    truncation), 1), Gaussian sampled at integer offsets, normalised to unit sum."""
    sigma = fwhm / (2 * np.sqrt(2 * np.log(2)))
    sigma_scaled = sigma / pixel_scale
    if supersampling_convolution:
        sigma_scaled *= supersampling_factor
    radius = max(round(sigma_scaled * truncation), 1)
    n = 2 * radius + 1
    g = np.arange(n) - (n - 1) / 2.0
    xx, yy = np.meshgrid(g, g)
    c = 1.0 / (2 * np.pi * sigma_scaled ** 2)
    k = c * np.exp(-(xx ** 2 / sigma_scaled ** 2 + yy ** 2 / sigma_scaled ** 2) / 2.0)
    return k / np.sum(k)
