"""PSF: the point spread function description used by ImageModel.

Follows the public behaviour of lenstronomy's ``PSF`` class, which the
reference imports (Util/class_creator.py:7, ImSim/image_model.py:64): PIXEL
kernels must be odd-sized and are normalised to unit sum; GAUSSIAN carries a
FWHM and a truncation (in units of sigma, default 5).  A PIXEL kernel given
with point_source_supersampling_factor = s > 1 is sampled on the supersampled
grid: it is kept (normalised) for supersampled convolution, and the regular
kernel is its s x s block mean (degrade_kernel, odd s).

A PIXEL kernel given at the regular resolution and asked for on a grid
supersampled by an odd factor s is interpolated there by subgrid_kernel
(lenstronomy's PSF.kernel_point_source_supersampled: kernel_util.subgrid_kernel
with odd=True and five iterations, then cut_psf to n s pixels), as
ImSim/Numerics/numerics.py:103-116 asks for supersampled convolution.

lenstronomy itself is not available here; degrade_kernel, split_kernel and
subgrid_kernel are restated from its documented behaviour and pinned by their
invariants (unit total flux, the hole under the supersampled cut, centring,
block means that reproduce the regular kernel), not by its output.
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
            self._kernel_super = None
            if point_source_supersampling_factor > 1:
                self._kernel_super = k / np.sum(k) if kernel_point_source_normalisation else k
                k = degrade_kernel(k, point_source_supersampling_factor)
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

    def kernel_point_source_supersampled(self, supersampling_factor):
        """The PIXEL kernel sampled on a grid supersampled by supersampling_factor:
        the kernel given at that factor (point_source_supersampling_factor), the
        regular kernel itself for a factor of 1, else the regular kernel
        interpolated onto the finer grid (subgrid_kernel, odd factors; the
        result is cached per factor)."""
        if supersampling_factor == 1:
            return self._kernel
        if self.psf_type != "PIXEL":
            raise ValueError("a supersampled kernel is only defined for PIXEL PSFs (psf_type %s)" % self.psf_type)
        if getattr(self, "_kernel_super", None) is not None and \
                self._point_source_supersampling_factor == supersampling_factor:
            return self._kernel_super
        if supersampling_factor % 2 == 0:
            raise NotImplementedError("interpolating a PIXEL kernel onto a grid supersampled by an even factor (%d) "
                                      "is not supported; give the supersampled kernel with "
                                      "point_source_supersampling_factor" % supersampling_factor)
        cache = self.__dict__.setdefault("_subgrid_cache", {})
        if supersampling_factor not in cache:
            k = subgrid_kernel(self._kernel, supersampling_factor, odd=True, num_iter=5)
            n_new = len(self._kernel) * supersampling_factor
            if n_new % 2 == 0:
                n_new -= 1
            k = cut_psf(k, n_new)
            cache[supersampling_factor] = k / np.sum(k)
        return cache[supersampling_factor]

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


def _resize_linear(values, n_out):
    """A square array sampled at the cell centres of the unit square, linearly
    interpolated (separably, in each axis) onto n_out x n_out cell centres;
    beyond the outermost input centres the edge values hold (lenstronomy
    image_util.re_size_array: linear interpolation of the cells' values,
    constant past the input range)."""
    n = values.shape[0]
    d = 1.0 / n
    c_in = np.linspace(d / 2, 1 - d / 2, n)
    d_out = 1.0 / n_out
    c_out = np.linspace(d_out / 2.0, 1 - d_out / 2.0, n_out)
    rows = np.array([np.interp(c_out, c_in, values[j]) for j in range(n)])  # along x, per input row
    return np.array([np.interp(c_out, c_in, rows[:, i]) for i in range(n_out)]).T  # along y


def subgrid_kernel(kernel, subgrid_res, odd=True, num_iter=5):
    """The kernel on a grid subgrid_res times finer (lenstronomy
    kernel_util.subgrid_kernel): the regular kernel is interpolated onto the
    finer grid and normalised; then, num_iter times, the finer kernel is
    block-averaged back to the regular grid, the residual against the regular
    kernel is added to the kernel being interpolated, and that is
    interpolated again -- so the fine kernel's block means approach the
    given kernel.  Odd subgrid_res only (the block mean of an odd factor is
    centred); odd=True makes the fine size odd."""
    subgrid_res = int(subgrid_res)
    if subgrid_res == 1:
        return kernel
    if subgrid_res % 2 == 0:
        raise NotImplementedError("even subgrid factors are not supported")
    kernel = np.asarray(kernel, dtype=np.float64)
    n = kernel.shape[0]
    if kernel.ndim != 2 or kernel.shape[1] != n:
        raise ValueError("subgrid_kernel takes a square kernel, not %s" % (kernel.shape,))
    n_new = n * subgrid_res
    if odd and n_new % 2 == 0:
        n_new -= 1
    source = kernel.copy()
    fine = _resize_linear(source, n_new)
    fine = fine / np.sum(fine)
    for _ in range(max(num_iter, 1)):
        coarse = fine.reshape(n, subgrid_res, n, subgrid_res).mean(axis=(1, 3))
        coarse = coarse / np.sum(coarse)  # both at unit flux: the residual is in shape only
        source = source + (kernel / np.sum(kernel) - coarse)
        fine = _resize_linear(source, n_new)
        fine = fine / np.sum(fine)
    return fine


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


def degrade_kernel(kernel_super, degrading_factor):
    """Kernel on a grid degrading_factor times coarser: zero-pad the (odd)
    kernel symmetrically to numPix * factor pixels (numPix odd) and take the
    mean of every factor x factor block (lenstronomy kernel_util.degrade_kernel,
    odd factors)."""
    if degrading_factor == 1:
        return kernel_super
    if degrading_factor % 2 == 0:
        raise NotImplementedError("even degrading factors are not supported")
    n_kernel = len(kernel_super)
    num_pix = int(round(n_kernel / degrading_factor + 0.5))
    if num_pix % 2 == 0:
        num_pix += 1
    n_high = num_pix * degrading_factor
    padded = np.zeros((n_high, n_high))
    i0 = (n_high - n_kernel) // 2
    padded[i0:i0 + n_kernel, i0:i0 + n_kernel] = kernel_super
    return padded.reshape(num_pix, degrading_factor, num_pix, degrading_factor).mean(axis=(1, 3))


def split_kernel(kernel_super, supersampling_kernel_size, supersampling_factor, normalized=True):
    """(low-resolution kernel with a hole, supersampled central cut) such that
    convolving the supersampled image with the cut and the pooled image with
    the holed kernel approximates the full supersampled convolution
    (lenstronomy kernel_util.split_kernel, used at
    ImSim/Numerics/convolution.py:117-131).  The cut is the central
    supersampling_kernel_size * s (made odd) pixels; the holed kernel is the
    supersampled kernel with that centre zeroed, degraded to the pixel grid and
    rescaled so that both parts together carry unit flux."""
    if supersampling_factor <= 1:
        raise ValueError("To split a kernel, the supersampling_factor needs to be > 1, given %s"
                         % supersampling_factor)
    if supersampling_kernel_size % 2 == 0:
        raise ValueError("supersampling_kernel_size needs to be an odd number, given %s" % supersampling_kernel_size)
    kernel_super = np.asarray(kernel_super, dtype=np.float64)
    n_super = len(kernel_super)
    n_sub = supersampling_kernel_size * supersampling_factor
    if n_sub % 2 == 0:
        n_sub += 1
    n_sub = min(n_sub, n_super)
    n_min = (n_super - 1) // 2 - (n_sub - 1) // 2
    n_max = (n_super - 1) // 2 + (n_sub - 1) // 2 + 1
    hole = kernel_super.copy()
    hole[n_min:n_max, n_min:n_max] = 0.0
    hole_low = degrade_kernel(hole, supersampling_factor)
    cut = kernel_super[n_min:n_max, n_min:n_max].copy()
    if normalized:
        flux_cut = np.sum(cut)
        flux_hole = np.sum(hole_low)
        if flux_hole > 0:
            hole_low = hole_low * (1.0 - flux_cut) / flux_hole
        else:
            cut = cut / np.sum(cut)
    else:
        hole_low = hole_low / supersampling_factor ** 2
    return hole_low, cut
