"""ImageModel: lensed images and their likelihood on the GPU.

Mirrors the reference's ``ImageModel`` (ImSim/image_model.py:23-694) for the
analytic single-band path: ray shooting -> source light at the deflected
supersampled grid -> lens light -> s x s pooling -> PSF convolution ->
x pixel_width^2 x flux_scaling, and the diagonal chi^2 likelihood.  Each call
is one ``lensim_image`` / ``lensim_loglike`` launch sequence on the current
stream.

Batches: parameter values may be 1-D arrays of length B (as under
``jax.vmap``); ``image`` then returns [B, ny, nx] and the likelihood [B].
``image_batch`` / ``likelihood_batch`` take an already packed device
parameter block [B, n_profiles, 8] (see include/lensim.h).
"""

import ctypes

import numpy as np
import torch

from .. import backend, profiles
from ..Data.psf import cut_psf, gaussian_pixel_kernel, split_kernel
from ..LensModel.lens_model import LensModel
from ..LensModel.lens_model_bulk import LensModelBulk
from ..LightModel.light_model import LightModel


class ImageModel(object):
    """This class uses functions of lens_model and source_model to make a lensed image."""

    def __init__(self, data_class, psf_class, lens_model_class=None, source_model_class=None,
                 lens_light_model_class=None, point_source_class=None, extinction_class=None, kwargs_numerics=None,
                 likelihood_mask=None, psf_error_map_bool_list=None, kwargs_pixelbased=None, precision=None):
        if kwargs_pixelbased is not None:
            raise ValueError("pixelbased solver not supported")
        if point_source_class is not None and len(getattr(point_source_class, "point_source_type_list", [])) > 0:
            raise NotImplementedError("point sources are not part of this library")
        if extinction_class is not None:
            raise NotImplementedError("differential extinction is not supported")
        self.type = "single-band"
        self.num_bands = 1
        self.PSF = psf_class
        self.Data = data_class
        self._flux_scaling = getattr(data_class, "flux_scaling", 1)
        self.PSF.set_pixel_size(self.Data.pixel_width)
        if lens_model_class is None:
            lens_model_class = LensModel(lens_model_list=[])
        if source_model_class is None:
            source_model_class = LightModel(light_model_list=[])
        if lens_light_model_class is None:
            lens_light_model_class = LightModel(light_model_list=[])
        self.LensModel = lens_model_class
        self.SourceModel = source_model_class
        self.LensLightModel = lens_light_model_class
        if precision is None:
            precision = getattr(lens_model_class, "precision", "fp64")
        if precision not in backend.PRECISIONS:
            raise ValueError("precision must be one of %s" % list(backend.PRECISIONS))
        self.precision = precision
        self._prec = backend.PRECISIONS[precision]
        if kwargs_numerics is None:
            kwargs_numerics = {}
        self._kwargs_numerics = kwargs_numerics
        self._init_numerics(**kwargs_numerics)

        nx, ny = data_class.num_pixel_axes
        if likelihood_mask is None:
            likelihood_mask = np.ones((nx, ny))
        self.likelihood_mask = np.array(likelihood_mask, dtype=bool)
        self.num_data_evaluate = int(np.sum(self.likelihood_mask))
        self._psf_error_map = False

        # static device state: grid, PSF kernel, mask, profile lists
        self._grid = backend.LensimGrid()
        self._grid.nx, self._grid.ny = int(nx), int(ny)
        self._grid.supersampling = self._s
        self._grid.ra_at_xy_0, self._grid.dec_at_xy_0 = data_class.radec_at_xy_0
        M = np.asarray(data_class.transform_pix2angle, dtype=np.float64)
        for i, v in enumerate([M[0, 0], M[0, 1], M[1, 0], M[1, 1]]):
            self._grid.pix2angle[i] = float(v)
        self._full_shape = (int(ny), int(nx))
        self._eval_mask_dev = None
        if self._frame is not None:
            x_min, y_min, n = self._frame
            self._grid.nx = self._grid.ny = n
            ra, dec = data_class.map_pix2coord(x_min, y_min)
            self._grid.ra_at_xy_0, self._grid.dec_at_xy_0 = float(ra), float(dec)
            self._eval_mask_dev = torch.as_tensor(self._eval_mask, device=backend.device())
            self._grid.d_eval_mask = self._eval_mask_dev.data_ptr()
        self._psf = backend.LensimPsf()
        self._kernel_dev = None
        if self._kernel is None:
            self._psf.type = backend.PSF_NONE
        else:
            self._kernel_dev = backend.to_device_f64(self._kernel)
            self._psf.type = backend.PSF_GAUSSIAN if self._psf_type == "GAUSSIAN" else backend.PSF_PIXEL
            self._psf.ky, self._psf.kx = self._kernel.shape
            self._psf.d_kernel = self._kernel_dev.data_ptr()
            if self._ss:
                self._psf.supersampled = 1
                if self._kernel_low is not None:
                    self._kernel_low_dev = backend.to_device_f64(self._kernel_low)
                    self._psf.ky_low, self._psf.kx_low = self._kernel_low.shape
                    self._psf.d_kernel_low = self._kernel_low_dev.data_ptr()
        self._mask_dev = None
        self._sub_masks = None  # sub-frame: likelihood mask inside / outside the rectangle
        self._sub_cache = None  # sub-frame: (device buffers of the ImageData, lensim_data, outside logL, keep)
        self._bulk = isinstance(self.LensModel, LensModelBulk)
        self._lens_types = [] if self._bulk else list(self.LensModel._types)
        self._src_types = list(self.SourceModel._types)
        self._light_types = list(self.LensLightModel._types)
        # slot layouts of every profile (see profiles.lens_layout): kwarg names,
        # or (name, value) for slots a lens profile's options fix
        self._names = ([] if self._bulk else list(self.LensModel._layouts)) + \
            list(self.SourceModel.param_name_list) + list(self.LensLightModel.param_name_list)
        self._model = backend.make_model(self._lens_types, self._src_types, self._light_types,
                                         alpha_scaling=self.LensModel.alpha_scaling, flux_scaling=self._flux_scaling)
        if not self._bulk:  # LensModel(multi_plane=True): its planes (kept alive by self.LensModel)
            self._model.multiplane = self.LensModel.multiplane_address() or None
        self._submodels = {}
        self._ws = backend.Workspace()

    # ------------------------------------------------------------ numerics --
    def _init_numerics(self, supersampling_factor=1, compute_mode="regular", supersampling_convolution=False,
                       supersampling_kernel_size=5, flux_evaluate_indexes=None, supersampled_indexes=None,
                       compute_indexes=None, point_source_supersampling_factor=1, convolution_kernel_size=None,
                       convolution_type="fft", truncation_conv=None):
        """Numerics.__init__, ImSim/Numerics/numerics.py:24-155."""
        if compute_mode != "regular":
            raise ValueError("compute_mode %s not supported; use 'regular'" % compute_mode)
        if not isinstance(supersampling_factor, int):
            raise TypeError("supersampling_factor needs to be an integer! Current type is %s"
                            % type(supersampling_factor))
        if supersampling_factor < 1:
            raise ValueError("supersampling_factor must be >= 1")
        if supersampling_factor == 1:
            supersampling_convolution = False
        self._frame = None
        self._eval_mask = None
        if flux_evaluate_indexes is not None:
            self._init_sub_frame(np.asarray(flux_evaluate_indexes, dtype=bool))
        if convolution_type not in ("fft", "grid", "fft_static"):
            raise ValueError("convolution_type %s not supported!" % convolution_type)
        self._s = supersampling_factor
        self._psf_type = self.PSF.psf_type
        self._ss = bool(supersampling_convolution) and self._psf_type != "NONE"
        self._kernel_low = None
        pixel_width = self.Data.pixel_width
        if self._ss and self._psf_type == "PIXEL":
            # SubgridKernelConvolution (ImSim/Numerics/convolution.py:100-166, numerics.py:103-116)
            kernel = np.asarray(self.PSF.kernel_point_source_supersampled(supersampling_factor), dtype=np.float64)
            if convolution_kernel_size is not None:
                size = convolution_kernel_size * supersampling_factor
                if size % 2 == 0:
                    size += 1
                kernel = cut_psf(kernel, size)
            if supersampling_kernel_size is not None:
                self._kernel_low, kernel = split_kernel(kernel, supersampling_kernel_size, supersampling_factor)
                self._kernel_low = np.ascontiguousarray(self._kernel_low)
            self._kernel = np.ascontiguousarray(kernel)
        elif self._ss and self._psf_type == "GAUSSIAN":
            # GaussianConvolution with supersampling_convolution (convolution.py:169-252)
            trunc = self.PSF.truncation if truncation_conv is None else truncation_conv
            self._kernel = gaussian_pixel_kernel(self.PSF.fwhm, pixel_width, trunc, supersampling_factor, True)
        elif self._psf_type == "PIXEL":
            kernel = np.asarray(self.PSF.kernel_point_source, dtype=np.float64)
            if convolution_kernel_size is not None:
                size = convolution_kernel_size
                if size % 2 == 0:
                    size += 1
                kernel = cut_psf(kernel, size)
            self._kernel = np.ascontiguousarray(kernel)
        elif self._psf_type == "GAUSSIAN":
            trunc = self.PSF.truncation if truncation_conv is None else truncation_conv
            self._kernel = gaussian_pixel_kernel(self.PSF.fwhm, pixel_width, trunc, supersampling_factor, False)
        elif self._psf_type == "NONE":
            self._kernel = None
        else:
            raise ValueError("psf_type %s not valid! Chose either NONE, GAUSSIAN or PIXEL." % self._psf_type)

    def _init_sub_frame(self, mask):
        """NumericsSubFrame (ImSim/Numerics/numerics_subframe.py:61-101, :135-191):
        the model is computed on the smallest rectangle holding every True pixel
        (rays shot only there, the convolution zero-padded at the rectangle's
        edge) and the rest of the frame is zero.  As in the reference, the
        rectangle's first-axis range [x_min, x_max] and second-axis range
        [y_min, y_max] index the image as [x_min:x_max+1, y_min:y_max+1] while
        the rectangle's grid starts at map_pix2coord(x_min, y_min), and the
        rectangle must be square (the reference assigns an (ny_sub, nx_sub)
        image into an (nx_sub, ny_sub) slot)."""
        nx, ny = self.Data.num_pixel_axes
        if mask.shape != (ny, nx) and mask.shape != (nx, ny):
            raise ValueError("flux_evaluate_indexes has shape %s for a %dx%d image" % (mask.shape, ny, nx))
        a, b = np.where(mask)
        if a.size == 0:
            raise ValueError("flux_evaluate_indexes selects no pixel")
        x_min, x_max, y_min, y_max = int(a.min()), int(a.max()), int(b.min()), int(b.max())
        n1, n2 = x_max - x_min + 1, y_max - y_min + 1
        if n1 != n2:
            raise ValueError("the sub-frame of flux_evaluate_indexes is %d x %d; the reference only handles square "
                             "sub-frames" % (n1, n2))
        self._frame = (x_min, y_min, n1)
        self._eval_mask = np.ascontiguousarray(mask[x_min:x_max + 1, y_min:y_max + 1].astype(np.uint8))

    @property
    def kernel(self):
        """The convolution kernel used for the PSF (None for psf_type NONE)."""
        return self._kernel

    # ------------------------------------------------------------- packing --
    def pack_bulk(self, kwargs_lens, kwargs_source=None, kwargs_lens_light=None):
        """For a LensModelBulk deflector: (params [B, n, 8], B, model, keep-alive) where
        kwargs_lens is the dict of LensModelBulk.prepare_ray_shooting_kwargs."""
        n_img, it, ip, bt, bp = self.LensModel.pack(
            kwargs_lens, max_inline=backend.MAX_PROFILES - len(self._src_types) - len(self._light_types))
        ks = kwargs_source or []
        kll = kwargs_lens_light or []
        nb = profiles.batch_size([ks, kll])
        if nb != 1 and n_img != 1 and nb != n_img:
            raise ValueError("inconsistent batch sizes %d and %d" % (n_img, nb))
        n = max(n_img, nb)
        light = profiles.pack(self._names, list(ks if self._src_types else []) +
                              list(kll if self._light_types else []), n)
        ip = np.broadcast_to(ip, (n,) + ip.shape[1:])
        bp = np.broadcast_to(bp, (n,) + bp.shape[1:])
        params = np.ascontiguousarray(np.concatenate([ip, bp, light], axis=1))
        m, keep = self.LensModel._model(it, bt, self._src_types, self._light_types, self._flux_scaling,
                                        ray_shooting_kwargs=kwargs_lens)
        return params, n, m, keep

    def pack(self, kwargs_lens=None, kwargs_source=None, kwargs_lens_light=None):
        """Parameter block [B, n_profiles, 8] (float64, host) of lists of kwargs."""
        if self._bulk:
            raise ValueError("use pack_bulk with a LensModelBulk deflector")
        kl = kwargs_lens or []
        ks = kwargs_source or []
        kll = kwargs_lens_light or []
        if len(kl) != len(self._lens_types) and len(self._lens_types) > 0:
            raise ValueError("expected %d lens kwargs, got %d" % (len(self._lens_types), len(kl)))
        if len(ks) != len(self._src_types) and len(self._src_types) > 0:
            raise ValueError("expected %d source kwargs, got %d" % (len(self._src_types), len(ks)))
        if len(kll) != len(self._light_types) and len(self._light_types) > 0:
            raise ValueError("expected %d lens light kwargs, got %d" % (len(self._light_types), len(kll)))
        kl = kl if self._lens_types else []
        ks = ks if self._src_types else []
        kll = kll if self._light_types else []
        n_img = profiles.batch_size([kl, ks, kll])
        return profiles.pack(self._names, list(kl) + list(ks) + list(kll), n_img), n_img

    def _submodel(self, lens_idx, src_idx, light_idx):
        key = (tuple(lens_idx), tuple(src_idx), tuple(light_idx))
        if key not in self._submodels:
            nl, ns = len(self._lens_types), len(self._src_types)
            rows = list(lens_idx) + [nl + i for i in src_idx] + [nl + ns + i for i in light_idx]
            m = backend.make_model([self._lens_types[i] for i in lens_idx], [self._src_types[i] for i in src_idx],
                                   [self._light_types[i] for i in light_idx],
                                   alpha_scaling=self.LensModel.alpha_scaling, flux_scaling=self._flux_scaling)
            if self._model.multiplane:  # the planes hold every deflector of the list
                if list(lens_idx) != list(range(nl)):
                    raise ValueError("no individual deflector can be evaluated in multi-plane lensing")
                m.multiplane = self._model.multiplane
            self._submodels[key] = (m, rows)
        return self._submodels[key]

    # ------------------------------------------------------------ batch API --
    def workspace_bytes(self, n_img, model=None):
        lib = backend.lib()
        m = self._model if model is None else model
        return lib.lensim_workspace_size(ctypes.byref(m), ctypes.byref(self._grid), ctypes.byref(self._psf), n_img,
                                         self._prec)

    def image_batch(self, params, n_img=None, flags=backend.IMAGE_SOURCE | backend.IMAGE_LENS_LIGHT, out=None,
                    model=None):
        """Images of a packed device parameter block [B, n_profiles, 8]; returns
        [B, ny, nx] in the compute dtype (float64 or float32)."""
        lib = backend.lib()
        m = self._model if model is None else model
        if n_img is None:
            n_img = params.shape[0]
        dtype = torch.float64 if self._prec == backend.FP64 else torch.float32
        if out is None:
            out = torch.empty((n_img,) + self._full_shape, dtype=dtype, device=params.device)
        sub = out
        if self._frame is not None:
            sub = torch.empty((n_img, self._grid.ny, self._grid.nx), dtype=dtype, device=params.device)
        ws = self._ws.get(self.workspace_bytes(n_img, m))
        rc = lib.lensim_image(ctypes.byref(m), ctypes.byref(self._grid), ctypes.byref(self._psf),
                              backend.ptr(params), n_img, self._prec, flags, backend.ptr(sub), backend.ptr(ws),
                              ws.numel(), backend.stream_handle())
        backend.check(rc, "lensim_image")
        if self._frame is not None:
            self._complete(sub, out)
        return out

    def _data_struct(self, data=None, noise_map=None):
        """lensim_data of the ImageData; ``data`` ([B, ny, nx] float64 device
        tensor) gives each batch image its own observation instead, and
        ``noise_map`` (same shape) its own noise map."""
        buf = self.Data.device_buffers()
        if self._mask_dev is None:
            self._mask_dev = backend.to_device_f64(self.likelihood_mask.astype(np.float64))
        d = backend.LensimData()
        d.d_data = buf["data"].data_ptr()
        d.d_noise_map = None if buf["noise"] is None else buf["noise"].data_ptr()
        d.d_exposure_map = None if buf["expo_map"] is None else buf["expo_map"].data_ptr()
        d.d_mask = self._mask_dev.data_ptr()
        d.exposure_time = buf["expo_time"]
        d.background_rms = self.Data.background_rms
        if data is not None:
            # observations cover the whole frame (with flux_evaluate_indexes the
            # likelihood runs on the completed frame, not on the sub-frame)
            ny, nx = self._full_shape
            for t in (data, noise_map):
                if t is not None and (t.dtype != torch.float64 or not t.is_contiguous() or
                                      tuple(t.shape) != tuple(data.shape) or tuple(t.shape[-2:]) != (ny, nx)):
                    raise ValueError("per-image data / noise maps must be contiguous float64 [B, %d, %d] tensors"
                                     % (ny, nx))
            if buf["noise"] is not None and noise_map is None:
                raise ValueError("per-image data with a noise map needs per-image noise maps")
            d.d_data = data.data_ptr()
            d.d_noise_map = None if noise_map is None else noise_map.data_ptr()
            d.image_stride = nx * ny
        elif noise_map is not None:
            raise ValueError("per-image noise maps need per-image data")
        return d

    def _sub_frame_data(self, data=None, noise_map=None, n_img=1):
        """(lensim_data cut to the sub-frame rectangle, float64 logL [1, or n_img
        with per-image data] of the zero model on the pixels outside it, tensors
        to keep alive).  The model is zero outside the rectangle
        (NumericsSubFrame._complete_frame, numerics_subframe.py:135-148), so those
        pixels add a term to ImageData.log_likelihood that no parameter changes,
        and nothing to the gradient or to the weighted normal equations of the
        linear solve (their response rows are zero): both run on the rectangle
        with the data, noise map, exposure map and likelihood mask cut to it."""
        lib = backend.lib()
        buf = self.Data.device_buffers()
        if data is None and self._sub_cache is not None and self._sub_cache[0] is buf:
            return self._sub_cache[1:]
        x0, y0, n = self._frame
        ny, nx = self._full_shape
        if self._sub_masks is None:
            inside = self.likelihood_mask[x0:x0 + n, y0:y0 + n]
            outside = self.likelihood_mask.copy()
            outside[x0:x0 + n, y0:y0 + n] = False
            self._sub_masks = (backend.to_device_f64(np.ascontiguousarray(inside, dtype=np.float64)),
                               backend.to_device_f64(outside.astype(np.float64)))
        # the outside term: ImageData.log_likelihood of a zero model under the outside mask
        full = self._data_struct(data, noise_map)
        full.d_mask = self._sub_masks[1].data_ptr()
        n_out = 1 if data is None else int(n_img)
        dtype = torch.float64 if self._prec == backend.FP64 else torch.float32
        zero = torch.zeros((n_out, ny, nx), dtype=dtype, device=backend.device())
        out_logl = torch.empty(n_out, dtype=torch.float64, device=backend.device())
        ws = self._ws.get(lib.lensim_data_loglike_workspace_size(nx, ny, n_out))
        rc = lib.lensim_data_loglike(ctypes.byref(full), nx, ny, backend.ptr(zero), n_out, self._prec,
                                     backend.ptr(out_logl), backend.ptr(ws), ws.numel(), backend.stream_handle())
        backend.check(rc, "lensim_data_loglike")

        def cut(t):
            return None if t is None else t[..., x0:x0 + n, y0:y0 + n].contiguous()
        keep = [cut(buf["data"] if data is None else data), cut(noise_map if data is not None else buf["noise"]),
                cut(buf["expo_map"]), self._sub_masks[0], zero]
        d = backend.LensimData()
        d.d_data = keep[0].data_ptr()
        d.d_noise_map = None if keep[1] is None else keep[1].data_ptr()
        d.d_exposure_map = None if keep[2] is None else keep[2].data_ptr()
        d.d_mask = keep[3].data_ptr()
        d.exposure_time = full.exposure_time
        d.background_rms = full.background_rms
        d.image_stride = n * n if data is not None else 0
        if data is None:
            self._sub_cache = (buf, d, out_logl, keep)
        return d, out_logl, keep

    def _complete(self, sub, full):
        """NumericsSubFrame._complete_frame (numerics_subframe.py:135-148): the
        sub-frame images [..., n, n] placed in zeroed full frames [..., ny, nx]."""
        x0, y0, n = self._frame
        full.zero_()
        full[..., x0:x0 + n, y0:y0 + n] = sub
        return full

    def likelihood_batch(self, params, n_img=None, model_out=None, logl=None, model=None, data=None,
                         noise_map=None):
        """log-likelihoods [B] (float64) of a packed device parameter block,
        against the ImageData or, with ``data`` [B, ny, nx] (and optionally
        ``noise_map`` [B, ny, nx]), one observation per batch image."""
        lib = backend.lib()
        m = self._model if model is None else model
        if n_img is None:
            n_img = params.shape[0]
        if data is not None and data.shape[0] < n_img:
            raise ValueError("per-image data holds %d images for a batch of %d" % (data.shape[0], n_img))
        if logl is None:
            logl = torch.empty(n_img, dtype=torch.float64, device=params.device)
        d = self._data_struct(data, noise_map)
        if self._frame is not None:
            # the model is zero outside the sub-frame, so the likelihood runs on
            # the completed frame (ImageData.log_likelihood)
            model_img = self.image_batch(params, n_img, model=m, out=model_out)
            ny, nx = self._full_shape
            nbytes = lib.lensim_data_loglike_workspace_size(nx, ny, n_img)
            ws = self._ws.get(nbytes)
            rc = lib.lensim_data_loglike(ctypes.byref(d), nx, ny, backend.ptr(model_img), n_img, self._prec,
                                         backend.ptr(logl), backend.ptr(ws), ws.numel(), backend.stream_handle())
            backend.check(rc, "lensim_data_loglike")
            return logl
        ws = self._ws.get(self.workspace_bytes(n_img, m))
        rc = lib.lensim_loglike(ctypes.byref(m), ctypes.byref(self._grid), ctypes.byref(self._psf),
                                ctypes.byref(d), backend.ptr(params), n_img, self._prec, backend.ptr(logl),
                                backend.ptr(model_out), backend.ptr(ws), ws.numel(), backend.stream_handle())
        backend.check(rc, "lensim_loglike")
        return logl

    def likelihood_and_grad_batch(self, params, n_img=None, model=None):
        """(logL [B], dlogL/dparams [B, n_profiles, 8]) of a packed device parameter
        block; the gradient slot layout is the parameter-block layout.  A
        LensModelBulk deflector passes the lensim_model of pack_bulk as model
        (single-plane and multi-plane lists alike)."""
        lib = backend.lib()
        if self._bulk and model is None:
            raise ValueError("a LensModelBulk deflector needs the model of pack_bulk")
        m = self._model if model is None else model
        if n_img is None:
            n_img = params.shape[0]
        params = params.contiguous()
        logl = torch.empty(n_img, dtype=torch.float64, device=params.device)
        grad = torch.empty((n_img, params.shape[1], backend.NPARAM), dtype=torch.float64, device=params.device)
        outside = None
        if self._frame is None:
            d = self._data_struct()
        else:  # the sub-frame rectangle, plus the parameter-independent outside term
            d, outside, _ = self._sub_frame_data()
        nbytes = lib.lensim_grad_workspace_size(ctypes.byref(m), ctypes.byref(self._grid),
                                                ctypes.byref(self._psf), n_img, self._prec)
        if nbytes == 0:
            raise ValueError("gradient workspace query failed: %s" % lib.lensim_last_error().decode())
        ws = self._ws.get(nbytes)
        rc = lib.lensim_loglike_grad(ctypes.byref(m), ctypes.byref(self._grid), ctypes.byref(self._psf),
                                     ctypes.byref(d), backend.ptr(params), n_img, self._prec, backend.ptr(logl),
                                     backend.ptr(grad), backend.ptr(ws), ws.numel(), backend.stream_handle())
        backend.check(rc, "lensim_loglike_grad")
        if outside is not None:
            logl += outside
        return logl, grad

    def log_likelihood_torch(self, params):
        """Differentiable logL [B] of a packed parameter tensor [B, n_profiles, 8]:
        torch.autograd calls the library's adjoint kernels in backward."""
        return _LogLikelihood.apply(params, self)

    def likelihood_gradient(self, kwargs_lens=None, kwargs_source=None, kwargs_lens_light=None):
        """logL and its gradient in kwargs form: (logL, [grad dicts for lens],
        [source], [lens light]), like jax.grad of likelihood_data_given_model
        (test/test_ImSim/test_image_model.py:274-335).  With a LensModelBulk
        deflector, kwargs_lens is its ray_shooting_kwargs and the lens gradient is
        a dict in the form of its "all_kwargs" (LensModelBulk.unpack_gradient)."""
        if self._bulk:
            params, n_img, m, keep = self.pack_bulk(kwargs_lens, kwargs_source, kwargs_lens_light)
            logl, grad = self.likelihood_and_grad_batch(backend.to_device_f64(params), n_img, model=m)
            del keep
            g = grad.cpu().numpy()
            n_defl = g.shape[1] - len(self._src_types) - len(self._light_types)
            glens = self.LensModel.unpack_gradient(kwargs_lens, g[:, :n_defl])
            g = g[:, n_defl:]
        else:
            params, n_img = self.pack(kwargs_lens, kwargs_source, kwargs_lens_light)
            logl, grad = self.likelihood_and_grad_batch(backend.to_device_f64(params), n_img)
            g = grad.cpu().numpy()
        out, k = [], 0
        for types in (self._lens_types, self._src_types, self._light_types):
            lst = []
            for _ in types:
                lst.append({nm: (g[:, k, i] if n_img > 1 else float(g[0, k, i]))
                            for i, nm in profiles.kwarg_slots(self._names[k])})
                k += 1
            out.append(lst)
        if self._bulk:
            out[0] = glens
        return (logl if n_img > 1 else logl[0]), out[0], out[1], out[2]

    # ------------------------------------------------------------ drop-in API --
    @staticmethod
    def _check_unsupported(kwargs_ps=None, kwargs_extinction=None):
        if kwargs_ps:
            raise NotImplementedError("point sources are not part of this library")
        if kwargs_extinction:
            raise ValueError("Extinction is not implemented")

    def _render(self, lens_idx, src_idx, light_idx, kwargs_lens, kwargs_source, kwargs_lens_light, flags):
        lens_idx, src_idx, light_idx = list(lens_idx), list(src_idx), list(light_idx)
        kl = [kwargs_lens[i] for i in lens_idx] if lens_idx else []
        ks = [kwargs_source[i] for i in src_idx] if src_idx else []
        kll = [kwargs_lens_light[i] for i in light_idx] if light_idx else []
        nl, ns = len(self._lens_types), len(self._src_types)
        names = ([self._names[i] for i in lens_idx] + [self._names[nl + i] for i in src_idx] +
                 [self._names[nl + ns + i] for i in light_idx])
        n_img = profiles.batch_size([kwargs_lens or [], kwargs_source or [], kwargs_lens_light or []])
        params = profiles.pack(names, kl + ks + kll, n_img)
        model, _ = self._submodel(lens_idx, src_idx, light_idx)
        pt = backend.to_device_f64(params)
        img = self.image_batch(pt, n_img, flags=flags, model=model).to(torch.float64)
        return img if n_img > 1 else img[0]

    def image(self, kwargs_lens=None, kwargs_source=None, kwargs_lens_light=None, kwargs_ps=None,
              kwargs_extinction=None, kwargs_special=None, unconvolved=False, source_add=True, lens_light_add=True,
              point_source_add=True):
        """ImageModel.image, ImSim/image_model.py:431-492."""
        self._check_unsupported(kwargs_ps, kwargs_extinction)
        flags = 0
        if source_add:
            flags |= backend.IMAGE_SOURCE
        if lens_light_add:
            flags |= backend.IMAGE_LENS_LIGHT
        if unconvolved:
            flags |= backend.IMAGE_UNCONVOLVED
        if self._bulk:
            params, n_img, m, keep = self.pack_bulk(kwargs_lens, kwargs_source, kwargs_lens_light)
            img = self.image_batch(backend.to_device_f64(params), n_img, flags=flags, model=m).to(torch.float64)
            del keep
            return img if n_img > 1 else img[0]
        nl, ns, nll = len(self._lens_types), len(self._src_types), len(self._light_types)
        return self._render(range(nl), range(ns), range(nll), kwargs_lens, kwargs_source, kwargs_lens_light, flags)

    def source_surface_brightness(self, kwargs_source, kwargs_lens=None, kwargs_extinction=None,
                                  kwargs_special=None, unconvolved=False, de_lensed=False, k=None,
                                  update_pixelbased_mapping=True):
        """ImageModel.source_surface_brightness, ImSim/image_model.py:223-368."""
        self._check_unsupported(None, kwargs_extinction)
        ns = len(self._src_types)
        if ns == 0:
            nx, ny = self.Data.num_pixel_axes
            return torch.zeros((nx, ny), dtype=torch.float64, device=backend.device())
        src_idx = profiles.select(k, ns)
        lens_idx = [] if de_lensed else range(len(self._lens_types))
        flags = backend.IMAGE_SOURCE | (backend.IMAGE_UNCONVOLVED if unconvolved else 0)
        return self._render(lens_idx, src_idx, [], kwargs_lens, kwargs_source, None, flags)

    def lens_surface_brightness(self, kwargs_lens_light, unconvolved=False, k=None):
        """ImageModel.lens_surface_brightness, ImSim/image_model.py:370-393."""
        nll = len(self._light_types)
        if nll == 0:
            nx, ny = self.Data.num_pixel_axes
            return torch.zeros((nx, ny), dtype=torch.float64, device=backend.device())
        light_idx = profiles.select(k, nll)
        flags = backend.IMAGE_LENS_LIGHT | (backend.IMAGE_UNCONVOLVED if unconvolved else 0)
        return self._render([], [], light_idx, None, None, kwargs_lens_light, flags)

    def likelihood_data_given_model(self, kwargs_lens=None, kwargs_source=None, kwargs_lens_light=None,
                                    kwargs_ps=None, kwargs_extinction=None, kwargs_special=None, source_marg=False,
                                    linear_prior=None, check_positive_flux=False):
        """ImageModel.likelihood_data_given_model, ImSim/image_model.py:162-221."""
        if check_positive_flux:
            raise ValueError("check positive flux is not supported")
        self._check_unsupported(kwargs_ps, kwargs_extinction)
        if self._bulk:
            params, n_img, m, keep = self.pack_bulk(kwargs_lens, kwargs_source, kwargs_lens_light)
            logl = self.likelihood_batch(backend.to_device_f64(params), n_img, model=m)
            del keep
            return logl if n_img > 1 else logl[0]
        params, n_img = self.pack(kwargs_lens, kwargs_source, kwargs_lens_light)
        pt = backend.to_device_f64(params)
        logl = self.likelihood_batch(pt, n_img)
        return logl if n_img > 1 else logl[0]

    def update_psf(self, psf_class):
        raise ValueError("Updating psf class not supported. Create a new instance of ImageModel instead.")

    def update_data(self, data_class):
        raise ValueError("Updating data class not supported. Create a new instance of ImageModel instead.")


class _LogLikelihood(torch.autograd.Function):
    """logL of packed parameters with the library's adjoint as the backward."""

    @staticmethod
    def forward(ctx, params, image_model):
        logl, grad = image_model.likelihood_and_grad_batch(params.detach().to(torch.float64))
        ctx.save_for_backward(grad)
        return logl

    @staticmethod
    def backward(ctx, grad_out):
        (grad,) = ctx.saved_tensors
        return grad_out[:, None, None] * grad, None
