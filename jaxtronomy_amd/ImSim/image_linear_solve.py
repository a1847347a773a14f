"""ImageLinearFit: ImageModel with the light amplitudes solved linearly.

Mirrors the reference's ``ImageLinearFit`` (ImSim/image_linear_solve.py:16-181,
:229-500) for the diagonal likelihood: every source and lens-light profile's
'amp' is a linear parameter; the library renders each profile's response at
amp = 1 (through ray shooting, supersampling and the PSF), forms the weighted
normal equations against the data, solves them per image of the batch
(de_lens.get_param_WLS: zero amplitudes when cond(M) >= 5 / eps) and evaluates
the likelihood of the solved model, all on the device (lensim_linear_solve).
"""

import ctypes

import numpy as np
import torch

from .. import backend
from .image_model import ImageModel


class ImageLinearFit(ImageModel):
    def __init__(self, data_class, psf_class=None, lens_model_class=None, source_model_class=None,
                 lens_light_model_class=None, point_source_class=None, extinction_class=None, kwargs_numerics=None,
                 likelihood_mask=None, psf_error_map_bool_list=None, kwargs_pixelbased=None, precision=None):
        super(ImageLinearFit, self).__init__(data_class, psf_class, lens_model_class, source_model_class,
                                             lens_light_model_class, point_source_class, extinction_class,
                                             kwargs_numerics, likelihood_mask, psf_error_map_bool_list,
                                             kwargs_pixelbased, precision)
        if getattr(self.Data, "likelihood_method", lambda: "diagonal")() != "diagonal":
            raise ValueError("only the diagonal likelihood method is supported")
        self._n_lin = len(self._src_types) + len(self._light_types)

    # ------------------------------------------------------------ batch API --
    def linear_solve_batch(self, params, n_img=None, source_marg=False, linear_prior=None, model=None,
                           want_model=False, want_cov=False, want_response=False, data=None):
        """Linear solve of a packed device parameter block [B, n_profiles, 8]
        (the light amplitudes in it are ignored).  Returns a dict with 'logl'
        [B], 'amp' [B, n_lin] and, on request, 'model' [B, ny, nx], 'cov'
        [B, n_lin, n_lin] (M^-1) and 'response' [n_lin, B, ny, nx]."""
        lib = backend.lib()
        m = self._model if model is None else model
        if n_img is None:
            n_img = params.shape[0]
        params = params.contiguous()
        dev = params.device
        dtype = torch.float64 if self._prec == backend.FP64 else torch.float32
        if data is not None and data.shape[0] < n_img:
            raise ValueError("per-image data holds %d images for a batch of %d" % (data.shape[0], n_img))
        # with a sub-frame the solve runs on its rectangle (the responses are zero
        # outside it), the images are completed and logL gains the outside term
        ny, nx = self._full_shape if self._frame is None else (self._grid.ny, self._grid.nx)
        out = {"logl": torch.empty(n_img, dtype=torch.float64, device=dev),
               "amp": torch.empty((n_img, self._n_lin), dtype=torch.float64, device=dev)}
        if want_model:
            out["model"] = torch.empty((n_img, ny, nx), dtype=dtype, device=dev)
        if want_cov or source_marg:
            out["cov"] = torch.empty((n_img, self._n_lin, self._n_lin), dtype=torch.float64, device=dev)
        if want_response:
            out["response"] = torch.empty((self._n_lin, n_img, ny, nx), dtype=dtype, device=dev)
        outside = None
        if self._frame is None:
            d = self._data_struct(data)
        else:
            d, outside, keep = self._sub_frame_data(data, None, n_img)
        nbytes = lib.lensim_linear_workspace_size(ctypes.byref(m), ctypes.byref(self._grid),
                                                  ctypes.byref(self._psf), n_img, self._prec)
        if nbytes == 0:
            raise ValueError("linear solve workspace query failed: %s" % lib.lensim_last_error().decode())
        ws = self._ws.get(nbytes)
        prior = -1.0 if linear_prior is None else float(linear_prior)
        rc = lib.lensim_linear_solve(ctypes.byref(m), ctypes.byref(self._grid), ctypes.byref(self._psf),
                                     ctypes.byref(d), backend.ptr(params), n_img, self._prec, int(bool(source_marg)),
                                     prior, backend.ptr(out["amp"]), backend.ptr(out["logl"]),
                                     backend.ptr(out.get("model")), backend.ptr(out.get("cov")),
                                     backend.ptr(out.get("response")), backend.ptr(ws), ws.numel(),
                                     backend.stream_handle())
        backend.check(rc, "lensim_linear_solve")
        if outside is not None:
            out["logl"] += outside
            fy, fx = self._full_shape
            for key in ("model", "response"):
                if key in out:
                    out[key] = self._complete(out[key], out[key].new_empty(out[key].shape[:-2] + (fy, fx)))
        return out

    def _packed(self, kwargs_lens, kwargs_source, kwargs_lens_light):
        """Parameters with placeholder amplitudes (the solve sets them)."""
        ks = [dict(k, amp=k.get("amp", 1.0)) for k in (kwargs_source or [])]
        kll = [dict(k, amp=k.get("amp", 1.0)) for k in (kwargs_lens_light or [])]
        if self._bulk:
            params, n_img, m, keep = self.pack_bulk(kwargs_lens, ks, kll)
            return backend.to_device_f64(params), n_img, m, keep
        params, n_img = self.pack(kwargs_lens, ks, kll)
        return backend.to_device_f64(params), n_img, None, None

    # ------------------------------------------------------------ drop-in API --
    def image_linear_solve(self, kwargs_lens=None, kwargs_source=None, kwargs_lens_light=None, kwargs_ps=None,
                           kwargs_extinction=None, kwargs_special=None, inv_bool=False):
        """ImageLinearFit.image_linear_solve, ImSim/image_linear_solve.py:98-181:
        (model image, model error (0: no PSF error map), M^-1 or None, amplitudes).
        The amplitudes are also written into kwargs_source / kwargs_lens_light
        (update_linear_kwargs)."""
        self._check_unsupported(kwargs_ps, kwargs_extinction)
        params, n_img, m, keep = self._packed(kwargs_lens, kwargs_source, kwargs_lens_light)
        r = self.linear_solve_batch(params, n_img, model=m, want_model=True, want_cov=inv_bool)
        del keep
        amp = r["amp"].cpu().numpy()
        model = r["model"].to(torch.float64)
        cov = r["cov"].cpu().numpy() if inv_bool else None
        if n_img == 1:
            amp, model = amp[0], model[0]
            cov = None if cov is None else cov[0]
        self.update_linear_kwargs(amp, kwargs_lens, kwargs_source, kwargs_lens_light, kwargs_ps)
        return model, 0, cov, amp

    def likelihood_data_given_model(self, kwargs_lens=None, kwargs_source=None, kwargs_lens_light=None,
                                    kwargs_ps=None, kwargs_extinction=None, kwargs_special=None, source_marg=False,
                                    linear_prior=None, check_positive_flux=False):
        """ImageLinearFit.likelihood_data_given_model, ImSim/image_linear_solve.py:229-340:
        (logL, amplitudes); with check_positive_flux a negative amplitude costs 1e8."""
        self._check_unsupported(kwargs_ps, kwargs_extinction)
        params, n_img, m, keep = self._packed(kwargs_lens, kwargs_source, kwargs_lens_light)
        r = self.linear_solve_batch(params, n_img, source_marg=source_marg, linear_prior=linear_prior, model=m)
        del keep
        logl = r["logl"]
        if check_positive_flux:
            logl = torch.where(torch.any(r["amp"] < 0, dim=1), logl - 1e8, logl)
        amp = r["amp"].cpu().numpy()
        if n_img == 1:
            return logl[0], amp[0]
        return logl, amp

    def linear_response_matrix(self, kwargs_lens=None, kwargs_source=None, kwargs_lens_light=None, kwargs_ps=None,
                               kwargs_extinction=None, kwargs_special=None, unconvolved=False):
        """ImageLinearFit.linear_response_matrix, ImSim/image_linear_solve.py:362-455:
        [n_lin, n_masked] (the responses at amp = 1, likelihood-mask pixels only)."""
        if unconvolved:
            raise NotImplementedError("unconvolved response matrices are not supported")
        self._check_unsupported(kwargs_ps, kwargs_extinction)
        params, n_img, m, keep = self._packed(kwargs_lens, kwargs_source, kwargs_lens_light)
        if n_img != 1:
            raise ValueError("linear_response_matrix takes one model; use linear_solve_batch for batches")
        r = self.linear_solve_batch(params, n_img, model=m, want_response=True)
        del keep
        resp = r["response"][:, 0].to(torch.float64)
        sel = torch.as_tensor(self.likelihood_mask.ravel(), device=resp.device)
        return resp.reshape(self._n_lin, -1)[:, sel]

    def num_param_linear(self, kwargs_lens=None, kwargs_source=None, kwargs_lens_light=None, kwargs_ps=None):
        """Number of linear amplitudes (one per light profile)."""
        return self._n_lin

    def update_linear_kwargs(self, param, kwargs_lens, kwargs_source, kwargs_lens_light, kwargs_ps=None):
        """Writes the amplitudes into the kwargs lists ('amp' of every light profile,
        sources first), image_linear_solve.py:458-482."""
        param = np.asarray(param)
        i = 0
        for kw in (kwargs_source or []):
            kw["amp"] = param[..., i]
            i += 1
        for kw in (kwargs_lens_light or []):
            kw["amp"] = param[..., i]
            i += 1
        return kwargs_lens, kwargs_source, kwargs_lens_light, kwargs_ps

    def linear_param_from_kwargs(self, kwargs_source, kwargs_lens_light, kwargs_ps=None):
        """The amplitudes of the kwargs lists, image_linear_solve.py:484-498."""
        return [kw["amp"] for kw in (kwargs_source or [])] + [kw["amp"] for kw in (kwargs_lens_light or [])]
