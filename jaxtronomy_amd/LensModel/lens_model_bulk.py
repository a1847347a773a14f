"""LensModelBulk: ray shooting through many deflectors (e.g. substructure).

Mirrors the reference's ``LensModelBulk`` / ``SinglePlaneBulk``
(LensModel/lens_model_bulk.py:10-128, LensModel/single_plane_bulk.py:11-131):
``prepare_ray_shooting_kwargs`` packs a lens list into per-parameter arrays plus
an index list into the unique profile list, padded with NULL deflectors; and
``ray_shooting`` evaluates the sum.  The device kernel loops over the
deflectors inside one thread per ray and keeps the deflection sum in
registers (float64), instead of stacking one deflection field per deflector.

EPL entries run as the kernel's inline deflectors (they carry per-image series
coefficients).  Of the other profiles, the most frequent type goes to the bulk
list and the rest run inline too while they fit (at most 64 inline profiles,
light profiles included), so that a substructure list (a main deflector plus
many subhalos of one type) keeps the kernel's single-type bulk loop.  A kwargs
array may also be 2-D, [B, n_deflectors], for B models at once.
"""

import ctypes

import numpy as np
import torch

from .. import backend, profiles

_BULK_OK = ("SIE", "NIE", "NFW", "TNFW", "SHEAR")


class LensModelBulk(object):
    def __init__(self, unique_lens_model_list, multi_plane=False, profile_kwargs_list=None, cosmo=None,
                 cosmology_model="FlatLambdaCDM", precision="fp64"):
        if multi_plane:
            raise NotImplementedError("multi-plane bulk ray shooting is not part of this library")
        if precision not in backend.PRECISIONS:
            raise ValueError("precision must be one of %s" % list(backend.PRECISIONS))
        for name in unique_lens_model_list:
            profiles.lens_type(name)
            if name not in _BULK_OK and name != "EPL":
                raise ValueError("%s is not supported in LensModelBulk" % name)
        self.unique_lens_model_list = list(unique_lens_model_list)
        self.param_name_list = [profiles.LENS_PROFILES[n][1] for n in self.unique_lens_model_list]
        self.type = "SinglePlaneBulk"
        self.precision = precision
        self.alpha_scaling = 1.0
        self.init_kwargs = {"unique_lens_model_list": self.unique_lens_model_list, "multi_plane": False,
                            "profile_kwargs_list": profile_kwargs_list, "cosmo": cosmo,
                            "cosmology_model": cosmology_model}
        self._ws = backend.Workspace()

    def prepare_ray_shooting_kwargs(self, lens_model_list, kwargs_lens, z_source=None, lens_redshift_list=None,
                                    num_deflectors=None):
        """SinglePlaneBulk.prepare_ray_shooting_kwargs, single_plane_bulk.py:47-106."""
        if len(lens_model_list) != len(kwargs_lens):
            raise ValueError("length of lens model list %d and length of kwargs_lens %d do not match"
                             % (len(lens_model_list), len(kwargs_lens)))
        if num_deflectors is None:
            num_deflectors = len(lens_model_list)
        elif num_deflectors < len(lens_model_list):
            raise ValueError("The provided num_deflectors %d is smaller than the length of lens_model_list %d"
                             % (num_deflectors, len(lens_model_list)))
        num_filler = num_deflectors - len(kwargs_lens)
        index_list = [self.unique_lens_model_list.index(m) for m in lens_model_list]
        index_list += [len(self.unique_lens_model_list)] * num_filler
        unique_kwargs = set(p for sub in self.param_name_list for p in sub)
        all_kwargs = {}
        for kwarg in unique_kwargs:
            vals = [kw.get(kwarg, 0) for kw in kwargs_lens] + [0] * num_filler
            all_kwargs[kwarg] = np.array(vals)
        return {"all_kwargs": all_kwargs, "index_list": np.array(index_list)}

    @staticmethod
    def stack_ray_shooting_kwargs(prepared):
        """Stacks prepare_ray_shooting_kwargs results of B models with the same
        index list into one with [B, n_deflectors] arrays (one batch)."""
        index_list = np.asarray(prepared[0]["index_list"])
        for r in prepared[1:]:
            if not np.array_equal(np.asarray(r["index_list"]), index_list):
                raise ValueError("stacked models must share the deflector index list")
        keys = prepared[0]["all_kwargs"].keys()
        all_kwargs = {k: np.stack([np.asarray(r["all_kwargs"][k], dtype=np.float64) for r in prepared])
                      for k in keys}
        return {"all_kwargs": all_kwargs, "index_list": index_list}

    # ------------------------------------------------------------ packing --
    def pack(self, ray_shooting_kwargs, max_inline=None):
        """(n_img, inline types, inline params [B][n_inline][8], bulk types, bulk params [B][n_bulk][8])."""
        if max_inline is None:
            max_inline = backend.MAX_PROFILES
        all_kwargs = ray_shooting_kwargs["all_kwargs"]
        index_list = np.asarray(ray_shooting_kwargs["index_list"], dtype=int)
        n_def = len(index_list)
        n_img = 1
        for v in all_kwargs.values():
            a = np.asarray(v.detach().cpu().numpy() if isinstance(v, torch.Tensor) else v, dtype=np.float64)
            if a.ndim == 2:
                n_img = a.shape[0]
                break
        cols = {}
        for k, v in all_kwargs.items():
            a = np.asarray(v.detach().cpu().numpy() if isinstance(v, torch.Tensor) else v, dtype=np.float64)
            cols[k] = np.broadcast_to(a, (n_img, n_def)) if a.ndim == 2 else np.broadcast_to(a, (n_img, n_def))
        n_unique = len(self.unique_lens_model_list)
        # NULL padding deflects by zero: it is dropped
        names = [self.unique_lens_model_list[u] if u < n_unique else None for u in index_list]
        inline_idx = [d for d, nm in enumerate(names) if nm == "EPL"]
        rest = [d for d, nm in enumerate(names) if nm not in ("EPL", None)]
        if len(inline_idx) > max_inline:
            raise ValueError("at most %d EPL deflectors per bulk model" % max_inline)
        counts = {}
        for d in rest:
            counts[names[d]] = counts.get(names[d], 0) + 1
        bulk_idx = rest
        if counts:
            major = max(sorted(counts), key=lambda nm: counts[nm])
            minor = [d for d in rest if names[d] != major]
            if len(inline_idx) + len(minor) <= max_inline:
                inline_idx = sorted(inline_idx + minor)
                bulk_idx = [d for d in rest if names[d] == major]

        def block(idx):
            p = np.zeros((n_img, len(idx), backend.NPARAM))
            types = np.zeros(len(idx), dtype=np.int32)
            for j, d in enumerate(idx):
                u = index_list[d]
                name = self.unique_lens_model_list[u]
                types[j] = profiles.LENS_PROFILES[name][0]
                for i, pn in enumerate(self.param_name_list[u]):
                    p[:, j, i] = cols[pn][:, d] if pn in cols else 0.0
            return types, p

        it, ip = block(inline_idx)
        bt, bp = block(bulk_idx)
        return n_img, it, ip, bt, bp

    def _model(self, inline_types, bulk_types, source_types=(), light_types=(), flux_scaling=1.0):
        d_types = None
        common = 0
        if len(bulk_types) > 0:
            u = np.unique(bulk_types)
            if len(u) == 1:
                common = int(u[0])
            else:
                d_types = torch.as_tensor(np.ascontiguousarray(bulk_types, dtype=np.int32), device=backend.device())
        m = backend.make_model(list(inline_types), list(source_types), list(light_types),
                               alpha_scaling=self.alpha_scaling, flux_scaling=flux_scaling,
                               bulk=(len(bulk_types), common, d_types))
        return m, d_types  # keep d_types alive while m is used

    def ray_shooting(self, x, y, ray_shooting_kwargs):
        """LensModelBulk.ray_shooting, lens_model_bulk.py:113-128."""
        lib = backend.lib()
        n_img, it, ip, bt, bp = self.pack(ray_shooting_kwargs)
        m, keep = self._model(it, bt)
        params = backend.to_device_f64(np.concatenate([ip, bp], axis=1))
        xt = backend.to_device_f64(x)
        yt = backend.to_device_f64(y)
        shape = xt.shape
        xt = xt.reshape(-1).contiguous()
        yt = yt.reshape(-1).contiguous()
        n_ray = xt.numel()
        bx = torch.empty((n_img, n_ray), dtype=torch.float64, device=xt.device)
        by = torch.empty_like(bx)
        prec = backend.PRECISIONS[self.precision]
        nbytes = lib.lensim_workspace_size(ctypes.byref(m), None, None, n_img, prec)
        ws = self._ws.get(nbytes)
        rc = lib.lensim_raytrace(ctypes.byref(m), backend.ptr(params), n_img, backend.ptr(xt), backend.ptr(yt),
                                 n_ray, prec, backend.ptr(bx), backend.ptr(by), backend.ptr(ws), ws.numel(),
                                 backend.stream_handle())
        backend.check(rc, "lensim_raytrace")
        del keep
        if n_img > 1:
            return bx.reshape((n_img,) + tuple(shape)), by.reshape((n_img,) + tuple(shape))
        return bx.reshape(shape), by.reshape(shape)
