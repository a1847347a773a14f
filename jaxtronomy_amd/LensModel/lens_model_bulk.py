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

With multi_plane=True it mirrors ``MultiPlaneBulk``
(LensModel/MultiPlane/multi_plane_bulk.py:11-275): the deflectors are sorted by
redshift, the transverse comoving distances and reduced-to-physical factors
come from a flat-LambdaCDM background (Util/cosmo.py), and the kernel traces
every ray plane by plane.
"""

import ctypes

import numpy as np
import torch

from .. import backend, profiles
from ..Util.cosmo import Background

_BULK_OK = ("SIE", "NIE", "NFW", "TNFW", "SHEAR", "CONVERGENCE")


class LensModelBulk(object):
    def __init__(self, unique_lens_model_list, multi_plane=False, profile_kwargs_list=None, cosmo=None,
                 cosmology_model="FlatLambdaCDM", precision="fp64"):
        if precision not in backend.PRECISIONS:
            raise ValueError("precision must be one of %s" % list(backend.PRECISIONS))
        for name in unique_lens_model_list:
            profiles.lens_type(name)
            if name not in _BULK_OK and name != "EPL":
                raise ValueError("%s is not supported in LensModelBulk" % name)
        self.unique_lens_model_list = list(unique_lens_model_list)
        if profile_kwargs_list is not None:
            # the constructor options of each unique profile (single_plane_bulk.py:30-36): the ones that
            # keep the profile's default evaluation are accepted, the others raise
            if len(profile_kwargs_list) != len(self.unique_lens_model_list):
                raise ValueError("profile_kwargs_list has %d entries for %d unique lens models"
                                 % (len(profile_kwargs_list), len(self.unique_lens_model_list)))
            for name, pk in zip(self.unique_lens_model_list, profile_kwargs_list):
                code, layout, names = profiles.lens_layout(name, pk)
                if code != profiles.LENS_PROFILES[name][0] or layout != names:
                    raise NotImplementedError("profile options %s of %s are not supported in LensModelBulk"
                                              % (pk, name))
        self.param_name_list = [profiles.LENS_PROFILES[n][1] for n in self.unique_lens_model_list]
        self.multi_plane = bool(multi_plane)
        self.type = "MultiPlaneBulk" if self.multi_plane else "SinglePlaneBulk"
        if self.multi_plane:
            if cosmology_model != "FlatLambdaCDM" and cosmo is None:
                raise ValueError("only FlatLambdaCDM cosmologies are available")
            self._cosmo_bkg = Background(cosmo)
        self.precision = precision
        self.alpha_scaling = 1.0
        self.init_kwargs = {"unique_lens_model_list": self.unique_lens_model_list, "multi_plane": self.multi_plane,
                            "profile_kwargs_list": profile_kwargs_list, "cosmo": cosmo,
                            "cosmology_model": cosmology_model}
        self._ws = backend.Workspace()

    def prepare_ray_shooting_kwargs(self, lens_model_list, kwargs_lens, z_source=None, lens_redshift_list=None,
                                    num_deflectors=None):
        """SinglePlaneBulk.prepare_ray_shooting_kwargs, single_plane_bulk.py:47-106, or for
        multi_plane MultiPlaneBulk.prepare_ray_shooting_kwargs, multi_plane_bulk.py:54-150."""
        if self.multi_plane:
            return self._prepare_multi_plane(lens_model_list, kwargs_lens, z_source, lens_redshift_list,
                                             num_deflectors)
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
        out = {"all_kwargs": all_kwargs, "index_list": index_list}
        for k in ("T_ij_list", "T_z_list", "reduced2physical_factor"):
            if k in prepared[0]:
                out[k] = np.stack([np.asarray(r[k], dtype=np.float64) for r in prepared])
        return out

    def _prepare_multi_plane(self, lens_model_list, kwargs_lens, z_source, lens_redshift_list, num_deflectors):
        if z_source is None or lens_redshift_list is None:
            raise ValueError("multi-plane ray shooting needs z_source and lens_redshift_list")
        if not (len(lens_model_list) == len(kwargs_lens) == len(lens_redshift_list)):
            raise ValueError("length of lens model list %d, length of kwargs_lens %d, and length of redshift list %d "
                             "should match" % (len(lens_model_list), len(kwargs_lens), len(lens_redshift_list)))
        if num_deflectors is None:
            num_deflectors = len(lens_model_list)
        elif num_deflectors < len(lens_model_list):
            raise ValueError("The provided num_deflectors %d is smaller than the length of lens_model_list %d"
                             % (num_deflectors, len(lens_model_list)))
        z = np.array(lens_redshift_list)
        order = np.argsort(z)  # as multi_plane_bulk.py:214-235
        lens_model_list = [lens_model_list[i] for i in order]
        kwargs_lens = [kwargs_lens[i] for i in order]
        z = z[order]
        num_filler = num_deflectors - len(lens_model_list)
        index_list = [self.unique_lens_model_list.index(m) for m in lens_model_list]
        index_list += [len(self.unique_lens_model_list)] * (num_filler + 1)
        unique_kwargs = set(p for sub in self.param_name_list for p in sub)
        all_kwargs = {}
        for kwarg in unique_kwargs:
            vals = [kw.get(kwarg, 0) for kw in kwargs_lens] + [0] * (num_filler + 1)
            all_kwargs[kwarg] = np.array(vals)
        # distances, multi_plane_bulk.py:237-275
        bkg = self._cosmo_bkg
        r2p = list(bkg.d_xy(0, z_source) / bkg.d_xy(z, np.ones(z.shape) * z_source)) if len(z) else []
        T_ij, T_z = [], []
        z_before, t_z = 0, 0
        for zi in np.append(z, z_source):
            if z_before == zi:
                delta = 0
            else:
                t_z = bkg.T_xy(0, zi)
                delta = bkg.T_xy(z_before, zi)
            T_ij.append(delta)
            T_z.append(t_z)
            z_before = zi
        T_ij += [0] * num_filler
        T_z += [T_z[-1]] * num_filler
        r2p += [0] * (num_filler + 1)
        return {"all_kwargs": all_kwargs, "index_list": np.array(index_list), "T_ij_list": np.array(T_ij),
                "T_z_list": np.array(T_z), "reduced2physical_factor": np.array(r2p)}

    def multiplane_arrays(self, ray_shooting_kwargs, order):
        """Device arrays of the lensim_multiplane descriptor: the n deflector planes
        (packed deflector index of each, in the dict's plane order) plus the
        source plane; trailing NULL planes (fillers, zero distance) are dropped."""
        idx = np.asarray(ray_shooting_kwargs["index_list"], dtype=int)
        n_unique = len(self.unique_lens_model_list)
        real = np.nonzero(idx < n_unique)[0]
        n = len(real)
        if n and real[-1] != n - 1:
            raise ValueError("NULL planes between deflectors are not supported")
        t_ij = np.atleast_2d(np.asarray(ray_shooting_kwargs["T_ij_list"], dtype=np.float64))[:, :n + 1]
        t_z = np.atleast_2d(np.asarray(ray_shooting_kwargs["T_z_list"], dtype=np.float64))[:, :n + 1]
        r2p = np.atleast_2d(np.asarray(ray_shooting_kwargs["reduced2physical_factor"], dtype=np.float64))[:, :n + 1]
        dev = backend.device()
        arrays = {"order": torch.as_tensor(np.ascontiguousarray(np.asarray(order, dtype=np.int32)), device=dev),
                  "tij": torch.as_tensor(np.ascontiguousarray(t_ij), device=dev),
                  "tz": torch.as_tensor(np.ascontiguousarray(t_z), device=dev),
                  "r2p": torch.as_tensor(np.ascontiguousarray(r2p), device=dev)}
        mp = backend.LensimMultiplane()
        mp.n_planes = n + 1
        mp.d_order = arrays["order"].data_ptr() if n else None
        mp.d_T_ij = arrays["tij"].data_ptr()
        mp.d_T_z = arrays["tz"].data_ptr()
        mp.d_r2p = arrays["r2p"].data_ptr()
        mp.stride = (n + 1) if t_ij.shape[0] > 1 else 0
        arrays["struct"] = mp
        return mp, arrays

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
            # one gather per (profile, parameter) over the deflectors of that profile
            idx = np.asarray(idx, dtype=int)
            p = np.zeros((n_img, len(idx), backend.NPARAM))
            types = np.zeros(len(idx), dtype=np.int32)
            us = index_list[idx] if len(idx) else idx
            for u in np.unique(us):
                sel = np.nonzero(us == u)[0]
                types[sel] = profiles.LENS_PROFILES[self.unique_lens_model_list[u]][0]
                for i, pn in enumerate(self.param_name_list[u]):
                    if pn in cols:
                        p[:, sel, i] = cols[pn][:, idx[sel]]
            return types, p

        it, ip = block(inline_idx)
        bt, bp = block(bulk_idx)
        packed = {d: j for j, d in enumerate(inline_idx)}
        packed.update({d: len(inline_idx) + j for j, d in enumerate(bulk_idx)})
        self._plane_order = [packed[d] for d in range(n_def) if d in packed]  # packed index in dict order
        self._packed = packed
        return n_img, it, ip, bt, bp

    def unpack_gradient(self, ray_shooting_kwargs, grad):
        """A gradient over the packed deflector rows of the last pack() (grad
        [B, n_inline + n_bulk, 8]) in the form of ray_shooting_kwargs["all_kwargs"]:
        {kwarg: [B, n_deflectors]} (or [n_deflectors] for B = 1), the entry of
        deflector d the derivative in its own kwarg; kwargs a deflector's profile
        does not have, and NULL padding, are zero -- the structure jax.grad gives
        for the dict argument of ray_shooting."""
        index_list = np.asarray(ray_shooting_kwargs["index_list"], dtype=int)
        keys = ray_shooting_kwargs["all_kwargs"].keys()
        g = np.asarray(grad)
        out = {kw: np.zeros((g.shape[0], len(index_list))) for kw in keys}
        for d, j in self._packed.items():
            for i, pn in enumerate(self.param_name_list[index_list[d]]):
                if pn in out:
                    out[pn][:, d] += g[:, j, i]
        if g.shape[0] == 1:
            out = {kw: v[0] for kw, v in out.items()}
        return out

    def _model(self, inline_types, bulk_types, source_types=(), light_types=(), flux_scaling=1.0,
               ray_shooting_kwargs=None):
        """lensim_model of a pack() result (call right after pack: the plane order
        of a multi-plane model comes from it) and the device arrays to keep alive."""
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
        keep = [d_types]
        if self.multi_plane:
            if ray_shooting_kwargs is None or "T_ij_list" not in ray_shooting_kwargs:
                raise ValueError("multi-plane models need the ray_shooting_kwargs of prepare_ray_shooting_kwargs")
            mp, arrays = self.multiplane_arrays(ray_shooting_kwargs, self._plane_order)
            m.multiplane = ctypes.addressof(mp)
            keep.append(arrays)
        return m, keep  # keep the device arrays alive while m is used

    def ray_shooting(self, x, y, ray_shooting_kwargs):
        """LensModelBulk.ray_shooting, lens_model_bulk.py:113-128."""
        lib = backend.lib()
        n_img, it, ip, bt, bp = self.pack(ray_shooting_kwargs)
        m, keep = self._model(it, bt, ray_shooting_kwargs=ray_shooting_kwargs)
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
