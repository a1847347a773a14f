"""LensModel: single- and multi-plane ray shooting on the GPU.

Mirrors the reference's ``LensModel`` (LensModel/lens_model.py:25-252) for the
profiles EPL, SIE, NIE, NFW, TNFW, SHEAR and CONVERGENCE.  ``ray_shooting``
runs ``lensim_raytrace`` (one HIP kernel over rays x parameter sets).
``profile_kwargs_list`` carries the profiles' constructor options
(profiles.lens_layout: SIE's NIE / s_scale / gamma, NFW's interpol), and a
source redshift other than the convention's scales the deflections by
alpha_scaling (lens_model.py:176-195).  With ``multi_plane=True`` the
deflectors sit on the planes of ``lens_redshift_list`` (MultiPlane,
LensModel/MultiPlane/multi_plane.py:20-280): the same plane recursion as
LensModelBulk's, with the distances fixed at construction.
Parameter values may be scalars (one model) or 1-D arrays of length B (a batch
of B models evaluated at the same coordinates, like ``jax.vmap`` over kwargs);
a batch adds a leading axis of length B to the outputs.
"""

import ctypes

import numpy as np
import torch

from .. import backend, profiles
from ..Util.cosmo import Background


class LensModel(object):
    """Class to handle an arbitrary list of (single-plane) lens models."""

    def __init__(self, lens_model_list, z_lens=None, z_source=None, lens_redshift_list=None, cosmo=None,
                 multi_plane=False, observed_convention_index=None, z_source_convention=None,
                 cosmo_interp=False, z_interp_stop=None, num_z_interp=100, profile_kwargs_list=None,
                 decouple_multi_plane=False, kwargs_multiplane_model=None, distance_ratio_sampling=False,
                 cosmology_sampling=False, cosmology_model="FlatLambdaCDM", precision="fp64"):
        # construction-time checks of LensModel/lens_model.py:76-84, :122-196
        if cosmology_sampling:
            raise ValueError("Cosmology sampling not supported")
        if distance_ratio_sampling:
            raise ValueError("Distance ratio sampling not supported")
        if multi_plane:
            # lens_model.py:123-135
            if lens_redshift_list is None:
                raise ValueError("In multi-plane lensing, you need to specify the redshifts of the lensing planes.")
            if z_source is None:
                raise ValueError("z_source needs to be set for multi-plane lens modelling.")
            if decouple_multi_plane:
                raise NotImplementedError("decoupled multi-plane lensing is not part of this library")
        if observed_convention_index is not None:
            raise NotImplementedError("observed_convention_index is not supported")
        if cosmology_model != "FlatLambdaCDM" and cosmo is None:
            raise ValueError("only FlatLambdaCDM cosmologies are available")
        if precision not in backend.PRECISIONS:
            raise ValueError("precision must be one of %s" % list(backend.PRECISIONS))
        self.lens_model_list = list(lens_model_list)
        self.z_lens = z_lens
        self.z_source = z_source
        self._z_source_convention = z_source_convention
        self.redshift_list = lens_redshift_list
        self.multi_plane = False
        self.type = "SinglePlane"
        self.precision = precision
        if profile_kwargs_list is None:
            profile_kwargs_list = [{} for _ in self.lens_model_list]
        if len(profile_kwargs_list) != len(self.lens_model_list):
            raise ValueError("profile_kwargs_list has %d entries for %d lens models"
                             % (len(profile_kwargs_list), len(self.lens_model_list)))
        self.profile_kwargs_list = profile_kwargs_list
        self.alpha_scaling = 1.0
        if (not multi_plane and z_source is not None and z_source_convention is not None
                and z_source != z_source_convention):
            if z_lens is None:
                raise ValueError("a lens redshift needs to provided when z_source != z_source_convention")
            self.alpha_scaling = beta_double_source_plane(Background(cosmo), z_lens, z_source, z_source_convention)
        self._types = []
        self._layouts = []
        self.param_name_list = []
        for name, pk in zip(self.lens_model_list, profile_kwargs_list):
            code, layout, names = profiles.lens_layout(name, pk)
            self._types.append(code)
            self._layouts.append(layout)
            self.param_name_list.append(names)
        self._ws = backend.Workspace()
        self._mp = None
        if multi_plane:
            self.multi_plane = True
            self.type = "MultiPlane"
            self._mp = self._plane_arrays(Background(cosmo), float(z_source),
                                          float(z_source if z_source_convention is None else z_source_convention))

    def _plane_arrays(self, bkg, z_source, z_convention):
        """The lensim_multiplane descriptor of the redshift-sorted deflectors and
        its device arrays (MultiPlaneBase.set_T_zs_and_T_ijs, multi_plane_base.py:89-116;
        MultiPlane._set_source_distances, multi_plane.py:190-204): per plane the
        transverse distance from the previous one (T_ij, 0 on a shared plane), from
        the observer (T_z) and the reduced-to-physical factor D_s / D_ds of the
        convention's source redshift; then the source plane."""
        z = np.asarray(self.redshift_list, dtype=np.float64)
        if len(z) != len(self.lens_model_list):
            raise ValueError("The length of lens_model_list does not correspond to redshift_list")
        if len(z) and z.max() >= z_convention:  # multi_plane_base.py:60-66
            raise ValueError("deflector redshifts higher or equal the source redshift convention (%s >= %s for the "
                             "reduced lens model quantities not allowed (leads to negative reduced deflection "
                             "angles!" % (z.max(), z_convention))
        if len(z) and z.max() > z_source:
            raise ValueError("deflectors beyond z_source (%s > %s) are not supported" % (z.max(), z_source))
        order = np.argsort(z)  # MultiPlaneBase._index_ordering, multi_plane_base.py:409-421
        tij, tz = [], []
        z_before, t_z = 0.0, 0.0
        for zl in z[order]:
            if zl == z_before:
                delta = 0.0
            else:
                t_z = bkg.T_xy(0.0, zl)
                delta = bkg.T_xy(z_before, zl)
            tij.append(delta)
            tz.append(t_z)
            z_before = zl
        r2p = [bkg.d_xy(0.0, z_convention) / bkg.d_xy(zl, z_convention) for zl in z[order]]
        tij.append(bkg.T_xy(z_before, z_source))  # transverse_distance_start_stop, multi_plane_base.py:303-330
        tz.append(bkg.T_xy(0.0, z_source))
        r2p.append(0.0)
        dev = backend.device()
        arrays = {"order": torch.as_tensor(np.ascontiguousarray(order, dtype=np.int32), device=dev),
                  "tij": backend.to_device_f64(np.asarray(tij)), "tz": backend.to_device_f64(np.asarray(tz)),
                  "r2p": backend.to_device_f64(np.asarray(r2p))}
        mp = backend.LensimMultiplane()
        mp.n_planes = len(z) + 1
        mp.d_order = arrays["order"].data_ptr() if len(z) else None
        mp.d_T_ij = arrays["tij"].data_ptr()
        mp.d_T_z = arrays["tz"].data_ptr()
        mp.d_r2p = arrays["r2p"].data_ptr()
        mp.stride = 0
        arrays["struct"] = mp
        arrays["lists"] = (np.asarray(tij), np.asarray(tz), np.asarray(r2p), order)
        return arrays

    def multiplane_address(self):
        """Address of the lensim_multiplane descriptor (0 on a single plane) for a
        lensim_model of this list's deflectors in their list order; valid while
        this LensModel lives."""
        return 0 if self._mp is None else ctypes.addressof(self._mp["struct"])

    @property
    def T_ij_list(self):
        """Transverse distances between consecutive planes, the source plane's last (Mpc)."""
        return None if self._mp is None else self._mp["lists"][0]

    @property
    def T_z_list(self):
        """Transverse distances from the observer to each plane, the source plane's last (Mpc)."""
        return None if self._mp is None else self._mp["lists"][1]

    @property
    def reduced2physical_factor(self):
        """D_s / D_ds of the convention's source redshift per plane (0 for the source plane)."""
        return None if self._mp is None else self._mp["lists"][2]

    def check_parameters(self, kwargs_list):
        """ProfileListBase.check_parameters, LensModel/profile_list_base.py:278-306."""
        if len(kwargs_list) != len(self.param_name_list):
            raise ValueError("length of input parameter list %s does not match number of lens models %s"
                             % (len(kwargs_list), len(self.param_name_list)))
        for i, names in enumerate(self.param_name_list):
            for key in kwargs_list[i]:
                if key not in names:
                    raise ValueError("parameter %s in lens model is not part of model %s (%s). Parameters allowed "
                                     "are %s" % (key, i, self.lens_model_list[i], names))
            for name in names:
                if name not in kwargs_list[i]:
                    raise ValueError("Lens model %s (%s) requires parameter %s which is not provided in input."
                                     % (i, self.lens_model_list[i], name))

    def _packed(self, kwargs, k):
        if self._mp is not None and k is not None:  # MultiPlane._check_raise, multi_plane.py:716-733
            raise ValueError("no individual deflector (k) can be evaluated in multi-plane lensing")
        idx = profiles.select(k, len(self._types))
        kwl = [kwargs[i] for i in idx]
        n_img = profiles.batch_size([kwargs])
        params = profiles.pack([self._layouts[i] for i in idx], kwl, n_img)
        model = backend.make_model([self._types[i] for i in idx], [], [], alpha_scaling=self.alpha_scaling)
        model.multiplane = self.multiplane_address() or None
        return model, params, n_img

    def ray_shooting(self, x, y, kwargs, k=None):
        """Maps image to source position (inverse deflection), lens_model.py:236-252.

        :return: beta_x, beta_y as float64 device tensors (shape of x, with a
            leading batch axis when the kwargs hold arrays of length B > 1)
        """
        model, params, n_img = self._packed(kwargs, k)
        return self._raytrace(model, params, n_img, x, y, batched=profiles.batch_size([kwargs]) > 1)

    def _raytrace(self, model, params, n_img, x, y, batched):
        lib = backend.lib()
        xt = backend.to_device_f64(x)
        yt = backend.to_device_f64(y)
        if xt.shape != yt.shape:
            raise ValueError("x and y must have the same shape")
        shape = xt.shape
        xt = xt.reshape(-1).contiguous()
        yt = yt.reshape(-1).contiguous()
        n_ray = xt.numel()
        pt = backend.to_device_f64(params)
        bx = torch.empty((n_img, n_ray), dtype=torch.float64, device=xt.device)
        by = torch.empty_like(bx)
        prec = backend.PRECISIONS[self.precision]
        nbytes = lib.lensim_workspace_size(ctypes_ref(model), None, None, n_img, prec)
        ws = self._ws.get(nbytes)
        rc = lib.lensim_raytrace(ctypes_ref(model), backend.ptr(pt), n_img, backend.ptr(xt), backend.ptr(yt),
                                 n_ray, prec, backend.ptr(bx), backend.ptr(by), backend.ptr(ws), ws.numel(),
                                 backend.stream_handle())
        backend.check(rc, "lensim_raytrace")
        if batched:
            return bx.reshape((n_img,) + tuple(shape)), by.reshape((n_img,) + tuple(shape))
        return bx.reshape(shape), by.reshape(shape)

    def alpha(self, x, y, kwargs, k=None):
        """Deflection angles alpha = theta - beta (SinglePlane.alpha, single_plane.py:112-139;
        MultiPlane.alpha, multi_plane.py:512-535)."""
        bx, by = self.ray_shooting(x, y, kwargs, k=k)
        xt = backend.to_device_f64(x)
        yt = backend.to_device_f64(y)
        return xt - bx, yt - by


def beta_double_source_plane(background, z_lens, z_source_1, z_source_2):
    """Ratio of the reduced deflections for sources at z_source_1 and at
    z_source_2 (lenstronomy LensCosmo.beta_double_source_plane, as called at
    LensModel/lens_model.py:180-189): D_ds1 / D_s1 * D_s2 / D_ds2, angular
    diameter distances of the background cosmology."""
    ds1 = background.d_xy(0.0, z_source_1)
    dds1 = background.d_xy(z_lens, z_source_1)
    ds2 = background.d_xy(0.0, z_source_2)
    dds2 = background.d_xy(z_lens, z_source_2)
    return float(dds1 / ds1 * ds2 / dds2)


def ctypes_ref(struct):
    import ctypes

    return ctypes.byref(struct)
