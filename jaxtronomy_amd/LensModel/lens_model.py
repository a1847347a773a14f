"""LensModel: single-plane ray shooting on the GPU.

Mirrors the reference's ``LensModel`` (LensModel/lens_model.py:25-252) for the
single-plane profiles EPL, SIE, NIE, NFW, TNFW and SHEAR.  ``ray_shooting``
runs ``lensim_raytrace`` (one HIP kernel over rays x parameter sets).
``profile_kwargs_list`` carries the profiles' constructor options
(profiles.lens_layout: SIE's NIE / s_scale / gamma, NFW's interpol), and a
source redshift other than the convention's scales the deflections by
alpha_scaling (lens_model.py:176-195).
Parameter values may be scalars (one model) or 1-D arrays of length B (a batch
of B models evaluated at the same coordinates, like ``jax.vmap`` over kwargs);
a batch adds a leading axis of length B to the outputs.
"""

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
            raise NotImplementedError("multi-plane ray shooting is not part of this library")
        if observed_convention_index is not None:
            raise NotImplementedError("observed_convention_index needs multi-plane lensing")
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
        if z_source is not None and z_source_convention is not None and z_source != z_source_convention:
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
        idx = profiles.select(k, len(self._types))
        kwl = [kwargs[i] for i in idx]
        n_img = profiles.batch_size([kwargs])
        params = profiles.pack([self._layouts[i] for i in idx], kwl, n_img)
        model = backend.make_model([self._types[i] for i in idx], [], [], alpha_scaling=self.alpha_scaling)
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
        """Deflection angles alpha = theta - beta (SinglePlane.alpha, single_plane.py:112-139)."""
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
