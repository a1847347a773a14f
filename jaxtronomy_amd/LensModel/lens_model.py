"""LensModel: single-plane ray shooting on the GPU.

Mirrors the reference's ``LensModel`` (LensModel/lens_model.py:25-252) for the
single-plane profiles EPL, SIE, NIE, NFW, TNFW and SHEAR.  ``ray_shooting``
runs ``lensim_raytrace`` (one HIP kernel over rays x parameter sets).
Parameter values may be scalars (one model) or 1-D arrays of length B (a batch
of B models evaluated at the same coordinates, like ``jax.vmap`` over kwargs);
a batch adds a leading axis of length B to the outputs.
"""

import torch

from .. import backend, profiles


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
        if z_source is not None and z_source_convention is not None and z_source != z_source_convention:
            raise NotImplementedError("alpha_scaling from a cosmology (z_source != z_source_convention) "
                                      "is not supported; pass alpha_scaling to the kwargs instead")
        if precision not in backend.PRECISIONS:
            raise ValueError("precision must be one of %s" % list(backend.PRECISIONS))
        self.lens_model_list = list(lens_model_list)
        self.z_lens = z_lens
        self.z_source = z_source
        self.redshift_list = lens_redshift_list
        self.multi_plane = False
        self.type = "SinglePlane"
        self.precision = precision
        self.alpha_scaling = 1.0
        self._types = []
        self.param_name_list = []
        for name in self.lens_model_list:
            code, names = profiles.lens_type(name)
            self._types.append(code)
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
        params = profiles.pack([self.param_name_list[i] for i in idx], kwl, n_img)
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


def ctypes_ref(struct):
    import ctypes

    return ctypes.byref(struct)
