"""LightModel: surface brightness on the GPU.

Mirrors the reference's ``LightModel`` / ``LightModelBase``
(LightModel/light_model.py:24-51, LightModel/light_model_base.py:63-252) for
SERSIC, SERSIC_ELLIPSE and GAUSSIAN; ``surface_brightness`` runs
``lensim_surface_brightness``.
"""

import ctypes

import torch

from .. import backend, profiles


class LightModel(object):
    """Class to handle source and lens light models."""

    def __init__(self, light_model_list, deflection_scaling_list=None, source_redshift_list=None,
                 profile_kwargs_list=None, precision="fp64"):
        if source_redshift_list is not None and source_redshift_list != []:
            raise ValueError("multi plane sources not supported, source_redshift_list must be None")
        if deflection_scaling_list is not None:
            raise NotImplementedError("deflection_scaling_list (several source planes) is not supported")
        if profile_kwargs_list is not None:
            for pk in profile_kwargs_list:
                if pk:
                    if set(pk) - {"smoothing", "sersic_major_axis"} or pk.get("smoothing", 0.0001) != 0.0001 \
                            or pk.get("sersic_major_axis", False):
                        raise NotImplementedError("non-default light profile settings %s are not supported" % pk)
        if precision not in backend.PRECISIONS:
            raise ValueError("precision must be one of %s" % list(backend.PRECISIONS))
        self.profile_type_list = list(light_model_list)
        self.deflection_scaling_list = None
        self.redshift_list = None
        self.precision = precision
        self._types = []
        self.param_name_list = []
        for name in self.profile_type_list:
            code, names = profiles.light_type(name)
            self._types.append(code)
            self.param_name_list.append(names)
        self._ws = backend.Workspace()

    def surface_brightness(self, x, y, kwargs_list, k=None):
        """LightModelBase.surface_brightness, light_model_base.py:230-252."""
        lib = backend.lib()
        idx = profiles.select(k, len(self._types))
        kwl = [kwargs_list[i] for i in idx]
        n_img = profiles.batch_size([kwargs_list])
        batched = n_img > 1
        params = profiles.pack([self.param_name_list[i] for i in idx], kwl, n_img)
        model = backend.make_model([], [self._types[i] for i in idx], [])
        xt = backend.to_device_f64(x)
        yt = backend.to_device_f64(y)
        shape = xt.shape
        xt = xt.reshape(-1).contiguous()
        yt = yt.reshape(-1).contiguous()
        n_ray = xt.numel()
        pt = backend.to_device_f64(params)
        flux = torch.empty((n_img, n_ray), dtype=torch.float64, device=xt.device)
        prec = backend.PRECISIONS[self.precision]
        nbytes = lib.lensim_workspace_size(ctypes.byref(model), None, None, n_img, prec)
        ws = self._ws.get(nbytes)
        rc = lib.lensim_surface_brightness(ctypes.byref(model), backend.ptr(pt), n_img, 0, backend.ptr(xt),
                                           backend.ptr(yt), n_ray, prec, backend.ptr(flux), backend.ptr(ws),
                                           ws.numel(), backend.stream_handle())
        backend.check(rc, "lensim_surface_brightness")
        if batched:
            return flux.reshape((n_img,) + tuple(shape))
        return flux.reshape(shape)
