"""Profile registry: reference profile names -> library type codes and the
ordered parameter names of each profile (the order of a parameter slot).

Names and parameter orders follow the reference's ``param_names``:
EPL LensModel/Profiles/epl.py:50, SIE sie.py:40, NIE nie.py, NFW nfw.py:48,
TNFW tnfw.py:27, SHEAR shear.py:16, CONVERGENCE convergence.py:15, SERSIC_ELLIPSE
LightModel/Profiles/sersic_ellipse.py:25, SERSIC sersic.py, GAUSSIAN gaussian.py.
"""

import numpy as np
import torch

from . import backend

LENS_PROFILES = {
    "EPL": (1, ["theta_E", "gamma", "e1", "e2", "center_x", "center_y"]),
    "SIE": (2, ["theta_E", "e1", "e2", "center_x", "center_y"]),
    "NIE": (3, ["theta_E", "e1", "e2", "s_scale", "center_x", "center_y"]),
    "NFW": (4, ["Rs", "alpha_Rs", "center_x", "center_y"]),
    "TNFW": (5, ["Rs", "alpha_Rs", "r_trunc", "center_x", "center_y"]),
    "SHEAR": (6, ["gamma1", "gamma2", "ra_0", "dec_0"]),
    "CONVERGENCE": (7, ["kappa", "ra_0", "dec_0"]),
}

LIGHT_PROFILES = {
    "SERSIC": (32, ["amp", "R_sersic", "n_sersic", "center_x", "center_y"]),
    "SERSIC_ELLIPSE": (33, ["amp", "R_sersic", "n_sersic", "e1", "e2", "center_x", "center_y"]),
    "GAUSSIAN": (34, ["amp", "sigma", "center_x", "center_y"]),
}

# keyword arguments with a default value in the reference's function signatures
_DEFAULTS = {"center_x": 0.0, "center_y": 0.0, "ra_0": 0.0, "dec_0": 0.0}

SIE_S_SCALE = 0.0000000001  # SIE's default core, LensModel/Profiles/sie.py:56


def lens_type(name):
    if name not in LENS_PROFILES:
        raise ValueError("%s is not a supported lens model. Supported are: %s." % (name, sorted(LENS_PROFILES)))
    return LENS_PROFILES[name]


def lens_layout(name, profile_kwargs=None):
    """Device type and parameter-slot layout of lens profile ``name`` built with
    ``profile_kwargs``, the constructor options of lens_class
    (LensModel/profile_list_base.py:309-765).

    Returns (type code, layout, param_names).  ``layout`` lists the slot
    contents in order: a kwarg name, or (name, value) for a slot that the
    option fixes to a constant.  Options:
      * SIE(s_scale, gamma, NIE) (sie.py:56-124): NIE=False evaluates the EPL
        with the fixed slope ``gamma`` (default 2); a non-default s_scale is
        the NIE with that core; otherwise the SIE;
      * NFW(interpol): interpol=True raises, as the reference does (nfw.py:60-63);
        its other keywords are accepted and unused, as there;
      * EPL / NIE (b, t|s, q, phi, static): static=True (the profile evaluated
        at stored constants instead of its kwargs) is not supported;
      * other profiles take no options (their constructors take none).
    """
    code, names = lens_type(name)
    pk = dict(profile_kwargs or {})
    layout = list(names)
    if name == "SIE":
        unknown = set(pk) - {"s_scale", "gamma", "NIE"}
        if unknown:
            raise TypeError("SIE got unexpected profile keyword arguments %s" % sorted(unknown))
        if not pk.get("NIE", True):
            layout = ["theta_E", ("gamma", float(pk.get("gamma", 2))), "e1", "e2", "center_x", "center_y"]
            code = LENS_PROFILES["EPL"][0]
        elif float(pk.get("s_scale", SIE_S_SCALE)) != SIE_S_SCALE:
            layout = ["theta_E", "e1", "e2", ("s_scale", float(pk["s_scale"])), "center_x", "center_y"]
            code = LENS_PROFILES["NIE"][0]
    elif name in ("EPL", "NIE"):
        allowed = {"b", "t", "q", "phi", "static"} if name == "EPL" else {"b", "s", "q", "phi", "static"}
        unknown = set(pk) - allowed
        if unknown:
            raise TypeError("%s got unexpected profile keyword arguments %s" % (name, sorted(unknown)))
        if pk.get("static", False):
            raise NotImplementedError("%s(static=True) is not supported: pass the parameters as kwargs" % name)
    elif name == "NFW":
        if pk.get("interpol", False):
            raise ValueError("This class no longer supports interpol functionality in JAXtronomy.")
    elif pk:
        raise TypeError("%s takes no profile keyword arguments (got %s)" % (name, sorted(pk)))
    return code, layout, list(names)


def kwarg_slots(layout):
    """(slot index, kwarg name) of the slots of a layout that hold kwargs."""
    return [(q, nm) for q, nm in enumerate(layout) if isinstance(nm, str)]


def light_type(name):
    if name not in LIGHT_PROFILES:
        raise ValueError("Light model of type %s not supported. Supported are %s" % (name, sorted(LIGHT_PROFILES)))
    return LIGHT_PROFILES[name]


def _batch_of(value):
    if isinstance(value, torch.Tensor):
        return 1 if value.dim() == 0 else int(value.numel())
    a = np.asarray(value)
    return 1 if a.ndim == 0 else int(a.size)


def batch_size(kwargs_lists):
    """Batch size of a set of kwargs lists whose values are scalars or 1-D arrays."""
    n = None
    for kwargs_list in kwargs_lists:
        for kw in kwargs_list or []:
            for v in kw.values():
                b = _batch_of(v)
                if b != 1:
                    if n is not None and n != b:
                        raise ValueError("inconsistent batch sizes %d and %d in kwargs" % (n, b))
                    n = b
    return 1 if n is None else n


def pack(names_lists, kwargs_lists, n_img):
    """Pack kwargs into a float64 [n_img][n_profiles][NPARAM] parameter block.

    names_lists: per profile, its slot layout: ordered parameter names, or
    (name, value) for a slot fixed to a constant (see lens_layout).  Values
    may be python scalars, numpy arrays or torch tensors (scalar or length
    n_img).  Missing parameters without a reference default raise TypeError,
    unknown ones raise ValueError, as the reference's profile calls do.
    """
    n_prof = len(names_lists)
    cols = []
    for names, kw in zip(names_lists, kwargs_lists):
        accepted = [nm for nm in names if isinstance(nm, str)]
        for key in kw:
            if key not in accepted:
                raise ValueError("parameter %s is not part of profile parameters %s" % (key, accepted))
        slot = []
        for name in names:
            if not isinstance(name, str):
                slot.append(np.full(n_img, float(name[1])))
                continue
            if name in kw:
                v = kw[name]
            elif name in _DEFAULTS:
                v = _DEFAULTS[name]
            else:
                raise TypeError("missing required parameter '%s'" % name)
            if isinstance(v, torch.Tensor):
                v = v.detach().to("cpu", torch.float64).numpy()
            v = np.asarray(v, dtype=np.float64).reshape(-1)
            if v.size == 1:
                v = np.full(n_img, v[0])
            elif v.size != n_img:
                raise ValueError("parameter %s has %d values for a batch of %d" % (name, v.size, n_img))
            slot.append(v)
        while len(slot) < backend.NPARAM:
            slot.append(np.zeros(n_img))
        cols.append(np.stack(slot, axis=1))
    if n_prof == 0:
        return np.zeros((n_img, 0, backend.NPARAM))
    return np.ascontiguousarray(np.stack(cols, axis=1))


def select(k, n):
    """convert_bool_list semantics: None -> all, int -> one, list -> those."""
    if k is None:
        return list(range(n))
    if isinstance(k, (int, np.integer)):
        if k < 0 or k >= n:
            raise ValueError("k=%d out of range for %d profiles" % (k, n))
        return [int(k)]
    idx = list(k)
    if all(isinstance(i, (bool, np.bool_)) for i in idx):
        if len(idx) != n:
            raise ValueError("bool list length %d does not match %d profiles" % (len(idx), n))
        return [i for i, b in enumerate(idx) if b]
    return [int(i) for i in idx]
