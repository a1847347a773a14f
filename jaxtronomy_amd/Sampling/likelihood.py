"""Sampler-facing batched likelihood: parameter vectors -> log-likelihoods.

Mirrors the reference's sampling entry points for the imaging likelihood:
``LikelihoodModule.logL`` (Sampling/likelihood.py:290-373: parameter vector
-> kwargs, hard bounds -> -1e18, the image likelihood with NaN -> -1e15 as in
ImageLikelihood.logL, Sampling/Likelihoods/image_likelihood.py:57-95, plus an
optional custom addition) and ``prepare_logL_func`` (Sampling/sampler.py:215-251),
which vectorises it over a set of walkers / particles.  Here the whole set is
one batched device call: the parameter vectors are cast into the packed
parameter block on the host in one vectorised step and evaluated by
ImageModel.likelihood_batch (or, with the linear solver, by
ImageLinearFit.linear_solve_batch).

Param orders the sampled parameters as lenstronomy's Param does for these
profiles: profile by profile (lens, source, lens light), each profile's
``param_names`` in order, leaving out the fixed ones and -- with the linear
solver -- the light amplitudes.
"""

import numpy as np
import torch

from .. import backend, profiles

BOUND_PENALTY = -1e18  # likelihood.py:312-315
NAN_LOGL = -1e15       # image_likelihood.py:93


class Param(object):
    def __init__(self, image_model, kwargs_fixed_lens=None, kwargs_fixed_source=None, kwargs_fixed_lens_light=None,
                 kwargs_lower_lens=None, kwargs_lower_source=None, kwargs_lower_lens_light=None,
                 kwargs_upper_lens=None, kwargs_upper_source=None, kwargs_upper_lens_light=None,
                 linear_solver=False):
        if image_model._bulk:
            raise NotImplementedError("Param covers plain lens model lists")
        self.linear_solver = bool(linear_solver)
        lists = [image_model._lens_types, image_model._src_types, image_model._light_types]
        fixed = [kwargs_fixed_lens, kwargs_fixed_source, kwargs_fixed_lens_light]
        lower = [kwargs_lower_lens, kwargs_lower_source, kwargs_lower_lens_light]
        upper = [kwargs_upper_lens, kwargs_upper_source, kwargs_upper_lens_light]
        self._n_prof = len(image_model._names)
        self._base = np.zeros((self._n_prof, backend.NPARAM))
        self._slots = []   # (profile row, slot) of every sampled parameter, in order
        self.names = []
        lo, hi = [], []
        k = 0
        for group, types in enumerate(lists):
            fx = fixed[group] or [{} for _ in types]
            lw = lower[group] or [{} for _ in types]
            up = upper[group] or [{} for _ in types]
            for i, _ in enumerate(types):
                names = image_model._names[k]
                for q, nm in enumerate(names):
                    if nm in fx[i]:
                        self._base[k, q] = float(fx[i][nm])
                        continue
                    if self.linear_solver and group > 0 and nm == "amp":
                        self._base[k, q] = 1.0  # set by the linear solve
                        continue
                    self._slots.append((k, q))
                    self.names.append("%s_%s%d" % (nm, ("lens", "source", "lens_light")[group], i))
                    lo.append(float(lw[i].get(nm, -np.inf)))
                    hi.append(float(up[i].get(nm, np.inf)))
                k += 1
        self._rows = np.array([s[0] for s in self._slots], dtype=int)
        self._cols = np.array([s[1] for s in self._slots], dtype=int)
        self._lower = np.array(lo)
        self._upper = np.array(hi)

    def num_param(self):
        """(number of sampled parameters, their names)."""
        return len(self._slots), list(self.names)

    def param_limits(self):
        return self._lower.copy(), self._upper.copy()

    def args2params(self, args):
        """[N, num_param] parameter vectors -> packed parameter block [N, n_profiles, 8]."""
        args = np.atleast_2d(np.asarray(args, dtype=np.float64))
        if args.shape[1] != len(self._slots):
            raise ValueError("expected %d parameters per vector, got %d" % (len(self._slots), args.shape[1]))
        out = np.repeat(self._base[None], args.shape[0], axis=0)
        out[:, self._rows, self._cols] = args
        return out

    def params2kwargs(self, params_row, image_model):
        """One packed parameter row [n_profiles, 8] -> dict of kwargs lists
        (kwargs_lens, kwargs_source, kwargs_lens_light), Param.args2kwargs layout."""
        out, k = [], 0
        for types in (image_model._lens_types, image_model._src_types, image_model._light_types):
            lst = []
            for _ in types:
                lst.append({nm: float(params_row[k, q]) for q, nm in enumerate(image_model._names[k])})
                k += 1
            out.append(lst)
        return {"kwargs_lens": out[0], "kwargs_source": out[1], "kwargs_lens_light": out[2], "kwargs_ps": None,
                "kwargs_special": None, "kwargs_extinction": None}

    def kwargs2args(self, kwargs_lens=None, kwargs_source=None, kwargs_lens_light=None, image_model=None):
        """kwargs lists -> parameter vector (the inverse of args2params)."""
        if image_model is None:
            raise ValueError("kwargs2args needs the image model (for its parameter names)")
        kl = list(kwargs_lens or [])
        ks = [dict(k, amp=k.get("amp", 1.0)) for k in (kwargs_source or [])]
        kll = [dict(k, amp=k.get("amp", 1.0)) for k in (kwargs_lens_light or [])]
        packed = profiles.pack(image_model._names, kl + ks + kll, 1)[0]
        return packed[self._rows, self._cols]


class LikelihoodModule(object):
    """Batched LikelihoodModule.logL for the imaging likelihood."""

    def __init__(self, image_model, param, check_bounds=True, source_marg=False, linear_prior=None,
                 check_positive_flux=False, custom_logL_addition=None):
        self.imSim = image_model
        self.param = param
        self._check_bounds = check_bounds
        self._source_marg = source_marg
        self._linear_prior = linear_prior
        self._check_positive_flux = check_positive_flux
        self._custom = custom_logL_addition
        self._lower, self._upper = param.param_limits()
        if param.linear_solver and not hasattr(image_model, "linear_solve_batch"):
            raise ValueError("the linear solver needs an ImageLinearFit")

    def logL(self, args):
        """log-likelihoods [N] (numpy) of [N, num_param] parameter vectors."""
        args = np.atleast_2d(np.asarray(args, dtype=np.float64))
        packed = self.param.args2params(args)
        params = backend.to_device_f64(packed)
        if self.param.linear_solver:
            r = self.imSim.linear_solve_batch(params, source_marg=self._source_marg, linear_prior=self._linear_prior)
            logl = r["logl"]
            if self._check_positive_flux:
                logl = torch.where(torch.any(r["amp"] < 0, dim=1), logl - 1e8, logl)
        else:
            logl = self.imSim.likelihood_batch(params)
        logl = torch.nan_to_num(logl, nan=NAN_LOGL).cpu().numpy()
        if self._custom is not None:  # custom_logL_addition(**kwargs_return), likelihood.py:349-351
            logl = logl + np.array([self._custom(**self.param.params2kwargs(row, self.imSim)) for row in packed])
        if self._check_bounds:
            hit = np.any((args < self._lower) | (args > self._upper), axis=1)
            logl = np.where(hit, BOUND_PENALTY, logl)
        return logl

    def __call__(self, a):
        return self.logL(a)


def prepare_logL_func(backend_name, logL_func, threadCount=1):
    """prepare_logL_func, Sampling/sampler.py:215-251: a callable taking a set of
    position vectors [N, num_param] and returning N log-likelihoods.  Both
    backends evaluate the set as one batched device call here."""
    if backend_name not in ("cpu", "gpu"):
        raise ValueError("backend must be either 'cpu' or 'gpu'")

    def new_logL_func(args):
        return np.asarray(logL_func(np.asarray(args)), dtype=np.float64).flatten()

    return new_logL_func
