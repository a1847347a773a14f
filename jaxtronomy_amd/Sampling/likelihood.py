"""Sampler-facing batched likelihood: parameter vectors -> log-likelihoods.

Mirrors the reference's sampling entry points for the imaging likelihood:
``LikelihoodModule.logL`` (Sampling/likelihood.py:290-373: parameter vector
-> kwargs, hard bounds -> -1e18; otherwise the custom addition, the Gaussian
priors (lenstronomy's PriorLikelihood, likelihood.py:153-172, :353-356) and
the image likelihood with NaN -> -1e15 as in ImageLikelihood.logL,
Sampling/Likelihoods/image_likelihood.py:57-95, summed in that order, and the
sum's NaN -> 1e-18, +-inf -> +-max float, likelihood.py:370) and
``prepare_logL_func`` (Sampling/sampler.py:215-251),
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
NAN_TOTAL = 1e-18      # likelihood.py:370


def _gaussian_priors(prior_list, group, image_model, packed):
    """sum over [index, param_name, mean, sigma] entries of -(value - mean)^2 / sigma^2 / 2
    (lenstronomy PriorLikelihood's Gaussian priors) for each packed row [N]."""
    out = np.zeros(packed.shape[0])
    if not prior_list:
        return out
    offset = [0, len(image_model._lens_types), len(image_model._lens_types) + len(image_model._src_types)][group]
    n_group = [len(image_model._lens_types), len(image_model._src_types), len(image_model._light_types)][group]
    for entry in prior_list:
        index, name, mean, sigma = entry
        if not 0 <= int(index) < n_group:
            raise ValueError("prior index %d out of range for %d profiles" % (index, n_group))
        k = offset + int(index)
        slots = dict((nm, q) for q, nm in profiles.kwarg_slots(image_model._names[k]))
        if name not in slots:
            raise ValueError("prior on %s, which profile %d does not have" % (name, index))
        value = packed[:, k, slots[name]]
        out -= (value - float(mean)) ** 2 / float(sigma) ** 2 / 2.0
    return out


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
                for q, nm in enumerate(image_model._names[k]):
                    if not isinstance(nm, str):  # a slot fixed by the profile's options
                        self._base[k, q] = float(nm[1])
                        continue
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
                lst.append({nm: float(params_row[k, q]) for q, nm in profiles.kwarg_slots(image_model._names[k])})
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
                 check_positive_flux=False, custom_logL_addition=None, prior_lens=None, prior_source=None,
                 prior_lens_light=None, prior_ps=None, prior_special=None, prior_extinction=None, **kwargs):
        if prior_ps or prior_special or prior_extinction:
            raise NotImplementedError("priors on point sources, special or extinction parameters: those models "
                                      "are not part of this library")
        for key, value in kwargs.items():
            if value is not None and value is not False:
                raise NotImplementedError("LikelihoodModule option %s is not supported" % key)
        self._priors = [(prior_lens, 0), (prior_source, 1), (prior_lens_light, 2)]
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
        # ImageLikelihood.logL: NaN -> -1e15, +-inf -> +-max float
        image = torch.nan_to_num(logl, nan=NAN_LOGL).cpu().numpy()
        # LikelihoodModule.log_likelihood (likelihood.py:346-371): custom addition, priors, image, in that order
        total = np.zeros(args.shape[0])
        if self._custom is not None:
            total = total + np.array([self._custom(**self.param.params2kwargs(row, self.imSim)) for row in packed],
                                     dtype=np.float64)
        prior = np.zeros(args.shape[0])
        for prior_list, group in self._priors:
            prior += _gaussian_priors(prior_list, group, self.imSim, packed)
        total = total + prior
        total = total + image
        total = np.nan_to_num(total, nan=NAN_TOTAL)
        if self._check_bounds:
            hit = np.any((args < self._lower) | (args > self._upper), axis=1)
            total = np.where(hit, BOUND_PENALTY, total)
        return total

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
