"""Host logic of LensModelBulk's packing (no GPU): which deflectors go to the
inline slots and which to the bulk loop, and the gradient's way back to the
form of ray_shooting_kwargs["all_kwargs"]."""

import numpy as np

from jaxtronomy_amd.LensModel.lens_model_bulk import LensModelBulk


def _kwargs(names, rng):
    out = []
    for nm in names:
        if nm == "EPL":
            out.append(dict(theta_E=1.2, gamma=2.0, e1=0.1, e2=-0.05, center_x=0.01, center_y=0.02))
        elif nm == "SHEAR":
            out.append(dict(gamma1=0.02, gamma2=-0.01))
        elif nm == "SIE":
            out.append(dict(theta_E=rng.uniform(0.01, 0.1), e1=0.1, e2=0.0, center_x=rng.uniform(-1, 1),
                            center_y=rng.uniform(-1, 1)))
        else:
            rs = rng.uniform(0.05, 0.2)
            kw = dict(Rs=rs, alpha_Rs=rng.uniform(0.01, 0.05), center_x=rng.uniform(-1, 1),
                      center_y=rng.uniform(-1, 1))
            if nm == "TNFW":
                kw["r_trunc"] = 5 * rs
            out.append(kw)
    return out


def _roundtrip(lm, names, kwargs, num_deflectors):
    rs = lm.prepare_ray_shooting_kwargs(names, kwargs, num_deflectors=num_deflectors)
    n_img, it, ip, bt, bp = lm.pack(rs)
    packed = np.concatenate([ip, bp], axis=1)
    # a "gradient" equal to the packed parameters maps back onto all_kwargs itself
    g = lm.unpack_gradient(rs, packed)
    for kw, col in rs["all_kwargs"].items():
        want = np.array(col, dtype=np.float64)
        for d, nm in enumerate(names):
            if kw not in lm.param_name_list[lm.unique_lens_model_list.index(nm)]:
                want[d] = 0.0  # a kwarg this deflector's profile does not have
        want[len(names):] = 0.0  # NULL padding
        np.testing.assert_array_equal(g[kw], want, err_msg=kw)
    return it, bt


def test_unpack_gradient_single_type_bulk():
    rng = np.random.default_rng(0)
    names = ["EPL", "SHEAR"] + ["TNFW"] * 20
    lm = LensModelBulk(["EPL", "SHEAR", "TNFW"])
    it, bt = _roundtrip(lm, names, _kwargs(names, rng), 32)
    assert list(it) == [1, 6] and list(bt) == [5] * 20


def test_unpack_gradient_mixed_bulk_and_batch():
    rng = np.random.default_rng(1)
    names = ["SIE", "EPL"] + ["TNFW", "NFW", "SIE"] * 40
    lm = LensModelBulk(["EPL", "TNFW", "NFW", "SIE"])
    it, bt = _roundtrip(lm, names, _kwargs(names, rng), 130)
    assert list(it) == [1] and len(bt) == 121 and set(bt) == {2, 4, 5}  # the minorities do not fit inline
    # a batch of two: [B, n_deflectors] columns
    rs = lm.stack_ray_shooting_kwargs([lm.prepare_ray_shooting_kwargs(names, _kwargs(names, rng))
                                       for _ in range(2)])
    n_img, it, ip, bt, bp = lm.pack(rs)
    g = lm.unpack_gradient(rs, np.concatenate([ip, bp], axis=1))
    assert n_img == 2 and g["theta_E"].shape == (2, len(names))
    np.testing.assert_array_equal(g["theta_E"][:, 1], rs["all_kwargs"]["theta_E"][:, 1])
    np.testing.assert_array_equal(g["r_trunc"][:, 2], rs["all_kwargs"]["r_trunc"][:, 2])
