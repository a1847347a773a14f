"""Synthetic workload generators: seeded, and the subhalo mass mapping obeys
the NFW identities it is built from."""

import numpy as np
from scipy import integrate

from jaxtronomy_amd import workloads as W


def test_comoving_distance_matches_quadrature():
    f = lambda z: 1.0 / np.sqrt(0.3 * (1 + z) ** 3 + 0.7)
    for z1, z2 in ((0.0, 0.5), (0.0, 1.5), (0.5, 1.5)):
        want = 299792.458 / 70.0 * integrate.quad(f, z1, z2, epsabs=0, epsrel=1e-13)[0]
        assert abs(W._comoving(z1, z2) - want) / want < 1e-12


def test_nfw_mapping_identities():
    """M200 = 4 pi rho0 Rs^3 (ln(1+c) - c/(1+c)) and M200 = 200 rho_c 4/3 pi r200^3
    hold for the physical quantities behind (Rs, alpha_Rs); alpha_Rs scales as
    rho0 Rs^2 at fixed lens geometry."""
    m = np.array([1e6, 1e7, 1e8, 1e9])
    c = W.concentration(m)
    rs_ang, a_ang = W.nfw_physical_to_angle(m, c, 0.5, 1.5)
    dd = W._comoving(0.0, 0.5) / 1.5
    rho_c = W._RHO_CRIT0 * (0.3 * 1.5 ** 3 + 0.7)
    rs = rs_ang * W._ARCSEC * dd
    r200 = rs * c
    np.testing.assert_allclose(200 * rho_c * 4 / 3 * np.pi * r200 ** 3, m, rtol=1e-12)
    rho0 = 200 / 3 * rho_c * c ** 3 / (np.log(1 + c) - c / (1 + c))
    np.testing.assert_allclose(4 * np.pi * rho0 * rs ** 3 * (np.log(1 + c) - c / (1 + c)), m, rtol=1e-12)
    ratio = a_ang / (rho0 * rs ** 2)
    np.testing.assert_allclose(ratio, ratio[0], rtol=1e-12)
    assert np.all(np.diff(rs_ang) > 0) and np.all(np.diff(a_ang) > 0)


def test_substructure_draw_is_seeded_and_in_range():
    kl, subs, ks = W.substructure_draw(4, n_sub=50)
    kl2, subs2, ks2 = W.substructure_draw(4, n_sub=50)
    assert subs == subs2 and kl == kl2 and ks == ks2
    assert len(subs) == 50
    for s in subs:
        assert abs(s["center_x"]) <= 3 and abs(s["center_y"]) <= 3
        assert s["r_trunc"] == 5 * s["Rs"] and s["Rs"] > 0 and s["alpha_Rs"] > 0
