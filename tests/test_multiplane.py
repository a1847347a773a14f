"""Multi-plane ray tracing: the background distances, the oracle's restatement
of MultiPlaneBulk.ray_shooting against closed-form limits, and the packing of
prepare_ray_shooting_kwargs."""

import numpy as np
import numpy.testing as npt
from scipy import integrate

from jaxtronomy_amd.LensModel.lens_model_bulk import LensModelBulk
from jaxtronomy_amd.Util.cosmo import Background, FlatLambdaCDM
from oracle import lensim_oracle as O


def test_background_distances_vs_quadrature():
    bkg = Background(FlatLambdaCDM(70.0, 0.3))
    f = lambda z: 1.0 / np.sqrt(0.3 * (1 + z) ** 3 + 0.7)
    dc = lambda z: 299792.458 / 70.0 * integrate.quad(f, 0, z, epsabs=0, epsrel=1e-13)[0]
    npt.assert_allclose(bkg.T_xy(0.3, 1.2), dc(1.2) - dc(0.3), rtol=1e-12)
    npt.assert_allclose(bkg.d_xy(0.3, 1.2), (dc(1.2) - dc(0.3)) / 2.2, rtol=1e-12)
    npt.assert_allclose(bkg.d_xy(0, 0.5), 1259.0836, rtol=1e-6)  # D_A(0.5) for H0=70, Om0=0.3


def _sis(theta_E):
    return dict(theta_E=theta_E, e1=0.0, e2=0.0, center_x=0.0, center_y=0.0)


def test_one_plane_equals_single_plane():
    """A single deflector plane: beta = theta - alpha(theta) (flat universe:
    T_s = T_l + T_ls and r2p T_ls / T_s = 1)."""
    lm = LensModelBulk(["EPL", "SHEAR"], multi_plane=True, cosmo=FlatLambdaCDM(70.0, 0.3))
    kw = [dict(theta_E=1.2, gamma=2.1, e1=0.1, e2=-0.05), dict(gamma1=0.03, gamma2=-0.01)]
    rs = lm.prepare_ray_shooting_kwargs(["EPL", "SHEAR"], kw, 1.5, [0.5, 0.5])
    x = np.linspace(-2, 2, 37)
    y = 0.7 * x[::-1] + 0.1
    bx, by = O.multi_plane_ray_shooting(x, y, ["EPL", "SHEAR"], kw, rs["T_ij_list"], rs["T_z_list"],
                                        rs["reduced2physical_factor"])
    wx, wy = O.ray_shooting(x, y, ["EPL", "SHEAR"], kw)
    npt.assert_allclose(bx, wx, rtol=1e-11, atol=1e-12)
    npt.assert_allclose(by, wy, rtol=1e-11, atol=1e-12)
    # same redshift: zero distance between the planes (multi_plane_bulk.py:263-264)
    assert rs["T_ij_list"][1] == 0


def test_two_planes_differ_from_single_plane_and_fillers_are_no_ops():
    lm = LensModelBulk(["SIE", "NFW"], multi_plane=True, cosmo=FlatLambdaCDM(70.0, 0.3))
    kw = [_sis(1.0), dict(Rs=0.3, alpha_Rs=0.2, center_x=0.4, center_y=-0.2)]
    names = ["SIE", "NFW"]
    x = np.linspace(-2, 2, 41)
    y = np.linspace(1.5, -1.0, 41)
    rs = lm.prepare_ray_shooting_kwargs(names, kw, 2.0, [0.9, 0.4], num_deflectors=6)
    # sorted by redshift: NFW (0.4) first, then SIE (0.9), then source and 4 fillers
    assert list(rs["index_list"]) == [1, 0, 2, 2, 2, 2, 2]
    assert len(rs["T_ij_list"]) == 7 and np.all(rs["T_ij_list"][3:] == 0)
    srt = [kw[1], kw[0]]
    bx, by = O.multi_plane_ray_shooting(x, y, ["NFW", "SIE"], srt, rs["T_ij_list"], rs["T_z_list"],
                                        rs["reduced2physical_factor"])
    rs0 = lm.prepare_ray_shooting_kwargs(names, kw, 2.0, [0.9, 0.4])
    bx0, by0 = O.multi_plane_ray_shooting(x, y, ["NFW", "SIE"], srt, rs0["T_ij_list"], rs0["T_z_list"],
                                          rs0["reduced2physical_factor"])
    npt.assert_allclose(bx, bx0, rtol=1e-14)
    wx, wy = O.ray_shooting(x, y, names, kw)
    assert np.max(np.abs(bx - wx)) > 1e-3  # the planes interact
    # a massless foreground plane changes nothing
    kw_zero = [_sis(1.0), dict(Rs=0.3, alpha_Rs=0.0, center_x=0.4, center_y=-0.2)]
    bz, _ = O.multi_plane_ray_shooting(x, y, ["NFW", "SIE"], [kw_zero[1], kw_zero[0]], rs0["T_ij_list"],
                                       rs0["T_z_list"], rs0["reduced2physical_factor"])
    rs1 = lm.prepare_ray_shooting_kwargs(["SIE"], [_sis(1.0)], 2.0, [0.9])
    b1, _ = O.multi_plane_ray_shooting(x, y, ["SIE"], [_sis(1.0)], rs1["T_ij_list"], rs1["T_z_list"],
                                       rs1["reduced2physical_factor"])
    npt.assert_allclose(bz, b1, rtol=1e-12, atol=1e-13)
