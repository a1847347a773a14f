"""The C-ABI library loads and exports every symbol include/lensim.h declares;
host-side argument validation returns the documented status codes.  No GPU:
every call here fails validation before any device work."""

import ctypes
import os
import re

import numpy as np
import pytest

from jaxtronomy_amd import backend, profiles

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    src = open(os.path.join(ROOT, "include", "lensim.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(lensim_[a-z_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = backend.lib()
    names = _declared_functions()
    assert "lensim_image" in names and "lensim_loglike" in names
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(backend._SIGNATURES), "ctypes signatures out of sync with the header"
    assert lib.lensim_abi_version() == backend.ABI_VERSION


def test_checked_library_exports_every_declared_symbol():
    """liblensim_checked.so (the bounds-checked debug build) has the same C ABI
    and says it is the checked build; the normal one says it is not."""
    checked = backend._load(backend.LIB_CHECKED_PATH)
    for name in _declared_functions():
        assert hasattr(checked, name), name
    assert checked.lensim_abi_version() == backend.ABI_VERSION
    assert checked.lensim_bounds_checked() == 1
    assert backend.lib().lensim_bounds_checked() == 0
    # the normal build refuses the self-test before any device work
    assert backend.lib().lensim_bounds_selftest(None, 16, None) == -2


def test_struct_layouts():
    assert ctypes.sizeof(backend.LensimModel) == 12 + 4 * 64 + 4 + 16 + 8 + 8 + 8  # padded to 8 before the doubles
    assert ctypes.sizeof(backend.LensimMultiplane) == 8 + 4 * 8 + 8
    assert ctypes.sizeof(backend.LensimGrid) == 16 + 16 + 32 + 8
    assert ctypes.sizeof(backend.LensimPsf) == 40
    assert ctypes.sizeof(backend.LensimData) == 56


def _grid(nx=10, ny=10, s=1):
    g = backend.LensimGrid()
    g.nx, g.ny, g.supersampling = nx, ny, s
    g.pix2angle[0], g.pix2angle[3] = -0.05, 0.05
    return g


def test_workspace_size_and_validation():
    lib = backend.lib()
    m = backend.make_model([1, 6], [33], [33])
    g = _grid(180, 180, 3)
    ws64 = lib.lensim_workspace_size(ctypes.byref(m), ctypes.byref(g), None, 64, 0)
    ws32 = lib.lensim_workspace_size(ctypes.byref(m), ctypes.byref(g), None, 64, 1)
    assert ws64 > ws32 > 180 * 180 * 64 * 4
    bad = backend.make_model([33], [], [])  # a light profile in the lens list
    assert lib.lensim_workspace_size(ctypes.byref(bad), ctypes.byref(g), None, 1, 0) == 0


def _call_image(m, g, psf, n_img, prec=0, flags=3, ws=0, out=1, params=1):
    lib = backend.lib()
    return lib.lensim_image(ctypes.byref(m), None if g is None else ctypes.byref(g),
                            None if psf is None else ctypes.byref(psf), ctypes.c_void_p(params), n_img, prec, flags,
                            ctypes.c_void_p(out), ctypes.c_void_p(0), ws, None)


def test_image_rejects_bad_arguments_before_device_work():
    lib = backend.lib()
    m = backend.make_model([2, 6], [33], [])
    assert _call_image(m, None, None, 1) == -1
    assert b"grid" in lib.lensim_last_error()
    g = _grid()
    g.supersampling = 0
    assert _call_image(m, g, None, 1) == -1
    g = _grid()
    psf = backend.LensimPsf()
    psf.type, psf.kx, psf.ky, psf.d_kernel = backend.PSF_PIXEL, 4, 5, 1
    assert _call_image(m, g, psf, 1) == -1  # even kernel width
    assert b"odd" in lib.lensim_last_error()
    psf.kx = 5
    psf.type = 7
    assert _call_image(m, g, psf, 1) == -2
    unsup = backend.make_model([40], [33], [])
    assert _call_image(unsup, g, None, 1) == -2
    assert _call_image(m, g, None, 1, prec=5, ws=1 << 30) == -1
    assert _call_image(m, g, None, 1, ws=16) == -4  # workspace too small
    assert b"workspace" in lib.lensim_last_error()
    assert _call_image(m, g, None, 1, out=0) == -1
    assert _call_image(m, g, None, 0) == 0  # empty batch: nothing to do
    assert _call_image(m, g, None, -1) == -1


def test_raytrace_and_loglike_validation():
    lib = backend.lib()
    m = backend.make_model([2], [], [])
    rc = lib.lensim_raytrace(ctypes.byref(m), ctypes.c_void_p(1), 1, None, None, 10, 0, None, None, None, 0, None)
    assert rc == -1
    rc = lib.lensim_raytrace(ctypes.byref(m), ctypes.c_void_p(1), 1, None, None, 0, 0, None, None, None, 0, None)
    assert rc == 0  # no rays
    g = _grid()
    d = backend.LensimData()
    rc = lib.lensim_loglike(ctypes.byref(m), ctypes.byref(g), None, ctypes.byref(d), ctypes.c_void_p(1), 1, 0,
                            ctypes.c_void_p(1), None, None, 0, None)
    assert rc == -1 and b"data" in lib.lensim_last_error()
    d.d_data = 1
    rc = lib.lensim_loglike(ctypes.byref(m), ctypes.byref(g), None, ctypes.byref(d), ctypes.c_void_p(1), 1, 0,
                            ctypes.c_void_p(1), None, None, 0, None)
    assert rc == -1 and b"exposure" in lib.lensim_last_error()
    rc = lib.lensim_surface_brightness(ctypes.byref(m), None, 1, 2, None, None, 1, 0, None, None, 0, None)
    assert rc == -1


def test_pack_follows_param_names_and_defaults():
    names = [profiles.LENS_PROFILES["EPL"][1], profiles.LENS_PROFILES["SHEAR"][1]]
    kw = [dict(theta_E=1.0, gamma=2.0, e1=0.1, e2=np.array([0.0, 0.2])), dict(gamma1=0.01, gamma2=-0.02)]
    p = profiles.pack(names, kw, 2)
    assert p.shape == (2, 2, 8)
    np.testing.assert_array_equal(p[1, 0, :6], [1.0, 2.0, 0.1, 0.2, 0.0, 0.0])
    np.testing.assert_array_equal(p[0, 1, :4], [0.01, -0.02, 0.0, 0.0])
    with pytest.raises(TypeError):
        profiles.pack(names, [dict(gamma=2.0, e1=0, e2=0), kw[1]], 1)
    with pytest.raises(ValueError):
        profiles.pack(names, [dict(kw[0], bogus=1.0), kw[1]], 2)
    assert profiles.batch_size([kw]) == 2


def test_profile_selection_semantics():
    assert profiles.select(None, 3) == [0, 1, 2]
    assert profiles.select(1, 3) == [1]
    assert profiles.select([0, 2], 3) == [0, 2]
    assert profiles.select([True, False, True], 3) == [0, 2]
    with pytest.raises(ValueError):
        profiles.select(5, 3)


def test_no_gpu_means_loud_failure(monkeypatch):
    import torch

    monkeypatch.setattr(torch.cuda, "is_available", lambda: False)
    with pytest.raises(backend.LensimError):
        backend.require_gpu()
