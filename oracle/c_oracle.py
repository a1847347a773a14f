"""Build and call the C restatement of the oracle (oracle/lensim_oracle.c).

TEST INFRASTRUCTURE ONLY (tests/, smoke() and bench.py's cpu_baseline leg).
The shared library is built with gcc into oracle/_build/ (git-ignored; it
travels to the GPU box with the tree).
"""

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "lensim_oracle.c")
OUT = os.path.join(HERE, "_build", "liblensim_oracle.so")

CFLAGS = ["-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-shared", "-std=gnu11"]

_lib = None


def build(verbose=False):
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    tmp = OUT + ".tmp%d" % os.getpid()
    res = subprocess.run(["gcc"] + CFLAGS + ["-o", tmp, SRC, "-lm"], stdout=subprocess.PIPE,
                         stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        raise RuntimeError("gcc failed:\n" + res.stdout)
    if verbose and res.stdout.strip():
        print(res.stdout)
    os.replace(tmp, OUT)
    return OUT


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(OUT):
            build()
        h = ctypes.CDLL(OUT)
        P = ctypes.c_void_p
        i, d = ctypes.c_int, ctypes.c_double
        h.oracle_image.argtypes = [i, i, i, P, P, i, i, i, i, d, d, P, i, i, i, P, d, d, ctypes.c_uint, P, i]
        h.oracle_image.restype = i
        h.oracle_loglike.argtypes = [i, i, i, P, P, i, i, i, i, d, d, P, i, i, i, P, d, d, P, P, P, d, d, P, P, i]
        h.oracle_loglike.restype = i
        _lib = h
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _common(n_lens, n_src, n_light, types, params, grid, kernel, psf_type):
    types = np.ascontiguousarray(np.asarray(types, dtype=np.int32))
    params = np.ascontiguousarray(np.asarray(params, dtype=np.float64))
    M = np.ascontiguousarray(np.asarray(grid["transform_pix2angle"], dtype=np.float64).reshape(4))
    if kernel is None or psf_type == 0:
        kernel = np.ones((1, 1))
        psf_type = 0
    kernel = np.ascontiguousarray(np.asarray(kernel, dtype=np.float64))
    return types, params, M, kernel, psf_type


def image(n_lens, n_src, n_light, types, params, grid, kernel=None, psf_type=1, alpha_scaling=1.0,
          flux_scaling=1.0, flags=3, n_threads=0):
    """params: [n_img][n_prof][8]; grid: dict(nx, ny, s, ra_at_xy_0, dec_at_xy_0, transform_pix2angle)."""
    types, params, M, kernel, psf_type = _common(n_lens, n_src, n_light, types, params, grid, kernel, psf_type)
    n_img = params.shape[0]
    out = np.zeros((n_img, grid["ny"], grid["nx"]))
    rc = lib().oracle_image(n_lens, n_src, n_light, _p(types), _p(params), n_img, grid["nx"], grid["ny"], grid["s"],
                            grid["ra_at_xy_0"], grid["dec_at_xy_0"], _p(M), psf_type, kernel.shape[1],
                            kernel.shape[0], _p(kernel), alpha_scaling, flux_scaling, flags, _p(out), n_threads)
    if rc:
        raise RuntimeError("oracle_image failed")
    return out


def loglike(n_lens, n_src, n_light, types, params, grid, data, kernel=None, psf_type=1, noise_map=None,
            exposure_map=None, exposure_time=1.0, background_rms=0.0, mask=None, alpha_scaling=1.0,
            flux_scaling=1.0, n_threads=0):
    types, params, M, kernel, psf_type = _common(n_lens, n_src, n_light, types, params, grid, kernel, psf_type)
    n_img = params.shape[0]
    c = lambda a: None if a is None else np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    data, noise_map, exposure_map, mask = c(data), c(noise_map), c(exposure_map), c(mask)
    logl = np.zeros(n_img)
    rc = lib().oracle_loglike(n_lens, n_src, n_light, _p(types), _p(params), n_img, grid["nx"], grid["ny"],
                              grid["s"], grid["ra_at_xy_0"], grid["dec_at_xy_0"], _p(M), psf_type, kernel.shape[1],
                              kernel.shape[0], _p(kernel), alpha_scaling, flux_scaling, _p(data), _p(noise_map),
                              _p(exposure_map), float(exposure_time), float(background_rms), _p(mask), _p(logl),
                              n_threads)
    if rc:
        raise RuntimeError("oracle_loglike failed")
    return logl
