"""ctypes binding of liblensim.so (the C ABI in include/lensim.h).

The library is the only compute path: nothing here falls back to a CPU or
PyTorch implementation.  ``lib()`` raises if the shared library is missing and
``require_gpu()`` raises if no HIP device is visible.  torch is used only for
device buffers and the current stream.
"""

import ctypes
import os

import numpy as np
import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "liblensim.so")
# the bounds-checked debug build: every kernel access to a caller or workspace
# buffer is tested against the call's buffers, and a call with an access
# outside them fails with LENSIM_EDEVICE naming the source line (each call
# also waits for its stream); LENSIM_CHECKED=1 loads it in place of LIB_PATH
LIB_CHECKED_PATH = os.path.join(_HERE, "_lib", "liblensim_checked.so")

ABI_VERSION = 6
NPARAM = 8
MAX_PROFILES = 64
EPL_NMAX = 500

FP64 = 0
MIXED = 1
PRECISIONS = {"fp64": FP64, "mixed": MIXED}

IMAGE_SOURCE = 1
IMAGE_LENS_LIGHT = 2
IMAGE_UNCONVOLVED = 4

PSF_NONE, PSF_PIXEL, PSF_GAUSSIAN = 0, 1, 2

STATUS = {0: "ok", -1: "invalid argument", -2: "unsupported", -3: "device error", -4: "workspace too small"}


class LensimModel(ctypes.Structure):
    _fields_ = [
        ("n_lens", ctypes.c_int32),
        ("n_source", ctypes.c_int32),
        ("n_lens_light", ctypes.c_int32),
        ("types", ctypes.c_int32 * MAX_PROFILES),
        ("alpha_scaling", ctypes.c_double),
        ("flux_scaling", ctypes.c_double),
        ("n_bulk", ctypes.c_int32),
        ("bulk_type", ctypes.c_int32),
        ("d_bulk_types", ctypes.c_void_p),
        ("multiplane", ctypes.c_void_p),
    ]


class LensimMultiplane(ctypes.Structure):
    _fields_ = [
        ("n_planes", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("d_order", ctypes.c_void_p),
        ("d_T_ij", ctypes.c_void_p),
        ("d_T_z", ctypes.c_void_p),
        ("d_r2p", ctypes.c_void_p),
        ("stride", ctypes.c_int64),
    ]


class LensimGrid(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32),
        ("ny", ctypes.c_int32),
        ("supersampling", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("ra_at_xy_0", ctypes.c_double),
        ("dec_at_xy_0", ctypes.c_double),
        ("pix2angle", ctypes.c_double * 4),
        ("d_eval_mask", ctypes.c_void_p),
    ]


class LensimPsf(ctypes.Structure):
    _fields_ = [
        ("type", ctypes.c_int32),
        ("kx", ctypes.c_int32),
        ("ky", ctypes.c_int32),
        ("supersampled", ctypes.c_int32),
        ("d_kernel", ctypes.c_void_p),
        ("kx_low", ctypes.c_int32),
        ("ky_low", ctypes.c_int32),
        ("d_kernel_low", ctypes.c_void_p),
    ]


class LensimData(ctypes.Structure):
    _fields_ = [
        ("d_data", ctypes.c_void_p),
        ("d_noise_map", ctypes.c_void_p),
        ("d_exposure_map", ctypes.c_void_p),
        ("d_mask", ctypes.c_void_p),
        ("exposure_time", ctypes.c_double),
        ("background_rms", ctypes.c_double),
        ("image_stride", ctypes.c_int64),
    ]


# exported symbol -> (restype, argtypes); every symbol declared in include/lensim.h
_P = ctypes.c_void_p
_SIGNATURES = {
    "lensim_abi_version": (ctypes.c_int, []),
    "lensim_last_error": (ctypes.c_char_p, []),
    "lensim_workspace_size": (ctypes.c_size_t, [_P, _P, _P, ctypes.c_int32, ctypes.c_int32]),
    "lensim_raytrace": (ctypes.c_int, [_P, _P, ctypes.c_int32, _P, _P, ctypes.c_int64, ctypes.c_int32, _P, _P, _P,
                                       ctypes.c_size_t, _P]),
    "lensim_surface_brightness": (ctypes.c_int, [_P, _P, ctypes.c_int32, ctypes.c_int32, _P, _P, ctypes.c_int64,
                                                 ctypes.c_int32, _P, _P, ctypes.c_size_t, _P]),
    "lensim_image": (ctypes.c_int, [_P, _P, _P, _P, ctypes.c_int32, ctypes.c_int32, ctypes.c_uint32, _P, _P,
                                    ctypes.c_size_t, _P]),
    "lensim_loglike": (ctypes.c_int, [_P, _P, _P, _P, _P, ctypes.c_int32, ctypes.c_int32, _P, _P, _P,
                                      ctypes.c_size_t, _P]),
    "lensim_data_loglike_workspace_size": (ctypes.c_size_t, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    "lensim_data_loglike": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32, _P, ctypes.c_int32, ctypes.c_int32,
                                           _P, _P, ctypes.c_size_t, _P]),
    "lensim_linear_workspace_size": (ctypes.c_size_t, [_P, _P, _P, ctypes.c_int32, ctypes.c_int32]),
    "lensim_linear_solve": (ctypes.c_int, [_P, _P, _P, _P, _P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                           ctypes.c_double, _P, _P, _P, _P, _P, _P, ctypes.c_size_t, _P]),
    "lensim_grad_workspace_size": (ctypes.c_size_t, [_P, _P, _P, ctypes.c_int32, ctypes.c_int32]),
    "lensim_loglike_grad": (ctypes.c_int, [_P, _P, _P, _P, _P, ctypes.c_int32, ctypes.c_int32, _P, _P, _P,
                                           ctypes.c_size_t, _P]),
    "lensim_noalloc_scratch_bytes": (ctypes.c_size_t, []),
    "lensim_bounds_checked": (ctypes.c_int, []),
    "lensim_bounds_selftest": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]),
    "lensim_loglike_noalloc": (ctypes.c_int, [_P, _P, _P, _P, _P, ctypes.c_int32, ctypes.c_int32, _P, _P, _P]),
    "lensim_graph_capture_loglike": (ctypes.c_int, [_P, _P, _P, _P, _P, ctypes.c_int32, ctypes.c_int32, _P, _P,
                                                    _P, ctypes.c_size_t, _P, ctypes.POINTER(ctypes.c_void_p)]),
    "lensim_graph_launch": (ctypes.c_int, [_P, _P]),
    "lensim_graph_destroy": (ctypes.c_int, [_P]),
    "lensim_timing_enable": (ctypes.c_int, [ctypes.c_int]),
    "lensim_timing_read": (ctypes.c_int, [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_int64)]),
}

_lib = None


class LensimError(RuntimeError):
    pass


def _load(path):
    if not os.path.exists(path):
        raise LensimError(
            "%s is not built (%s); run __graft_entry__.build() first" % (os.path.basename(path), path))
    handle = ctypes.CDLL(path)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(handle, name)
        fn.restype = res
        fn.argtypes = args
    if handle.lensim_abi_version() != ABI_VERSION:
        raise LensimError("liblensim ABI version mismatch")
    return handle


def lib():
    """Load liblensim.so (built in-tree by ``__graft_entry__.build()``), or
    liblensim_checked.so when LENSIM_CHECKED=1."""
    global _lib
    if _lib is None:
        _lib = _load(LIB_CHECKED_PATH if os.environ.get("LENSIM_CHECKED") == "1" else LIB_PATH)
    return _lib


def use_checked(on=True):
    """Route every later call through the bounds-checked build (on=True) or
    the normal one; returns the handle in use before, for restore_lib."""
    global _lib
    before = _lib
    _lib = _load(LIB_CHECKED_PATH if on else LIB_PATH)
    return before


def restore_lib(handle):
    global _lib
    _lib = handle


def require_gpu():
    if not torch.cuda.is_available():
        raise LensimError("no HIP device visible: the lensim kernels need an MI355X (gfx950)")


def device():
    require_gpu()
    return torch.device("cuda", torch.cuda.current_device())


def check(rc, what):
    if rc != 0:
        msg = lib().lensim_last_error().decode(errors="replace")
        if rc in (-1, -2):
            raise ValueError("%s failed (%s): %s" % (what, STATUS.get(rc, rc), msg))
        raise LensimError("%s failed (%s): %s" % (what, STATUS.get(rc, rc), msg))


def stream_handle():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


class Workspace:
    """Grow-only device scratch buffer (256-byte aligned by torch's allocator)."""

    def __init__(self):
        self._buf = None

    def get(self, nbytes):
        if self._buf is None or self._buf.numel() < nbytes:
            self._buf = torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=device())
        return self._buf


def make_model(lens_types, source_types, light_types, alpha_scaling=1.0, flux_scaling=1.0, bulk=None):
    """lensim_model for inline profile lists; bulk = (n_bulk, common_type, device int32 types or None)."""
    types = list(lens_types) + list(source_types) + list(light_types)
    if len(types) > MAX_PROFILES:
        raise ValueError("at most %d profiles per model" % MAX_PROFILES)
    m = LensimModel()
    m.n_lens = len(lens_types)
    m.n_source = len(source_types)
    m.n_lens_light = len(light_types)
    for i, t in enumerate(types):
        m.types[i] = t
    m.alpha_scaling = float(alpha_scaling)
    m.flux_scaling = float(flux_scaling)
    if bulk is not None:
        n_bulk, common, d_types = bulk
        m.n_bulk = int(n_bulk)
        m.bulk_type = int(common)
        m.d_bulk_types = None if d_types is None else d_types.data_ptr()
    return m


def to_device_f64(a):
    if isinstance(a, torch.Tensor):
        return a.to(device=device(), dtype=torch.float64).contiguous()
    return torch.as_tensor(np.ascontiguousarray(np.asarray(a, dtype=np.float64)), device=device())


def timing_enable(on=True):
    lib().lensim_timing_enable(1 if on else 0)


def timing_read():
    """(render_ms, conv_ms, calls) summed since the last read (see lensim_timing_read)."""
    r, c, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
    check(lib().lensim_timing_read(ctypes.byref(r), ctypes.byref(c), ctypes.byref(n)), "lensim_timing_read")
    return r.value, c.value, n.value
