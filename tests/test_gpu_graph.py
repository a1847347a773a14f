"""lensim_graph_capture_loglike / lensim_graph_launch through the C ABI: a
captured log-likelihood call replayed with new parameters written in place
between launches (and eager work between the replays) equals the eager
lensim_loglike call on the same parameters, bit for bit, in both forms."""

import ctypes

import numpy as np
import pytest
import torch

from jaxtronomy_amd import backend
from jaxtronomy_amd import workloads as W
from jaxtronomy_amd.Data.imaging_data import ImageData
from jaxtronomy_amd.Data.psf import PSF
from jaxtronomy_amd.ImSim.image_model import ImageModel
from jaxtronomy_amd.LensModel.lens_model import LensModel
from jaxtronomy_amd.LightModel.light_model import LightModel

pytestmark = pytest.mark.gpu


def _model(precision):
    cfg = dict(W.EPL180, numPix=64, supersampling=2)
    kernel = W.moffat_kernel(21, cfg["psf_fwhm"] / cfg["deltaPix"], cfg["psf_beta"])
    kl, ks, kll = W.epl180_draw(0)
    probe = ImageModel(ImageData(**W.data_kwargs(cfg)), PSF("PIXEL", kernel_point_source=kernel),
                       LensModel(["EPL", "SHEAR"]), LightModel(["SERSIC_ELLIPSE"]), LightModel(["SERSIC_ELLIPSE"]),
                       kwargs_numerics={"supersampling_factor": 2})
    data = W.noisy(probe.image(kl, ks, kll).cpu().numpy(), cfg["exposure_time"], cfg["background_rms"],
                   np.random.default_rng(5))
    return ImageModel(ImageData(**W.data_kwargs(cfg, data)), PSF("PIXEL", kernel_point_source=kernel),
                      LensModel(["EPL", "SHEAR"]), LightModel(["SERSIC_ELLIPSE"]), LightModel(["SERSIC_ELLIPSE"]),
                      kwargs_numerics={"supersampling_factor": 2}, precision=precision)


def _params(im, draws):
    p, _ = im.pack(W.stack_kwargs([d[0] for d in draws]), W.stack_kwargs([d[1] for d in draws]),
                   W.stack_kwargs([d[2] for d in draws]))
    return p


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_capture_replays_equal_eager(precision):
    lib = backend.lib()
    im = _model(precision)
    n = 6
    batches = [_params(im, [W.epl180_draw(10 * r + i) for i in range(n)]) for r in range(4)]
    d = im._data_struct()
    nbytes = im.workspace_bytes(n)
    ws_graph = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    params = torch.as_tensor(batches[0], device="cuda")
    logl = torch.full((n,), np.nan, dtype=torch.float64, device="cuda")
    st = torch.cuda.Stream()
    handle = ctypes.c_void_p()
    with torch.cuda.stream(st):
        rc = lib.lensim_graph_capture_loglike(ctypes.byref(im._model), ctypes.byref(im._grid),
                                              ctypes.byref(im._psf), ctypes.byref(d), backend.ptr(params), n,
                                              im._prec, backend.ptr(logl), None, backend.ptr(ws_graph), nbytes,
                                              ctypes.c_void_p(st.cuda_stream), ctypes.byref(handle))
    backend.check(rc, "lensim_graph_capture_loglike")
    assert handle.value
    st.synchronize()
    assert torch.isnan(logl).all()  # capture runs nothing
    try:
        for r, b in enumerate(batches):
            with torch.cuda.stream(st):
                params.copy_(torch.as_tensor(b, device="cuda"))  # new parameters, same buffer
                backend.check(lib.lensim_graph_launch(handle, ctypes.c_void_p(st.cuda_stream)), "lensim_graph_launch")
            st.synchronize()
            got = logl.cpu().numpy().copy()
            # eager work between the replays, then the eager call on its own workspace
            want = im.likelihood_batch(torch.as_tensor(b, device="cuda")).cpu().numpy()
            torch.cuda.synchronize()
            np.testing.assert_array_equal(got, want, err_msg="replay %d" % r)
            assert np.all(np.isfinite(got))
    finally:
        assert lib.lensim_graph_destroy(handle) == 0


def test_graph_capture_rejects_null_stream_and_bad_calls():
    lib = backend.lib()
    im = _model("fp64")
    d = im._data_struct()
    handle = ctypes.c_void_p(123)
    params = torch.zeros((1, 4, 8), dtype=torch.float64, device="cuda")
    logl = torch.empty(1, dtype=torch.float64, device="cuda")
    rc = lib.lensim_graph_capture_loglike(ctypes.byref(im._model), ctypes.byref(im._grid), ctypes.byref(im._psf),
                                          ctypes.byref(d), backend.ptr(params), 1, 0, backend.ptr(logl), None,
                                          None, 0, None, ctypes.byref(handle))
    assert rc == -1 and handle.value is None
    st = torch.cuda.Stream()
    # a workspace that is too small: the capture ends, nothing is returned, the stream stays usable
    rc = lib.lensim_graph_capture_loglike(ctypes.byref(im._model), ctypes.byref(im._grid), ctypes.byref(im._psf),
                                          ctypes.byref(d), backend.ptr(params), 1, 0, backend.ptr(logl), None,
                                          None, 0, ctypes.c_void_p(st.cuda_stream), ctypes.byref(handle))
    assert rc == -4 and handle.value is None
    with torch.cuda.stream(st):
        x = torch.ones(4, device="cuda") * 2
    st.synchronize()
    assert float(x.sum()) == 8.0
    assert lib.lensim_graph_destroy(None) == 0
