"""Request server: each request's result is the same whether computed alone or
batched with others (continuous batching, padding, graph replay), and graph
replays with eager work between them match eager launches."""

import time

import numpy as np
import pytest
import torch

from jaxtronomy_amd import workloads as W
from jaxtronomy_amd.Data.imaging_data import ImageData
from jaxtronomy_amd.Data.psf import PSF
from jaxtronomy_amd.ImSim.image_model import ImageModel
from jaxtronomy_amd.LensModel.lens_model import LensModel
from jaxtronomy_amd.LightModel.light_model import LightModel
from jaxtronomy_amd.serving import LensServer, latency_percentiles

pytestmark = pytest.mark.gpu


def _image_model(precision, n=64, s=2, k=13):
    cfg = dict(W.EPL180, numPix=n, supersampling=s)
    kernel = W.moffat_kernel(k, cfg["psf_fwhm"] / cfg["deltaPix"], cfg["psf_beta"])
    im = ImageModel(ImageData(**W.data_kwargs(cfg, np.zeros((n, n)))), PSF("PIXEL", kernel_point_source=kernel),
                    LensModel(["EPL", "SHEAR"]), LightModel(["SERSIC_ELLIPSE"]), LightModel(["SERSIC_ELLIPSE"]),
                    kwargs_numerics={"supersampling_factor": s}, precision=precision)
    return im, cfg


def _requests(im, n_req, seed):
    params = []
    for d in range(n_req):
        p, _ = im.pack(*W.epl180_draw(seed + d))
        params.append(p[0])
    return np.stack(params)


def _observations(im, cfg, n_obs, seed):
    obs = []
    for d in range(n_obs):
        p, _ = im.pack(*W.epl180_draw(1000 + seed + d))
        clean = im.image_batch(torch.as_tensor(p, device="cuda")).double().cpu().numpy()[0]
        obs.append(W.noisy(clean, cfg["exposure_time"], cfg["background_rms"], np.random.default_rng(seed + d)))
    return obs


def _alone_logl(im, p, obs):
    data = torch.as_tensor(obs[None], device="cuda").contiguous()
    return im.likelihood_batch(torch.as_tensor(p[None], device="cuda"), data=data).item()


def _alone_image(im, p):
    return im.image_batch(torch.as_tensor(p[None], device="cuda")).double().cpu()[0]


@pytest.mark.parametrize("precision,use_graphs", [("fp64", True), ("mixed", True), ("mixed", False)])
def test_mix_equals_alone(precision, use_graphs):
    im, cfg = _image_model(precision)
    srv = LensServer(n_streams=3, use_graphs=use_graphs)
    srv.register("a", im, max_batch=16, n_slots=4, n_lanes=2)
    obs = _observations(im, cfg, 3, 7)
    slots = [srv.put_data("a", o) for o in obs]
    params = _requests(im, 37, 11)
    futs = []
    # arrivals in uneven waves: the scheduler forms batches of varied sizes
    # (padded to powers of two) while earlier batches are in flight
    waves = [1, 5, 16, 3, 9, 2, 1]
    i = 0
    for w in waves:
        for _ in range(w):
            kind = "image" if i % 4 == 3 else "loglike"
            futs.append((i, kind, srv.submit("a", params[i], kind=kind, slot=slots[i % 3])))
            i += 1
        srv.step()
    assert i == len(params)
    srv.drain()
    assert srv.requests_done == len(params)
    assert len(srv.batch_sizes) > 1  # really batched in different mixes
    for i, kind, f in futs:
        assert f.done()
        if kind == "loglike":
            ref = _alone_logl(im, params[i], obs[i % 3])
            assert f.result() == ref, (i, f.result(), ref)
        else:
            ref = _alone_image(im, params[i])
            assert torch.equal(f.result().double(), ref), i
    p = latency_percentiles([f for _, _, f in futs])
    assert p[50] is not None and p[99] >= p[50]


def test_two_buckets_and_background_thread():
    im64, cfg = _image_model("fp64", n=48, s=1, k=9)
    imx, _ = _image_model("mixed", n=64, s=2, k=13)
    srv = LensServer(n_streams=4)
    srv.register("fp64", im64, max_batch=8, n_slots=2)
    srv.register("mixed", imx, max_batch=8, n_slots=2)
    o64 = _observations(im64, cfg, 1, 3)[0]
    ox = _observations(imx, cfg, 1, 4)[0]
    s64 = srv.put_data("fp64", o64)
    sx = srv.put_data("mixed", ox)
    p64 = _requests(im64, 20, 21)
    px = _requests(imx, 20, 41)
    srv.start()
    try:
        f64 = [srv.submit("fp64", p, slot=s64) for p in p64]
        fx = [srv.submit("mixed", p, slot=sx) for p in px]
        r64 = [f.result(timeout=60) for f in f64]
        rx = [f.result(timeout=60) for f in fx]
    finally:
        srv.stop()
    for p, r in zip(p64, r64):
        assert r == _alone_logl(im64, p, o64)
    for p, r in zip(px, rx):
        assert r == _alone_logl(imx, p, ox)


def test_slot_pool_and_budget():
    im, cfg = _image_model("mixed", n=32, s=1, k=5)
    srv = LensServer(n_streams=1)
    b = srv.register("a", im, max_batch=4, n_slots=2, n_lanes=1)
    z = np.zeros((32, 32))
    s0, s1 = srv.put_data("a", z), srv.put_data("a", z)
    assert {s0, s1} == {0, 1} and srv.put_data("a", z) is None
    srv.release_data("a", s0)
    assert srv.put_data("a", z) == s0
    with pytest.raises(ValueError):
        srv.submit("a", _requests(im, 1, 0)[0], kind="loglike", slot=None)
    with pytest.raises(ValueError):
        srv.put_data("a", np.zeros((31, 32)))
    with pytest.raises(MemoryError):
        LensServer(n_streams=1, hbm_budget=b.nbytes - 1).register("a", im, max_batch=4, n_slots=2, n_lanes=1)


def test_failed_batch_fails_only_its_requests():
    """A lane whose workspace is too small: the library refuses its batches
    before any device work (EWORKSPACE), their requests end with status "error"
    and the library's message, the lane and stream return to the pool, and the
    other lane's batches are served as usual."""
    im, cfg = _image_model("mixed", n=48, s=2, k=9)
    srv = LensServer(n_streams=2, use_graphs=False)
    b = srv.register("a", im, max_batch=4, n_slots=2, n_lanes=2)
    obs = _observations(im, cfg, 1, 5)[0]
    slot = srv.put_data("a", obs)
    bad = b.lanes[1]
    bad.ws = torch.empty(256, dtype=torch.uint8, device="cuda")
    params = _requests(im, 24, 60)
    # the first batch takes lane 0; while it runs, the next ones take lane 1 and fail
    futs = srv.submit_many("a", params, slots=[slot] * 24)
    srv.drain(timeout=60)
    st = [f.status for f in futs]
    assert "ok" in st and "error" in st and set(st) == {"ok", "error"}
    for p, f in zip(params, futs):
        if f.status == "ok":
            assert f.result() == _alone_logl(im, p, obs)
        else:
            with pytest.raises(RuntimeError, match="workspace"):
                f.result()
    h = srv.health()
    assert h["failed_batches"] > 0 and h["requests_failed"] == st.count("error") and h["free_streams"] == 2
    assert h["requests_done"] == st.count("ok") and "workspace" in h["last_error"]


class _HeldEvent(object):
    """A batch's completion event that reports 'not done' until released: the
    host-side stand-in for a slow batch (the GPU work itself ends as usual)."""

    def __init__(self, ev):
        self.ev = ev
        self.held = True

    def query(self):
        return (not self.held) and self.ev.query()


def test_watchdog_fails_overdue_batch_and_recovers_its_lane():
    im, cfg = _image_model("mixed", n=48, s=2, k=9)
    srv = LensServer(n_streams=1, use_graphs=False, watchdog=0.05)
    b = srv.register("a", im, max_batch=4, n_slots=1, n_lanes=1)
    obs = _observations(im, cfg, 1, 6)[0]
    slot = srv.put_data("a", obs)
    run, held = b.run, []

    def run_held(*a):
        ev, images = run(*a)
        held.append(_HeldEvent(ev))
        return held[-1], images

    b.run = run_held
    params = _requests(im, 8, 80)
    f1 = srv.submit_many("a", params[:4], slots=[slot] * 4)
    srv.step()  # launches the first batch
    f2 = srv.submit_many("a", params[4:], slots=[slot] * 4)
    time.sleep(0.1)
    srv.step()  # the watchdog fails the first batch; its lane and stream stay held
    assert [f.status for f in f1] == ["timeout"] * 4 and all(f.status == "pending" for f in f2)
    with pytest.raises(TimeoutError, match="still running"):
        f1[0].result()
    h = srv.health()
    assert h["watchdog_trips"] == 1 and h["overdue_batches"] == 1 and h["free_streams"] == 0
    held[0].ev.synchronize()
    held[0].held = False  # the batch finishes: lane and stream come back, the queue moves on
    b.run = run
    srv.drain(timeout=60)
    assert [f.status for f in f2] == ["ok"] * 4
    for p, f in zip(params[4:], f2):
        assert f.result() == _alone_logl(im, p, obs)
    h = srv.health()
    assert h["overdue_batches"] == 0 and h["free_streams"] == 1 and h["requests_done"] == 4


def test_scheduler_error_fails_outstanding_requests():
    im, cfg = _image_model("mixed", n=32, s=1, k=5)
    srv = LensServer(n_streams=1, use_graphs=False)
    srv.register("a", im, max_batch=4, n_slots=1, n_lanes=1)
    slot = srv.put_data("a", np.zeros((32, 32)))

    def broken():
        raise RuntimeError("scheduler failure")

    srv._dispatch = broken
    futs = srv.submit_many("a", _requests(im, 6, 90), slots=[slot] * 6)
    srv.start()
    for f in futs:
        with pytest.raises(RuntimeError, match="scheduler failure"):
            f.result(timeout=30)
    srv.stop()
    assert srv.health()["requests_failed"] == 6 and srv.idle()


def test_graph_replays_with_eager_work_between():
    """Every replay of a captured batch matches the eager launch, with unrelated
    eager device work and synchronisations between the replays."""
    im, cfg = _image_model("mixed")
    obs = _observations(im, cfg, 2, 5)
    params = _requests(im, 8, 77)
    eager = LensServer(n_streams=1, use_graphs=False)
    eager.register("a", im, max_batch=8, n_slots=2, n_lanes=1)
    graph = LensServer(n_streams=1, use_graphs=True)
    graph.register("a", im, max_batch=8, n_slots=2, n_lanes=1)
    for srv in (eager, graph):
        for o in obs:
            srv.put_data("a", o)
    ref = [eager.submit("a", p, slot=j % 2) for j, p in enumerate(params)]
    eager.drain()
    ref = [f.result() for f in ref]
    scratch = torch.randn(1 << 20, device="cuda")
    for rep in range(4):
        futs = [graph.submit("a", p, slot=j % 2) for j, p in enumerate(params)]
        graph.drain()
        assert [f.result() for f in futs] == ref, rep
        scratch = (scratch * 1.0001 + 1.0).sum() + scratch  # eager work between replays
        torch.cuda.synchronize()
    assert len(graph.buckets["a"].lanes[0].graphs) == 1


def test_noise_map_slots():
    """Observations with their own noise maps: server results equal the same
    requests evaluated alone, and the noise-map variance is really used."""
    im, cfg = _image_model("mixed", n=48, s=1, k=9)
    srv = LensServer(n_streams=2)
    srv.register("n", im, max_batch=8, n_slots=3, noise_maps=True)
    obs = _observations(im, cfg, 3, 13)
    rng = np.random.default_rng(3)
    noise = [np.full((48, 48), 0.01) + 0.005 * rng.random((48, 48)) for _ in obs]
    slots = [srv.put_data("n", o, noise_map=nm) for o, nm in zip(obs, noise)]
    with pytest.raises(ValueError):
        srv.put_data("n", obs[0])
    params = _requests(im, 11, 5)
    futs = srv.submit_many("n", params, "loglike", [slots[i % 3] for i in range(11)])
    srv.drain()
    for i, f in enumerate(futs):
        d = torch.as_tensor(obs[i % 3][None], device="cuda").contiguous()
        nm = torch.as_tensor(noise[i % 3][None], device="cuda").contiguous()
        alone = im.likelihood_batch(torch.as_tensor(params[i][None], device="cuda"), data=d, noise_map=nm).item()
        assert f.result() == alone
        plain = im.likelihood_batch(torch.as_tensor(params[i][None], device="cuda"), data=d).item()
        assert alone != plain


@pytest.mark.parametrize("threaded", [True, False])
def test_poisson_load_latency(threaded):
    """Both loop forms: the scheduler in its own thread, and the load loop
    driving the scheduler itself (one event loop)."""
    from jaxtronomy_amd.serving import poisson_load

    im, cfg = _image_model("mixed", n=64, s=2, k=13)
    srv = LensServer(n_streams=4)
    srv.register("a", im, max_batch=32, n_slots=2, n_lanes=3)
    obs = _observations(im, cfg, 1, 2)[0]
    slot = srv.put_data("a", obs)
    params = _requests(im, 16, 3)
    srv.warmup("a", params[0])
    if threaded:
        srv.start()
    try:
        futs, wall = poisson_load(srv, "a", params, rate=5000.0, duration=0.3, slots=[slot] * 16)
    finally:
        srv.stop()
    assert len(futs) == 1500 and all(f.done() for f in futs)
    p = latency_percentiles(futs)
    assert 0 < p[50] <= p[99] < 1e4
    assert srv.batches < len(futs)  # requests were batched
    # the results equal the same requests evaluated alone
    for i in range(16):
        assert futs[i].value == _alone_logl(im, params[i], obs)


def test_serving_mix_buckets_equal_alone():
    """Poisson arrivals over buckets of different grids, supersampling factors,
    deflector sets and PSF sizes (log-likelihood with resident noise maps, and
    simulate requests): every result equals the same request alone."""
    from jaxtronomy_amd.serving import poisson_load_mix

    m = W.SERVE_MIX
    combos = [(64, 3, ("SIE", "SHEAR"), 11), (100, 1, ("NFW", "SIE"), 21), (48, 5, ("EPL", "SHEAR"), 31)]
    srv = LensServer(n_streams=4)
    streams, ref = [], {}
    for bi, (n, s, defl, k) in enumerate(combos):
        cfg = dict(numPix=n, deltaPix=m["deltaPix"], supersampling=s, exposure_time=m["exposure_time"],
                   background_rms=m["background_rms"])
        kernel = W.moffat_kernel(k, m["psf_fwhm"] / m["deltaPix"], m["psf_beta"])
        im = ImageModel(ImageData(**W.data_kwargs(cfg, np.zeros((n, n)))), PSF("PIXEL", kernel_point_source=kernel),
                        LensModel(list(defl)), LightModel(["SERSIC_ELLIPSE"]), LightModel(["SERSIC_ELLIPSE"]),
                        kwargs_numerics={"supersampling_factor": s}, precision="mixed")
        key = "b%d" % bi
        srv.register(key, im, max_batch=8, n_slots=3, n_lanes=2, noise_maps=True)
        rows = np.stack([im.pack(*W.mix_draw(defl, 10 * bi + d))[0][0] for d in range(6)])
        clean = im.image_batch(torch.as_tensor(rows[:3], device="cuda")).double().cpu().numpy()
        rng = np.random.default_rng(bi)
        obs, noise, slots = [], [], []
        for o in range(3):
            obs.append(W.noisy(clean[o], cfg["exposure_time"], cfg["background_rms"], rng))
            noise.append(np.sqrt(cfg["background_rms"] ** 2 + np.maximum(clean[o], 0) / cfg["exposure_time"]))
            slots.append(srv.put_data(key, obs[-1], noise[-1]))
        srv.warmup(key, rows[0], kinds=("loglike", "image"))
        streams.append(dict(key=key, params=rows, slots=np.array(slots)[np.arange(6) % 3], kind="loglike"))
        streams.append(dict(key=key, params=rows, kind="image", weight=0.5))
        ref[key] = (im, rows, obs, noise)
    futs, wall, peak, (which, row) = poisson_load_mix(srv, streams, rate=3000.0, duration=0.2, seed=5)
    srv.stop()
    assert len(futs) == 600 and all(f.done() for f in futs) and peak >= 1
    assert len(srv.batch_sizes) > 1 and len(set(which.tolist())) == len(streams)
    for f, st, r in zip(futs, which, row):
        s = streams[st]
        im, rows, obs, noise = ref[s["key"]]
        p = torch.as_tensor(rows[r][None], device="cuda")
        if s["kind"] == "loglike":
            o = r % 3  # slots were filled in order 0, 1, 2 and cycle with the rows
            want = im.likelihood_batch(p, data=torch.as_tensor(obs[o][None], device="cuda"),
                                       noise_map=torch.as_tensor(noise[o][None], device="cuda")).item()
            assert f.value == want
        else:
            assert torch.equal(f.value, im.image_batch(p).cpu()[0])


def test_serving_at_stated_scale():
    """The serving configuration at its stated scale: buckets of 240^2 at
    supersampling 5 and 180^2 at supersampling 3 with 41x41 PSFs beside
    100^2 / 64^2 buckets with 31x31 / 11x11 PSFs and the three deflector sets,
    resident observations and noise maps, and Poisson arrivals fast enough that
    more than 2000 requests are in flight at once.  A sample of every stream's
    results equals the same requests run alone, bit for bit, and one
    log-likelihood request per bucket equals the oracle's (1e-10 in the FP64
    bucket, 1e-6 relative in the mixed ones)."""
    from jaxtronomy_amd import profiles
    from jaxtronomy_amd.serving import poisson_load_mix
    from oracle import lensim_oracle as O

    m = W.SERVE_MIX
    combos = [("big", 240, 5, ("EPL", "SHEAR"), 41, "mixed"), ("c2", 180, 3, ("SIE", "SHEAR"), 41, "mixed"),
              ("mid", 100, 3, ("NFW", "SIE"), 31, "mixed"), ("small", 64, 1, ("SIE", "SHEAR"), 11, "mixed"),
              ("big64", 240, 5, ("EPL", "SHEAR"), 41, "fp64")]
    srv = LensServer(n_streams=4)
    streams, ref = [], {}
    for bi, (key, n, s, defl, k, precision) in enumerate(combos):
        cfg = dict(numPix=n, deltaPix=m["deltaPix"], supersampling=s, exposure_time=m["exposure_time"],
                   background_rms=m["background_rms"])
        kernel = W.moffat_kernel(k, m["psf_fwhm"] / m["deltaPix"], m["psf_beta"])
        im = ImageModel(ImageData(**W.data_kwargs(cfg, np.zeros((n, n)))), PSF("PIXEL", kernel_point_source=kernel),
                        LensModel(list(defl)), LightModel(["SERSIC_ELLIPSE"]), LightModel(["SERSIC_ELLIPSE"]),
                        kwargs_numerics={"supersampling_factor": s}, precision=precision)
        srv.register(key, im, max_batch=32, n_slots=4, n_lanes=2, noise_maps=True)
        rows = np.stack([im.pack(*W.mix_draw(defl, 20 * bi + d))[0][0] for d in range(8)])
        clean = im.image_batch(torch.as_tensor(rows[:4], device="cuda")).double().cpu().numpy()
        rng = np.random.default_rng(100 + bi)
        obs, noise, slots = [], [], []
        for o in range(4):
            obs.append(W.noisy(clean[o], cfg["exposure_time"], cfg["background_rms"], rng))
            noise.append(np.sqrt(cfg["background_rms"] ** 2 + np.maximum(clean[o], 0) / cfg["exposure_time"]))
            slots.append(srv.put_data(key, obs[-1], noise[-1]))
        srv.warmup(key, rows[0], kinds=("loglike", "image"))
        w = 0.1 if precision == "fp64" else 1.0
        streams.append(dict(key=key, params=rows, slots=np.array(slots)[np.arange(8) % 4], kind="loglike",
                            weight=0.8 * w))
        streams.append(dict(key=key, params=rows, kind="image", weight=0.2 * w))
        ref[key] = (im, rows, obs, noise, cfg, kernel, defl, precision)
    futs, wall, peak, (which, row) = poisson_load_mix(srv, streams, rate=2.0e6, duration=0.0025, seed=11)
    srv.stop()
    assert len(futs) == 5000 and all(f.done() for f in futs)
    assert peak >= 2000, peak
    assert len(srv.batch_sizes) > 1
    # a sample of every stream: equal to the request alone
    rng = np.random.default_rng(0)
    for st in range(len(streams)):
        idx = np.nonzero(which == st)[0]
        assert idx.size > 0
        for q in rng.choice(idx, size=min(6, idx.size), replace=False):
            s = streams[st]
            im, rows, obs, noise = ref[s["key"]][:4]
            p = torch.as_tensor(rows[row[q]][None], device="cuda")
            if s["kind"] == "loglike":
                o = row[q] % 4
                want = im.likelihood_batch(p, data=torch.as_tensor(obs[o][None], device="cuda"),
                                           noise_map=torch.as_tensor(noise[o][None], device="cuda")).item()
                assert futs[q].value == want, (s["key"], q)
            else:
                assert torch.equal(futs[q].value, im.image_batch(p).cpu()[0]), (s["key"], q)
    # one log-likelihood request per bucket against the oracle
    for st in range(0, len(streams), 2):
        s = streams[st]
        im, rows, obs, noise, cfg, kernel, defl, precision = ref[s["key"]]
        q = int(np.nonzero(which == st)[0][0])
        r = int(row[q])
        kw = [{nm: float(rows[r][j][i]) for i, nm in profiles.kwarg_slots(im._names[j])}
              for j in range(len(im._names))]
        nd = len(defl)
        want_img = O.image(W.grid_dict(cfg), list(defl), kw[:nd], ["SERSIC_ELLIPSE"], kw[nd:nd + 1],
                           ["SERSIC_ELLIPSE"], kw[nd + 1:], kernel=kernel)
        want = O.log_likelihood(want_img, obs[r % 4], noise_map=noise[r % 4])
        tol = 1e-10 if precision == "fp64" else 1e-6
        assert abs(futs[q].value - want) < tol * abs(want), (s["key"], futs[q].value, want)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_sub_frame_bucket(precision):
    """A model with flux_evaluate_indexes: each observation is cut to the square
    around the evaluated pixels when it is put (with the zero model's logL on
    the pixels outside), images come back completed to the frame.  logL equals
    likelihood_batch on the whole frame to rounding, images are bit-equal."""
    n = 48
    cfg = dict(W.EPL180, numPix=n, supersampling=2)
    kernel = W.moffat_kernel(9, cfg["psf_fwhm"] / cfg["deltaPix"], cfg["psf_beta"])
    ev = np.zeros((n, n), bool)
    yy, xx = np.mgrid[0:n, 0:n]
    ev[(yy - 22) ** 2 + (xx - 25) ** 2 < 15 ** 2] = True
    lmask = np.ones((n, n))
    lmask[:, :12] = 0
    im = ImageModel(ImageData(**W.data_kwargs(cfg, np.zeros((n, n)))), PSF("PIXEL", kernel_point_source=kernel),
                    LensModel(["EPL", "SHEAR"]), LightModel(["SERSIC_ELLIPSE"]), LightModel(["SERSIC_ELLIPSE"]),
                    kwargs_numerics={"supersampling_factor": 2, "flux_evaluate_indexes": ev},
                    likelihood_mask=lmask, precision=precision)
    assert im._frame is not None and im._frame[2] < n
    srv = LensServer(n_streams=2, use_graphs=True)
    srv.register("sf", im, max_batch=8, n_slots=3, n_lanes=2)
    obs = _observations(im, cfg, 3, 5)
    slots = [srv.put_data("sf", o) for o in obs]
    params = _requests(im, 13, 31)
    futs = []
    for i in range(len(params)):
        kind = "image" if i % 4 == 1 else "loglike"
        futs.append((i, kind, srv.submit("sf", params[i], kind=kind, slot=slots[i % 3])))
        if i % 5 == 4:
            srv.step()
    srv.drain()
    for i, kind, f in futs:
        if kind == "loglike":
            ref = _alone_logl(im, params[i], obs[i % 3])
            assert abs(f.result() - ref) <= 1e-12 * abs(ref), (i, f.result(), ref)
        else:
            got = f.result().double()
            assert tuple(got.shape) == (n, n)
            assert torch.equal(got, _alone_image(im, params[i])), i


@pytest.mark.parametrize("multi_plane", [False, True])
def test_bulk_deflector_bucket(multi_plane):
    """An ImageModel with a LensModelBulk deflector (EPL + SHEAR inline, 30
    TNFW in the bulk loop; on one plane or on their own planes): the bucket's
    deflector structure is fixed at registration, requests come in its
    pack_bulk layout, and every result equals the same request alone."""
    from jaxtronomy_amd.LensModel.lens_model_bulk import LensModelBulk
    from jaxtronomy_amd.Util.cosmo import FlatLambdaCDM

    n = 48
    cfg = dict(W.SUBSTRUCTURE, numPix=n, deltaPix=0.08, supersampling=2)
    kernel = W.moffat_kernel(9, 1.5, 3.0)
    kl, subs, ks = W.substructure_draw(2, n_sub=30)
    names = ["EPL", "SHEAR"] + ["TNFW"] * 30
    lm = LensModelBulk(["EPL", "SHEAR", "TNFW"], multi_plane=multi_plane, cosmo=FlatLambdaCDM(70.0, 0.3),
                       precision="mixed")
    im = ImageModel(ImageData(**W.data_kwargs(cfg, np.zeros((n, n)))), PSF("PIXEL", kernel_point_source=kernel), lm,
                    LightModel(["SERSIC_ELLIPSE"]), kwargs_numerics={"supersampling_factor": 2}, precision="mixed")
    if multi_plane:
        z = [0.5, 0.5] + list(np.random.default_rng(8).uniform(0.2, 1.3, 30))
        rs = lm.prepare_ray_shooting_kwargs(names, kl + subs, 1.5, z)
    else:
        rs = lm.prepare_ray_shooting_kwargs(names, kl + subs)
    srv = LensServer(n_streams=2, use_graphs=True)
    with pytest.raises(ValueError):
        srv.register("no", im)
    b = srv.register("bulk", im, max_batch=8, n_slots=2, n_lanes=2, bulk=(rs, ks, None))
    base, _, _, _ = im.pack_bulk(rs, ks)
    rng = np.random.default_rng(4)
    params = np.repeat(base, 9, axis=0)
    params[:, :, :5] *= 1 + 0.01 * rng.normal(size=params[:, :, :5].shape)
    obs = [np.random.default_rng(s).normal(0, 0.05, (n, n)) for s in (1, 2)]
    slots = [srv.put_data("bulk", o) for o in obs]
    futs = []
    for i in range(len(params)):
        kind = "image" if i % 3 == 1 else "loglike"
        futs.append((i, kind, srv.submit("bulk", params[i], kind=kind, slot=slots[i % 2])))
    srv.drain()
    for i, kind, f in futs:
        p = torch.as_tensor(params[i][None], device="cuda")
        if kind == "loglike":
            data = torch.as_tensor(obs[i % 2][None], device="cuda").contiguous()
            assert f.result() == im.likelihood_batch(p, model=b.model, data=data).item(), i
        else:
            ref = im.image_batch(p, model=b.model).double().cpu()[0]
            assert torch.equal(f.result().double(), ref), i
