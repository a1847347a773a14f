"""Request server: continuous batching of simulate / log-likelihood requests.

A *bucket* is one request class: an ImageModel configuration (grid,
supersampling, PSF, profile lists, precision).  It owns

* a pool of resident data slots: a log-likelihood request names the slot that
  holds its observation (and, for buckets registered with ``noise_maps``, its
  noise map), which stays on the device between requests;
* ``n_lanes`` lanes, each a full set of static device buffers for up to
  ``max_batch`` parameter sets (parameters, slot indices, gathered
  observations, workspace, results, pinned host staging) and one captured HIP
  graph per (request kind, padded batch size): the gather of the batch's
  observations plus the library's launch sequence, replayed for every batch.

The scheduler keeps up to ``n_streams`` batches in flight.  Whenever a stream
and a lane of the oldest waiting bucket are free, it admits every request of
that bucket that has arrived so far (up to ``max_batch``), so requests that
arrive while other batches run join the next batch, and batches of different
buckets (or lanes of one bucket) overlap on the device.  Results are copied to
pinned host memory on the lane's stream before its completion event (log-
likelihoods to the lane's staging, images to a pinned buffer of the batch
that the results view), so retiring a batch reads host memory only.  A
captured batch is launched with hipGraphLaunch on its stream directly.

Each image of a batch is computed from its own parameters (and observation)
alone, so a request's result does not depend on the requests it was batched
with; batches are padded to a power of two with copies of their first request
so that a handful of graphs covers every batch size.

A bucket of an ImageModel with a LensModelBulk deflector fixes its deflector
structure at registration (``register(..., bulk=...)``); its requests are
parameter arrays in that structure's ImageModel.pack_bulk layout.

A model with a ``flux_evaluate_indexes`` sub-frame runs on its square: an
observation is cut to it when it is put, together with the log-likelihood of
a zero model on the pixels outside (which no parameter changes), added to each
request's log-likelihood; images are completed to the frame (zero outside).

Admission: every bucket's device memory is charged against an HBM budget at
registration; ``put_data`` returns None while all of a bucket's slots are in
use.  All device work is on streams of the default priority.

Failures are per request: a batch whose launch fails (a library error, a
failed graph launch) fails its own requests -- their Futures carry the
exception, ``result()`` raises it and ``status`` is "error" -- and the lane and
stream go back to the pool.  A watchdog fails the requests of a batch still
running ``watchdog`` seconds after its launch (status "timeout", result()
raises TimeoutError) and keeps its lane and stream out of use until the batch
does finish, then returns them; ``health()`` reports the counts.  An error in
the scheduler itself fails every outstanding request and stops the
background thread.
"""

import collections
import ctypes
import threading
import time

import numpy as np
import torch

from . import backend

_HIP = None


def _hip():
    """The HIP runtime already loaded by torch: graph launches and async copies
    straight from the scheduler, without torch's stream context per batch."""
    global _HIP
    if _HIP is None:
        h = ctypes.CDLL("libamdhip64.so")
        h.hipGraphLaunch.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        h.hipGraphLaunch.restype = ctypes.c_int
        h.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int,
                                     ctypes.c_void_p]
        h.hipMemcpyAsync.restype = ctypes.c_int
        _HIP = h
    return _HIP


_D2H = 2  # hipMemcpyDeviceToHost


class Future(object):
    """Result handle of one request (set by the scheduler thread)."""

    __slots__ = ("value", "error", "t_submit", "t_done", "_cond")

    def __init__(self, cond, t_submit):
        self.value = None
        self.error = None
        self.t_submit = t_submit
        self.t_done = None
        self._cond = cond

    def done(self):
        return self.t_done is not None

    @property
    def status(self):
        """"pending", "ok", "error" (the batch failed: see ``error``) or
        "timeout" (the watchdog failed it)."""
        if self.t_done is None:
            return "pending"
        if self.error is None:
            return "ok"
        return "timeout" if isinstance(self.error, TimeoutError) else "error"

    def result(self, timeout=None):
        """The value; raises the batch's exception for a failed request."""
        if self.t_done is None:
            with self._cond:
                if not self._cond.wait_for(self.done, timeout):
                    raise TimeoutError("request not completed")
        if self.error is not None:
            raise self.error
        return self.value

    @property
    def latency(self):
        """Seconds from arrival (submit time unless given) to completion."""
        return None if self.t_done is None else self.t_done - self.t_submit


def _pow2_at_least(n, cap):
    m = 1
    while m < n:
        m *= 2
    return min(m, cap)


class _Lane(object):
    """One set of static buffers of a bucket (and the graphs captured on them)."""

    def __init__(self, b):
        dev = backend.device()
        ny, nx = b.shape
        mb = b.max_batch
        self.busy = False
        self.params = torch.zeros((mb, b.n_prof, backend.NPARAM), dtype=torch.float64, device=dev)
        self.params_host = torch.zeros((mb, b.n_prof, backend.NPARAM), dtype=torch.float64).pin_memory()
        self.params_np = self.params_host.numpy()
        self.slot_idx = torch.zeros(mb, dtype=torch.long, device=dev)
        self.slot_host = torch.zeros(mb, dtype=torch.long).pin_memory()
        self.slot_np = self.slot_host.numpy()
        self.batch_data = torch.zeros((mb, ny, nx), dtype=torch.float64, device=dev)
        self.batch_noise = torch.zeros((mb, ny, nx), dtype=torch.float64, device=dev) if b.noise_maps else None
        self.logl = torch.zeros(mb, dtype=torch.float64, device=dev)
        self.logl_host = torch.zeros(mb, dtype=torch.float64).pin_memory()
        self.images = torch.zeros((mb, ny, nx), dtype=b.dtype, device=dev)
        # a sub-framed model's images completed to the frame (zero outside the square, written once)
        self.full_images = (torch.zeros((mb,) + b.full_shape, dtype=b.dtype, device=dev) if b.frame is not None
                            else self.images)
        self.ws = torch.empty(b.ws_bytes, dtype=torch.uint8, device=dev)
        self.data = backend.LensimData()
        ctypes.memmove(ctypes.byref(self.data), ctypes.byref(b.data_template), ctypes.sizeof(self.data))
        self.data.d_data = self.batch_data.data_ptr()
        if self.batch_noise is not None:
            self.data.d_noise_map = self.batch_noise.data_ptr()
        self.data.image_stride = nx * ny
        self.graphs = {}
        self.nbytes = sum(t.numel() * t.element_size() for t in (self.params, self.slot_idx, self.batch_data,
                                                                  self.logl, self.images, self.ws))
        if b.frame is not None:
            self.nbytes += self.full_images.numel() * self.full_images.element_size()
        if self.batch_noise is not None:
            self.nbytes += self.batch_noise.numel() * 8


class _Bucket(object):
    def __init__(self, key, image_model, max_batch, n_slots, n_lanes, use_graphs, noise_maps, bulk=None):
        self.key = key
        self.im = image_model
        # the lensim_model of the bucket's requests: the ImageModel's own, or for a LensModelBulk deflector the
        # one of ImageModel.pack_bulk for the bucket's deflector structure (requests in that packed layout)
        self.model, self._bulk_keep = image_model._model, None
        if image_model._bulk:
            if bulk is None:
                raise ValueError("a LensModelBulk bucket needs bulk=(ray_shooting_kwargs, kwargs_source, "
                                 "kwargs_lens_light) fixing its deflector structure")
            params, _, self.model, self._bulk_keep = image_model.pack_bulk(*bulk)
            self.n_prof = params.shape[1]
        elif bulk is not None:
            raise ValueError("bulk is for ImageModels with a LensModelBulk deflector")
        else:
            self.n_prof = len(image_model._names)
        self.max_batch = int(max_batch)
        if not 1 <= self.max_batch <= 65535:
            raise ValueError("max_batch must be in [1, 65535]")
        self.use_graphs = use_graphs
        self.noise_maps = bool(noise_maps)
        # a sub-framed model (flux_evaluate_indexes) runs on its square: observations are cut to it when they
        # are put, each with the logL of a zero model outside the square, and images are completed to the frame
        self.shape = (image_model._grid.ny, image_model._grid.nx)
        self.full_shape = tuple(image_model._full_shape)
        self.frame = image_model._frame
        self.dtype = torch.float64 if image_model._prec == backend.FP64 else torch.float32
        dev = backend.device()
        buf = image_model.Data.device_buffers()
        if buf["noise"] is not None and not self.noise_maps:
            raise ValueError("the ImageData carries a noise map: register the bucket with noise_maps=True and "
                             "give each observation its noise map")
        self.outside = None
        if self.frame is None:
            self.mask = backend.to_device_f64(image_model.likelihood_mask.astype(np.float64))
            self.expo_map = buf["expo_map"]
        else:  # the likelihood mask and exposure map cut to the square
            sub, _, self._sub_keep = image_model._sub_frame_data()
            x0, y0, n = self.frame
            self.mask = self._sub_keep[3]
            self.expo_map = self._sub_keep[2]
            self.outside = torch.zeros(int(n_slots), dtype=torch.float64, device=dev)
        d = backend.LensimData()
        d.d_mask = self.mask.data_ptr()
        d.d_exposure_map = None if self.expo_map is None else self.expo_map.data_ptr()
        d.exposure_time = buf["expo_time"]
        d.background_rms = image_model.Data.background_rms
        self.data_template = d
        self.ws_bytes = max(image_model.workspace_bytes(self.max_batch, self.model), 256)
        self.slots = torch.zeros((int(n_slots),) + self.shape, dtype=torch.float64, device=dev)
        self.noise_slots = torch.ones_like(self.slots) if self.noise_maps else None
        self.free_slots = list(range(int(n_slots)))[::-1]
        self.lanes = [_Lane(self) for _ in range(int(n_lanes))]
        self.nbytes = self.slots.numel() * 8 * (2 if self.noise_maps else 1) + sum(l.nbytes for l in self.lanes)

    def free_lane(self):
        for lane in self.lanes:
            if not lane.busy:
                return lane
        return None

    def _launch(self, lane, kind, n):
        """The launch sequence of one batch on the current stream."""
        im = self.im
        lib = backend.lib()
        st = backend.stream_handle()
        if kind == "loglike":
            idx = lane.slot_idx[:n]
            torch.index_select(self.slots, 0, idx, out=lane.batch_data[:n])
            if self.noise_maps:
                torch.index_select(self.noise_slots, 0, idx, out=lane.batch_noise[:n])
            rc = lib.lensim_loglike(ctypes.byref(self.model), ctypes.byref(im._grid), ctypes.byref(im._psf),
                                    ctypes.byref(lane.data), backend.ptr(lane.params), n, im._prec,
                                    backend.ptr(lane.logl), None, backend.ptr(lane.ws), lane.ws.numel(), st)
            backend.check(rc, "lensim_loglike")
            if self.outside is not None:  # the observation's pixels outside the square
                lane.logl[:n] += torch.index_select(self.outside, 0, idx)
        else:
            rc = lib.lensim_image(ctypes.byref(self.model), ctypes.byref(im._grid), ctypes.byref(im._psf),
                                  backend.ptr(lane.params), n, im._prec,
                                  backend.IMAGE_SOURCE | backend.IMAGE_LENS_LIGHT, backend.ptr(lane.images),
                                  backend.ptr(lane.ws), lane.ws.numel(), st)
            backend.check(rc, "lensim_image")
            if self.frame is not None:
                x0, y0, m = self.frame
                lane.full_images[:n, x0:x0 + m, y0:y0 + m] = lane.images[:n]

    def _batch(self, lane, kind, npad):
        """One batch of npad rows: staging in, launches and, for log-likelihoods,
        the results out (the pinned host buffers are fixed, so a captured graph
        holds those copies too).  Images go out per batch (``run``)."""
        lane.params[:npad].copy_(lane.params_host[:npad], non_blocking=True)
        lane.slot_idx[:npad].copy_(lane.slot_host[:npad], non_blocking=True)
        self._launch(lane, kind, npad)
        if kind == "loglike":
            lane.logl_host[:npad].copy_(lane.logl[:npad], non_blocking=True)

    def run(self, lane, kind, n, stream):
        """Launch a batch whose first n rows of the lane's host staging are filled;
        returns (completion event, pinned host images [n, ny, nx] or None).  A
        captured batch is one graph launch on the stream; simulate results then
        go to a pinned buffer of their own, which the results view, so the lane
        can take its next batch at once."""
        npad = _pow2_at_least(n, self.max_batch)
        if npad > n:  # padding rows: copies of the first request, results ignored
            lane.params_np[n:npad] = lane.params_np[0]
            lane.slot_np[n:npad] = lane.slot_np[0]
        ent = lane.graphs.get((kind, npad)) if self.use_graphs else None
        if ent is not None:
            rc = _hip().hipGraphLaunch(ctypes.c_void_p(ent[1]), ctypes.c_void_p(stream.cuda_stream))
            if rc != 0:
                raise RuntimeError("hipGraphLaunch failed (%d)" % rc)
        else:
            with torch.cuda.stream(stream):
                self._batch(lane, kind, npad)  # eager, or the warm-up launch before a capture
                if self.use_graphs:
                    stream.synchronize()
                    g = torch.cuda.CUDAGraph()
                    with torch.cuda.graph(g, stream=stream):
                        self._batch(lane, kind, npad)
                    lane.graphs[(kind, npad)] = (g, g.raw_cuda_graph_exec())
        images = None
        if kind == "image":
            # padded size: few distinct sizes keep torch's pinned-memory cache warm
            images = torch.empty((npad,) + self.full_shape, dtype=self.dtype, pin_memory=True)[:n]
            rc = _hip().hipMemcpyAsync(ctypes.c_void_p(images.data_ptr()),
                                       ctypes.c_void_p(lane.full_images.data_ptr()),
                                       images.numel() * images.element_size(), _D2H,
                                       ctypes.c_void_p(stream.cuda_stream))
            if rc != 0:
                raise RuntimeError("hipMemcpyAsync failed (%d)" % rc)
        ev = torch.cuda.Event()
        ev.record(stream)
        return ev, images


class _Block(object):
    """Queued requests of one kind for one bucket (consumed from the front)."""

    __slots__ = ("kind", "params", "slots", "futures", "pos")

    def __init__(self, kind, params, slots, futures):
        self.kind = kind
        self.params = params
        self.slots = slots
        self.futures = futures
        self.pos = 0


class LensServer(object):
    """Continuous-batching server over registered request classes (buckets).

    ``register(key, image_model)`` sets up a bucket; ``put_data`` makes an
    observation resident; ``submit`` / ``submit_many`` queue requests and return
    Futures; ``step()`` / ``drain()`` drive the scheduler from the caller's
    thread and ``start()`` / ``stop()`` from a background thread.
    """

    def __init__(self, n_streams=4, hbm_budget=64 << 30, use_graphs=True, watchdog=30.0):
        backend.require_gpu()
        backend.timing_enable(False)  # event timing synchronises; it cannot sit in a graph
        self.streams = [torch.cuda.Stream() for _ in range(int(n_streams))]
        self.free_streams = list(range(int(n_streams)))
        self.hbm_budget = int(hbm_budget)
        self.hbm_used = 0
        self.use_graphs = use_graphs
        self.buckets = {}
        self.pending = collections.OrderedDict()  # (key, kind) -> deque of _Block
        self.inflight = []
        self.lock = threading.Lock()
        self.cond = threading.Condition()
        self.batches = 0
        self.requests_done = 0
        self.batch_sizes = collections.Counter()
        self.watchdog = float(watchdog)  # seconds a batch may run before its requests fail
        self.overdue = []  # batches the watchdog failed, holding their lane and stream until they finish
        self.failed_batches = 0
        self.watchdog_trips = 0
        self.requests_failed = 0
        self.last_error = None
        self._thread = None
        self._stop = False

    # ---------------------------------------------------------- admission --
    def register(self, key, image_model, max_batch=64, n_slots=16, n_lanes=2, noise_maps=False, bulk=None):
        """Register a bucket.  For an ImageModel with a LensModelBulk deflector,
        ``bulk`` = (ray_shooting_kwargs, kwargs_source, kwargs_lens_light) fixes
        the deflector structure (types, count, planes); its requests are then
        [n_profiles, 8] parameter arrays in that structure's ImageModel.pack_bulk
        layout."""
        if key in self.buckets:
            raise ValueError("bucket %r is already registered" % (key,))
        b = _Bucket(key, image_model, max_batch, n_slots, n_lanes, self.use_graphs, noise_maps, bulk)
        if self.hbm_used + b.nbytes > self.hbm_budget:
            raise MemoryError("bucket %r needs %d B; %d of %d B already used" % (key, b.nbytes, self.hbm_used,
                                                                              self.hbm_budget))
        self.hbm_used += b.nbytes
        self.buckets[key] = b
        return b

    def warmup(self, key, example_params, kinds=("loglike",)):
        """Capture every lane's graphs (each request kind, each padded batch size)
        ahead of load, with ``example_params`` [n_profiles, 8] as the
        parameters, so that no capture stalls a batch under load."""
        b = self.buckets[key]
        ex = np.asarray(example_params, dtype=np.float64).reshape(b.n_prof, backend.NPARAM)
        with self.lock:
            if any(l.busy for l in b.lanes) or any(self.pending.get((key, k)) for k in ("loglike", "image")):
                raise RuntimeError("warm up a bucket before submitting requests to it")
        sizes = []
        n = 1
        while n < b.max_batch:
            sizes.append(n)
            n *= 2
        sizes.append(b.max_batch)
        stream = self.streams[0]
        for lane in b.lanes:
            lane.params_np[:] = ex
            lane.slot_np[:] = 0
            for kind in kinds:
                for n in sizes:
                    b.run(lane, kind, n, stream)  # capture
                    b.run(lane, kind, n, stream)  # first replay (uploads the graph)
        stream.synchronize()

    def put_data(self, key, image, noise_map=None):
        """Make an observation (and its noise map, for buckets registered with
        noise_maps) resident; returns its slot id, or None when all of the
        bucket's slots are in use."""
        b = self.buckets[key]
        img = np.asarray(image, dtype=np.float64)
        if img.shape != b.full_shape:
            raise ValueError("observation of shape %s for a %s bucket" % (img.shape, b.full_shape))
        if b.noise_maps:
            if noise_map is None:
                raise ValueError("this bucket needs a noise map with each observation")
            noise_map = np.asarray(noise_map, dtype=np.float64)
            if noise_map.shape != b.full_shape:
                raise ValueError("noise map of shape %s for a %s bucket" % (noise_map.shape, b.full_shape))
        elif noise_map is not None:
            raise ValueError("bucket %r was registered without noise maps" % (key,))
        with self.lock:
            if not b.free_slots:
                return None
            slot = b.free_slots.pop()
        if b.frame is None:
            b.slots[slot].copy_(torch.from_numpy(np.ascontiguousarray(img)))
            if b.noise_maps:
                b.noise_slots[slot].copy_(torch.from_numpy(np.ascontiguousarray(noise_map)))
        else:  # cut to the square, with the zero model's logL outside it
            dev = backend.device()
            obs = torch.as_tensor(np.ascontiguousarray(img[None]), device=dev)
            nm = None if noise_map is None else torch.as_tensor(np.ascontiguousarray(noise_map[None]), device=dev)
            _, out_logl, keep = b.im._sub_frame_data(obs, nm, 1)
            b.slots[slot].copy_(keep[0][0])
            if b.noise_maps:
                b.noise_slots[slot].copy_(keep[1][0])
            b.outside[slot:slot + 1].copy_(out_logl)
        torch.cuda.current_stream().synchronize()
        return slot

    def release_data(self, key, slot):
        """Return a slot to the pool; requests naming it must have completed."""
        with self.lock:
            self.buckets[key].free_slots.append(int(slot))

    def submit(self, key, params, kind="loglike", slot=None):
        """Queue one request: ``params`` is a [n_profiles, 8] parameter array in
        the ImageModel.pack layout; ``kind`` is "loglike" (needs ``slot``) or
        "image".  Returns a Future whose value is the float logL or the
        [ny, nx] image (a CPU tensor)."""
        p = np.asarray(params, dtype=np.float64)[None]
        return self.submit_many(key, p, kind, None if slot is None else [slot])[0]

    def submit_many(self, key, params, kind="loglike", slots=None, t_arrival=None):
        """Queue B requests at once: params [B, n_profiles, 8]; slots [B] for
        log-likelihood requests; t_arrival optional [B] arrival times on the
        time.perf_counter clock (latencies are measured from them)."""
        if kind not in ("loglike", "image"):
            raise ValueError("kind must be 'loglike' or 'image'")
        b = self.buckets[key]
        params = np.ascontiguousarray(params, dtype=np.float64).reshape(-1, b.n_prof, backend.NPARAM)
        n = params.shape[0]
        if kind == "loglike":
            if slots is None:
                raise ValueError("log-likelihood requests need data slots")
            slots = np.asarray(slots, dtype=np.int64).reshape(n)
            if n and (int(slots.min()) < 0 or int(slots.max()) >= b.slots.shape[0]):
                raise ValueError("data slot out of range")
        else:
            slots = np.zeros(n, dtype=np.int64)
        now = time.perf_counter()
        ts = [now] * n if t_arrival is None else list(np.asarray(t_arrival, dtype=np.float64))
        futs = [Future(self.cond, t) for t in ts]
        with self.lock:
            self.pending.setdefault((key, kind), collections.deque()).append(_Block(kind, params, slots, futs))
        return futs

    # ---------------------------------------------------------- scheduling --
    def _fail(self, futs, exc):
        now = time.perf_counter()
        for f in futs:
            f.error = exc
            f.t_done = now
        self.requests_failed += len(futs)
        self.last_error = exc
        with self.cond:
            self.cond.notify_all()

    def _retire(self):
        if self.overdue:  # batches the watchdog gave up on: their lane and stream return once they finish
            still = []
            for item in self.overdue:
                if item[0].query():
                    item[3].busy = False
                    self.free_streams.append(item[6])
                else:
                    still.append(item)
            self.overdue = still
        if not self.inflight:
            return
        still, done, late = [], [], []
        now = time.perf_counter()
        for item in self.inflight:
            if item[0].query():
                done.append(item)
            elif now - item[7] > self.watchdog:
                late.append(item)
            else:
                still.append(item)
        if not done and not late:
            return
        self.inflight = still
        for item in late:
            self.watchdog_trips += 1
            self.overdue.append(item)
            self._fail(item[5], TimeoutError("batch of %d %s requests for bucket %r still running %.1f s after its "
                                             "launch" % (len(item[5]), item[4], item[2].key, now - item[7])))
        for ev, images, b, lane, kind, futs, si, t0 in done:
            n = len(futs)
            if kind == "loglike":
                vals = lane.logl_host[:n].numpy().tolist()
            else:
                vals = list(images.unbind(0))  # views of the batch's own pinned buffer
            for f, v in zip(futs, vals):
                f.value = v
                f.t_done = now
            lane.busy = False
            self.free_streams.append(si)
            self.requests_done += n
        with self.cond:
            self.cond.notify_all()

    def _take(self, q, max_n):
        """Up to max_n requests from the front of q, one (bucket, kind) queue:
        (kind, params, slots, futures)."""
        kind = q[0].kind
        ps, ss, fs = [], [], []
        need = max_n
        while q and need > 0 and q[0].kind == kind:
            blk = q[0]
            k = min(need, len(blk.futures) - blk.pos)
            ps.append(blk.params[blk.pos:blk.pos + k])
            ss.append(blk.slots[blk.pos:blk.pos + k])
            fs.extend(blk.futures[blk.pos:blk.pos + k])
            blk.pos += k
            need -= k
            if blk.pos == len(blk.futures):
                q.popleft()
        return kind, ps, ss, fs

    def _dispatch(self):
        launched = False
        while self.free_streams:
            with self.lock:
                pick = lane = None
                for qk, q in self.pending.items():
                    if q:
                        lane = self.buckets[qk[0]].free_lane()
                        if lane is not None:
                            pick = qk
                            break
                if pick is None:
                    return launched
                b = self.buckets[pick[0]]
                kind, ps, ss, futs = self._take(self.pending[pick], b.max_batch)
                self.pending.move_to_end(pick)  # round-robin between (bucket, kind) queues
            n = len(futs)
            np.concatenate(ps, axis=0, out=lane.params_np[:n])
            np.concatenate(ss, axis=0, out=lane.slot_np[:n])
            si = self.free_streams.pop(0)
            lane.busy = True
            try:
                ev, images = b.run(lane, kind, n, self.streams[si])
            except Exception as exc:  # the batch did not launch: its own requests fail
                lane.busy = False
                self.free_streams.append(si)
                self.failed_batches += 1
                self._fail(futs, exc)
                continue
            self.inflight.append((ev, images, b, lane, kind, futs, si, time.perf_counter()))
            self.batches += 1
            self.batch_sizes[n] += 1
            launched = True
        return launched

    def step(self):
        """One scheduler iteration: retire finished batches, launch new ones.
        Returns True when it launched a batch."""
        self._retire()
        return self._dispatch()

    def idle(self):
        with self.lock:
            return not self.inflight and not any(self.pending.values())

    def health(self):
        """Counters of the scheduler: batches launched and failed, watchdog
        trips, requests done and failed, batches still holding a lane after a
        watchdog trip, and the last error (None if there was none)."""
        return {"batches": self.batches, "failed_batches": self.failed_batches,
                "watchdog_trips": self.watchdog_trips, "overdue_batches": len(self.overdue),
                "requests_done": self.requests_done, "requests_failed": self.requests_failed,
                "free_streams": len(self.free_streams), "last_error": None if self.last_error is None
                else "%s: %s" % (type(self.last_error).__name__, self.last_error)}

    def _fail_all(self, exc):
        """Fail every queued and in-flight request (an error in the scheduler)."""
        with self.lock:
            futs = []
            for q in self.pending.values():
                while q:
                    blk = q.popleft()
                    futs.extend(blk.futures[blk.pos:])
            for item in self.inflight:
                futs.extend(item[5])
            self.overdue.extend(self.inflight)
            self.inflight = []
        self._fail(futs, exc)

    def drain(self, timeout=120.0):
        """Run the scheduler until every queued request has completed."""
        t0 = time.perf_counter()
        while not self.idle():
            if not self.step():
                time.sleep(20e-6)
            if time.perf_counter() - t0 > timeout:
                raise TimeoutError("server did not drain in %.0f s" % timeout)

    def start(self):
        """Run the scheduler in a background thread."""
        if self._thread is not None:
            return
        self._stop = False
        device = self.streams[0].device

        def loop():
            torch.cuda.set_device(device)
            while not self._stop:
                try:
                    if not self.step():
                        time.sleep(20e-6)
                except Exception as exc:  # the scheduler itself failed: nothing outstanding is left waiting
                    self._fail_all(exc)
                    self._stop = True

        self._thread = threading.Thread(target=loop, daemon=True)
        self._thread.start()

    def stop(self):
        """Stop the background scheduler and finish the batches in flight."""
        self._stop = True
        if self._thread is not None:
            self._thread.join()
            self._thread = None
        for s in self.streams:
            s.synchronize()
        self._retire()


def latency_percentiles(futures, qs=(50, 99)):
    """Latency percentiles in milliseconds of completed futures."""
    lat = np.array([f.latency for f in futures if f.latency is not None]) * 1e3
    if lat.size == 0:
        return {q: None for q in qs}
    return {q: float(np.percentile(lat, q)) for q in qs}


def poisson_load(server, key, params, rate, duration, kind="loglike", slots=None, seed=0, tick=50e-6):
    """Offer requests to a server as a Poisson process of ``rate`` requests/s
    for ``duration`` seconds, cycling through ``params`` [M, n, 8] (and
    ``slots`` [M]); waits for completion and returns (futures, wall seconds
    from the first arrival to the last completion).  See ``poisson_load_mix``
    for the two loop forms."""
    futs, wall, _, _ = poisson_load_mix(server, [dict(key=key, params=params, slots=slots, kind=kind)], rate,
                                        duration, seed=seed, tick=tick)
    return futs, wall


def poisson_load_mix(server, streams, rate, duration, seed=0, tick=50e-6):
    """Poisson arrivals of ``rate`` requests/s for ``duration`` seconds over
    several request streams.  Each stream is a dict with ``key`` (bucket),
    ``params`` [M, n, 8], ``slots`` [M] (log-likelihood requests) and ``kind``,
    and optionally ``weight``; every arrival picks a stream with probability
    proportional to its weight and takes that stream's next parameter row.
    Waits for completion; returns (futures in arrival order, wall seconds from
    the first arrival to the last completion, the largest number of requests
    in flight (submitted, not yet completed), (stream index, row) arrays of
    the arrivals).

    With the server's background thread running (``start``), arrivals are
    submitted from this thread and the scheduler runs in its own.  Without it,
    this thread is the server's event loop: it submits the arrivals that are
    due and steps the scheduler in turn, never sleeping.  The second form
    avoids two Python threads contending for the interpreter lock: a loop that
    releases and retakes the lock every few tens of microseconds (as an
    arrival loop does between arrivals) can keep a waiting thread from taking
    it for tens of milliseconds."""
    import gc

    driven = server._thread is None
    rng = np.random.default_rng(seed)
    n_req = max(1, int(rate * duration))
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_req))  # offsets, shifted to the clock below
    w = np.array([float(s.get("weight", 1.0)) for s in streams])
    which = rng.choice(len(streams), size=n_req, p=w / w.sum())
    row = np.zeros(n_req, dtype=np.int64)  # each stream cycles through its own rows
    for k, s in enumerate(streams):
        sel = np.nonzero(which == k)[0]
        row[sel] = np.arange(sel.size) % s["params"].shape[0]
    slots = [None if s.get("slots") is None else np.asarray(s["slots"]) for s in streams]
    futs = [None] * n_req
    i = 0
    done0 = server.requests_done
    peak = 0
    # the cyclic garbage collector is paused while load is offered: the
    # server's objects form no reference cycles, and a full collection over
    # the outstanding requests would stall the loop for tens of milliseconds
    gc.collect()
    gc.disable()
    arrivals += time.perf_counter() + 1e-3
    try:
        while i < n_req:
            now = time.perf_counter()
            j = int(np.searchsorted(arrivals, now, side="right"))
            if j > i:
                due = np.arange(i, j)
                for k in np.unique(which[due]):
                    idx = due[which[due] == k]
                    s = streams[k]
                    r = row[idx]
                    got = server.submit_many(s["key"], s["params"][r], s.get("kind", "loglike"),
                                             None if slots[k] is None else slots[k][r], t_arrival=arrivals[idx])
                    for q, f in zip(idx, got):
                        futs[q] = f
                i = j
                peak = max(peak, i - (server.requests_done - done0))
            if driven:
                server.step()
            elif j <= i:
                time.sleep(tick)
        if driven:
            server.drain()
        for f in futs:
            f.result(timeout=120)
    finally:
        gc.enable()
    wall = max(f.t_done for f in futs) - arrivals[0]
    return futs, wall, peak, (which, row)
