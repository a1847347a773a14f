#!/usr/bin/env python
"""Benchmark: images/s of the batched image + log-likelihood evaluation.

Workload (one step): log-likelihoods of a batch of 64 parameter sets of the
EPL + SHEAR deflector with SERSIC_ELLIPSE lens light and source, 180x180 pixels
at 0.04"/pix, supersampling 3 (540x540 rays), 41x41 Moffat PSF, against one
mock data image; parameters and data are resident in HBM before timing.  The
headline value is the mixed-precision form (float32 profiles and convolution,
float64 coordinates and chi^2 sums); the plain FP64 form is timed beside it.

Prints one JSON line.  --gpus above 1 is refused: this build runs on one GPU.
"""

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3  # MI355X published FP32 vector peak
PEAK_FP64_TFLOPS = 78.6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_models(batch, in_flight=3):
    import torch

    from jaxtronomy_amd import workloads as W
    from jaxtronomy_amd.Data.imaging_data import ImageData
    from jaxtronomy_amd.Data.psf import PSF
    from jaxtronomy_amd.ImSim.image_model import ImageModel
    from jaxtronomy_amd.LensModel.lens_model import LensModel
    from jaxtronomy_amd.LightModel.light_model import LightModel

    cfg = W.EPL180
    kernel = W.moffat_kernel(cfg["psf_size"], cfg["psf_fwhm"] / cfg["deltaPix"], cfg["psf_beta"])
    draws = [W.epl180_draw(d) for d in range(batch)]

    def model(precision, data=None):
        return ImageModel(ImageData(**W.data_kwargs(cfg, data)), PSF("PIXEL", kernel_point_source=kernel),
                          LensModel(["EPL", "SHEAR"]), LightModel(["SERSIC_ELLIPSE"]), LightModel(["SERSIC_ELLIPSE"]),
                          kwargs_numerics={"supersampling_factor": cfg["supersampling"]}, precision=precision)

    probe = model("fp64")
    truth = probe.image(*draws[0])
    data = W.noisy(truth.cpu().numpy(), cfg["exposure_time"], cfg["background_rms"], np.random.default_rng(1))
    im64 = model("fp64", data)
    im32 = model("mixed", data)
    # further instances (own workspaces) for the other batches in flight
    im64.twins = [model("fp64", data) for _ in range(in_flight - 1)]
    im32.twins = [model("mixed", data) for _ in range(in_flight - 1)]
    kl = W.stack_kwargs([d[0] for d in draws])
    ks = W.stack_kwargs([d[1] for d in draws])
    kll = W.stack_kwargs([d[2] for d in draws])
    params_np, n = im64.pack(kl, ks, kll)
    params = torch.as_tensor(params_np, device="cuda")
    return cfg, kernel, draws, data, params, params_np, im64, im32


def time_steps(im, params, steps, warmup, backend, model=None):
    import torch

    logl = torch.empty(params.shape[0], dtype=torch.float64, device=params.device)
    for _ in range(warmup):
        im.likelihood_batch(params, logl=logl, model=model)
    torch.cuda.synchronize()
    backend.timing_read()  # drop warm-up records
    backend.timing_enable(True)
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    evs[0].record()
    for i in range(steps):
        im.likelihood_batch(params, logl=logl, model=model)
        evs[i + 1].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    backend.timing_enable(False)
    render_ms, conv_ms, calls = backend.timing_read()
    per_step = np.array([evs[i].elapsed_time(evs[i + 1]) for i in range(steps)])
    total_ms = evs[0].elapsed_time(evs[-1])
    return dict(total_ms=total_ms, wall_s=wall, per_step=per_step, render_ms=render_ms / max(calls, 1),
                conv_ms=conv_ms / max(calls, 1), logl=logl.cpu().numpy())


def time_in_flight(im, params, steps, warmup, backend):
    """Throughput with k = 1 + len(im.twins) batches in flight: step i runs on
    stream i % k with its own workspace (``im`` and its twins, instances of
    the same model), as the request server keeps several batches in flight on
    its streams.  One batch's render (vector ALU) then overlaps another's MFMA
    convolution.  Returns the total time of ``steps`` steps, the mean render
    and convolution durations seen on their streams while batches overlap, and
    the log-likelihoods of the last step of each lane."""
    import torch

    lanes = [im] + list(im.twins)
    k = len(lanes)
    streams = [torch.cuda.Stream() for _ in lanes]
    logls = [torch.empty(params.shape[0], dtype=torch.float64, device=params.device) for _ in lanes]
    main = torch.cuda.current_stream()

    def issue(n):
        for s in streams:
            s.wait_stream(main)
        for i in range(n):
            with torch.cuda.stream(streams[i % k]):
                lanes[i % k].likelihood_batch(params, logl=logls[i % k])
        for s in streams:
            main.wait_stream(s)

    issue(warmup)
    torch.cuda.synchronize()
    backend.timing_read()  # drop earlier records
    backend.timing_enable(True)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    issue(steps)
    e1.record()
    torch.cuda.synchronize()
    backend.timing_enable(False)
    render_ms, conv_ms, calls = backend.timing_read()
    return dict(total_ms=e0.elapsed_time(e1), render_ms=render_ms / max(calls, 1), conv_ms=conv_ms / max(calls, 1),
                logl=[x.cpu().numpy() for x in logls])


def clock_settle(im, params, seconds):
    """Untimed sustained load before the timed phases: batches three in flight
    (as the headline phase runs them) until ``seconds`` have passed, so the
    timed steps start at settled clocks (an idle GPU takes ~15 ms of load to
    reach them).  Returns the seconds spent and the batches run."""
    import torch

    lanes = [im] + list(im.twins)
    streams = [torch.cuda.Stream() for _ in lanes]
    logl = torch.empty(params.shape[0], dtype=torch.float64, device=params.device)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    n = 0
    while time.perf_counter() - t0 < seconds:
        for _ in range(3 * len(lanes)):
            with torch.cuda.stream(streams[n % len(lanes)]):
                lanes[n % len(lanes)].likelihood_batch(params, logl=logl)
            n += 1
        streams[(n - 1) % len(lanes)].synchronize()
    torch.cuda.synchronize()
    return time.perf_counter() - t0, n


def measured_traffic(name):
    """HBM bytes per launch of a kernel from its committed rocprofv3 PMC summary
    (FETCH_SIZE and WRITE_SIZE runs scaled by a plain copy's readings), or None."""
    import re

    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    m = re.search(r"^traffic_bytes_per_launch\s*=\s*([0-9.eE+]+)", open(path).read(), re.M)
    return float(m.group(1)) if m else None


def render_work(params_np, n_ray, lens_light=True):
    """Algorithmic flops and special-function evaluations of one render launch:
    per ray EPL 72 + 8 (N_terms - 1) flops and 6 specials (N_terms = the series
    length chosen from the axis ratio), SHEAR 10, deflected position 2, two
    SERSIC_ELLIPSE (source and lens light) 17 + 4 specials each, pooling 1."""
    from jaxtronomy_amd.Util.param_util import ellipticity2phi_q

    F = S = 0.0
    for p in params_np:
        _, q = ellipticity2phi_q(p[0, 2], p[0, 3])
        f = (1 - q) / (1 + q)
        nterms = 1 if f == 0 else (10 if f < 0.12 else 20 if f < 0.4 else 35 if f < 0.6 else 50 if f < 0.73
                                   else 100 if f < 0.86 else 200 if f < 0.92 else 500)
        nl = 2 if lens_light else 1
        F += n_ray * (72 + 8 * (nterms - 1) + 10 + 2 + 17 * nl + 1)
        S += n_ray * (6 + 4 * nl)
    return F, S


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def cgroup_cpus():
    """The CPU quota of this process's cgroup (cpu.max, in CPUs), or None."""
    try:
        quota, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        return None if quota == "max" else float(quota) / float(period)
    except (OSError, ValueError):
        return None


def cpu_baseline(cfg, kernel, params_np, data, repeats=20, warmups=3):
    """The C FP64 oracle timed on the host, as the measurement plan asks: 3
    warm-up repeats, then the median of ``repeats`` timed repeats
    (time.perf_counter, a steady clock), each repeat one log-likelihood per
    thread (draws of the benchmark batch in turn), on the affinity set's
    threads (capped at the cgroup's CPU quota, which otherwise throttles the
    extra threads) and on one thread.  Returns the baseline record and the
    log-likelihoods of the first draws."""
    from oracle import c_oracle
    from jaxtronomy_amd import workloads as W

    affinity = len(os.sched_getaffinity(0))
    quota = cgroup_cpus()
    threads = affinity if quota is None else max(1, min(affinity, int(quota)))
    g = W.grid_dict(cfg)
    c_oracle.lib()

    def run(p, n_threads):
        t0 = time.perf_counter()
        ll = c_oracle.loglike(2, 1, 1, [1, 6, 33, 33], p, g, data, kernel=kernel, psf_type=1,
                              exposure_time=cfg["exposure_time"], background_rms=cfg["background_rms"],
                              n_threads=n_threads)
        return ll, time.perf_counter() - t0

    def median_rate(n_threads):
        times, first = [], None
        for r in range(warmups + repeats):
            idx = (np.arange(n_threads) + r * n_threads) % params_np.shape[0]
            ll, dt = run(params_np[idx], n_threads)
            if first is None:
                first = ll
            if r >= warmups:
                times.append(dt)
        med = float(np.median(times))
        return n_threads / med, med, float(np.sum(times)), first

    rate, med, tot, ll = median_rate(threads)
    rate1, med1, tot1, _ = median_rate(1)
    rec = dict(value=rate, unit="images/s", cores=threads, kind="port",
               sample="median of %d repeats (after %d warm-ups) of %d log-likelihood evaluations, one per thread, of "
                      "the benchmark's draws; C FP64 restatement of the reference algorithm, OpenMP; %.3f s per repeat, "
                      "%.1f s timed; host %s, affinity set %d CPUs, cgroup quota %s CPUs"
                      % (repeats, warmups, threads, med, tot, cpu_model(), affinity,
                         "none" if quota is None else "%g" % quota),
               single_thread={"value": rate1, "sample": "median of %d repeats (after %d warm-ups) of one evaluation, "
                                                        "%.3f s each" % (repeats, warmups, med1)})
    return rec, ll


def build_substructure(batch, n_sub):
    import torch

    from jaxtronomy_amd import workloads as W
    from jaxtronomy_amd.Data.imaging_data import ImageData
    from jaxtronomy_amd.Data.psf import PSF
    from jaxtronomy_amd.ImSim.image_model import ImageModel
    from jaxtronomy_amd.LensModel.lens_model_bulk import LensModelBulk
    from jaxtronomy_amd.LightModel.light_model import LightModel

    cfg = dict(W.SUBSTRUCTURE, n_sub=n_sub)
    kernel = W.moffat_kernel(cfg["psf_size"], cfg["psf_fwhm"] / cfg["deltaPix"], cfg["psf_beta"])
    draws = [W.substructure_draw(d, cfg=cfg) for d in range(batch)]
    names = ["EPL", "SHEAR"] + ["TNFW"] * n_sub

    def model(precision, data=None):
        return ImageModel(ImageData(**W.data_kwargs(cfg, data)), PSF("PIXEL", kernel_point_source=kernel),
                          LensModelBulk(["EPL", "SHEAR", "TNFW"], precision=precision),
                          LightModel(["SERSIC_ELLIPSE"]),
                          kwargs_numerics={"supersampling_factor": cfg["supersampling"]}, precision=precision)

    probe = model("fp64")
    kl0, subs0, ks0 = draws[0]
    truth = probe.image(probe.LensModel.prepare_ray_shooting_kwargs(names, kl0 + subs0), ks0)
    data = W.noisy(truth.cpu().numpy(), cfg["exposure_time"], cfg["background_rms"], np.random.default_rng(1))
    im64 = model("fp64", data)
    im32 = model("mixed", data)
    lm = im64.LensModel
    rsk = lm.stack_ray_shooting_kwargs([lm.prepare_ray_shooting_kwargs(names, d[0] + d[1]) for d in draws])
    ks = W.stack_kwargs([d[2] for d in draws])
    params_np, n, m, keep = im64.pack_bulk(rsk, ks)
    params = torch.as_tensor(params_np, device="cuda")
    return cfg, kernel, names, data, params, params_np, (m, keep), im64, im32


def substructure_work(params_np, n_ray, n_sub):
    """Per ray: EPL 72 + 8 (N_terms - 1) flops and 6 specials, SHEAR 10, n_sub
    TNFW at 36 flops and 7 specials each, deflected position 2,
    SERSIC_ELLIPSE 17 + 4 specials, pooling 1."""
    F, S = render_work(params_np, n_ray, lens_light=False)
    n_img = params_np.shape[0]
    return F + n_img * n_ray * 36.0 * n_sub, S + n_img * n_ray * 7.0 * n_sub


def serve_measure(im, params_np, data, capacity, seconds, fractions, label=""):
    """Served load on the request server: Poisson arrivals of log-likelihood
    requests at fractions of the batched capacity (a fraction above 1 offers
    more than the server can take, so its achieved rate is the saturated
    throughput); achieved rate and latency percentiles per offered rate."""
    from jaxtronomy_amd.serving import LensServer, latency_percentiles, poisson_load

    srv = LensServer(n_streams=4)
    srv.register("epl180", im, max_batch=64, n_slots=2, n_lanes=3)
    slot = srv.put_data("epl180", data)
    slots = np.full(params_np.shape[0], slot)
    srv.warmup("epl180", params_np[0])
    points = []  # no background thread: the load loop steps the scheduler itself
    try:
        # steady state first: the lowest offered rate for half a second, not measured
        poisson_load(srv, "epl180", params_np, fractions[0] * capacity, 0.5, slots=slots, seed=99)
        for i, frac in enumerate(fractions):
            rate = frac * capacity
            b0, n0 = srv.batches, srv.requests_done
            futs, wall = poisson_load(srv, "epl180", params_np, rate, seconds, slots=slots, seed=i)
            pct = latency_percentiles(futs)
            nb = srv.batches - b0
            points.append({"offered_per_s": round(rate, 1), "achieved_per_s": round(float(len(futs) / wall), 1),
                           "p50_ms": round(pct[50], 4), "p99_ms": round(pct[99], 4),
                           "mean_batch": round((srv.requests_done - n0) / max(nb, 1), 2)})
            del futs
            log("bench: served %s %.0f/s offered -> %s" % (label, rate, points[-1]))
    finally:
        srv.stop()
    return points


def main_substructure(args):
    import torch

    from jaxtronomy_amd import backend

    log("bench: building substructure models (%d subhalos, batch %d)" % (args.n_sub, args.batch))
    cfg, kernel, names, data, params, params_np, mk, im64, im32 = build_substructure(args.batch, args.n_sub)
    m = mk[0]
    n_ray = (cfg["numPix"] * cfg["supersampling"]) ** 2
    log("bench: timing mixed precision (%d steps)" % args.steps)
    mixed = time_steps(im32, params, args.steps, args.warmup, backend, model=m)
    log("bench: timing fp64 (%d steps)" % args.fp64_steps)
    fp64 = time_steps(im64, params, args.fp64_steps, max(1, args.warmup // 2), backend, model=m)
    ll_rel = float(np.max(np.abs(mixed["logl"] - fp64["logl"]) / np.abs(fp64["logl"])))
    value = args.batch * args.steps / (mixed["total_ms"] / 1e3)
    value64 = args.batch * args.fp64_steps / (fp64["total_ms"] / 1e3)
    F, S = substructure_work(params_np, n_ray, args.n_sub)
    achieved = (F + 8 * S) / (mixed["render_ms"] / 1e3) / 1e12
    achieved64 = (F + 20 * S) / (fp64["render_ms"] / 1e3) / 1e12
    line = {
        "metric": "realisations/s, EPL+SHEAR + %d TNFW subhalos, 300^2 ss3, mixed precision" % args.n_sub,
        "value": round(value, 2),
        "unit": "images/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mixed["total_ms"] / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "EPL+SHEAR + %d TNFW (log10 M ~ U[6,9], positions U[-3,3]\", r_trunc = 5 Rs), "
                               "SERSIC_ELLIPSE source, 300x300 @0.03\", supersampling 3, 41x41 Moffat PSF, "
                               "image + chi2 logL" % args.n_sub, "batch": args.batch},
        "p50_ms": round(float(np.percentile(mixed["per_step"], 50)), 4),
        "p99_ms": round(float(np.percentile(mixed["per_step"], 99)), 4),
        "fp64": {"value": round(value64, 2), "ms_per_step": round(fp64["total_ms"] / args.fp64_steps, 4),
                 "render_ms": round(fp64["render_ms"], 4), "render_tflops_equiv": round(achieved64, 3)},
        "accuracy": {"mixed_logl_max_rel_vs_fp64": ll_rel},
        "kernels_ms": {"render": round(mixed["render_ms"], 4), "conv": round(mixed["conv_ms"], 4)},
        "roofline": {"bound": "valu", "kernel": "render_kernel<float>", "achieved": round(achieved, 3),
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                     "traffic": None,
                     "work": "per launch %.3g flops + 8 x %.3g special-function evaluations" % (F, S)},
    }
    if not args.no_cpu:
        from oracle import c_oracle
        from jaxtronomy_amd import workloads as W

        threads = min(16, len(os.sched_getaffinity(0)))
        g = W.grid_dict(cfg)
        types = [1, 6] + [5] * args.n_sub + [33]
        cp = params_np[:1]  # [EPL, SHEAR] inline, TNFW bulk, source: the oracle's lens-list order
        c_oracle.lib()
        t0 = time.perf_counter()
        ll = c_oracle.loglike(2 + args.n_sub, 1, 0, types, cp, g, data, kernel=kernel, psf_type=1,
                              exposure_time=cfg["exposure_time"], background_rms=cfg["background_rms"],
                              n_threads=threads)
        dt = time.perf_counter() - t0
        line["cpu_baseline"] = dict(value=1.0 / dt, unit="images/s", cores=threads, kind="port",
                                    sample="1 log-likelihood evaluation (C FP64 restatement, OpenMP, %.1f s)" % dt)
        line["accuracy"]["fp64_logl_rel_vs_cpu_oracle"] = float(abs(fp64["logl"][0] - ll[0]) / abs(ll[0]))
    print(json.dumps(line), flush=True)


GRAD_FREE = [(0, 0), (0, 1), (0, 2), (0, 3), (0, 4), (0, 5), (1, 0), (1, 1), (2, 0), (2, 1), (2, 2), (2, 3), (2, 4),
             (3, 0), (3, 1), (3, 2), (3, 3), (3, 4)]  # EPL 6, SHEAR 2, source 5, lens light 5 (centres fixed)


def main_grad(args):
    """Gradient config: dlogL/dtheta of 18 parameters at 256 perturbed points,
    120x120 grid, supersampling 2, 21x21 Moffat PSF, mock data from a seeded
    truth; mixed form timed, FP64 form beside it."""
    import torch

    from jaxtronomy_amd import workloads as W
    from jaxtronomy_amd.Data.imaging_data import ImageData
    from jaxtronomy_amd.Data.psf import PSF
    from jaxtronomy_amd.ImSim.image_model import ImageModel
    from jaxtronomy_amd.LensModel.lens_model import LensModel
    from jaxtronomy_amd.LightModel.light_model import LightModel

    cfg = dict(W.EPL180, numPix=120, supersampling=2)
    kernel = W.moffat_kernel(21, cfg["psf_fwhm"] / cfg["deltaPix"], cfg["psf_beta"])
    truth = W.epl180_draw(0)

    def model(precision, data=None):
        return ImageModel(ImageData(**W.data_kwargs(cfg, data)), PSF("PIXEL", kernel_point_source=kernel),
                          LensModel(["EPL", "SHEAR"]), LightModel(["SERSIC_ELLIPSE"]), LightModel(["SERSIC_ELLIPSE"]),
                          kwargs_numerics={"supersampling_factor": cfg["supersampling"]}, precision=precision)

    clean = model("fp64").image(*truth).cpu().numpy()
    data = W.noisy(clean, cfg["exposure_time"], cfg["background_rms"], np.random.default_rng(3))
    im64, im32 = model("fp64", data), model("mixed", data)
    base, _ = im64.pack(*truth)
    rng = np.random.default_rng(W.seed(3, 0))
    p = np.repeat(base, args.batch, axis=0)
    for (k, i) in GRAD_FREE:  # truth + N(0, 0.02 x scale) on every free parameter
        p[:, k, i] += rng.normal(0, 0.02, args.batch) * max(abs(base[0, k, i]), 0.1)
    params = torch.as_tensor(p, device="cuda")

    def timed(im, steps, warmup):
        for _ in range(warmup):
            im.likelihood_and_grad_batch(params)
        torch.cuda.synchronize()
        evs = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
        evs[0].record()
        for i in range(steps):
            ll, g = im.likelihood_and_grad_batch(params)
            evs[i + 1].record()
        torch.cuda.synchronize()
        per = np.array([evs[i].elapsed_time(evs[i + 1]) for i in range(steps)])
        return per, ll.cpu().numpy(), g.cpu().numpy()

    log("bench: timing mixed gradients (%d steps)" % args.steps)
    per32, ll32, g32 = timed(im32, args.steps, args.warmup)
    log("bench: timing fp64 gradients (%d steps)" % args.fp64_steps)
    per64, ll64, g64 = timed(im64, args.fp64_steps, max(1, args.warmup // 2))
    scale = np.array([np.max(np.abs(g64[:, k, i])) for (k, i) in GRAD_FREE])
    err = np.array([np.max(np.abs(g32[:, k, i] - g64[:, k, i])) for (k, i) in GRAD_FREE]) / scale
    ms32, ms64 = float(np.sum(per32)) / args.steps, float(np.sum(per64)) / args.fp64_steps
    line = {
        "metric": "gradient evaluations/s (logL + 18-parameter gradient), 120^2 ss2, 21x21 PSF, mixed precision",
        "value": round(args.batch / (ms32 / 1e3), 1),
        "unit": "gradients/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms32, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "EPL+SHEAR, SERSIC_ELLIPSE source + lens light, 120x120 @0.04\", supersampling 2, "
                               "21x21 Moffat PSF, mock data from a seeded truth, points = truth + N(0, 0.02 scale)",
                   "batch": args.batch, "free_parameters": len(GRAD_FREE)},
        "p50_ms": round(float(np.percentile(per32, 50)), 4),
        "p99_ms": round(float(np.percentile(per32, 99)), 4),
        "fp64": {"value": round(args.batch / (ms64 / 1e3), 1), "ms_per_step": round(ms64, 4)},
        "accuracy": {"mixed_grad_max_err_rel_to_param_scale_vs_fp64": float(np.max(err)),
                     "mixed_logl_max_rel_vs_fp64": float(np.max(np.abs(ll32 - ll64) / np.abs(ll64)))},
    }
    print(json.dumps(line), flush=True)


def main_serve_mix(args):
    """Serving mix: Poisson arrivals of log-likelihood (80%) and simulate (20%)
    requests over 12 buckets spanning grids {64, 100, 180, 240}^2, supersampling
    {1, 3, 5}, deflectors {SIE+SHEAR, EPL+SHEAR, NFW+SIE} and PSFs 11-41, each
    request naming one of its bucket's resident observations and noise maps."""
    import torch

    from jaxtronomy_amd import workloads as W
    from jaxtronomy_amd.Data.imaging_data import ImageData
    from jaxtronomy_amd.Data.psf import PSF
    from jaxtronomy_amd.ImSim.image_model import ImageModel
    from jaxtronomy_amd.LensModel.lens_model import LensModel
    from jaxtronomy_amd.LightModel.light_model import LightModel
    from jaxtronomy_amd.serving import LensServer, latency_percentiles, poisson_load_mix

    m = W.SERVE_MIX
    combos = W.mix_buckets(12)
    n_obs, n_req_rows, max_batch = args.mix_slots, 16, 32
    srv = LensServer(n_streams=4)
    streams, cost, names, rays = [], [], [], []
    log("bench: registering %d buckets (%d resident observations each)" % (len(combos), n_obs))
    for bi, (n, s, defl, k) in enumerate(combos):
        cfg = dict(numPix=n, deltaPix=m["deltaPix"], supersampling=s, exposure_time=m["exposure_time"],
                   background_rms=m["background_rms"])
        kernel = W.moffat_kernel(k, m["psf_fwhm"] / m["deltaPix"], m["psf_beta"])
        im = ImageModel(ImageData(**W.data_kwargs(cfg, np.zeros((n, n)))), PSF("PIXEL", kernel_point_source=kernel),
                        LensModel(list(defl)), LightModel(["SERSIC_ELLIPSE"]), LightModel(["SERSIC_ELLIPSE"]),
                        kwargs_numerics={"supersampling_factor": s}, precision="mixed")
        key = "b%d" % bi
        srv.register(key, im, max_batch=max_batch, n_slots=n_obs, n_lanes=2, noise_maps=True)
        rows = np.stack([im.pack(*W.mix_draw(defl, 100 * bi + d))[0][0] for d in range(n_req_rows)])
        truth = np.stack([im.pack(*W.mix_draw(defl, 100 * bi + 50 + d))[0][0] for d in range(4)])
        clean = im.image_batch(torch.as_tensor(truth, device="cuda")).double().cpu().numpy()
        slots = []
        rng = np.random.default_rng(W.seed(4, 2000 + bi))
        for o in range(n_obs):
            c = clean[o % 4]
            obs = W.noisy(c, cfg["exposure_time"], cfg["background_rms"], rng)
            noise = np.sqrt(cfg["background_rms"] ** 2 + np.maximum(c, 0) / cfg["exposure_time"])
            slots.append(srv.put_data(key, obs, noise))
        srv.warmup(key, rows[0], kinds=("loglike", "image"))
        # cost of a full batch on this bucket (for the mix's capacity estimate)
        b = srv.buckets[key]
        lane = b.lanes[0]
        lane.params_np[:] = rows[np.arange(max_batch) % n_req_rows]
        lane.slot_np[:] = 0
        st = srv.streams[0]
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        t0.record(st)
        for _ in range(5):
            b.run(lane, "loglike", max_batch, st)
        t1.record(st)
        st.synchronize()
        cost.append(t0.elapsed_time(t1) / 5 / max_batch / 1e3)
        # log-likelihood requests cycle through every resident observation
        streams.append(dict(key=key, params=rows[np.arange(n_obs) % n_req_rows], slots=np.array(slots),
                            kind="loglike", weight=0.8))
        streams.append(dict(key=key, params=rows, slots=None, kind="image", weight=0.2))
        names.append("%d^2 s%d %s K%d" % (n, s, "+".join(defl), k))
        rays.append((n * s) ** 2)
    capacity = 1.0 / float(np.mean(cost))  # requests/s when every request is batched at max_batch
    log("bench: mix capacity estimate %.0f requests/s" % capacity)
    points = []
    poisson_load_mix(srv, streams, 0.3 * capacity, 0.5, seed=99)  # steady state first, not measured
    for i, frac in enumerate((0.5, 0.8, 2.0)):
        b0, n0 = srv.batches, srv.requests_done
        futs, wall, peak, _ = poisson_load_mix(srv, streams, frac * capacity, args.serve_seconds, seed=i)
        kinds = np.array([0 if isinstance(f.value, float) else 1 for f in futs])
        pall = latency_percentiles(futs)
        pll = latency_percentiles([f for f, k in zip(futs, kinds) if k == 0])
        pim = latency_percentiles([f for f, k in zip(futs, kinds) if k == 1])
        points.append({"offered_per_s": round(frac * capacity, 1), "achieved_per_s": round(len(futs) / wall, 1),
                       "p50_ms": round(pall[50], 4), "p99_ms": round(pall[99], 4),
                       "loglike_p99_ms": round(pll[99], 4), "image_p99_ms": round(pim[99], 4),
                       "peak_in_flight": int(peak),
                       "mean_batch": round((srv.requests_done - n0) / max(srv.batches - b0, 1), 2)})
        log("bench: mix offered %.0f/s -> %s" % (frac * capacity, points[-1]))
        del futs
    srv.stop()
    sat = points[-1]["achieved_per_s"]
    line = {
        "metric": "requests/s + p50/p99 latency, serving mix (grids 64-240^2, ss 1/3/5, 3 deflector sets, PSF 11-41)",
        "value": sat,
        "unit": "requests/s",
        "n_gpus": 1,
        "steps": len(points),
        "warmup": 1,
        "ms_per_step": round(1e3 / sat, 6),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "Poisson arrivals, 80% log-likelihood / 20% simulate, uniform over 12 buckets: " +
                               "; ".join(names),
                   "resident_observations": n_obs * len(combos), "max_batch": max_batch,
                   "capacity_estimate_per_s": round(capacity, 1)},
        "p50_ms": points[-1]["p50_ms"],
        "p99_ms": points[-1]["p99_ms"],
        "points": points,
    }
    print(json.dumps(line), flush=True)


def main_sie60(args):
    """The small scene (SIE+SHEAR, SERSIC_ELLIPSE source, 60x60 at 0.05"/pix,
    supersampling 1, 11x11 Gaussian PSF of FWHM 0.1", batch 1).  One image is
    far too little work to fill the GPU, so its figure is latency: a single
    log-likelihood call with data resident, as the caller sees it (host clock
    from the call to its completion), issued eagerly and as a replay of a
    captured HIP graph, in both precision forms; the C FP64 oracle's time for
    the same call on one host thread is reported beside it."""
    import torch

    from jaxtronomy_amd import backend
    from jaxtronomy_amd import workloads as W
    from jaxtronomy_amd.Data.imaging_data import ImageData
    from jaxtronomy_amd.Data.psf import PSF
    from jaxtronomy_amd.ImSim.image_model import ImageModel
    from jaxtronomy_amd.LensModel.lens_model import LensModel
    from jaxtronomy_amd.LightModel.light_model import LightModel
    from jaxtronomy_amd.serving import _hip

    cfg = W.SMALL
    kl, ks = W.small_draw(0)

    def model(precision, data=None):
        return ImageModel(ImageData(**W.data_kwargs(cfg, data)),
                          PSF("GAUSSIAN", fwhm=cfg["psf_fwhm"], truncation=cfg["psf_truncation"]),
                          LensModel(["SIE", "SHEAR"]), LightModel(["SERSIC_ELLIPSE"]),
                          kwargs_numerics={"supersampling_factor": cfg["supersampling"]}, precision=precision)

    truth = model("fp64").image(kl, ks)
    data = W.noisy(truth.cpu().numpy(), cfg["exposure_time"], cfg["background_rms"], np.random.default_rng(1))
    backend.timing_enable(False)  # the library's own event records synchronise; none inside a graph
    forms, logls = {}, {}
    for precision in ("fp64", "mixed"):
        im = model(precision, data)
        params_np, _ = im.pack(kl, ks)
        params = torch.as_tensor(params_np, device="cuda")
        logl = torch.empty(1, dtype=torch.float64, device="cuda")
        st = torch.cuda.Stream()
        torch.cuda.synchronize()

        def call():
            im.likelihood_batch(params, logl=logl)

        rec = {}
        with torch.cuda.stream(st):
            for _ in range(args.warmup):
                call()
            st.synchronize()
            host = np.empty(args.steps)
            for i in range(args.steps):
                t0 = time.perf_counter()
                call()
                st.synchronize()
                host[i] = time.perf_counter() - t0
            eager_ll = float(logl.item())
        rec["eager"] = {"p50_ms": round(1e3 * float(np.percentile(host, 50)), 4),
                        "p99_ms": round(1e3 * float(np.percentile(host, 99)), 4)}
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=st):
            call()
        # launched as the request server launches its graphs: hipGraphLaunch on
        # the stream, without torch's replay bookkeeping
        exe, sp = ctypes.c_void_p(g.raw_cuda_graph_exec()), ctypes.c_void_p(st.cuda_stream)
        hip = _hip()

        def replay():
            rc = hip.hipGraphLaunch(exe, sp)
            if rc != 0:
                raise RuntimeError("hipGraphLaunch failed (%d)" % rc)

        with torch.cuda.stream(st):
            for _ in range(args.warmup):
                replay()
            st.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            dev = np.empty(args.steps)
            for i in range(args.steps):
                t0 = time.perf_counter()
                e0.record(st)
                replay()
                e1.record(st)
                st.synchronize()
                host[i] = time.perf_counter() - t0
                dev[i] = e0.elapsed_time(e1)
            graph_ll = float(logl.item())
        if graph_ll != eager_ll:
            raise RuntimeError("graph replay logL %r differs from the eager call's %r" % (graph_ll, eager_ll))
        rec["graph"] = {"p50_ms": round(1e3 * float(np.percentile(host, 50)), 4),
                        "p99_ms": round(1e3 * float(np.percentile(host, 99)), 4),
                        "device_p50_ms": round(float(np.percentile(dev, 50)), 4)}
        forms[precision] = rec
        logls[precision] = eager_ll
        log("bench: %s %s" % (precision, rec))
        del g
    line = {
        "metric": "latency of one log-likelihood call (p50, graph replay, FP64), SIE+SHEAR+Sersic 60^2 ss1",
        "value": forms["fp64"]["graph"]["p50_ms"],
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": forms["fp64"]["graph"]["p50_ms"],
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "SIE+SHEAR deflector, SERSIC_ELLIPSE source, 60x60 @0.05\", supersampling 1, "
                               "11x11 Gaussian PSF (FWHM 0.1\", truncation 6), image + chi2 logL, batch 1",
                   "timing": "host clock from the call to its completion on the stream; device_p50_ms: HIP events "
                             "around the graph replay"},
        "p50_ms": forms["fp64"]["graph"]["p50_ms"],
        "p99_ms": forms["fp64"]["graph"]["p99_ms"],
        "fp64": forms["fp64"],
        "mixed": forms["mixed"],
        "accuracy": {"mixed_logl_rel_vs_fp64": abs(logls["mixed"] - logls["fp64"]) / abs(logls["fp64"])},
    }
    if not args.no_cpu:
        from oracle import c_oracle

        g = W.grid_dict(cfg)
        im = model("fp64", data)
        p = im.pack(kl, ks)[0]
        c_oracle.lib()
        run = lambda: c_oracle.loglike(2, 1, 0, [2, 6, 33], p, g, data, kernel=im.kernel, psf_type=2,
                                       exposure_time=cfg["exposure_time"], background_rms=cfg["background_rms"],
                                       n_threads=1)
        ll_cpu = run()[0]
        n = 200
        t0 = time.perf_counter()
        for _ in range(n):
            run()
        dt = (time.perf_counter() - t0) / n
        line["cpu_baseline"] = {"value": round(1e3 * dt, 4), "unit": "ms", "cores": 1, "kind": "port",
                                "sample": "%d calls of the same log-likelihood, C FP64 restatement of the reference "
                                          "algorithm, one thread, on %s" % (n, cpu_model())}
        line["accuracy"]["fp64_logl_rel_vs_cpu_oracle"] = abs(logls["fp64"] - ll_cpu) / abs(ll_cpu)
    print(json.dumps(line), flush=True)


LOS_PUBLISHED_S = 0.4684  # the reference's GPU time for this ray-shooting call (notebook cell 24)


def main_los(args):
    """Multi-plane ray shooting through a line-of-sight halo realisation
    (workloads.los_halo_draw: 5130 TNFW + 126 CONVERGENCE sheets on 126
    planes, 5500 deflector slots, z_source 2.5) at the 400 x 400 grid over
    +-10", the call the reference times on its GPU.  value: seconds per
    LensModelBulk.ray_shooting call from host arrays to device results
    (the drop-in call as the reference's notebook times it), FP64 as the
    reference runs it; the mixed form and the device-only kernel time beside."""
    import torch

    from jaxtronomy_amd import backend
    from jaxtronomy_amd import workloads as W
    from jaxtronomy_amd.LensModel.lens_model_bulk import LensModelBulk

    cfg = W.LOS
    names, kw, z = W.los_halo_draw(0)
    x, y = W.los_grid()
    n_ray = x.size
    n_tnfw, n_conv = names.count("TNFW"), names.count("CONVERGENCE")
    lib = backend.lib()
    res = {}
    for prec in ("fp64", "mixed"):
        lm = LensModelBulk(["TNFW", "CONVERGENCE"], multi_plane=True, precision=prec)
        rs = lm.prepare_ray_shooting_kwargs(names, kw, z_source=cfg["z_source"], lens_redshift_list=z,
                                            num_deflectors=cfg["num_deflectors"])
        steps = args.steps if prec == "mixed" else args.fp64_steps
        log("bench: los %s: drop-in calls (%d)" % (prec, steps))
        for _ in range(max(1, args.warmup)):
            lm.ray_shooting(x, y, rs)
        torch.cuda.synchronize()
        per = []
        for _ in range(steps):
            t0 = time.perf_counter()
            bx, by = lm.ray_shooting(x, y, rs)
            torch.cuda.synchronize()
            per.append(time.perf_counter() - t0)
        # device-resident: the raytrace launch alone, inputs already in HBM, HIP events on its stream
        n_img, it, ip, bt, bp = lm.pack(rs)
        m, keep = lm._model(it, bt, ray_shooting_kwargs=rs)
        params = backend.to_device_f64(np.concatenate([ip, bp], axis=1))
        xt, yt = backend.to_device_f64(x), backend.to_device_f64(y)
        ox = torch.empty((1, n_ray), dtype=torch.float64, device="cuda")
        oy = torch.empty_like(ox)
        pr = backend.PRECISIONS[prec]
        ws = torch.empty(max(256, lib.lensim_workspace_size(ctypes.byref(m), None, None, 1, pr)), dtype=torch.uint8,
                         device="cuda")

        def launch():
            backend.check(lib.lensim_raytrace(ctypes.byref(m), backend.ptr(params), 1, backend.ptr(xt),
                                              backend.ptr(yt), n_ray, pr, backend.ptr(ox), backend.ptr(oy),
                                              backend.ptr(ws), ws.numel(), backend.stream_handle()),
                          "lensim_raytrace")

        launch()
        torch.cuda.synchronize()
        evs = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
        evs[0].record()
        for i in range(steps):
            launch()
            evs[i + 1].record()
        torch.cuda.synchronize()
        kern = [evs[i].elapsed_time(evs[i + 1]) for i in range(steps)]
        del keep
        res[prec] = dict(call_s=float(np.median(per)), call_p99_s=float(np.percentile(per, 99)),
                         kernel_ms=float(np.mean(kern)), bx=ox.cpu().numpy()[0], by=oy.cpu().numpy()[0],
                         drop_in_equal=bool(np.array_equal(bx.cpu().numpy(), ox.cpu().numpy()[0])))
    # algorithmic work per call: per (ray, TNFW) 36 flops + 7 special functions, per (ray, sheet) 4 flops,
    # per (ray, plane) 10 flops of the plane recursion
    F = n_ray * (36.0 * n_tnfw + 4.0 * n_conv + 10.0 * (n_tnfw + n_conv))
    S = n_ray * 7.0 * n_tnfw
    k64 = res["fp64"]["kernel_ms"] / 1e3
    achieved = (F + 20 * S) / k64 / 1e12
    scale = float(np.max(np.abs(np.concatenate([x, y]))))
    mixed_err = max(float(np.max(np.abs(res["mixed"]["bx"] - res["fp64"]["bx"]))),
                    float(np.max(np.abs(res["mixed"]["by"] - res["fp64"]["by"])))) / scale
    line = {
        "metric": "seconds per multi-plane ray-shooting call, 400^2 rays through 5130 TNFW + 126 CONVERGENCE "
                  "planes (5500 deflector slots), FP64",
        "value": round(res["fp64"]["call_s"], 6),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.fp64_steps,
        "warmup": args.warmup,
        "ms_per_step": round(res["fp64"]["call_s"] * 1e3, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": round(res["fp64"]["call_s"] / LOS_PUBLISHED_S, 5),
        "dtype": "f64",
        "data": "synthetic: a seeded realisation with the published counts (pyHalo is not available)",
        "config": {"workload": "LensModelBulk(['TNFW', 'CONVERGENCE'], multi_plane=True).ray_shooting: z_lens 0.5, "
                               "z_source 2.5, 126 planes, 400x400 grid over +-10\", host arrays in, device arrays "
                               "out", "rays": n_ray, "deflectors": n_tnfw + n_conv,
                   "published_reference_s": LOS_PUBLISHED_S},
        "p99_ms": round(res["fp64"]["call_p99_s"] * 1e3, 4),
        "kernel_ms": round(res["fp64"]["kernel_ms"], 4),
        "mixed": {"value": round(res["mixed"]["call_s"], 6), "kernel_ms": round(res["mixed"]["kernel_ms"], 4),
                  "max_err_rel_to_coordinate_scale_vs_fp64": mixed_err},
        "accuracy": {"drop_in_equals_device_call": res["fp64"]["drop_in_equal"] and res["mixed"]["drop_in_equal"]},
        "roofline": {"bound": "valu", "kernel": "raytrace_kernel<double> (multi-plane)", "achieved": round(achieved, 3),
                     "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP64_TFLOPS, 4),
                     "traffic": None, "duration": "mean raytrace launch over the timed steps, HIP events on its "
                     "stream", "work": "per call %.3g flops + 20 x %.3g special-function evaluations" % (F, S)},
    }
    if not args.no_cpu:
        from oracle import lensim_oracle as O

        order = np.argsort(np.array(z))
        rs = LensModelBulk(["TNFW", "CONVERGENCE"], multi_plane=True).prepare_ray_shooting_kwargs(
            names, kw, z_source=cfg["z_source"], lens_redshift_list=z)
        rng = np.random.default_rng(0)
        sel = rng.choice(n_ray, 2000, replace=False)
        t0 = time.perf_counter()
        wx, wy = O.multi_plane_ray_shooting(x[sel], y[sel], [names[i] for i in order], [kw[i] for i in order],
                                            rs["T_ij_list"], rs["T_z_list"], rs["reduced2physical_factor"])
        dt = time.perf_counter() - t0
        line["cpu_baseline"] = dict(value=dt * n_ray / sel.size, unit="s", cores=1, kind="port",
                                    sample="2000 of the 160000 rays through the numpy oracle's plane scan "
                                           "(%.1f s), scaled to the full grid" % dt)
        err = max(float(np.max(np.abs(res["fp64"]["bx"][sel] - wx))), float(np.max(np.abs(res["fp64"]["by"][sel] -
                                                                                         wy)))) / scale
        line["accuracy"]["fp64_max_err_rel_to_coordinate_scale_vs_oracle"] = err
    print(json.dumps(line), flush=True)


def main_los_grad(args):
    """logL and its gradient with respect to every deflector and source
    parameter of the line-of-sight scene (workloads.los_halo_draw: 5130 TNFW +
    126 CONVERGENCE sheets on 126 planes, z_source 2.5) for a SERSIC_ELLIPSE
    source on a 100 x 100 grid (0.2"/pixel, no supersampling, 5x5 Moffat PSF):
    jax.grad of likelihood_data_given_model through
    LensModelBulk(multi_plane=True) (MultiPlaneBulk.ray_shooting,
    LensModel/MultiPlane/multi_plane_bulk.py:152-212, differentiated in reverse).
    value: milliseconds per call of one image, FP64; the mixed form and a batch
    of 8 images beside it."""
    import torch

    from jaxtronomy_amd import backend
    from jaxtronomy_amd import workloads as W
    from jaxtronomy_amd.Data.imaging_data import ImageData
    from jaxtronomy_amd.Data.psf import PSF
    from jaxtronomy_amd.ImSim.image_model import ImageModel
    from jaxtronomy_amd.LensModel.lens_model_bulk import LensModelBulk
    from jaxtronomy_amd.LightModel.light_model import LightModel

    names, kw, z = W.los_halo_draw(1)
    ks = [dict(amp=40.0, R_sersic=0.8, n_sersic=1.5, e1=0.1, e2=-0.05, center_x=0.3, center_y=-0.2)]
    npix = 100
    cfg = dict(W.EPL180, numPix=npix, deltaPix=20.0 / npix, supersampling=1)
    kernel = W.moffat_kernel(5, 1.2, 3.0)

    def model(prec, data):
        lm = LensModelBulk(["TNFW", "CONVERGENCE"], multi_plane=True, precision=prec)
        im = ImageModel(ImageData(**W.data_kwargs(cfg, data)), PSF("PIXEL", kernel_point_source=kernel), lm,
                        LightModel(["SERSIC_ELLIPSE"]), kwargs_numerics={"supersampling_factor": 1}, precision=prec)
        rs = lm.prepare_ray_shooting_kwargs(names, kw, z_source=W.LOS["z_source"], lens_redshift_list=z)
        return im, rs

    im0, rs0 = model("fp64", np.zeros((npix, npix)))
    clean = im0.image(rs0, ks).cpu().numpy()
    data = W.noisy(clean, cfg["exposure_time"], cfg["background_rms"], np.random.default_rng(3))
    res = {}
    for prec in ("fp64", "mixed"):
        im, rs = model(prec, data)
        params, _, m, keep = im.pack_bulk(rs, ks)
        for B in (1, 8):
            pt = backend.to_device_f64(np.repeat(params, B, axis=0))
            log("bench: los-grad %s, batch %d (%d steps)" % (prec, B, args.steps))
            for _ in range(max(1, args.warmup)):
                im.likelihood_and_grad_batch(pt, B, model=m)
            torch.cuda.synchronize()
            evs = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
            evs[0].record()
            for i in range(args.steps):
                ll, g = im.likelihood_and_grad_batch(pt, B, model=m)
                evs[i + 1].record()
            torch.cuda.synchronize()
            per = np.array([evs[i].elapsed_time(evs[i + 1]) for i in range(args.steps)])
            res[(prec, B)] = dict(ms=float(np.mean(per)), p99=float(np.percentile(per, 99)),
                                  ll=ll.cpu().numpy(), g=g.cpu().numpy())
        del keep
    g64, g32 = res[("fp64", 1)]["g"], res[("mixed", 1)]["g"]
    scale = np.maximum(np.max(np.abs(g64), axis=(0, 1)), 1e-300)  # per parameter slot, over all profiles
    err = float(np.max(np.max(np.abs(g32 - g64), axis=(0, 1)) / scale))
    ll64, ll32 = res[("fp64", 1)]["ll"], res[("mixed", 1)]["ll"]
    n_par = int(np.count_nonzero(g64[0]))
    line = {
        "metric": "ms per logL + gradient call, line-of-sight multi-plane scene (5130 TNFW + 126 CONVERGENCE "
                  "planes), 100^2, one image, FP64",
        "value": round(res[("fp64", 1)]["ms"], 4),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(res[("fp64", 1)]["ms"], 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: a seeded line-of-sight realisation and mock data from it",
        "config": {"workload": "LensModelBulk(['TNFW', 'CONVERGENCE'], multi_plane=True) + SERSIC_ELLIPSE source, "
                               "100x100 @0.2\", supersampling 1, 5x5 Moffat PSF, logL and its gradient in every "
                               "deflector and source parameter", "batch": 1, "nonzero_gradient_entries": n_par},
        "p99_ms": round(res[("fp64", 1)]["p99"], 4),
        "batch8_ms": round(res[("fp64", 8)]["ms"], 4),
        "mixed": {"ms_per_step": round(res[("mixed", 1)]["ms"], 4), "batch8_ms": round(res[("mixed", 8)]["ms"], 4)},
        "accuracy": {"mixed_grad_max_err_rel_to_slot_scale_vs_fp64": err,
                     "mixed_logl_max_rel_vs_fp64": float(np.max(np.abs(ll32 - ll64) / np.abs(ll64)))},
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=100, help="untimed steps before each timed phase (the GPU's "
                    "clocks settle over the first ~15 ms of sustained load)")
    ap.add_argument("--batch", type=int, default=None,
                    help="batch per step (default: 64 epl180, 16 substructure, 256 grad, as each workload is specified)")
    ap.add_argument("--fp64-steps", type=int, default=60)
    ap.add_argument("--settle", type=float, default=0.5, help="seconds of untimed sustained load before the timed "
                    "phases (epl180)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=("epl180", "substructure", "grad", "serve-mix", "sie60", "los", "los-grad"),
                    default="epl180")
    ap.add_argument("--in-flight", type=int, default=3, help="batches in flight for the headline value (epl180)")
    ap.add_argument("--mix-slots", type=int, default=256, help="resident observations per bucket (serve-mix)")
    ap.add_argument("--n-sub", type=int, default=1000)
    ap.add_argument("--serve-seconds", type=float, default=1.0)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB three times (the reference for HBM counters)")
    args = ap.parse_args()
    if args.batch is None:
        args.batch = {"substructure": 16, "grad": 256}.get(args.workload, 64)
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported (multi-GPU sharding is a design "
              "note in DESIGN.md)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch

    from jaxtronomy_amd import backend

    backend.require_gpu()
    torch.cuda.set_device(0)
    if args.copy_calibration:
        n = args.copy_calibration << 18  # float32 elements
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "launches": 3}), flush=True)
        return
    if args.workload == "substructure":
        return main_substructure(args)
    if args.workload == "grad":
        return main_grad(args)
    if args.workload == "serve-mix":
        return main_serve_mix(args)
    if args.workload == "sie60":
        return main_sie60(args)
    if args.workload == "los":
        return main_los(args)
    if args.workload == "los-grad":
        return main_los_grad(args)
    log("bench: building models")
    cfg, kernel, draws, data, params, params_np, im64, im32 = build_models(args.batch, args.in_flight)
    n_ray = (cfg["numPix"] * cfg["supersampling"]) ** 2
    settle_s, settle_n = clock_settle(im32, params, args.settle)
    log("bench: clocks settled (%.2f s, %d batches)" % (settle_s, settle_n))

    log("bench: timing mixed precision, one batch at a time (%d steps)" % args.steps)
    mixed = time_steps(im32, params, args.steps, args.warmup, backend)
    log("bench: timing mixed precision, %d batches in flight (%d steps)" % (args.in_flight, args.steps))
    mixed_fl = time_in_flight(im32, params, args.steps, args.warmup, backend)
    log("bench: timing fp64, one batch at a time (%d steps)" % args.fp64_steps)
    fp64 = time_steps(im64, params, args.fp64_steps, max(2, args.warmup // 2), backend)
    log("bench: timing fp64, %d batches in flight (%d steps)" % (args.in_flight, args.fp64_steps))
    fp64_fl = time_in_flight(im64, params, args.fp64_steps, max(2, args.warmup // 2), backend)
    fl_diff = max(float(np.max(np.abs(x - mixed["logl"]))) for x in mixed_fl["logl"])
    fl_diff64 = max(float(np.max(np.abs(x - fp64["logl"]))) for x in fp64_fl["logl"])

    # accuracy of the mixed form against the FP64 form of the same batch (the
    # FP64 form itself is held to 1e-10 of the oracle by tests/test_gpu_parity.py)
    ll_rel = float(np.max(np.abs(mixed["logl"] - fp64["logl"]) / np.abs(fp64["logl"])))
    img64 = im64.image_batch(params[:4]).cpu().numpy()
    img32 = im32.image_batch(params[:4]).double().cpu().numpy()
    img_rel = float(max(np.max(np.abs(img32[i] - img64[i])) / np.max(np.abs(img64[i])) for i in range(4)))

    value = args.batch * args.steps / (mixed_fl["total_ms"] / 1e3)
    value64 = args.batch * args.fp64_steps / (fp64_fl["total_ms"] / 1e3)
    alone = args.batch * args.steps / (mixed["total_ms"] / 1e3)
    alone64 = args.batch * args.fp64_steps / (fp64["total_ms"] / 1e3)
    F, S = render_work(params_np, n_ray)
    achieved = (F + 8 * S) / (mixed["render_ms"] / 1e3) / 1e12
    achieved64 = (F + 20 * S) / (fp64["render_ms"] / 1e3) / 1e12
    conv_flops = 2.0 * kernel.size * cfg["numPix"] ** 2 * args.batch
    line = {
        "metric": "images/s + p99 latency, EPL+Sersic 180^2 ss3, served mixed-precision vs plain FP64",
        "value": round(value, 1),
        "unit": "images/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mixed_fl["total_ms"] / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "EPL+SHEAR deflector, SERSIC_ELLIPSE lens light + source, 180x180 @0.04\", "
                        "supersampling 3, 41x41 Moffat PSF, image + chi2 logL",
            "batch": args.batch,
            "precision": "mixed: f32 profiles/convolution, f64 coordinates and chi2 sums",
            "in_flight": "%d batches, on %d streams with a workspace each (value, ms_per_step); "
                         "p50/p99 and one_batch_at_a_time: one batch alone" % (args.in_flight, args.in_flight),
        },
        "p50_ms": round(float(np.percentile(mixed["per_step"], 50)), 4),
        "p99_ms": round(float(np.percentile(mixed["per_step"], 99)), 4),
        "latency": "p50_ms / p99_ms: one batch of 64 alone (HIP events); the served request latencies are under "
                   "serving.mixed / serving.fp64 (p50_ms / p99_ms of the served line at 75% of capacity)",
        "clock_settle": {"seconds": round(settle_s, 3), "batches": settle_n,
                         "note": "untimed sustained load (three batches in flight) before the warm-up steps"},
        "one_batch_at_a_time": {"value": round(alone, 1), "ms_per_step": round(mixed["total_ms"] / args.steps, 4),
                                "p50_ms": round(float(np.percentile(mixed["per_step"], 50)), 4),
                                "p99_ms": round(float(np.percentile(mixed["per_step"], 99)), 4)},
        "fp64": {"value": round(value64, 1), "ms_per_step": round(fp64_fl["total_ms"] / args.fp64_steps, 4),
                 "one_batch_at_a_time": {"value": round(alone64, 1),
                                         "ms_per_step": round(fp64["total_ms"] / args.fp64_steps, 4)},
                 "p99_ms": round(float(np.percentile(fp64["per_step"], 99)), 4),
                 "render_ms": round(fp64["render_ms"], 4), "conv_ms": round(fp64["conv_ms"], 4),
                 "render_tflops_equiv": round(achieved64, 3)},
        "mixed_over_fp64": round(value / value64, 2),
        "accuracy": {"mixed_image_max_rel_vs_fp64": img_rel, "mixed_logl_max_rel_vs_fp64": ll_rel,
                     "in_flight_logl_max_abs_diff_vs_alone": fl_diff, "in_flight_fp64_logl_max_abs_diff": fl_diff64},
        "kernels_ms": {"render": round(mixed["render_ms"], 4), "conv": round(mixed["conv_ms"], 4),
                       "in_flight_render": round(mixed_fl["render_ms"], 4),
                       "in_flight_conv": round(mixed_fl["conv_ms"], 4)},
        "roofline": {"bound": "valu", "kernel": "render_fast_kernel<float, EPL, SHEAR, 1, 1>",
                     "achieved": round(achieved, 3),
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                     "traffic": measured_traffic("render_fast_float_epl180_b64_pmc.txt"),
                     "work": "per launch %.3g flops + 8 x %.3g special-function evaluations" % (F, S),
                     "algorithmic_bytes": 4 * n_ray // cfg["supersampling"] ** 2 * args.batch,
                     "conv_tflops": round(conv_flops / (mixed["conv_ms"] / 1e3) / 1e12, 3),
                     "duration": "mean render launch over the %d timed one-batch-at-a-time steps: HIP events that "
                                 "the library records on the launch stream around the render kernel (lensim_timing_*)"
                                 "; the same command's rocprofv3 kernel trace is profiles/epl180_b64_kernel_trace.txt"
                                 % args.steps,
                     "in_flight_achieved": round((F + 8 * S) / (mixed_fl["render_ms"] / 1e3) / 1e12, 3)},
    }
    if args.serve_seconds > 0:
        log("bench: served load")
        served = {}
        for form, im_, cap in (("mixed", im32, value), ("fp64", im64, value64)):
            pts = serve_measure(im_, params_np, data, cap, args.serve_seconds, (0.25, 0.5, 0.75, 2.0), label=form)
            load = pts[2]
            served[form] = {"saturated_per_s": pts[-1]["achieved_per_s"],
                            "p50_ms": load["p50_ms"], "p99_ms": load["p99_ms"],
                            "stated_load": "75%% of the form's in-flight capacity (%.0f requests/s offered)"
                                           % load["offered_per_s"],
                            "points": pts}
        line["serving"] = {
            "workload": "Poisson arrivals of single log-likelihood requests (same model and data as above) on the "
                        "continuous-batching server: max batch 64, 3 graph lanes, 4 streams; arrivals and "
                        "scheduler in one event loop; offered at 25/50/75% of each form's in-flight capacity and "
                        "at 200% (its achieved rate is the saturated throughput)",
            "mixed": served["mixed"],
            "fp64": served["fp64"],
            "mixed_over_fp64_saturated": round(served["mixed"]["saturated_per_s"] /
                                               served["fp64"]["saturated_per_s"], 2),
        }
    if not args.no_cpu:
        log("bench: cpu baseline")
        cb, ll_cpu = cpu_baseline(cfg, kernel, params_np, data)
        line["cpu_baseline"] = cb
        line["accuracy"]["fp64_logl_max_rel_vs_cpu_oracle"] = float(
            np.max(np.abs(fp64["logl"][:len(ll_cpu)] - ll_cpu) / np.abs(ll_cpu)))
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
