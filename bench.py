#!/usr/bin/env python3
"""Benchmark of the ARMA hot path on one MI355X.

One step = one pass of the hot path over one synthetic surface:
white noise (dynamic-creator MT streams, one per partition) -> 3-D AR
recurrence (persistent wavefront kernel) -> post-rescale -> linear velocity
potential (batched FFTs).  The workload is the AR + linear potential
configuration: propagating-wave ACF exp(-0.06(|t|+|x|+|y|)) cos(0.8t - 0.8x)
on a 10^3 unit grid, Choi coefficients of order 7, surface 1024x256x256,
potential on 256x256 at 16 z levels in [-10, 0], float64.

Prints ONE JSON line (rank 0).  value = (surface points + potential points)
per second of the whole step; the two stage rates are reported beside it.
The roofline object is computed for the kernel with the largest share of the
timed region, from HIP events recorded on the stream the kernels run on.
The CPU baseline is the FP64 oracle (a single-threaded C restatement of the
reference's CPU path) timed on a bounded sample of the same workload.
"""
import argparse
import ctypes as C
import json
import math
import os
import shutil
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3

WORKLOADS = {
    # AR + linear potential, float64 (the configuration the metric is quoted on)
    "ar-potential-1024x256x256-fp64": dict(
        acf=("propagating", 0.8, (-0.8, 0.0)), shape=(1024, 256, 256), order=7,
        partition=(16, 32, 32), seed=0, depth=12.0, lbound=(0.0, -10.0),
        ubound=(1023.0, 0.0), npoints=(1024, 16), precision="fp64"),
    # the same surface under the high-amplitude potential solver
    # (velocity_potential_solver = high_amplitude)
    "ar-potential-ha-1024x256x256-fp64": dict(
        acf=("propagating", 0.8, (-0.8, 0.0)), shape=(1024, 256, 256), order=7,
        partition=(16, 32, 32), seed=0, depth=12.0, lbound=(0.0, -10.0),
        ubound=(1023.0, 0.0), npoints=(1024, 16), precision="fp64", solver="high_amplitude"),
    # the AR + potential step with output=binary: zeta.bin (big-endian) is
    # streamed by the writer while the potential computes; the step ends when
    # the file is written
    "ar-potential-zeta-bin-1024x256x256-fp64": dict(
        acf=("propagating", 0.8, (-0.8, 0.0)), shape=(1024, 256, 256), order=7,
        partition=(16, 32, 32), seed=0, depth=12.0, lbound=(0.0, -10.0),
        ubound=(1023.0, 0.0), npoints=(1024, 16), precision="fp64", write_surface=True),
    # surface with transform = nit: Gram-Charlier target with the parameters of
    # the reference's input/nit-standing-gramcharlier (skewness 3.25, kurtosis
    # 2.4, cdf_solver interval [-5,5], absolute error 1e-6, 100 iterations)
    "ar-nit-1024x256x256-fp64": dict(
        acf=("propagating", 0.8, (-0.8, 0.0)), shape=(1024, 256, 256), order=7,
        partition=(16, 32, 32), seed=0, depth=None, precision="fp64",
        nit=dict(distribution=0, skewness=3.25, kurtosis=2.4, interval=(-5.0, 5.0))),
    # precision variant: FP32 storage, compensated accumulation, FP64 setup
    "ar-potential-2048x512x512-fp32": dict(
        acf=("propagating", 0.8, (-0.8, 0.0)), shape=(2048, 512, 512), order=7,
        partition=(32, 64, 64), seed=1, depth=12.0, lbound=(0.0, -10.0),
        ubound=(2047.0, 0.0), npoints=(2048, 32), precision="fp32"),
    # the same surface and potential in FP64: with --ar plain this is the
    # plain GPU path the FP32 path is compared with at equal size
    "ar-potential-2048x512x512-fp64": dict(
        acf=("propagating", 0.8, (-0.8, 0.0)), shape=(2048, 512, 512), order=7,
        partition=(32, 64, 64), seed=1, depth=12.0, lbound=(0.0, -10.0),
        ubound=(2047.0, 0.0), npoints=(2048, 32), precision="fp64"),
    # precision check on a stable model: the 2048x512x512 FP32 path with the
    # standing-wave ACF (the propagating ACF's Choi coefficients grow without
    # bound, so relative errors there measure the growth, not the precision)
    "ar-standing-2048x512x512-fp32": dict(
        acf=("standing", 0.8, (0.8, 0.8)), shape=(2048, 512, 512), order=7,
        partition=(32, 64, 64), seed=1, depth=12.0, lbound=(0.0, -10.0),
        ubound=(2047.0, 0.0), npoints=(2048, 32), precision="fp32"),
    # many uneven surfaces in one persistent launch: 64 surfaces, seed 42,
    # t ~ U[64,1024], x,y ~ U[32,384], order per surface from {3,5,7,9},
    # propagating ACF with alpha ~ U[0.04,0.1], beta ~ U[0.5,1.2]; FP32,
    # autotuned tile shape / schedule / residency, launches replayed as a graph
    "batch-64-fp32": dict(kind="batch", count=64, seed=42, t=(64, 1024), xy=(32, 384),
                          orders=(3, 5, 7, 9), alpha=(0.04, 0.1), beta=(0.5, 1.2),
                          partition=(32, 48, 48), precision="fp32"),
    # serving: Poisson arrivals at 100 requests/s for 20 s (seed 7); surface
    # volumes log-uniform in [32^3, 512x384x384]; a quarter of the requests
    # are 16-slice chunks of 8 realtime streams; a fifth of the surfaces also
    # ask for the linear potential at 2 z levels; FP32; 200 GB HBM budget
    "serve-100qps-fp32": dict(kind="serve", qps=100.0, duration=20.0, seed=7,
                              volume=(32 ** 3, 512 * 384 * 384), t=(32, 512), xy=(32, 384),
                              stream_share=0.25, streams=8, chunk=16, potential_share=0.2,
                              nz=2, budget=200 << 30, lanes=4, max_batch=32,
                              partition=(32, 48, 48), precision="fp32"),
    # CPU-runnable plain AR case (no potential)
    "ar-standing-128x64x64-fp64": dict(
        acf=("standing", 0.8, (0.8, 0.8)), shape=(128, 64, 64), order=7,
        partition=(16, 16, 16), seed=0, depth=None, precision="fp64"),
}
DEFAULT = "ar-potential-1024x256x256-fp64"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def setup_model(arma, w, ctx):
    """FP64 setup, as the reference generator does before do_generate: the
    closed-form ACF on the host, the Choi solve and the white-noise variance
    in FP64 on the GPU."""
    kind, velocity, beta = w["acf"]
    func = arma._abi.ACF_STANDING_WAVE if kind == "standing" else arma._abi.ACF_PROPAGATING_WAVE
    acf = arma._abi.acf(func, (10, 10, 10), velocity=velocity, alpha=(0.06, 0.06, 0.06), beta=beta)
    phi, _ = ctx.yule_walker_choi(acf, w["order"])
    varwn = ctx.white_noise_variance(acf, phi)
    return acf, phi, varwn


BATCH_CANDIDATES = 32  # ACF parameter draws per batch member, drawn up front


def batch_members(arma, w, ctx):
    """The surfaces of a batch workload, drawn from a seeded generator: the
    shape and order of every member, then a fixed list of ACF parameter
    candidates.  A member takes its first candidate whose Choi coefficients
    give a positive white-noise variance (the reference rejects the others
    with std::invalid_argument).  The Choi solves of all undecided members
    run together, one GPU launch per candidate round."""
    import numpy as np
    rng = np.random.default_rng(w["seed"])
    specs = []
    for _ in range(w["count"]):
        nt = int(rng.integers(w["t"][0], w["t"][1] + 1))
        nx = int(rng.integers(w["xy"][0], w["xy"][1] + 1))
        ny = int(rng.integers(w["xy"][0], w["xy"][1] + 1))
        order = int(rng.choice(w["orders"]))
        cands = [(float(rng.uniform(*w["alpha"])), float(rng.uniform(*w["beta"])))
                 for _ in range(BATCH_CANDIDATES)]
        specs.append((nt, nx, ny, order, cands))

    def acf_of(a, b):
        return arma._abi.acf(arma._abi.ACF_PROPAGATING_WAVE, (10, 10, 10), velocity=b,
                             alpha=(a, a, a), beta=(-b, 0.0))

    out = [None] * len(specs)
    for k in range(BATCH_CANDIDATES):
        todo = [i for i in range(len(specs)) if out[i] is None]
        if not todo:
            break
        acfs = [acf_of(*specs[i][4][k]) for i in todo]
        sols = ctx.yule_walker_choi_many(acfs, [specs[i][3] for i in todo])
        for i, acf, sol in zip(todo, acfs, sols):
            if sol is None:
                continue
            phi = sol[0]
            varwn = ctx.white_noise_variance(acf, phi)
            if varwn > 0:
                nt, nx, ny, order, cands = specs[i]
                out[i] = dict(shape=(nt, nx, ny), order=order, alpha=cands[k][0],
                              beta=cands[k][1], phi=phi, varwn=varwn, gamma0=float(acf[0, 0, 0]))
    if any(m is None for m in out):
        raise RuntimeError("a batch member found no stationary ACF among its candidates")
    return out


def run_batch(args, w):
    """Batching workload: noise, recurrence and rescale of every surface,
    captured once as a graph and replayed per step.  value = surface points
    per second.  Per-kernel times come from an eager pass of the same work
    with per-kernel events, right after the timed replays."""
    import numpy as np
    import torch
    import arma_amd
    from arma_amd import _abi
    t0 = time.time()
    ctx = arma_amd.Context(0, torch_stream=False)
    mem = batch_members(arma_amd, w, ctx)
    setup_s = time.time() - t0
    part = w["partition"]
    nstreams = [math.prod(math.ceil(m["shape"][d] / part[d]) for d in range(3)) for m in mem]
    params = _abi.dcmt_p521_params(max(nstreams))
    log(f"bench: {len(mem)} surfaces set up in {setup_s:.2f}s, "
        f"{sum(nstreams)} noise streams")
    prec = arma_amd.FP32 if w["precision"] == "fp32" else arma_amd.FP64
    tdt = torch.float32 if prec == arma_amd.FP32 else torch.float64
    descs = [ctx.ar_desc(m["shape"], m["phi"], m["varwn"], params,
                         seeds=np.full(nstreams[i], w["seed"] + i, np.uint32),
                         noise_mode=arma_amd.NOISE_PARTITIONED, partition=part, precision=prec,
                         algorithm=arma_amd.AR_WAVEFRONT) for i, m in enumerate(mem)]
    batch = ctx.batch(descs, [m["gamma0"] for m in mem], prec)
    arena = torch.empty(batch.arena_elements, dtype=tdt, device="cuda:0")
    torch.cuda.synchronize()
    t0 = time.time()
    trials = batch.tune(arena, reps=1)
    cfg, grid = batch.config()
    log(f"bench: tuned in {time.time() - t0:.1f}s over {len(trials)} candidates: "
        f"tile_x {cfg[0]}, schedule {cfg[1]}, slots/CU {cfg[2]}, grid {grid}")
    batch.capture(arena)
    for i in range(args.warmup):
        batch.replay()
        ctx.synchronize()
        log(f"bench: warmup step {i + 1}/{args.warmup} done")
    t_start = time.perf_counter()
    for i in range(args.steps):
        batch.replay()
        ctx.synchronize()
        log(f"bench: step {i + 1}/{args.steps} done")
    t_total = time.perf_counter() - t_start
    # eager pass with per-kernel events
    _abi.check(_abi.lib().arma_hip_profile(ctx.h, 1))
    parse_profile(ctx)
    t_e = time.perf_counter()
    for i in range(args.steps):
        batch.generate(arena)
        ctx.synchronize()
    t_eager = time.perf_counter() - t_e
    prof = parse_profile(ctx)
    _abi.check(_abi.lib().arma_hip_profile(ctx.h, 0))
    pts = sum(math.prod(m["shape"]) for m in mem)
    flops = sum(2.0 * m["phi"].size * math.prod(m["shape"]) for m in mem)
    esize = 4 if prec == arma_amd.FP32 else 8
    kernels = {}
    for name, (us, cnt) in prof.items():
        mean_s = us / cnt * 1e-6
        e = {"mean_us": us / cnt, "launches": cnt, "share": us / 1e6 / t_eager}
        if name == "ar_batch":
            e.update(bound="valu", unit="TFLOP/s", achieved=flops / mean_s / 1e12,
                     peak=FP32_PEAK_TFLOPS if prec == arma_amd.FP32 else FP64_PEAK_TFLOPS)
        elif name == "noise_batch":
            e.update(bound="hbm", unit="GB/s", achieved=pts * esize / mean_s / 1e9,
                     peak=HBM_PEAK_GBS)
        kernels[name] = e
    dom = max(kernels, key=lambda k: kernels[k]["share"])
    k = kernels[dom]
    roofline = None
    if "achieved" in k:
        roofline = {"bound": k["bound"], "achieved": round(k["achieved"], 3), "peak": k["peak"],
                    "unit": k["unit"], "frac": round(k["achieved"] / k["peak"], 4),
                    "traffic": None, "kernel": dom}
    ms = 1000.0 * t_total / args.steps
    line = {
        "metric": "surface pts/s + potential pts/s (FP32 vs FP64) at RMS<=1e-5; p50/p99 latency",
        "value": pts / (t_total / args.steps), "unit": "pts/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if prec == arma_amd.FP32 else "f64", "data": "synthetic",
        "config": {"workload": args.workload, "surfaces": len(mem), "points": pts,
                   "orders": {str(o): sum(1 for m in mem if m["order"] == o)
                              for o in w["orders"]},
                   "noise_streams": sum(nstreams), "tile_x": cfg[0], "schedule": cfg[1],
                   "slots_per_cu": cfg[2], "grid": grid, "graph": True,
                   "parallelism": "single-gpu"},
        "eager_ms_per_step": 1000.0 * t_eager / args.steps,
        "setup_s": setup_s,
        "tuning": [{"tile_x": c[0], "schedule": c[1], "slots_per_cu": c[2], "grid": g,
                    "ar_ms": round(t, 3)} for c, g, t in trials],
        "roofline": roofline, "kernels": kernels, "cpu_baseline": None,
    }
    batch.close()
    ctx.close()
    return line


def serve_requests(w):
    """The request trace of a serving workload: arrival times and shapes."""
    import numpy as np
    rng = np.random.default_rng(w["seed"])
    t, out = 0.0, []
    lo, hi = math.log(w["volume"][0]), math.log(w["volume"][1])
    while True:
        t += float(rng.exponential(1.0 / w["qps"]))
        if t > w["duration"]:
            break
        if rng.random() < w["stream_share"]:
            out.append(dict(t=t, stream=int(rng.integers(w["streams"]))))
            continue
        v = math.exp(rng.uniform(lo, hi))
        pot = bool(rng.random() < w["potential_share"])
        # aspect uniform on the feasible set: log nt, then log nx, the rest is ny
        tl, th = w["t"]
        xl, xh = w["xy"]
        nt_lo, nt_hi = max(tl, v / (xh * xh)), min(th, v / (xl * xl))
        nt = int(round(math.exp(rng.uniform(math.log(nt_lo), math.log(nt_hi)))))
        rest = v / nt
        nx_lo, nx_hi = max(xl, rest / xh), min(xh, rest / xl)
        nx = int(round(math.exp(rng.uniform(math.log(nx_lo), math.log(max(nx_lo, nx_hi))))))
        ny = int(min(xh, max(xl, round(rest / nx))))
        if pot:  # the FFT potential takes power-of-two (x, y) extents
            nx = 1 << int(round(math.log2(min(256, max(32, nx)))))
            ny = 1 << int(round(math.log2(min(256, max(32, ny)))))
        out.append(dict(t=t, shape=(nt, nx, ny), potential=pot))
    return out


def run_serve(args, w, qps=None, duration=None):
    """Serving workload: a client thread submits the trace at its arrival
    times; the main thread polls completions and releases them at once.
    value = (surface + potential points) per second from the first arrival
    to the last completion; latency = completion - arrival.  qps and
    duration override the workload's."""
    w = dict(w)
    if qps is not None:
        w["qps"] = qps
    if duration is not None:
        w["duration"] = duration
    import threading
    import numpy as np
    import arma_amd
    from arma_amd import _abi
    trace = serve_requests(w)
    wl = dict(acf=("propagating", 0.8, (-0.8, 0.0)), order=7)
    sctx = arma_amd.Context(0, torch_stream=False)
    acf, phi, varwn = setup_model(arma_amd, wl, sctx)
    sctx.close()
    gamma0 = float(acf[0, 0, 0])
    part = w["partition"]
    prec = arma_amd.FP32
    ctx_desc = arma_amd.Context  # descriptors only: built with the static helper
    params = _abi.dcmt_p521_params(4096)
    rng = np.random.default_rng(w["seed"] + 1)
    log(f"bench: {len(trace)} requests over {w['duration']} s")
    srv = arma_amd.Server(0, lanes=w["lanes"], budget=w["budget"], precision=prec,
                          max_batch=w["max_batch"])
    sids = []
    for i in range(w["streams"]):
        nx, ny = (int(v) for v in rng.integers(w["xy"][0], w["xy"][1] + 1, 2))
        ncol = math.ceil(nx / part[1]) * math.ceil(ny / part[2])
        d = ctx_desc.ar_desc((1, nx, ny), phi, varwn, [params[j] for j in range(ncol)],
                             seeds=[10000 + 1000 * i], noise_mode=arma_amd.NOISE_PARTITIONED,
                             partition=(w["chunk"], part[1], part[2]), precision=prec,
                             algorithm=arma_amd.AR_WAVEFRONT)
        sids.append((srv.open_stream(d, w["chunk"]), nx * ny))
    descs = []
    for i, r in enumerate(trace):
        if "shape" not in r:
            descs.append(None)
            continue
        n = math.prod(math.ceil(r["shape"][k] / part[k]) for k in range(3))
        d = ctx_desc.ar_desc(r["shape"], phi, varwn, [params[j % 4096] for j in range(n)],
                             seeds=np.full(n, i, np.uint32),
                             noise_mode=arma_amd.NOISE_PARTITIONED, partition=part,
                             precision=prec, algorithm=arma_amd.AR_WAVEFRONT)
        pd = None
        if r["potential"]:
            pd = ctx_desc.pot_desc(r["shape"], (2 * math.pi,) * 3, 12.0, (0.0, -10.0),
                                   (float(r["shape"][0] - 1), 0.0), (r["shape"][0], w["nz"]),
                                   prec)
        descs.append((d, pd))
    pts = {}
    errors = []

    def client(t0):
        for i, r in enumerate(trace):
            delay = t0 + r["t"] - arma_amd.Server.clock()
            if delay > 0:
                time.sleep(delay)
            try:
                if descs[i] is None:
                    sid, area = sids[r["stream"]]
                    rid = srv.submit_chunk(sid, arrival=t0 + r["t"])
                    pts[rid] = w["chunk"] * area
                else:
                    d, pd = descs[i]
                    rid = srv.submit_surface(d, gamma0, pd, arrival=t0 + r["t"])
                    n = math.prod(r["shape"])
                    pts[rid] = n + (n * w["nz"] if pd is not None else 0)
            except Exception as e:
                errors.append(str(e))

    t0 = arma_amd.Server.clock() + 0.5
    th = threading.Thread(target=client, args=(t0,))
    th.start()
    lat, sizes, done_pts, t_last, last_log = [], [], 0, t0, t0
    while len(lat) + len(errors) < len(trace):
        for c in srv.poll(256, wait_ms=1000):
            if c.status != 0:
                errors.append(f"request {c.id} status {c.status}")
            else:
                lat.append(c.t_done - c.t_arrival)
                sizes.append(c.batch_size)
                t_last = max(t_last, c.t_done)
                done_pts += pts.get(c.id, 0)
            srv.release(c.id)
        now = arma_amd.Server.clock()
        if now - last_log > 5:
            log(f"bench: {len(lat)} of {len(trace)} requests served")
            last_log = now
    th.join()
    st = srv.stats()
    for sid, _ in sids:
        srv.close_stream(sid)
    srv.close()
    lat = np.array(lat) * 1000.0
    wall = t_last - t0
    line = {
        "metric": "surface pts/s + potential pts/s (FP32 vs FP64) at RMS<=1e-5; p50/p99 latency",
        "value": done_pts / wall, "unit": "pts/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": wall * 1000.0, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": args.workload, "qps_offered": w["qps"],
                   "duration_s": w["duration"], "requests": len(trace),
                   "lanes": w["lanes"], "max_batch": w["max_batch"],
                   "hbm_budget_bytes": w["budget"], "parallelism": "single-gpu"},
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "mean": float(lat.mean()), "max": float(lat.max())},
        "qps_served": len(lat) / wall, "errors": errors[:5], "n_errors": len(errors),
        "batch_size": {"mean": float(np.mean(sizes)), "max": int(np.max(sizes))},
        "server": st, "roofline": None, "cpu_baseline": None,
    }
    return line




def cpu_threads():
    """Host threads this process may use: OMP_NUM_THREADS when set (the share
    of the host a GPU comes with), else the CPUs it is allowed to run on."""
    try:
        n = int(os.environ.get("OMP_NUM_THREADS", "0"))
    except ValueError:
        n = 0
    return n if n > 0 else len(os.sched_getaffinity(0))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def sample_shape(w, big=True):
    """The bounded sample the CPU oracle runs: the leading slices of the
    surface and the potential of the first time slices.  big: the CPU
    baseline's sample (8 partition layers, 64 potential slices, about 5 s of
    one core on the 1024x256x256 surface); else the accuracy sample (2
    partition layers, 4 potential slices)."""
    nt, nx, ny = w["shape"]
    layers, tp = (8, 64) if big else (2, 4)
    ts = min(nt, layers * w["partition"][0] if nx * ny >= 65536 else nt)
    if w.get("depth") is not None:
        tp = min(tp, ts - 1, w["npoints"][0])
    return ts, tp


def errors(got, ref):
    """Max and RMS error relative to the reference's max and RMS magnitude.
    When either array holds non-finite values (an unstable model overflowing
    its precision), the errors cover the leading slices along axis 0 where
    both are finite, and the count of those slices is reported."""
    import numpy as np
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    out = {}
    fin = np.isfinite(got) & np.isfinite(ref)
    if not fin.all():
        ok = fin.reshape(fin.shape[0], -1).all(axis=1)
        k = int(np.argmin(ok)) if not ok.all() else len(ok)
        out.update(non_finite=int((~fin).sum()), finite_slices=k)
        got, ref = got[:k], ref[:k]
    if got.size == 0 or not np.any(ref):
        out.update(max_rel=None, rms_rel=None)
        return out
    d = got - ref
    out.update(max_rel=float(np.max(np.abs(d)) / np.max(np.abs(ref))),
               rms_rel=float(np.sqrt(np.mean(d * d) / np.mean(ref * ref))))
    return out


def oracle_sample(w, phi, varwn, params, seeds, ts, tp, threads):
    """The FP64 oracle on the sample: partitioned noise and the recurrence
    over the partition DAG for the leading ts slices (causal in t, so they
    depend on the leading noise only), then the potential of the first tp
    time slices.  threads > 1 runs the OpenMP variants (partitions, DAG
    levels and z levels in parallel, as the reference's framework=openmp
    path does); the values do not depend on it.  Returns the surface, the
    potential (or None) and the two wall times."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    nt, nx, ny = w["shape"]
    part = w["partition"]
    sample = (ts, nx, ny)
    nstreams = math.prod(math.ceil(sample[d] / part[d]) for d in range(3))
    plist = [O.MtParams(*params[i].as_tuple()) for i in range(nstreams)]
    t0 = time.perf_counter()
    eps = O.noise_partitioned(plist, seeds[:nstreams], sample, part, math.sqrt(varwn),
                              threads=threads)
    zeta = O.ar_generate_dag(eps, phi, part, threads=threads)
    t_surf = time.perf_counter() - t0
    pot, t_pot = None, 0.0
    if w.get("depth") is not None:
        nz = w["npoints"][1]
        args = ((2 * math.pi,) * 3, w["depth"], (0.0, w["lbound"][1]),
                (float(tp - 1), w["ubound"][1]), (tp, nz))
        solver = w.get("solver", "linear")
        t0 = time.perf_counter()
        if threads > 1 and solver == "linear":
            pot = O.linear_potential_omp(zeta[:tp + 1], *args, threads)
        else:
            pot = O.linear_potential(zeta[:tp + 1], *args, solver=solver)
        t_pot = time.perf_counter() - t0
    return zeta, pot, t_surf, t_pot


def gpu_sample(ctx, w, desc, zeta, phi_out, prec, solver, ts, tp):
    """The GPU's own output on the sample (untimed): the surface before the
    post-rescale, and the potential of the first time slices."""
    ctx.ar_generate(desc, zeta)
    out = {"zeta": zeta[:ts].double().cpu().numpy()}
    if w.get("depth") is not None:
        nz = w["npoints"][1]
        sdesc = ctx.pot_desc(w["shape"], (2 * math.pi,) * 3, w["depth"], (0.0, w["lbound"][1]),
                             (float(tp - 1), w["ubound"][1]), (tp, nz), prec)
        ctx.linear_potential(sdesc, zeta, phi_out, solver=solver)
        out["potential"] = phi_out[:tp].double().cpu().numpy()
    return out


def cpu_baseline(w, phi, varwn, params, seeds, gsample):
    """The oracle timed on a bounded sample of the workload, on one thread
    and on every host thread this process may use (OpenMP), on the same
    inputs; its output is the FP64 reference for the GPU's output on the
    sample (same seeds, before the post-rescale)."""
    import numpy as np
    nt, nx, ny = w["shape"]
    ts, tp = sample_shape(w, big=True)
    threads = cpu_threads()
    z1, p1, s1, q1 = oracle_sample(w, phi, varwn, params, seeds, ts, tp, 1)
    zn, pn, sn, qn = oracle_sample(w, phi, varwn, params, seeds, ts, tp, threads)
    same = bool(np.array_equal(z1, zn) and (p1 is None or np.array_equal(p1, pn)))
    surf_pts = ts * nx * ny
    pot_pts = tp * w["npoints"][1] * nx * ny if p1 is not None else 0

    def rates(ts_, tp_):
        r = {"surface_pts_per_s": surf_pts / ts_, "value": (surf_pts + pot_pts) / (ts_ + tp_)}
        if pot_pts:
            r["potential_pts_per_s"] = pot_pts / tp_
        return r

    one, alln = rates(s1, q1), rates(sn, qn)
    desc = f"noise+AR on {ts}x{nx}x{ny}"
    if pot_pts:
        desc += f"; potential of {tp} time slices x {w['npoints'][1]} z levels"
    acc = {"surface": errors(gsample["zeta"][:ts], z1)}
    if pot_pts:
        acc["potential"] = errors(gsample["potential"][:tp], p1)
    acc["against"] = "FP64 CPU oracle, same seeds, on the sample"
    return {"value": alln["value"], "unit": "pts/s", "cores": threads, "kind": "port",
            "sample": desc, "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "build": "gcc -O3 -fopenmp, no fast-math",
            "all_core": {**alln, "threads": threads},
            "single_thread": {**one, "threads": 1},
            "same_values_all_core": same, "accuracy": acc}


def kernel_source_sha(name):
    import hashlib
    path = os.path.join(ROOT, "arma_amd", "csrc", name)
    with open(path, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()[:16]


def profiled_traffic(workload, kernel):
    """HBM bytes per launch of a kernel from the committed counter summary
    (profiles/hbm_traffic_*.txt: FETCH_SIZE and WRITE_SIZE runs scaled by a
    known-size copy).  Each figure there carries the hash of its kernel's
    source file; when that source has changed since, the figure is stale and
    None is returned with the reason."""
    import glob
    for path in glob.glob(os.path.join(ROOT, "profiles", "hbm_traffic_*.txt")):
        rel = os.path.relpath(path, ROOT)
        for line in open(path):
            f = line.split()
            if len(f) >= 4 and f[0] == "traffic" and f[1] == workload and f[2] == kernel:
                if len(f) < 6:
                    return None, f"{rel}: no source hash recorded"
                if kernel_source_sha(f[4]) != f[5]:
                    return None, f"{rel}: {f[4]} changed since it was measured"
                return int(f[3]), (f"{rel} (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE; "
                                   f"{f[4]} {f[5]})")
    return None, "not measured for this workload"


def parse_profile(ctx):
    from arma_amd import _abi
    buf = C.create_string_buffer(1 << 16)
    _abi.check(_abi.lib().arma_hip_profile_report(ctx.h, buf, len(buf), 1))
    out = {}
    for line in buf.value.decode().splitlines():
        parts = line.split()
        if len(parts) == 4 and parts[0] == "prfl":
            out[parts[1]] = (float(parts[2].rstrip("us")), int(parts[3]))
    return out


def run_single(name, steps, warmup, ar="wavefront", cpu=True, accuracy=False):
    """One surface workload: `steps` timed steps of noise + AR + rescale (+
    NIT) (+ potential) after `warmup` untimed ones.  cpu: time the oracle on
    a sample (the CPU baseline); accuracy: only compare the GPU with the
    oracle on the small accuracy sample."""
    import numpy as np
    import torch
    import arma_amd
    from arma_amd import _abi

    w = WORKLOADS[name]
    fp32 = w["precision"] == "fp32"
    prec = arma_amd.FP32 if fp32 else arma_amd.FP64
    tdt = torch.float32 if fp32 else torch.float64
    esize = 4 if fp32 else 8
    nt, nx, ny = w["shape"]
    log(f"bench: workload {name}")
    ctx = arma_amd.Context(0)
    acf, phi, varwn = setup_model(arma_amd, w, ctx)
    log(f"bench: phi shape {phi.shape}, white-noise variance {varwn:.6g}")
    part = w["partition"]
    nstreams = math.prod(math.ceil(w["shape"][d] / part[d]) for d in range(3))
    t0 = time.time()
    params = arma_amd._abi.dcmt_p521_params(nstreams, min(16, os.cpu_count() or 1))
    log(f"bench: {nstreams} MT parameter sets ready in {time.time() - t0:.1f}s")
    seeds = np.full(nstreams, w["seed"], np.uint32)

    algo = arma_amd.AR_WAVEFRONT if ar == "wavefront" else arma_amd.AR_PLAIN
    desc = ctx.ar_desc(w["shape"], phi, varwn, params, seeds=seeds,
                       noise_mode=arma_amd.NOISE_PARTITIONED, partition=part, precision=prec,
                       algorithm=algo)
    zeta = torch.empty(w["shape"], dtype=tdt, device="cuda:0")
    has_pot = w.get("depth") is not None
    solver = w.get("solver", "linear")
    nz = 0
    phi_out = None
    if has_pot:
        nz = w["npoints"][1]
        pdesc = ctx.pot_desc(w["shape"], (2 * math.pi,) * 3, w["depth"], w["lbound"],
                             w["ubound"], w["npoints"], prec)
        phi_out = torch.empty((w["npoints"][0], nz, nx, ny), dtype=tdt, device="cuda:0")
    gamma0 = float(acf[0, 0, 0])
    nit = None
    if w.get("nit"):
        spec = dict(w["nit"])
        nit = arma_amd.nit_desc(spec.pop("distribution"), **spec)

    ev = {k: (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for k in ("surface", "potential")}
    stage_ms = {"surface": 0.0, "potential": 0.0}
    writes = {"bytes": 0, "busy_s": 0.0, "steps": 0}
    out_dir = (tempfile.mkdtemp(prefix="arma_bench_", dir=os.environ.get("ARMA_BENCH_OUT_DIR"))
               if w.get("write_surface") else None)

    def step(timed):
        if timed:
            ev["surface"][0].record()
        ctx.ar_generate(desc, zeta)
        ctx.rescale(zeta, w["shape"], prec, gamma0)
        if nit is not None:
            ctx.nit_realisation(zeta, w["shape"], prec, gamma0, nit)
        if timed:
            ev["surface"][1].record()
        writer = None
        if out_dir:
            # zeta.bin streams out while the potential computes
            # a new file per step: truncating the last one would be timed too
            writes["n"] = writes.get("n", 0) + 1
            writer = arma_amd.Writer(ctx, os.path.join(out_dir, f"zeta{writes['n']}.bin"), esize)
            writer.append(zeta)
        if has_pot:
            if timed:
                ev["potential"][0].record()
            ctx.linear_potential(pdesc, zeta, phi_out, solver=solver)
            if timed:
                ev["potential"][1].record()
        if writer is not None:
            nbytes, busy = writer.close()
            if timed:
                writes["bytes"] += nbytes
                writes["busy_s"] += busy
                writes["steps"] += 1

    for i in range(warmup):
        step(False)
        torch.cuda.synchronize()
        log(f"bench: warmup step {i + 1}/{warmup} done")
    _abi.check(_abi.lib().arma_hip_profile(ctx.h, 1))
    parse_profile(ctx)  # drop anything recorded before the timed region
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for i in range(steps):
        step(True)
        torch.cuda.synchronize()
        stage_ms["surface"] += ev["surface"][0].elapsed_time(ev["surface"][1])
        if has_pot:
            stage_ms["potential"] += ev["potential"][0].elapsed_time(ev["potential"][1])
        log(f"bench: step {i + 1}/{steps} done")
    torch.cuda.synchronize()
    t_total = time.perf_counter() - t_start
    prof = parse_profile(ctx)
    _abi.check(_abi.lib().arma_hip_profile(ctx.h, 0))
    if out_dir:
        shutil.rmtree(out_dir, ignore_errors=True)

    ms_per_step = 1000.0 * t_total / steps
    surf_pts = nt * nx * ny
    pot_pts = (w["npoints"][0] * nz * nx * ny) if has_pot else 0
    value = (surf_pts + pot_pts) / (t_total / steps)
    surf_rate = surf_pts / (stage_ms["surface"] / 1000.0 / steps)
    pot_rate = pot_pts / (stage_ms["potential"] / 1000.0 / steps) if has_pot else None

    # per-kernel roofline figures (algorithmic work per launch / mean launch time)
    taps = int(np.prod(phi.shape))
    kernels = {}
    for kname, (us, cnt) in prof.items():
        mean_s = us / cnt * 1e-6
        entry = {"mean_us": us / cnt, "launches": cnt, "share": us / 1e3 / (t_total * 1e3)}
        if kname == "ar_wavefront" or kname == "ar_plain":
            # FP64/FP32 vector FMAs (no matrix cores): the VALU peak
            flops = 2.0 * taps * surf_pts
            peak = FP32_PEAK_TFLOPS if fp32 else FP64_PEAK_TFLOPS
            entry.update(bound="valu", unit="TFLOP/s", achieved=flops / mean_s / 1e12, peak=peak)
        elif kname == "noise_partitioned" or kname == "noise_sequential":
            entry.update(bound="hbm", unit="GB/s", achieved=surf_pts * esize / mean_s / 1e9,
                         peak=HBM_PEAK_GBS)
        elif kname in ("pot_inv_rows", "pot_inv_cols"):
            # each pass of the 2-D inverse produces one phi-sized array per chunk
            chunk_pts = pot_pts / cnt * steps
            entry.update(bound="hbm", unit="GB/s", achieved=chunk_pts * esize / mean_s / 1e9,
                         peak=HBM_PEAK_GBS)
        elif kname == "rescale_scale":
            entry.update(bound="hbm", unit="GB/s", achieved=2 * surf_pts * esize / mean_s / 1e9,
                         peak=HBM_PEAK_GBS)
        kernels[kname] = entry
    dom = max(kernels, key=lambda k: kernels[k]["share"]) if kernels else None
    roofline = None
    if dom and "achieved" in kernels[dom]:
        k = kernels[dom]
        traffic, src = profiled_traffic(name, dom)
        roofline = {"bound": k["bound"], "achieved": round(k["achieved"], 3), "peak": k["peak"],
                    "unit": k["unit"], "frac": round(k["achieved"] / k["peak"], 4),
                    "traffic": traffic, "traffic_source": src, "kernel": dom}

    cpu_line = acc = None
    if cpu or accuracy:
        ts, tp = sample_shape(w, big=cpu)
        gs = gpu_sample(ctx, w, desc, zeta, phi_out, prec, solver, ts, tp)
        if cpu:
            log("bench: timing the CPU oracle on a sample (one thread, then all threads)")
            try:
                cpu_line = cpu_baseline(w, phi, varwn, params, seeds, gs)
            except Exception as e:  # report, never fail the bench line for it
                cpu_line = {"error": str(e)}
        else:
            log("bench: the FP64 oracle on the accuracy sample")
            zr, pr, _, _ = oracle_sample(w, phi, varwn, params, seeds, ts, tp, cpu_threads())
            acc = {"surface": errors(gs["zeta"], zr),
                   "sample": f"first {ts} slices before the post-rescale"}
            if pr is not None:
                acc["potential"] = errors(gs["potential"], pr)
                acc["sample"] += f"; potential of {tp} time slices x {nz} levels"
            acc["against"] = "FP64 CPU oracle, same seeds"

    line = {
        "metric": "surface pts/s + potential pts/s (FP32 vs FP64) at RMS<=1e-5; p50/p99 latency",
        "value": value, "unit": "pts/s", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32" if fp32 else "f64",
        "data": "synthetic",
        "config": {"workload": name, "surface": list(w["shape"]),
                   "phi_shape": list(phi.shape), "partition": list(part),
                   "noise_streams": nstreams, "ar_kernel": ar,
                   "potential": list(w["npoints"]) + [nx, ny] if has_pot else None,
                   "solver": solver if has_pot else None,
                   "transform": "nit gram_charlier" if nit is not None else "none",
                   "setup": "FP64 Yule-Walker (Choi) and white-noise variance on the GPU",
                   "parallelism": "single-gpu"},
        "surface_pts_per_s": surf_rate, "potential_pts_per_s": pot_rate,
        "stage_ms_per_step": {k: v / steps for k, v in stage_ms.items()},
        "zeta_bin_writer": ({"bytes_per_step": writes["bytes"] / writes["steps"],
                             "writer_busy_ms_per_step": 1e3 * writes["busy_s"] / writes["steps"],
                             "writer_gb_per_s": writes["bytes"] / writes["busy_s"] / 1e9}
                            if writes["steps"] else None),
        "roofline": roofline, "kernels": kernels, "cpu_baseline": cpu_line,
    }
    if acc is not None:
        line["accuracy"] = acc
    del zeta, phi_out
    ctx.close()
    torch.cuda.empty_cache()
    return line


def brief(line, keys=("value", "ms_per_step", "surface_pts_per_s", "potential_pts_per_s",
                      "stage_ms_per_step", "accuracy")):
    out = {"workload": line["config"]["workload"], "dtype": line["dtype"],
           "steps": line["steps"], "warmup": line["warmup"]}
    out.update({k: line[k] for k in keys if k in line})
    k = line.get("kernels", {})
    out["kernels_mean_us"] = {n: round(v["mean_us"], 1) for n, v in k.items()}
    return out


def precision_leg():
    """The precision configuration inside the default run: the FP32 path
    (compensated sums, FP64 setup and twiddles) on 2048x512x512 with the
    potential at 32 levels, its accuracy against the FP64 oracle (the
    configured propagating model, and the standing-wave model of the same
    size, which stays finite), and the FP64 wavefront path at the same size
    for the FP32/FP64 ratio."""
    r32 = run_single("ar-potential-2048x512x512-fp32", 3, 1, accuracy=True)
    st = run_single("ar-standing-2048x512x512-fp32", 1, 1, accuracy=True)
    r64 = run_single("ar-potential-2048x512x512-fp64", 2, 1)
    out = {"fp32": brief(r32), "fp32_standing_wave": brief(st), "fp64_wavefront": brief(r64),
           "fp32_over_fp64_wavefront": r64["ms_per_step"] / r32["ms_per_step"]}
    plain = os.path.join(ROOT, "profiles", "ar_plain_2048x512x512_fp64.txt")
    if os.path.exists(plain):
        for line in open(plain):
            f = line.split()
            if len(f) == 2 and f[0] == "ms_per_step":
                out["fp64_plain"] = {"ms_per_step": float(f[1]),
                                     "source": os.path.relpath(plain, ROOT)}
                out["fp32_over_fp64_plain"] = float(f[1]) / r32["ms_per_step"]
    return out


def serving_leg(args):
    """Serving inside the default run: a shortened 100-QPS trace, then rising
    offered loads until p99 latency is at least twice the 100-QPS p99."""
    w = WORKLOADS["serve-100qps-fp32"]
    base = run_serve(args, w, duration=6.0)
    p99 = base["latency_ms"]["p99"]

    def row(r):
        return {"qps_offered": r["config"]["qps_offered"], "qps_served": r["qps_served"],
                "p50_ms": r["latency_ms"]["p50"], "p99_ms": r["latency_ms"]["p99"],
                "pts_per_s": r["value"], "errors": r["n_errors"]}

    sweep, doubled = [row(base)], None
    for q in (200.0, 400.0, 800.0, 1600.0):
        r = run_serve(args, w, qps=q, duration=3.0)
        sweep.append(row(r))
        log(f"bench: serving at {q:.0f} req/s: p99 {r['latency_ms']['p99']:.1f} ms")
        if r["latency_ms"]["p99"] >= 2.0 * p99 or r["n_errors"]:
            doubled = q
            break
    return {"workload": "serve-100qps-fp32 (6 s trace)", "latency_ms": base["latency_ms"],
            "qps_served": base["qps_served"], "pts_per_s": base["value"],
            "batch_size": base["batch_size"], "n_errors": base["n_errors"],
            "p99_doubles_at_qps": doubled, "sweep": sweep}


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default=DEFAULT, choices=sorted(WORKLOADS))
    ap.add_argument("--ar", default="wavefront", choices=["wavefront", "plain"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="default workload only: skip the precision and serving legs")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="after the run, one device-to-device copy of this many MiB "
                         "(a known-size kernel for scaling hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the AR path shards by "
              "x-bands and the potential by time slices, neither is built for several GPUs here.",
              file=sys.stderr)
        sys.exit(2)

    w = WORKLOADS[args.workload]
    if w.get("kind") == "batch":
        line = run_batch(args, w)
    elif w.get("kind") == "serve":
        line = run_serve(args, w)
    else:
        line = run_single(args.workload, args.steps, args.warmup, args.ar,
                          cpu=not args.no_cpu_baseline)
        if args.workload == DEFAULT and args.ar == "wavefront" and not args.no_extras:
            line["precision"] = precision_leg()
            line["serving"] = serving_leg(args)
    print(json.dumps(line), flush=True)
    if args.copy_calibration_mb:
        import torch
        n = args.copy_calibration_mb << 20
        src = torch.ones(n // 4, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"bench: copied {args.copy_calibration_mb} MiB device to device")


if __name__ == "__main__":
    main()
