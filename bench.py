#!/usr/bin/env python3
"""Throughput of the batched contact solve on one MI355X.

Workload (one step): a new case for each of 2048 Hertzian steady-rolling contact patches --
grid sides drawn log-uniform in [20, 256], Hertz semi-axis ratio log-uniform in [0.2, 5],
Fn uniform 20-150 kN, creepages uniform +-0.005, spin uniform +-0.5 rad/m, friction uniform
0.2-0.4, seed 2024 -- set through the cntc_* calls and submitted at once to the
continuous-batching loop (cntc_submit), which runs all of them through shared batched launches
(NormCG + GDsteady with the SteadyGS fall-back, eps 1e-8) with FP32 FFT products inside the FP64
refinement.

Reported beside the throughput:
* the plain FP64 path (one cntc_calculate per patch) on a systematic subset, and the force
  error of the mixed-precision batch against it;
* per-patch latency (batch submit to result) p50/p99;
* the dominant kernel's achieved bandwidth from HIP events over the timed steps;
* the CPU restatement timed on the host cores (a baseline, not the target).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import ctypes
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor, FIRST_COMPLETED, wait

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# The engine runs SteadyGS launches (one work-group per patch, seconds long on the largest
# patches) on side streams beside the main stream; with HIP's default of 4 hardware queues per
# process only 3 of them run at once.  Eight queues let 7 side streams overlap (set before the
# HIP runtime starts; an explicit larger setting is kept).
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < 8:
    os.environ["GPU_MAX_HW_QUEUES"] = "8"

IRE_MIXED, IRE_FP64 = 1, 301


def ids(k, base):
    return base + k // 9, 1 + k % 9


HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def kernel_name(op):
    conv = {"conv32_fwd": "k_rowfwd<float>", "conv32_mul": "k_colmul<float>",
            "conv32_inv": "k_rowinv<float>", "conv_fwd": "k_rowfwd<double>",
            "conv_mul": "k_colmul<double>", "conv_inv": "k_rowinv<double>",
            "stdygs_staged": "k_stdygs2", "stdygs_overlap": "k_stdygs3"}
    return conv.get(op, "k_" + op)


def roofline_of(name, sec, amount, nl, total_sec):
    """HBM roofline for the FFT and element kernels (amount = algorithmic bytes); FP64 roofline
    for SteadyGS (amount = algorithmic flops of its direct influence sums)."""
    common = {"kernel": kernel_name(name), "avg_launch_us": round(sec / nl * 1e6, 2),
              "launches": nl, "share_of_device_time": round(sec / total_sec, 3)}
    if name.startswith("stdygs"):
        achieved = amount / sec / 1e12
        r = {"bound": "mfma", "achieved": round(achieved, 4), "peak": FP64_PEAK_TFLOPS,
             "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 5), "traffic": None,
             "flops_per_launch": round(amount / nl),
             "note": "FP64 FMA work of the direct sums (vector FMA; MI355X FP64 vector and matrix peaks are equal)"}
    else:
        achieved = amount / sec / 1e9
        r = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
             "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
             "bytes_per_launch": round(amount / nl)}
    r.update(common)
    return r


# ---- CPU baseline: the oracle on the host cores ------------------------------------------

def cpu_leg(cases, seconds, cores, max_elems=3000, seed=7):
    """Solve seeded-random patches of the batch (N <= max_elems) with the CPU restatement on
    `cores` threads for about `seconds`; returns rates and the forces of the solved patches."""
    from oracle import pyoracle as O
    O.lib()
    rng = np.random.default_rng(seed)
    order = [int(k) for k in rng.permutation(len(cases))
             if cases[k]["mx"] * cases[k]["my"] <= max_elems]

    def one(k):
        c = cases[k]
        P = O.Problem()
        inp, h = O.case_input(c)
        t0 = time.perf_counter()
        rc = P.calculate(inp, h)
        dt = time.perf_counter() - t0
        itnorm, ittang, itcg, itgs = P.iters()
        return k, rc, dt, c["mx"] * c["my"] * (itcg + itgs), P.forces()

    done = []
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=cores) as ex:
        pending, pos = set(), 0
        while pos < len(order) and len(pending) < cores:
            pending.add(ex.submit(one, order[pos]))
            pos += 1
        while pending:
            fin, pending = wait(pending, return_when=FIRST_COMPLETED)
            done.extend(f.result() for f in fin)
            while (pos < len(order) and len(pending) < cores
                   and time.perf_counter() - t0 < seconds):
                pending.add(ex.submit(one, order[pos]))
                pos += 1
    wall = time.perf_counter() - t0
    ok = [d for d in done if d[1] == 0]
    return dict(wall=wall, n=len(ok), cpu_seconds=sum(d[2] for d in ok),
                elem_iters=float(sum(d[3] for d in ok)),
                forces={d[0]: d[4] for d in ok}, max_elems=max_elems)


# ---- GPU legs -----------------------------------------------------------------------------

def new_case(cntc, ire, icp, c):
    # the first set-call after a calculate starts a new case (src/test_caddon.f90:28-34)
    cntc.setnormalforce(ire, icp, c["fn"])
    cntc.setcreepages(ire, icp, c["cksi"], c["ceta"], c["cphi"])


def device_sync():
    hip = ctypes.CDLL("libamdhip64.so")
    rc = hip.hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError("hipDeviceSynchronize failed: %d" % rc)


def serve_leg(cntc, W, seconds, rate, nchan=64, seed=77, ire0=801, sens_frac=0.1):
    """Serving: `nchan` wheel channels with fixed grids from the batch's size distribution,
    each stepping T=2 transient rolling (dq = dx, warm start from its previous tractions),
    requests as Poisson arrivals at `rate` per second on random channels with creepages as
    seeded random walks.  A fraction `sens_frac` of the requests also asks the sensitivities
    (CNTC_ic_sens = 3: one extra NormCG and three extra tangential solves).  A channel's
    request waits for the channel's previous one.  Latency = arrival to result."""
    rng = np.random.default_rng(seed)
    chans = W.hertz_batch(n=nchan, seed=seed)
    ids = [(ire0 + k // 9, 1 + k % 9) for k in range(nchan)]
    state = []
    for (ire, icp), c in zip(ids, chans):
        c = dict(c)
        c.update(tang=2, eps=1e-8, maxgs=500, maxin=50, pvtime=2, iestim=0)
        W.apply_case(cntc, ire, icp, c)
        state.append(c)
    # first step of every channel (initiation), not timed
    err = cntc.calculatebatch([i for i, _ in ids], [j for _, j in ids])
    if np.any(err != 0):
        raise SystemExit("serving warm-up failed: %s" % sorted(set(err)))
    for c in state:
        c.update(pvtime=0, iestim=3)
    cntc.servestart(200e9)
    n_req = max(1, int(seconds * rate))
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_req))
    chan_of = rng.integers(0, nchan, n_req)
    walk = rng.normal(0.0, 2e-4, (n_req, 2))
    want_sens = rng.random(n_req) < sens_frac
    waiting = [[] for _ in range(nchan)]
    inflight = {}
    lat, is_sens, t0, nxt, done = [], [], time.perf_counter(), 0, 0

    def start(ch, t_arr, r):
        c = state[ch]
        c["cksi"] = float(np.clip(c["cksi"] + walk[r, 0], -0.005, 0.005))
        c["ceta"] = float(np.clip(c["ceta"] + walk[r, 1], -0.005, 0.005))
        ire, icp = ids[ch]
        cntc.setcreepages(ire, icp, c["cksi"], c["ceta"], c["cphi"])
        cntc.setflags(ire, icp, [1988], [3 if want_sens[r] else 1])
        t = cntc.submit(ire, icp)
        if t <= 0:
            raise SystemExit("serving submit failed: %d" % t)
        inflight[t] = (ch, t_arr, r)

    while done < n_req:
        now = time.perf_counter() - t0
        while nxt < n_req and arrivals[nxt] <= now:
            ch = int(chan_of[nxt])
            busy = any(v[0] == ch for v in inflight.values())
            if busy:
                waiting[ch].append((arrivals[nxt], nxt))
            else:
                start(ch, arrivals[nxt], nxt)
            nxt += 1
        for t in list(inflight):
            d, rc, _ = cntc.poll(t)
            if d != 1:
                continue
            ch, t_arr, r = inflight.pop(t)
            if rc != 0:
                raise SystemExit("serving request failed: %d" % rc)
            lat.append(time.perf_counter() - t0 - t_arr)
            is_sens.append(bool(want_sens[r]))
            if want_sens[r]:
                S = cntc.getsensitivities(*ids[ch])
                if not (np.all(np.isfinite(S)) and S[0, 0] > 0.0 and S[1, 1] < 0.0):
                    raise SystemExit("serving sensitivities out of range: %s" % S)
            done += 1
            if waiting[ch]:
                ta, r = waiting[ch].pop(0)
                start(ch, ta, r)
        time.sleep(0.0002)
    wall = time.perf_counter() - t0
    lat = np.array(lat)
    is_sens = np.array(is_sens)
    pct = lambda a, q: round(float(np.percentile(a, q)) * 1e3, 2) if len(a) else None
    return {"workload": "%d channels, T=2 transient rolling, Poisson arrivals, %d%% with "
                        "sensitivities" % (nchan, round(100 * sens_frac)),
            "offered_rate": rate, "requests": n_req, "achieved_rate": round(n_req / wall, 1),
            "latency_ms_p50": pct(lat, 50), "latency_ms_p99": pct(lat, 99),
            "sens_requests": int(is_sens.sum()),
            "sens_latency_ms_p50": pct(lat[is_sens], 50), "sens_latency_ms_p99": pct(lat[is_sens], 99),
            "elements_per_channel_mean": round(float(np.mean([c["mx"] * c["my"] for c in chans])), 1)}


def forces_of(cntc, idl):
    return np.array([cntc.getcontactforces(i, j) for i, j in idl])


def solve_batch(cntc, idl, label):
    """Submit every problem of `idl` to the serving loop at once and wait for all of them,
    reporting progress as patches complete.  Returns (error codes, latencies in s)."""
    t0 = time.perf_counter()
    tick = {}
    for k, (i, j) in enumerate(idl):
        t = cntc.submit(i, j)
        if t <= 0:
            raise SystemExit("%s: submit failed (%d)" % (label, t))
        tick[t] = k
    err = np.zeros(len(idl), dtype=np.int64)
    lat = np.zeros(len(idl))
    last = t0
    while tick:
        for t in list(tick):
            d, rc, l = cntc.poll(t)
            if d == 1:
                k = tick.pop(t)
                err[k], lat[k] = rc, l
        now = time.perf_counter()
        if now - last >= 15.0 and tick:
            last = now
            print("# %s: %d/%d patches done at %.0f s" % (label, len(idl) - len(tick), len(idl), now - t0),
                  file=sys.stderr, flush=True)
        time.sleep(0.002)
    return err, lat


def rel_err(a, b):
    """Per-component relative error |a - b| / |b| (Fn, Fx, Fy)."""
    return np.abs(a[:, :3] - b[:, :3]) / np.maximum(np.abs(b[:, :3]), 1e-300)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--patches", type=int, default=2048)
    ap.add_argument("--plain", type=int, default=32, help="patches in the one-at-a-time subset")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-cores", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=4.0)
    ap.add_argument("--serve-rate", type=float, default=250.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: patches are independent, so a multi-GPU run would "
              "split the batch by patch with no exchange; --gpus %d is not supported" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    from contact_amd import workloads as W
    cases_fp64 = W.hertz_batch(n=args.patches, seed=2024, eps=1e-8, precision=0)
    cases = W.hertz_batch(n=args.patches, seed=2024, eps=1e-8, precision=1)
    npatch = len(cases)
    n_elem = np.array([c["mx"] * c["my"] for c in cases], dtype=np.float64)

    # CPU leg first: it runs threads only and must not share the process with an active GPU run
    cpu = None
    if not args.no_cpu:
        print("# cpu baseline: %d threads, %.0f s budget" % (args.cpu_cores, args.cpu_seconds),
              file=sys.stderr, flush=True)
        cpu = cpu_leg(cases_fp64, args.cpu_seconds, args.cpu_cores)
        print("# cpu baseline: %d patches in %.1f s" % (cpu["n"], cpu["wall"]), file=sys.stderr,
              flush=True)

    from contact_amd import cntc
    cntc.lib()
    mixed = [ids(k, IRE_MIXED) for k in range(npatch)]
    fp64 = [ids(k, IRE_FP64) for k in range(npatch)]
    for k in range(npatch):
        W.apply_case(cntc, *mixed[k], cases[k])

    # all solves go through the continuous-batching loop: a batch is every patch submitted at once
    cntc.servestart(200e9)

    def step(label):
        for k in range(npatch):
            new_case(cntc, *mixed[k], cases[k])
        e, l = solve_batch(cntc, mixed, label)
        if np.any(e != 0):
            raise SystemExit("mixed batch: %d patches failed (codes %s)" % (np.sum(e != 0), sorted(set(e))))
        return l

    t0 = time.perf_counter()
    for w in range(args.warmup):
        step("warmup %d" % w)
    t_first = time.perf_counter() - t0

    cntc.amdprofile(True)
    device_sync()
    t0 = time.perf_counter()
    for s in range(args.steps):
        lat = step("step %d" % s)
        print("# step %d done at %.2f s" % (s, time.perf_counter() - t0), file=sys.stderr, flush=True)
    device_sync()
    elapsed = time.perf_counter() - t0
    prof = cntc.amdprofileget()
    cntc.amdprofile(False)

    FM = forces_of(cntc, mixed)
    its = np.array([cntc.getiterations(i, j) for i, j in mixed])
    elem_iters = float(np.sum(n_elem * (its[:, 1] + its[:, 3])))
    flags = its[:, 4]

    # plain FP64 path: one cntc_calculate per patch on a systematic subset
    sub = list(range(0, npatch, max(1, npatch // max(1, args.plain))))[:args.plain] if args.plain > 0 else []
    print("# plain fp64 subset: %d patches" % len(sub), file=sys.stderr, flush=True)
    t_plain = []
    FP = []
    for k in sub:
        W.apply_case(cntc, *fp64[k], cases_fp64[k])
        t1 = time.perf_counter()
        rc = cntc.calculate(*fp64[k])
        t_plain.append(time.perf_counter() - t1)
        if rc != 0:
            raise SystemExit("plain solve of patch %d failed: %d" % (k, rc))
        FP.append(cntc.getcontactforces(*fp64[k]))
    t_plain = np.array(t_plain) if sub else np.array([np.nan])
    # the accuracy of the mixed-precision batch: against the plain FP64 solves of the subset
    e_mixed = rel_err(FM[sub], np.array(FP)) if sub else np.array([[np.nan]])

    serving = None
    if args.serve_seconds > 0:
        print("# serving leg", file=sys.stderr, flush=True)
        serving = serve_leg(cntc, W, args.serve_seconds, args.serve_rate)
    cntc.servestop()

    value = npatch * args.steps / elapsed
    plain_rate = len(sub) / float(np.sum(t_plain)) if sub else float("nan")

    # dominant kernel (device time) over the timed steps, and the dominant one on the main stream
    kernel_sec = sum(v[0] for v in prof.values())
    name, (sec, amt, nl) = max(prof.items(), key=lambda kv: kv[1][0])
    roofline = roofline_of(name, sec, amt, nl, kernel_sec)
    main = {k: v for k, v in prof.items() if not k.startswith("stdygs")}
    roofline_main = None
    if main:
        mname, (msec, mamt, mnl) = max(main.items(), key=lambda kv: kv[1][0])
        roofline_main = roofline_of(mname, msec, mamt, mnl, kernel_sec)
    top = sorted(prof.items(), key=lambda kv: -kv[1][0])[:8]
    kernels = [{"kernel": kernel_name(k), "seconds": round(v[0], 3), "launches": v[2]} for k, v in top]

    cpu_baseline = None
    if cpu is not None and cpu["n"] > 0:
        per_patch = elem_iters / npatch            # element-iterations per patch of this mix
        rate = cpu["elem_iters"] / cpu["cpu_seconds"] * args.cpu_cores
        cpu_baseline = {
            "value": round(rate / per_patch, 4), "unit": "patches/s", "cores": args.cpu_cores,
            "kind": "port",
            "sample": ("CPU restatement (oracle/cntc_oracle.c, FP64) on %d threads for %.1f s: %d "
                       "seeded-random patches of the batch with N<=%d, %.3g element-iterations "
                       "in %.1f thread-seconds; patches/s of the full mix = that element-iteration "
                       "rate / the mix's %.3g element-iterations per patch (GPU count)"
                       % (args.cpu_cores, cpu["wall"], cpu["n"], cpu["max_elems"],
                          cpu["elem_iters"], cpu["cpu_seconds"], per_patch)),
        }
        ks = sorted(cpu["forces"])
        FO = np.array([cpu["forces"][k] for k in ks])
        cpu_baseline["mixed_vs_oracle_force_err_max"] = float(np.max(rel_err(FM[ks], FO)))

    out = {
        "metric": "contact patches solved/s at 1e-6 force err vs FP64 plain; p50/p99 latency",
        "value": round(value, 2),
        "unit": "patches/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32",
        "data": "synthetic",
        "config": {
            "workload": "hertz_batch_2048_mixed_precision" if npatch == 2048 else "hertz_batch_%d_mixed_precision" % npatch,
            "patches": npatch, "elements": int(n_elem.sum()), "grid_sides": [20, 256],
            "solver": "NormCG + GDsteady, eps 1e-8", "precision": "FP32 FFT products in FP64 refinement",
            "seed": 2024,
        },
        "force_err_max": float(np.max(e_mixed)),
        "force_err_p99": float(np.percentile(np.max(e_mixed, axis=1), 99)),
        "force_err_ok": bool(np.max(e_mixed) <= 1e-6),
        "force_err_what": "max over Fn, Fx, Fy of |mixed - FP64 plain| / |FP64 plain|, %d-patch systematic subset" % len(sub),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 2),
                       "p99": round(float(np.percentile(lat, 99)) * 1e3, 2),
                       "what": "batch submit to patch result, last timed step"},
        "steadygs_fallbacks": int(np.sum(flags >= 2)),
        "element_iterations_per_s": round(elem_iters * args.steps / elapsed, 0),
        "plain_fp64": {"patches_per_s": round(plain_rate, 3), "patches": len(sub),
                       "latency_ms_p50": round(float(np.percentile(t_plain, 50)) * 1e3, 2),
                       "latency_ms_p99": round(float(np.percentile(t_plain, 99)) * 1e3, 2)},
        "speedup_vs_plain_fp64": round(value / plain_rate, 2),
        "first_call_patches_per_s": round(npatch * max(1, args.warmup) / t_first, 2) if args.warmup else None,
        "device_time_per_wall": round(kernel_sec / elapsed, 3),
        "kernels": kernels,
        "serving": serving,
        "roofline": roofline,
        "roofline_main_stream": roofline_main,
        "cpu_baseline": cpu_baseline,
    }
    def clean(v):
        if isinstance(v, dict):
            return {k: clean(x) for k, x in v.items()}
        if isinstance(v, list):
            return [clean(x) for x in v]
        if isinstance(v, float) and not np.isfinite(v):
            return None
        return v

    print(json.dumps(clean(out)), flush=True)


if __name__ == "__main__":
    main()
