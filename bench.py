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
* the dominant kernel's achieved flops or bandwidth from HIP events over the timed steps
  (events only around the SteadyGS, GD probe and FFT product kernels), with its HBM traffic
  per launch from the committed rocprofv3 counter summary;
* the CPU restatement timed on the host cores, sampled by size stratum (a baseline, not the
  target);
* a serving leg: 64 transient-rolling channels with Poisson arrivals.

A wall-clock budget (--budget, default 480 s) bounds the whole run: warm-up and timed steps
stop after the last whole step that fits, and the JSON line reports the steps actually run
next to the steps requested.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--budget SECONDS]
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

# The engine runs SteadyGS launches (one work-group per patch, seconds long on the largest
# patches) on side streams beside the main stream; with HIP's default of 4 hardware queues per
# process only 3 of them run at once.  Eight queues let 7 side streams overlap (set before the
# HIP runtime starts; an explicit larger setting is kept).
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < 8:
    os.environ["GPU_MAX_HW_QUEUES"] = "8"

IRE_MIXED, IRE_FP64 = 1, 700       # mixed-precision sets of 228 result elements, then FP64


def side_streams():
    """The engine's SteadyGS side streams: CNTC_SIDE_STREAMS, else 7, at most one less than
    the hardware queues."""
    hwq = max(2, int(os.environ.get("GPU_MAX_HW_QUEUES", "8")))
    v = int(os.environ.get("CNTC_SIDE_STREAMS", "0"))
    return max(1, min(hwq - 1, v)) if v > 0 else min(7, hwq - 1)


def ids(k, base):
    return base + k // 9, 1 + k % 9


HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def kernel_name(op):
    conv = {"conv32_fwd": "k_rowfwd<float>", "conv32_mul": "k_colmul<float>",
            "conv32_inv": "k_rowinv<float>", "conv_fwd": "k_rowfwd<double>",
            "conv_mul": "k_colmul<double>", "conv_inv": "k_rowinv<double>",
            "stdygs_staged": "k_stdygs2", "stdygs_overlap": "k_stdygs3",
            "stdygs_block": "steadygs sweeps by blocks (k_gsb_walk, FFT products, k_gsb_check)"}
    return conv.get(op, "k_" + op)


def roofline_of(name, sec, amount, nl, total_sec):
    """HBM roofline for the FFT and element kernels (amount = algorithmic bytes); FP64 roofline
    for SteadyGS (amount = algorithmic flops of its direct influence sums)."""
    common = {"kernel": kernel_name(name), "avg_launch_us": round(sec / nl * 1e6, 2),
              "launches": nl, "share_of_device_time": round(sec / total_sec, 3)}
    if name.startswith("stdygs") or name == "gsb_walk":
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

def _cpu_worker(cases, order, q):
    """Solve the patches of `order` one after another with the CPU restatement, reporting the
    start and the outcome of each on `q` (a separate process: the parent ends it at the
    deadline, possibly in the middle of a patch)."""
    from oracle import pyoracle as O
    O.lib()
    for k in order:
        c = cases[k]
        q.put(("start", k, time.perf_counter()))
        P = O.Problem()
        inp, h = O.case_input(c)
        t0 = time.perf_counter()
        rc = P.calculate(inp, h)
        dt = time.perf_counter() - t0
        itnorm, ittang, itcg, itgs = P.iters()
        q.put(("done", k, rc, dt, c["mx"] * c["my"] * (itcg + itgs), tuple(P.forces())))


def cpu_leg(cases, seconds, cores, nstrata=8, seed=7):
    """The CPU restatement on `cores` processes for about `seconds`, sampled by size stratum.

    The patches are sorted by element count and split into `nstrata` strata of equal count;
    cores/nstrata processes work through seeded-random patches of each stratum until the
    deadline.  A patch still running at the deadline counts with its elapsed time as a lower
    bound of its solve time, so a stratum whose patches outlast the sample (the largest ones)
    gets a lower bound of its mean time and the mix rate is an upper bound of the CPU rate.
    Mean time per patch of the mix = the mean over the strata (equal counts)."""
    import multiprocessing as mp
    ctx = mp.get_context("fork")
    rng = np.random.default_rng(seed)
    n_elem = np.array([c["mx"] * c["my"] for c in cases])
    ranked = np.argsort(n_elem, kind="stable")
    strata = np.array_split(ranked, nstrata)
    per = max(1, cores // nstrata)
    q = ctx.Queue()
    procs, who = [], []
    for si, st in enumerate(strata):
        order = [int(k) for k in rng.permutation(st)]
        for w in range(per):
            pr = ctx.Process(target=_cpu_worker, args=(cases, order[w::per], q), daemon=True)
            pr.start()
            procs.append(pr)
            who.append(si)
    t0 = time.perf_counter()
    running, done = {}, []
    deadline = t0 + seconds
    while time.perf_counter() < deadline:
        try:
            m = q.get(timeout=max(0.01, min(0.5, deadline - time.perf_counter())))
        except Exception:
            continue
        if m[0] == "start":
            running[m[1]] = m[2]
        else:
            running.pop(m[1], None)
            done.append(m[1:])
    t_end = time.perf_counter()
    for pr in procs:
        pr.kill()
    for pr in procs:
        pr.join()
    while True:     # messages sent before the kill
        try:
            m = q.get_nowait()
        except Exception:
            break
        if m[0] == "done":
            running.pop(m[1], None)
            done.append(m[1:])
    stratum_of = {int(k): si for si, st in enumerate(strata) for k in st}
    rows = []
    for si, st in enumerate(strata):
        fin = [d for d in done if stratum_of[d[0]] == si and d[1] == 0]
        cen = [t_end - t for k, t in running.items() if stratum_of[k] == si]
        tsum = sum(d[2] for d in fin) + sum(cen)
        cnt = len(fin) + len(cen)
        rows.append({"stratum": si, "elements": [int(n_elem[st].min()), int(n_elem[st].max())],
                     "solved": len(fin), "censored": len(cen),
                     "mean_s": round(tsum / cnt, 4) if cnt else None,
                     "lower_bound": bool(cen) and not fin})
    ok = [d for d in done if d[1] == 0]
    means = [r["mean_s"] for r in rows if r["mean_s"] is not None]
    mix = float(np.mean(means)) if len(means) == nstrata else float("nan")
    return dict(wall=t_end - t0, n=len(ok), procs=len(procs), strata=rows, mix_s=mix,
                elem_iters=float(sum(d[3] for d in ok)), cpu_seconds=sum(d[2] for d in ok),
                forces={d[0]: d[4] for d in ok})


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
    # later steps continue from the previous one (P=0) and start from its tractions (I=3)
    for (ire, icp), c in zip(ids, state):
        c.update(pvtime=0, iestim=3)
        cntc.setflags(ire, icp, [1970, 1983], [0, 3])
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


def sens_leg(cntc, W, n=256, seed=31, ire0=920):
    """Sensitivities of small patches: `n` Hertzian steady-rolling patches of at most 2500
    elements (the batch's size distribution below that bound) solved in one batch with
    CNTC_ic_sens = 3, so that one NormCG and three GDsteady perturbation solves per patch run
    their influence products on the dense FP64-MFMA kernel.  Reports the batch's patches/s
    and that kernel's FP64 rate from HIP events on its stream."""
    cases = [c for c in W.hertz_batch(n=4096, seed=seed) if c["mx"] * c["my"] <= 2500][:n]
    ids = [(ire0 + k // 9, 1 + k % 9) for k in range(len(cases))]
    for (ire, icp), c in zip(ids, cases):
        W.apply_case(cntc, ire, icp, dict(c, sens=3))
    cntc.amdprofile(2)
    t0 = time.perf_counter()
    err = cntc.calculatebatch([i for i, _ in ids], [j for _, j in ids])
    dt = time.perf_counter() - t0
    prof = cntc.amdprofileget()
    cntc.amdprofile(False)
    if np.any(err != 0):
        raise SystemExit("sensitivity leg failed: %s" % sorted(set(err)))
    out = {"workload": "%d Hertzian steady-rolling patches of 80-2500 elements with sensitivities "
                       "(S=3), one batch" % len(cases),
           "patches": len(cases), "seconds": round(dt, 3), "patches_per_s": round(len(cases) / dt, 1),
           "elements_mean": round(float(np.mean([c["mx"] * c["my"] for c in cases])), 1)}
    if "dense" in prof:
        sec, flops, nl = prof["dense"]
        achieved = flops / sec / 1e12
        out["roofline_dense"] = {
            "bound": "mfma", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / FP64_PEAK_TFLOPS, 4), "kernel": "k_dense",
            "flops_per_launch": round(flops / nl), "avg_launch_us": round(sec / nl * 1e6, 2), "launches": nl,
            "note": "algorithmic FP64 flops of the dense products (2 n^2 per nonzero 2x2 or normal "
                    "block) over the launch time; v_mfma_f64_16x16x4f64"}
    return out


def rough_table(cntc, W):
    """The single rough patches alone (512^2 FP64, and 1024^2 with FP32 products in the FP64
    refinement): after a first solve on the grid (coefficient spectra built), a second solve
    with every kernel timed by HIP events; prints per kernel the launches, time, algorithmic
    bytes per launch and GB/s against the 8 TB/s HBM peak."""
    for n, prec, ire in ((512, 0, 961), (1024, 1, 962)):
        c = W.rough_patch(n=n, precision=prec)
        W.apply_case(cntc, ire, 1, c)
        if cntc.calculate(ire, 1) != 0:
            raise SystemExit("rough patch %d failed" % n)
        W.apply_case(cntc, ire, 1, c)
        cntc.amdprofile(True)
        t0 = time.perf_counter()
        rc = cntc.calculate(ire, 1)
        wall = time.perf_counter() - t0
        prof = cntc.amdprofileget()
        cntc.amdprofile(False)
        print("== %d^2 %s: rc %d, wall %.3f s, iterations (itnorm, itcg, ittang, itgs, flags) %s"
              % (n, "mixed" if prec else "fp64", rc, wall, cntc.getiterations(ire, 1)))
        print("%-16s %8s %10s %13s %9s %6s" % ("kernel", "launches", "total ms", "alg MB/launch", "GB/s", "frac"))
        for k, (sec, byt, nl) in sorted(prof.items(), key=lambda kv: -kv[1][0]):
            if nl == 0 or sec <= 0:
                continue
            gbs = byt / sec / 1e9
            print("%-16s %8d %10.3f %13.3f %9.1f %6.3f" % (k, nl, sec * 1e3, byt / nl / 1e6, gbs, gbs / HBM_PEAK_GBS))
        print("device time in the timed kernels %.3f s" % sum(v[0] for v in prof.values()), flush=True)


def forces_of(cntc, idl):
    return np.array([cntc.getcontactforces(i, j) for i, j in idl])


def submit_batch(cntc, idl, label):
    """Submit every problem of `idl` to the serving loop; returns the open batch."""
    t0 = time.perf_counter()
    tick = {}
    for k, (i, j) in enumerate(idl):
        t = cntc.submit(i, j)
        if t <= 0:
            raise SystemExit("%s: submit failed (%d)" % (label, t))
        tick[t] = k
    return dict(label=label, t0=t0, tick=tick, n=len(idl))


def finish_batch(cntc, b):
    """Wait for every patch of an open batch, reporting progress as patches complete.
    Returns (error codes, latencies in s, seconds from submit to the last result)."""
    tick, n = b["tick"], b["n"]
    err = np.zeros(n, dtype=np.int64)
    lat = np.zeros(n)
    last = time.perf_counter()
    while tick:
        for t in list(tick):
            d, rc, l = cntc.poll(t)
            if d == 1:
                k = tick.pop(t)
                err[k], lat[k] = rc, l
        now = time.perf_counter()
        if now - last >= 15.0 and tick:
            last = now
            print("# %s: %d/%d patches done at %.0f s" % (b["label"], n - len(tick), n, now - b["t0"]),
                  file=sys.stderr, flush=True)
        time.sleep(0.002)
    if np.any(err != 0):
        raise SystemExit("%s: %d patches failed (codes %s)" % (b["label"], np.sum(err != 0), sorted(set(err))))
    return err, lat, time.perf_counter() - b["t0"]


def rel_err(a, b):
    """Per-component relative error |a - b| / |b| (Fn, Fx, Fy)."""
    return np.abs(a[:, :3] - b[:, :3]) / np.maximum(np.abs(b[:, :3]), 1e-300)


def read_traffic(path=os.path.join(ROOT, "profiles", "traffic_per_dispatch.txt")):
    """HBM bytes per dispatch by kernel from the committed rocprofv3 counter summary:
    FETCH_SIZE and WRITE_SIZE averages (KiB per dispatch) scaled by the factors a plain copy
    of known size gives.  {} when the summary is absent."""
    if not os.path.exists(path):
        return {}
    ff = wf = 1.0
    out = {}
    with open(path) as f:
        for line in f:
            w = line.split()
            if not w or w[0].startswith("#"):
                continue
            if w[0] == "fetch_factor":
                ff = float(w[1])
            elif w[0] == "write_factor":
                wf = float(w[1])
            elif len(w) >= 3:
                out[w[0]] = (float(w[1]), float(w[2]))
    return {k: (f * ff + wr * wf) * 1024.0 for k, (f, wr) in out.items()}


def main():
    t_start = time.perf_counter()
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--patches", type=int, default=2048)
    ap.add_argument("--plain", type=int, default=32, help="patches in the one-at-a-time subset")
    ap.add_argument("--cpu-seconds", type=float, default=25.0)
    ap.add_argument("--cpu-cores", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--inflight", type=int, default=3,
                    help="batches in flight (each on its own set of 2048 problems, 1-3)")
    ap.add_argument("--serve-seconds", type=float, default=2.0)
    ap.add_argument("--rough-table", action="store_true",
                    help="only the single rough patches (512^2 FP64, 1024^2 mixed): per-kernel GB/s")
    ap.add_argument("--sens-patches", type=int, default=256,
                    help="small patches in the sensitivity leg (dense FP64-MFMA products); 0 skips it")
    ap.add_argument("--serve-rate", type=float, default=1000.0)
    ap.add_argument("--budget", type=float, default=480.0,
                    help="wall-clock seconds for the whole run: warm-up and timed steps stop "
                         "after the last whole step that fits, and the JSON line reports the "
                         "steps actually run")
    args = ap.parse_args()
    if args.rough_table:
        from contact_amd import cntc, workloads as W
        rough_table(cntc, W)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: patches are independent, so a multi-GPU run would "
              "split the batch by patch with no exchange; --gpus %d is not supported" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    def left():
        return args.budget - (time.perf_counter() - t_start)

    from contact_amd import workloads as W
    cases_fp64 = W.hertz_batch(n=args.patches, seed=2024, eps=1e-8, precision=0)
    cases = W.hertz_batch(n=args.patches, seed=2024, eps=1e-8, precision=1)
    npatch = len(cases)
    n_elem = np.array([c["mx"] * c["my"] for c in cases], dtype=np.float64)

    # CPU leg first: it forks worker processes, which must happen before the GPU is touched
    cpu = None
    if not args.no_cpu:
        print("# cpu baseline: %d processes, %.0f s sample" % (args.cpu_cores, args.cpu_seconds),
              file=sys.stderr, flush=True)
        cpu = cpu_leg(cases_fp64, args.cpu_seconds, args.cpu_cores)
        print("# cpu baseline: %d patches in %.1f s" % (cpu["n"], cpu["wall"]), file=sys.stderr,
              flush=True)

    from contact_amd import cntc
    cntc.lib()
    # one set of problems per batch in flight, so that a batch can be submitted while the
    # previous ones finish
    nfl = max(1, min(3, args.inflight))
    sets = [[ids(k, IRE_MIXED + 229 * h) for k in range(npatch)] for h in range(nfl)]
    mixed = sets[0]
    for h in range(nfl):
        for k in range(npatch):
            W.apply_case(cntc, *sets[h][k], cases[k])

    # all solves go through the continuous-batching loop: a step is one batch of every patch,
    # submitted at once; the next batch is submitted as soon as this one is (at most two in
    # flight), so the loop stays full while a batch's slowest patches finish
    cntc.servestart(200e9)

    def submit_step(s, label):
        idl = sets[s % nfl]
        for k in range(npatch):       # waits for this set's previous batch, nfl steps back
            new_case(cntc, *idl[k], cases[k])
        return submit_batch(cntc, idl, label)

    # time kept back for the legs after the timed steps (plain FP64 subset, serving, report)
    reserve = 75.0 if args.plain > 0 or args.serve_seconds > 0 or args.sens_patches > 0 else 5.0

    def run_steps(nsteps, label, can_start):
        """Up to `nsteps` batches, nfl in flight; `can_start(done_latencies)` decides whether
        another batch may be submitted.  Returns (steps run, per-batch latencies, last
        batch's per-patch latencies)."""
        open_, done, last_lat = [], [], np.array([np.nan])
        for s in range(nsteps):
            if not can_start(done, s):
                break
            if len(open_) == nfl:
                _, last_lat, tb = finish_batch(cntc, open_.pop(0))
                done.append(tb)
            open_.append(submit_step(s, "%s %d" % (label, s)))
        while open_:
            _, last_lat, tb = finish_batch(cntc, open_.pop(0))
            done.append(tb)
            print("# %s: batch done (%.1f s from submit)" % (label, tb), file=sys.stderr, flush=True)
        return len(done), done, last_lat

    t0 = time.perf_counter()
    # warm-up: both sets once (their coefficient spectra are built on first use); more only
    # while everything asked for still fits the budget
    def warm_ok(done, s):
        if s < nfl:
            return True
        per = (time.perf_counter() - t0) / max(1, s)
        return left() - reserve >= (args.steps + args.warmup - s) * per * 1.1
    warm_done, t_warm, _ = run_steps(max(nfl, args.warmup), "warmup", warm_ok)
    t_first = time.perf_counter() - t0

    cntc.amdprofile(2)
    device_sync()
    t0 = time.perf_counter()

    def timed_ok(done, s):
        # whole batches only: submit another while it can finish inside the budget (a batch
        # takes at most the longest batch latency seen, counted from now)
        if s == 0:
            return True
        est = max(done + t_warm) if (done or t_warm) else 0.0
        if left() - reserve < est * 1.1:
            print("# stopping after %d of %d timed steps: the wall-clock budget (%.0f s) is spent"
                  % (s, args.steps, args.budget), file=sys.stderr, flush=True)
            return False
        return True
    steps_done, t_step, lat = run_steps(args.steps, "step", timed_ok)
    device_sync()
    elapsed = time.perf_counter() - t0
    prof = cntc.amdprofileget()
    cntc.amdprofile(False)
    mixed = sets[(steps_done - 1) % nfl]

    FM = forces_of(cntc, mixed)
    its = np.array([cntc.getiterations(i, j) for i, j in mixed])
    elem_iters = float(np.sum(n_elem * (its[:, 1] + its[:, 3])))
    flags = its[:, 4]

    # plain FP64 path: one cntc_calculate per patch on a systematic subset (while time is left)
    sub_all = list(range(0, npatch, max(1, npatch // max(1, args.plain))))[:args.plain] if args.plain > 0 else []
    print("# plain fp64 subset: up to %d patches" % len(sub_all), file=sys.stderr, flush=True)
    t_plain, FP, sub = [], [], []
    for k in sub_all:
        # keep time for the serving leg
        if left() < (45.0 if args.serve_seconds > 0 else 25.0):
            break
        fp = ids(len(t_plain), IRE_FP64)
        W.apply_case(cntc, *fp, cases_fp64[k])
        t1 = time.perf_counter()
        rc = cntc.calculate(*fp)
        t_plain.append(time.perf_counter() - t1)
        if rc != 0:
            raise SystemExit("plain solve of patch %d failed: %d" % (k, rc))
        FP.append(cntc.getcontactforces(*fp))
        sub.append(k)
    t_plain = np.array(t_plain) if sub else np.array([np.nan])
    # the accuracy of the mixed-precision batch: against the plain FP64 solves of the subset
    e_mixed = rel_err(FM[sub], np.array(FP)) if sub else np.array([[np.nan]])

    serving = None
    if args.serve_seconds > 0 and left() > 15.0:
        print("# serving leg", file=sys.stderr, flush=True)
        serving = serve_leg(cntc, W, args.serve_seconds, args.serve_rate)
    cntc.servestop()
    sens = None
    if args.sens_patches > 0 and left() > 12.0:
        print("# sensitivity leg", file=sys.stderr, flush=True)
        sens = sens_leg(cntc, W, n=args.sens_patches)

    value = npatch * steps_done / elapsed
    plain_rate = len(sub) / float(np.sum(t_plain)) if sub else float("nan")

    # dominant kernel (device time) over the timed steps, and the dominant one on the main stream
    traffic = read_traffic()
    kernel_sec = sum(v[0] for k, v in prof.items() if k != "stdygs_block")

    def roof(name, v):
        r = roofline_of(name, v[0], v[1], v[2], kernel_sec)
        t = traffic.get(r["kernel"])
        r["traffic"] = round(t) if t is not None else None
        r["traffic_unit"] = "HBM bytes per launch (rocprofv3 FETCH_SIZE/WRITE_SIZE, profiles/traffic_per_dispatch.txt)"
        return r

    # "stdygs_block" times whole sweeps of the block variant (walk, FFT products, tests): not a
    # kernel; its walk kernel is recorded on its own as "gsb_walk"
    single = {k: v for k, v in prof.items() if k != "stdygs_block"}
    name, top1 = max(single.items(), key=lambda kv: kv[1][0])
    roofline = roof(name, top1)
    main_k = {k: v for k, v in prof.items() if not k.startswith("stdygs") and not k.startswith("gsb")}
    roofline_main = roof(*max(main_k.items(), key=lambda kv: kv[1][0])) if main_k else None
    fft = {k: v for k, v in prof.items() if k.startswith("conv")}
    roofline_fft = [roof(k, v) for k, v in sorted(fft.items())]
    kernels = [{"kernel": kernel_name(k), "seconds": round(v[0], 3), "launches": v[2]}
               for k, v in sorted(prof.items(), key=lambda kv: -kv[1][0])]

    cpu_baseline = None
    if cpu is not None and np.isfinite(cpu["mix_s"]):
        cpu_baseline = {
            "value": round(cpu["procs"] / cpu["mix_s"], 4), "unit": "patches/s", "cores": cpu["procs"],
            "kind": "port",
            "sample": ("CPU restatement (oracle/cntc_oracle.c, FP64, eps 1e-8) in %d processes for "
                       "%.1f s on the same 2048-patch mix, stratified by element count into %d "
                       "equal-count strata (%d processes each, seeded-random patches); mean solve "
                       "time per patch of the mix = mean over strata; a patch unfinished at the "
                       "deadline counts with its elapsed time, so strata marked lower_bound make "
                       "this an upper bound of the CPU rate"
                       % (cpu["procs"], cpu["wall"], len(cpu["strata"]), cpu["procs"] // len(cpu["strata"]))),
            "strata": cpu["strata"],
            "patches_solved": cpu["n"],
        }
        ks = sorted(cpu["forces"])
        if ks:
            FO = np.array([cpu["forces"][k] for k in ks])
            cpu_baseline["mixed_vs_oracle_force_err_max"] = float(np.max(rel_err(FM[ks], FO)))

    t_step_a = np.array(t_step) if t_step else np.array([np.nan])
    out = {
        "metric": "contact patches solved/s at 1e-6 force err vs FP64 plain; p50/p99 latency",
        "value": round(value, 2),
        "unit": "patches/s",
        "n_gpus": 1,
        "steps": steps_done,
        "warmup": warm_done,
        "steps_requested": args.steps,
        "warmup_requested": args.warmup,
        "budget_s": args.budget,
        "ms_per_step": round(elapsed / max(1, steps_done) * 1e3, 2),
        "batch_latency_ms": {"mean": round(float(np.mean(t_step_a)) * 1e3, 1),
                             "min": round(float(np.min(t_step_a)) * 1e3, 1),
                             "max": round(float(np.max(t_step_a)) * 1e3, 1),
                             "what": "submit of a batch to its last result; up to %d batches in flight" % nfl},
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32",
        "data": "synthetic",
        "config": {
            "workload": "hertz_batch_2048_mixed_precision" if npatch == 2048 else "hertz_batch_%d_mixed_precision" % npatch,
            "patches": npatch, "elements": int(n_elem.sum()), "grid_sides": [20, 256],
            "solver": "NormCG + GDsteady (SteadyGS fall-back), eps 1e-8",
            "precision": "FP32 FFT products in FP64 refinement", "seed": 2024,
            "batches_in_flight": nfl,
            "side_streams": side_streams(),
        },
        "force_err_max": float(np.max(e_mixed)),
        "force_err_p99": float(np.percentile(np.max(e_mixed, axis=1), 99)),
        "force_err_ok": bool(np.max(e_mixed) <= 1e-6),
        "force_err_what": "max over Fn, Fx, Fy of |mixed - FP64 plain| / |FP64 plain|, %d-patch systematic subset" % len(sub),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 2),
                       "p99": round(float(np.percentile(lat, 99)) * 1e3, 2),
                       "what": "batch submit to patch result, last timed step"},
        "steadygs_fallbacks": int(np.sum((flags & 2) != 0)),
        "element_iterations_per_s": round(elem_iters * steps_done / elapsed, 0),
        "plain_fp64": {"patches_per_s": round(plain_rate, 3), "patches": len(sub),
                       "latency_ms_p50": round(float(np.percentile(t_plain, 50)) * 1e3, 2),
                       "latency_ms_p99": round(float(np.percentile(t_plain, 99)) * 1e3, 2)},
        "speedup_vs_plain_fp64": round(value / plain_rate, 2),
        "warmup_patches_per_s": round(npatch * max(1, warm_done) / t_first, 2) if warm_done else None,
        "device_time_per_wall": round(kernel_sec / elapsed, 3),
        "device_time_what": "HIP-event time of the recorded kernels (SteadyGS walk, GD probe, FFT products) / wall",
        "kernels": kernels,
        "serving": serving,
        "sensitivity_leg": sens,
        "roofline": roofline,
        "roofline_main_stream": roofline_main,
        "roofline_fft": roofline_fft,
        "cpu_baseline": cpu_baseline,
        "wall_s": round(time.perf_counter() - t_start, 1),
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
