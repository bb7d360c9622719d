"""Benchmark: MIS solves/sec of the sliced n=300 3-regular network.

One step = contract_slices over every slice of the workload (BATCHED mode)
plus the maximum, i.e. one MIS solve, with the slices already planned-for and
their inputs on the host side of the C ABI (the backend plans and uploads
inside the step).  Workload: random 3-regular graph, n = 300, seed 2, sliced
by the host slicer to sc_target = 25, Tropical Float32, unit weights.

Before the timed steps the autotuner (tbh_autotune) times up to 12 candidate
planner / launch settings on the workload and installs the fastest (untimed,
like a warmup); the choice is printed.

Printed on one JSON line: solves/sec (value), the plain per-slice GPU loop on
the same slices, the dominant kernel's algorithmic throughput against its
roofline (HIP events on the backend's stream), a standalone max-plus GEMM
against the VALU ceiling, the serving legs, the other benchmark
configurations (batched against plain), and the CPU restatement timed on a
bounded sample.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MIS solves/sec + p50/p99 latency; tropical GOP/s vs roofline, n=300 3-regular"
HBM_PEAK_GBS = 8000.0
# MI355X published FP32 vector peak (FMA = 2 flops); the max-plus GEMM is
# quoted against the measured ceiling of its own instruction mix
# (HIPBackend.valu_peak) and, beside it, against this figure.
FP32_VECTOR_TFLOPS = 157.3


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def build_workload(tb, n, seed, sc_target, trials):
    t = time.time()
    g = tb.random_regular_graph(n, 3, seed=seed)
    br = tb.slice_graph(g, None, tb.ContractionTreeSlicer(sc_target=sc_target,
                                                          order_trials=trials), seed=seed)
    log(f"workload: 3-regular n={n} seed={seed} -> {len(br)} slices "
        f"(sc <= {sc_target}) in {time.time() - t:.1f}s")
    return g, br


def cpu_baseline(br, gpu_results, seconds, threads):
    """The oracle (the C restatement of the reference's CPU path: the same
    OMEinsum lowering over a plain max-plus loop nest) timed on a stratified
    sample of the workload's slices, on `threads` host threads and on one.

    Slices are split into four strata by tc (quartiles of the slice count);
    from each stratum slices are drawn at random (seed 0) and contracted --
    on all threads, slices of tc >= 33 one at a time with the threads inside
    each contraction, smaller ones one per thread -- until the stratum's
    share of the time budget is spent (at least one draw per stratum).  A
    stratum's time is its total work (sum of 2^tc) divided by the rate
    (work per second) measured on its sample; the solve's time is the sum
    over strata.  The one-thread figure samples the same way with a quarter
    of the budget, skipping slices that would alone exceed the stratum's
    share but always contracting at least the stratum's smallest slice, so
    every stratum's rate is measured on one thread too.  Every sampled result
    is checked equal to the GPU's."""
    from oracle import oracle
    import __graft_entry__
    tb = __graft_entry__.package()
    tcs = np.array([tb.complexity(b).tc if b.code is not None else -np.inf for b in br])
    live = np.flatnonzero(np.isfinite(tcs))
    work = np.where(np.isfinite(tcs), 2.0 ** tcs, 0.0)
    order = live[np.argsort(tcs[live], kind="stable")]
    strata = [s for s in np.array_split(order, 4) if len(s)]

    def run(nthreads, budget, guide=None):
        rng = np.random.default_rng(0)
        per, sampled, t_all = [], 0, 0.0
        for si, st in enumerate(strata):
            pick = rng.permutation(st)
            if guide is not None:
                # one thread: skip slices whose time alone, predicted from the
                # stratum's rate on all threads over their count, exceeds the
                # stratum's share of the budget; a stratum keeps at least its
                # smallest slice (strata are in ascending tc)
                r1 = guide[si] / threads if guide[si] else None
                keep = [k for k in pick if r1 and work[k] / r1 <= budget / len(strata)]
                pick = keep if keep else [st[0]]
            t_used, w_used, i = 0.0, 0.0, 0
            while i < len(pick) and (i == 0 or t_used < budget / len(strata)):
                chunk = pick[i:i + max(1, nthreads if tcs[pick[i]] < 33 else 1)]
                i += len(chunk)
                t0 = time.perf_counter()
                res = oracle.contract_slices([br[k] for k in chunk], np.float32,
                                             nthreads=nthreads)
                t_used += time.perf_counter() - t0
                w_used += work[chunk].sum()
                if not np.array_equal(res, gpu_results[np.asarray(chunk)]):
                    raise RuntimeError("CPU restatement and GPU differ on a sampled slice")
                sampled += len(chunk)
            per.append({"tc": [round(float(tcs[st[0]]), 2), round(float(tcs[st[-1]]), 2)],
                        "slices": int(len(st)), "sampled": int(i),
                        "rate": (w_used / t_used) if i else None,
                        "work": float(work[st].sum())})
            t_all += t_used
        rates = [p["rate"] for p in per]
        est = 0.0
        for k, p in enumerate(per):
            r = p["rate"]
            if r is None:  # nearest sampled stratum
                near = sorted((abs(j - k), j) for j, x in enumerate(rates) if x)
                r = rates[near[0][1]]
                p["rate_from_stratum"] = near[0][1]
            est += p["work"] / r
        return est, per, sampled, t_all

    est_all, per_all, n_all, t_all = run(threads, seconds)
    est_one, per_one, n_one, t_one = run(1, seconds / 4, guide=[p["rate"] for p in per_all])
    fmt = lambda per: "; ".join(
        f"tc {p['tc'][0]}-{p['tc'][1]}: {p['sampled']}/{p['slices']}"
        + (f" (rate of stratum {p['rate_from_stratum']})" if "rate_from_stratum" in p else "")
        for p in per)
    return {
        "value": 1.0 / est_all,
        "unit": "MIS solves/sec",
        "cores": threads,
        "host_cpu_count": os.cpu_count(),
        "kind": "port",
        "one_core_solves_per_s": 1.0 / est_one,
        "sample": (f"stratified by tc (quartiles of the {len(live)} slices): on {threads} "
                   f"threads {n_all} slices in {t_all:.1f}s [{fmt(per_all)}]; on 1 thread "
                   f"{n_one} slices in {t_one:.1f}s [{fmt(per_one)}]; each stratum's time = "
                   "its work (sum of 2^tc) / its sampled rate; results checked equal to the "
                   "GPU's"),
    }


def work_accounting(tb, br):
    """Per-solve work of the workload's slices: the trees' own (MACs = sum of
    2^tc, bytes = sum of 4 x 2^rwc, every label of size 2) against what the
    BATCHED and PLAIN plans execute (sum over their steps of the configurations
    each step visits and of 4 x (|A| + |B| + |C|)), from the planner's step
    records (tbh_plan_steps)."""
    tree_macs = tree_bytes = 0.0
    ex = {m: {"ew_macs": 0.0, "ew_bytes": 0.0, "gemm_macs": 0.0, "gemm_bytes": 0.0}
          for m in ("batched", "plain")}
    for b in br:
        if b.code is None:
            continue
        c = tb.complexity(b)
        tree_macs += 2.0 ** c.tc
        tree_bytes += 4.0 * 2.0 ** c.rwc
        for m in ex:
            for st in tb.plan_steps(b, np.float32, m):
                k = "gemm" if st["kind"] == 1 else "ew"
                ex[m][k + "_macs"] += st["macs"]
                ex[m][k + "_bytes"] += st["bytes"]
    out = {"tree": {"macs": tree_macs, "bytes": tree_bytes}}
    for m, v in ex.items():
        macs = v["ew_macs"] + v["gemm_macs"]
        byts = v["ew_bytes"] + v["gemm_bytes"]
        out[m] = dict({k: float("%.4g" % x) for k, x in v.items()},
                      macs=float("%.4g" % macs), bytes=float("%.4g" % byts),
                      macs_over_tree=round(macs / tree_macs, 3),
                      bytes_over_tree=round(byts / tree_bytes, 3))
    out["tree"] = {k: float("%.4g" % x) for k, x in out["tree"].items()}
    return out


def _pct(xs, q):
    xs = sorted(xs)
    return xs[min(len(xs) - 1, max(0, math.ceil(q * len(xs)) - 1))]


def _drain(srv, tickets):
    """Polls every ticket to completion: {ticket: (value, t_submit, t_done)}."""
    out = {}
    pending = list(tickets)
    while pending:
        rest = []
        for t in pending:
            st, v, t0, t1 = srv.poll(t)
            if st == "done":
                out[t] = (v, t0, t1)
            elif st == "failed":
                raise RuntimeError(f"served request {t} failed")
            else:
                rest.append(t)
        pending = rest
        if pending:
            time.sleep(0.0005)
    return out


def _fixtures_module():
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "make_fixtures", os.path.join(ROOT, "tests", "golden", "make_fixtures.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def serving_mix(tb):
    """The serving configuration's request mix with the oracle's optima: the
    seeded draw of tests/golden/serving_mix.json, regenerated (digests
    checked) so that every served value is compared with the oracle."""
    mf = _fixtures_module()
    fx = json.load(open(os.path.join(ROOT, "tests", "golden", "serving_mix.json")))
    reqs, refused = mf.serving_draw(tb)
    if refused != fx["refused"] or len(reqs) != len(fx["requests"]):
        raise RuntimeError("serving mix differs from its fixture")
    want = []
    for q, f in zip(reqs, fx["requests"]):
        if mf.sliceset_digest(q.branches) != f["digest"]:
            raise RuntimeError(f"serving mix request {q.desc} differs from its fixture")
        want.append(float(f["optimum"]))
    p = fx["params"]
    desc = (f"serving_requests({p['count']}, seed={p['seed']}): {fx['mix']}; draws refused: "
            f"{refused['width']} for unsliced width > sc_target + {p['max_width_gap']}, "
            f"{refused['slices']} for > {p['max_slices']} slices, {refused['work']} for "
            f"> 2^{p['max_log2_macs']} MACs; oracle-checked values")
    return reqs, want, desc


def _served(srv, reqs, order, want, rate=None, rng=None):
    """Submits reqs[k] for k in order (Poisson arrivals at `rate`, or all at
    once), polls them to completion and checks each value; returns
    (latencies, window seconds)."""
    tick = []
    start = time.perf_counter()
    due = 0.0
    for k in order:
        if rate:
            due += rng.exponential(1.0 / rate)
            dt = start + due - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
        tick.append((srv.submit(reqs[k].branches, reqs[k].dtype), k))
    res = _drain(srv, [t for t, _ in tick])
    lat = []
    for t, k in tick:
        if float(res[t][0]) != want[k]:
            raise RuntimeError(f"served value of {reqs[k].desc} differs from the oracle's")
        lat.append(res[t][2] - res[t][1])
    t0 = min(r[1] for r in res.values())
    t1 = max(r[2] for r in res.values())
    return lat, t1 - t0


def serving_leg(tb, be, br, mis, copies, seed):
    """The continuous-batching scheduler over the 200 GB pool: (a) `copies`
    requests of the main workload submitted at once (served throughput);
    (b) the serving configuration's request mix (serving_mix) first as a
    burst (capacity), then as Poisson arrivals at 70% of that capacity
    (latency p50 / p99 from the scheduler's submit and completion stamps),
    then at 30%, 50% and 90% (rate_sweep), every value checked against the
    oracle's; (c) the latency lane: a one-slice priority request submitted
    while a main-workload request is being served."""
    t = time.time()
    mix, want, desc = serving_mix(tb)
    log(f"serving mix: {len(mix)} requests, {sum(len(q.branches) for q in mix)} slices, "
        f"generated in {time.time() - t:.1f}s")
    out = {}
    with tb.Server(be) as srv:
        # (a) main workload, burst
        tick = [srv.submit(br) for _ in range(copies)]
        res = _drain(srv, tick)
        t0 = min(r[1] for r in res.values())
        t1 = max(r[2] for r in res.values())
        if any(float(r[0]) != float(mis) for r in res.values()):
            raise RuntimeError("served MIS differs from the batched call")
        out["main_burst"] = {"requests": copies, "solves_per_s": round(copies / (t1 - t0), 4),
                             "window_s": round(t1 - t0, 3)}
        # (b) the mix: burst for capacity, then Poisson arrivals
        lat, window = _served(srv, mix, range(len(mix)), want)
        cap = len(mix) / window

        def poisson(frac, stream_seed):
            rate = frac * cap
            order = list(range(len(mix))) * 2
            lat, window = _served(srv, mix, order, want, rate, np.random.default_rng(stream_seed))
            return {"load": frac, "poisson_rate_req_per_s": round(rate, 3),
                    "arrivals": len(order), "served_req_per_s": round(len(order) / window, 3),
                    "latency_ms": {"p50": round(1e3 * _pct(lat, 0.5), 2),
                                   "p99": round(1e3 * _pct(lat, 0.99), 2),
                                   "max": round(1e3 * max(lat), 2)}}

        main = poisson(0.7, seed + 1)
        out["mixed_stream"] = {
            "requests": len(mix), "slices": sum(len(q.branches) for q in mix),
            "burst_capacity_req_per_s": round(cap, 3),
            "poisson_rate_req_per_s": main["poisson_rate_req_per_s"],
            "arrivals": main["arrivals"], "served_req_per_s": main["served_req_per_s"],
            "latency_ms": main["latency_ms"], "mix": desc}
        out["rate_sweep"] = [poisson(f, seed + 2 + i) for i, f in enumerate((0.3, 0.5, 0.9))]
        # (c) latency lane behind a main-workload request in flight
        one = [b for b in br if b.code is not None][:1]
        want_one = float(be.contract_slices(one, np.float32, "batched")[0])
        t_big = srv.submit(br)
        time.sleep(0.03)
        t_fast = srv.submit(one, priority=1)
        res = _drain(srv, [t_big, t_fast])
        if float(res[t_fast][0]) != want_one or float(res[t_big][0]) != float(mis):
            raise RuntimeError("latency-lane values differ from the batched call")
        out["latency_lane"] = {
            "priority_request_ms": round(1e3 * (res[t_fast][2] - res[t_fast][1]), 2),
            "main_request_ms": round(1e3 * (res[t_big][2] - res[t_big][1]), 2),
            "note": "one slice (priority 1) submitted 30 ms after a main-workload request"}
        out["scheduler"] = srv.stats()
    return out


def configs_leg(tb, be):
    """The other benchmark configurations, each solved BATCHED (best of 3
    wall times, results checked equal to the plain per-slice loop's):
    3-regular n=100 seed 0 sliced to sc 10; the unsliced 3-regular n=200
    seed 1 network (one contraction, sc 26); the 40 x 40 King's subgraph
    (filling 0.8, seed 3) with Int32 weights in [1, 100] sliced to sc 25,
    also checked equal to Float32 on the same weights."""
    def run(br, dt):
        be.contract_slices(br, dt, "batched")
        best, r = 1e30, None
        for _ in range(3):
            t = time.perf_counter()
            r = be.contract_slices(br, dt, "batched")
            best = min(best, time.perf_counter() - t)
        t = time.perf_counter()
        rp = be.contract_slices(br, dt, "plain")
        t_plain = time.perf_counter() - t
        if not np.array_equal(r, rp):
            raise RuntimeError("batched and plain results differ")
        ops = sum(2.0 * 2.0 ** tb.complexity(b).tc for b in br if b.code is not None)
        return r, {"slices": len(br), "batched_ms": round(1e3 * best, 2),
                   "plain_ms": round(1e3 * t_plain, 1),
                   "batched_speedup": round(t_plain / best, 2),
                   "tropical_GOPs": round(ops / best / 1e9, 1)}
    out = {}
    g = tb.random_regular_graph(100, 3, seed=0)
    br = tb.slice_graph(g, None, tb.ContractionTreeSlicer(sc_target=10, order_trials=20), seed=0)
    r, out["3regular_n100_sc10_f32"] = run(br, np.float32)
    out["3regular_n100_sc10_f32"]["mis"] = float(r.max())
    g = tb.random_regular_graph(200, 3, seed=1)
    b = tb.SlicedBranch(tb.MISProblem(g), tb.initialize_code(g, order_trials=100, seed=1), 0)
    r, out["3regular_n200_unsliced_f32"] = run([b], np.float32)
    out["3regular_n200_unsliced_f32"].update(mis=float(r.max()), sc=tb.complexity(b).sc,
                                              tc=round(tb.complexity(b).tc, 2))
    g = tb.random_ksg(40, 40, 0.8, seed=3)
    w = np.random.default_rng(3).integers(1, 101, g.nv).astype(np.int32)
    br = tb.slice_graph(g, w, tb.ContractionTreeSlicer(sc_target=25, order_trials=50), seed=3)
    r, out["ksg40_fill0.8_sc25_i32"] = run(br, np.int32)
    r32 = be.contract_slices(br, np.float32, "batched")
    if not np.array_equal(r32.astype(np.int64), r.astype(np.int64)):
        raise RuntimeError("Int32 and Float32 results differ on integer weights")
    out["ksg40_fill0.8_sc25_i32"]["mis"] = int(r.max())
    return out


def calibrate_copy():
    """A plain copy-like kernel of known size for scaling the FETCH_SIZE /
    WRITE_SIZE counters: y = -x over 2^28 floats (reads 1 GiB, writes 1 GiB),
    three times.  Runs in its own process (torch only)."""
    import torch
    n = 1 << 28
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        torch.neg(x, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "torch.neg", "read_bytes": 4 * n, "write_bytes": 4 * n,
                      "reps": 3}))


def _pmc_rows(d):
    import csv
    import glob
    rows = []
    for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        with open(f) as fh:
            rows.extend(csv.DictReader(fh))
    if not rows:
        raise RuntimeError(f"no counter_collection.csv under {d}")
    return rows


def summarize_pmc(fetch_dir, write_dir, copy_fetch_dir, copy_write_dir, out_json):
    """Per-kernel HBM bytes per dispatch from two rocprofv3 --pmc runs
    (FETCH_SIZE, WRITE_SIZE: KiB per dispatch), each divided by the factor
    the same counter shows on the calibration copy (measured / true bytes)."""
    factors = {}
    for name, d, true_key in (("FETCH_SIZE", copy_fetch_dir, "read_bytes"),
                              ("WRITE_SIZE", copy_write_dir, "write_bytes")):
        v = [float(r["Counter_Value"]) for r in _pmc_rows(d)
             if r["Counter_Name"] == name and "neg" in r["Kernel_Name"]]
        if not v:
            raise RuntimeError(f"no calibration dispatch with {name} in {d}")
        factors[name] = (sum(v) / len(v)) * 1024 / (4 << 28)
    per = {}
    for name, d in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        for r in _pmc_rows(d):
            if r["Counter_Name"] != name:
                continue
            k = r["Kernel_Name"].split("(")[0].replace("void ", "")
            e = per.setdefault(k, {})
            e.setdefault(name, []).append(float(r["Counter_Value"]) * 1024 / factors[name])
    kernels = {}
    for k, e in per.items():
        f, w = e.get("FETCH_SIZE", []), e.get("WRITE_SIZE", [])
        if not f or not w:
            continue
        kernels[k] = {"dispatches": len(f),
                      "fetch_bytes_per_dispatch": sum(f) / len(f),
                      "write_bytes_per_dispatch": sum(w) / len(w),
                      "traffic_bytes_per_dispatch": sum(f) / len(f) + sum(w) / len(w)}
    res = {"factors": factors, "kernels": kernels}
    with open(out_json, "w") as fh:
        json.dump(res, fh, indent=1, sort_keys=True)
    print(json.dumps(res))


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_3regular_n300_sc25.json")


def measured_traffic(kernel):
    """HBM bytes per dispatch of `kernel` on this workload from the committed
    PMC summary, or None."""
    try:
        with open(TRAFFIC_FILE) as fh:
            return json.load(fh)["kernels"][kernel]["traffic_bytes_per_dispatch"]
    except (OSError, KeyError, ValueError):
        return None


def main():
    if "--calibrate-copy" in sys.argv:
        return calibrate_copy()
    if "--summarize-pmc" in sys.argv:
        i = sys.argv.index("--summarize-pmc")
        return summarize_pmc(*sys.argv[i + 1:i + 6])
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=300)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--sc-target", type=int, default=25)
    ap.add_argument("--order-trials", type=int, default=200)
    ap.add_argument("--pool-gb", type=float, default=200.0)
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--autotune-trials", type=int, default=12,
                    help="contractions the autotuner may time before the timed steps (0: off)")
    ap.add_argument("--serve-copies", type=int, default=3)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the other configurations (n=100, unsliced n=200, KSG Int32)")
    ap.add_argument("--batched-only", action="store_true",
                    help="time only the batched steps (for profiler runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path "
              f"(--gpus {args.gpus} requested); slices shard independently, see the "
              f"design notes", file=sys.stderr)
        sys.exit(2)

    import __graft_entry__
    tb = __graft_entry__.package()
    g, br = build_workload(tb, args.n, args.seed, args.sc_target, args.order_trials)
    be = tb.HIPBackend(0, pool_bytes=int(args.pool_gb * (1 << 30)))

    # warmup
    ref = None
    for _ in range(args.warmup):
        r = be.contract_slices(br, np.float32, "batched")
        ref = r if ref is None else ref
    # autotuner (untimed): planner / launch parameters for this workload
    tune = None
    if args.autotune_trials > 0 and not args.batched_only:
        t_tune = time.perf_counter()
        tuning, tuned_ms = be.autotune(br, np.float32, max_trials=args.autotune_trials)
        tune = {"trials": args.autotune_trials, "chosen": tuning, "ms": round(tuned_ms, 1),
                "seconds": round(time.perf_counter() - t_tune, 1)}
        log(f"autotune: {tune}")
    # timed steps (the call is synchronous: results are on the host at return)
    be.reset_stats()
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        r = be.contract_slices(br, np.float32, "batched")
        mis = r.max()
        lat.append(time.perf_counter() - ts)
        if ref is not None and not np.array_equal(r, ref):
            raise RuntimeError("batched results differ between steps")
        ref = r
    t_total = time.perf_counter() - t0
    st_timed = be.stats()
    solves_per_s = args.steps / t_total
    log(f"batched: {1e3 * t_total / args.steps:.1f} ms/solve, MIS = {mis}")

    # plain per-slice GPU loop on the same slices
    t_plain = None
    if not args.batched_only:
        t0 = time.perf_counter()
        for _ in range(args.plain_steps):
            rp = be.contract_slices(br, np.float32, "plain")
        t_plain = (time.perf_counter() - t0) / args.plain_steps
        if not np.array_equal(rp, ref):
            raise RuntimeError("plain and batched per-slice results differ")
        log(f"plain: {1e3 * t_plain:.1f} ms/solve")

    # one profiled step: kernel time per kind from HIP events on the stream
    be.set_profiling(True)
    be.reset_stats()
    be.contract_slices(br, np.float32, "batched")
    sp = be.stats()
    be.set_profiling(False)
    kinds = ["ew_direct", "ew_tiled", "gemm"]
    per_kind = {}
    for k, name in enumerate(kinds):
        ms = sp["kind_ms"][k]
        n = sp["kind_launches"][k]
        per_kind[name] = {
            "launches": n, "ms": round(ms, 3),
            "avg_launch_ms": round(ms / n, 4) if n else None,
            "algo_GB": round(sp["kind_bytes"][k] / 1e9, 3),
            "achieved_GBs": round(sp["kind_bytes"][k] / ms / 1e6, 1) if ms else None,
            "achieved_TOPs": round(2 * sp["kind_macs"][k] / ms / 1e9, 3) if ms else None,
        }
    dom = max(range(2), key=lambda k: sp["kind_ms"][k])  # EW kernels are HBM-bound
    dom_name = kinds[dom]
    achieved = sp["kind_bytes"][dom] / sp["kind_ms"][dom] / 1e6
    kname = f"tbh::tbh_ew_{'direct_' if dom == 0 else ''}kernel<float>"
    roofline = {
        "bound": "hbm", "kernel": kname,
        "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(kname),
        "per_launch_algo_bytes": sp["kind_bytes"][dom] / max(1, sp["kind_launches"][dom]),
        "avg_launch_ms": sp["kind_ms"][dom] / max(1, sp["kind_launches"][dom]),
    }

    # standalone max-plus GEMM (VALU-bound), M = N = K = 4096: host operands,
    # the kernel alone timed with HIP events on the backend's stream
    gemm = None
    if not args.batched_only:
        n = 4096
        rng = np.random.default_rng(0)
        A = np.asfortranarray(rng.integers(0, 1024, (n, n)).astype(np.float32))
        B = np.asfortranarray(rng.integers(0, 1024, (n, n)).astype(np.float32))
        be.maxplus_gemm(A, B)
        be.set_profiling(True)
        be.reset_stats()
        reps = 5
        for _ in range(reps):
            be.maxplus_gemm(A, B)
        sg = be.stats()
        be.set_profiling(False)
        ms = sg["kind_ms"][2] / reps
        tops = 2 * n ** 3 / ms / 1e9
        peak = be.valu_peak(np.float32) / 1e12
        gemm = {"M": n, "N": n, "K": n, "dtype": "f32", "kernel_ms": round(ms, 3),
                "TOPs": round(tops, 2),
                "valu_maxplus_ceiling_TOPs": round(peak, 1),
                "ceiling": "measured: tbh_valu_peak, the GEMM's instruction mix (2 adds + "
                           "1 max3 per 2 MACs) on register operands over all CUs",
                "frac_of_ceiling": round(tops / peak, 3),
                "fp32_vector_published_TFLOPs": FP32_VECTOR_TFLOPS,
                "frac_of_fp32_vector_published": round(tops / FP32_VECTOR_TFLOPS, 3)}

    serving = None
    if not args.no_serve and not args.batched_only:
        serving = serving_leg(tb, be, br, mis, args.serve_copies, seed=7)
        log(f"serving: {serving}")

    configs = None
    if not args.no_configs and not args.batched_only:
        configs = configs_leg(tb, be)
        log(f"configs: {configs}")

    cpu = None
    if not args.no_cpu and not args.batched_only:
        # the threads the host gives this process: OMP_NUM_THREADS where it
        # is set (the GPU box allots 16 host threads per GPU), else every core
        threads = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS") or 0) or \
            (os.cpu_count() or 1)
        cpu = cpu_baseline(br, ref, args.cpu_seconds, threads)

    work = None if args.batched_only else work_accounting(tb, br)

    lat_ms = sorted(1e3 * x for x in lat)
    out = {
        "metric": METRIC,
        "value": round(solves_per_s, 4),
        "unit": "MIS solves/sec",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * t_total / args.steps, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"random 3-regular n={args.n} seed={args.seed}, sliced to "
                        f"sc_target={args.sc_target} ({len(br)} slices), unit weights, "
                        "Tropical Float32, one MIS solve per step",
            "slices": len(br), "mis": float(mis), "pool_gb": args.pool_gb,
            "parallelism": "1 GPU",
        },
        "latency_ms": {"p50": round(_pct(lat_ms, 0.5), 2), "p99": round(_pct(lat_ms, 0.99), 2)},
        "plain_gpu_loop": None if t_plain is None else {
            "solves_per_s": round(1.0 / t_plain, 4), "ms_per_solve": round(1e3 * t_plain, 1),
            "batched_speedup": round(solves_per_s * t_plain, 2)},
        "autotune": tune,
        "serving": serving,
        "tropical_GOPs": round(st_timed["total_algo_ops"] / t_total / 1e9, 1),
        "roofline": roofline,
        "kernels": per_kind,
        "work_per_solve": work,
        "gemm_4096": gemm,
        "other_configs": configs,
        "cpu_baseline": cpu,
    }
    be.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
