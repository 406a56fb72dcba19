"""Benchmark: 100M-row scan + filter + high-cardinality GROUP BY on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-join]

Workload (one step): SELECT key, COUNT(*), SUM(v1), MIN(v1), MAX(v1), SUM(v2)
FROM t WHERE f < 40 GROUP BY key over 100M device-resident rows (key int64 with
1M distinct values drawn Zipf(1.1), v1 int64 uniform [-1e6, 1e6], v2 float64
N(0,1), f int32 uniform [0, 100)).  The group table, the fused scan and the
emission of every group stay on the device; `value` is input rows per second
of the whole operator with the table already in HBM.  The 1M x 100M join
(INNER / LEFT / SEMI / ANTI fused with SUM + COUNT) is reported beside it.

The JSON line carries a roofline block for the dominant kernel (algorithmic
bytes / its average duration from HIP events on the stream it runs on; traffic
= its measured HBM bytes per launch, read from the rocprofv3 FETCH_SIZE /
WRITE_SIZE summary committed under profiles/ for this same command) and a
CPU baseline: the oracle's restatement of the reference's rayon operators on
the GPU box's host cores over a bounded sample.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "rows/s + HBM GB/s (% roofline): 100M group-by, 1M⋈100M join; batched p99"
HBM_PEAK_GBS = 8000.0
C2_ROWS = 100_000_000
C2_KEYS = 1_000_000
C3_BUILD = 1_000_000
C3_PROBE = 100_000_000


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def groupby_call(lib, ctx, table, pred, spec):
    h = C.c_void_p()
    rc = lib.fgpu_groupby(ctx.handle, table.handle, C.byref(pred) if pred else None,
                          C.byref(spec), C.byref(h))
    if rc:
        raise RuntimeError(lib.fgpu_last_error().decode())
    n = lib.fgpu_result_rows(h)
    lib.fgpu_result_free(h)
    return n


def join_call(lib, ctx, probe, build, idx, spec):
    h = C.c_void_p()
    rc = lib.fgpu_join_agg(ctx.handle, probe.handle, None, build.handle, idx.handle,
                           C.byref(spec), C.byref(h))
    if rc:
        raise RuntimeError(lib.fgpu_last_error().decode())
    lib.fgpu_result_free(h)


def bench_batched(ctx, lib, steps, trials=18):
    """256 filter + GROUP BY queries over batches of log-uniform size (the C4
    recipe): the same queries as 256 plain fgpu_groupby calls, then as one
    prepared batch (one graph replay per run, autotuned over `trials`
    configurations; 1 = the first candidate only)."""
    import falcondb_amd as fa
    from falcondb_amd import _abi
    from falcondb_amd import datagen as dg
    from falcondb_amd._abi import COUNT, MAX, MIN, SUM
    from falcondb_amd.operators import AggExpr, _aggs_struct, col
    qs = dg.c4_queries()
    tabs = [fa.Table(ctx, dg.c4_batch(q["rows"], q["card"], q["seed"])) for q in qs]
    aggs = [AggExpr(COUNT), AggExpr(SUM, "val"), AggExpr(MIN, "val"), AggExpr(MAX, "val"),
            AggExpr(SUM, "score")]
    queries = [(t, "key", aggs, col("score") < q["theta"]) for q, t in zip(qs, tabs)]
    rows = sum(q["rows"] for q in qs)
    # parity: every query, batched against plain (results copied back)
    ctx.set_tuning("keep_on_device", 0)
    batch = fa.BatchGroupBy(ctx, queries)
    same = True
    groups = 0
    for (t, key, ag, f), got in zip(queries, batch.run()):
        plain = fa.GroupByExec([key], ag).execute(ctx, t, f).sort_by(key)
        got = got.sort_by(key)
        groups += got.num_rows
        same &= got.num_rows == plain.num_rows and all(
            np.array_equal(a, b) for a, b in zip(got.columns, plain.columns))
    ctx.set_tuning("keep_on_device", 1)
    specs = []
    for t, key, ag, f in queries:
        p, keep = f._compile(t)
        arr, keep2 = _aggs_struct(ag, lambda side, name, t=t: t.column_index(name))
        specs.append((t, p, keep, arr, keep2, _abi.GroupBySpec(t.column_index(key), len(ag),
                                                               C.cast(arr, C.POINTER(_abi.Agg)))))

    def plain_all():
        for t, p, keep, arr, keep2, spec in specs:
            groupby_call(lib, ctx, t, p, spec)

    plain_all()
    t0 = time.perf_counter()
    for _ in range(steps):
        plain_all()
    plain_ms = (time.perf_counter() - t0) * 1e3 / steps
    tune_info = batch.tune(trials)

    def batched():
        for h in batch.run_raw():
            lib.fgpu_result_free(h)

    batched()
    dev_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        batched()
        dev_ms.append(ctx.last_timing()["main_ms"])
    batch_ms = (time.perf_counter() - t0) * 1e3 / steps
    alg = rows * (4 + 8 + 8) + groups * 48  # key, val, score read once; every group row written
    log(f"[bench] C4: plain {plain_ms:.2f} ms, batched {batch_ms:.2f} ms ({np.mean(dev_ms):.2f} ms on "
        f"the device), parity {same}; {tune_info}")
    out = {"queries": len(qs), "rows": rows, "groups": groups, "algorithmic_bytes": alg,
           "plain_ms": plain_ms, "batched_ms": batch_ms,
           "batched_device_ms": float(np.mean(dev_ms)), "speedup": plain_ms / batch_ms,
           "rows_per_s": rows / (batch_ms * 1e-3), "achieved_GBs": alg / (np.mean(dev_ms) * 1e-3) / 1e9,
           "frac": alg / (np.mean(dev_ms) * 1e-3) / 1e9 / HBM_PEAK_GBS, "parity_vs_plain": bool(same),
           "config": tune_info, "fallback_queries": batch.fallbacks()}
    batch.free()
    for t in tabs:
        t.free()
    return out


def bench_zone_maps(ctx, rows, reps=10):
    """COUNT(*) WHERE ts < X over a sorted int64 column (a clustered column such
    as a load timestamp; 2 % of its batches hold qualifying rows), with the zone
    maps on and off: pruned batches are not read."""
    import falcondb_amd as fa
    from falcondb_amd.operators import col
    t = fa.Table(ctx, {"ts": np.arange(rows, dtype=np.int64)})
    pred = col("ts") < rows // 50
    out = {}
    for name, on in (("on", 1), ("off", 0)):
        ctx.set_tuning("zone_maps", on)
        n = fa.count_star(ctx, t, pred)
        t0 = time.perf_counter()
        for _ in range(reps):
            fa.count_star(ctx, t, pred)
        out[f"{name}_ms"] = (time.perf_counter() - t0) * 1e3 / reps
        out["count"] = int(n)
    ctx.set_tuning("zone_maps", 1)
    out["plan"] = ctx.last_plan()
    t.free()
    log(f"[bench] zone maps: COUNT(*) WHERE ts < {rows // 50} over {rows} sorted rows: "
        f"{out['on_ms']:.3f} ms pruned, {out['off_ms']:.3f} ms full scan")
    return out


def _serve_stream(srv, reqs, kinds, qps, seed_):
    """Serves `reqs` on `srv` with Poisson arrivals at `qps` (seeded); returns
    latency percentiles and throughput.  A waiter thread collects results while
    the main thread submits on schedule."""
    import queue as _q
    import threading
    rng = np.random.default_rng(seed_)
    gaps = rng.exponential(1.0 / qps, len(reqs))
    done, tq = [], _q.Queue()
    st0 = srv.stats()

    def waiter():
        while True:
            t = tq.get()
            if t is None:
                return
            res, st = srv.wait(t)
            done.append(st)

    th = threading.Thread(target=waiter)
    th.start()
    t0 = time.perf_counter()
    due = t0
    for r, g in zip(reqs, gaps):
        due += g
        while True:
            now = time.perf_counter()
            if now >= due:
                break
            time.sleep(min(due - now, 0.002))
        tq.put(srv.submit(r))
    tq.put(None)
    th.join()
    wall = time.perf_counter() - t0
    st1 = srv.stats()
    lat = np.array([d["latency_ms"] for d in done])
    by_kind = {}
    for k, d in zip(kinds, done):
        by_kind.setdefault(k, []).append(d["latency_ms"])
    return {"qps_offered": qps, "requests": len(reqs), "qps_served": len(reqs) / wall,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "rows_per_s": float(sum(d["rows_in"] for d in done) / wall),
            "p99_ms_by_kind": {k: float(np.percentile(v, 99)) for k, v in by_kind.items()},
            "shared_requests": int(st1["shared_requests"] - st0["shared_requests"]),
            "peak_running": int(st1["peak_running"]), "peak_held_GB": st1["peak_held_bytes"] / 1e9,
            "failed": int(st1["failed"] - st0["failed"])}


SERVE_FORMS = {  # name -> (slots, shared COUNT(*) scans, overlap at any load)
    "served": (4, True, False),
    "served_overlap_always": (4, True, True),
    "plain": (1, False, False),
}


def bench_serving(ctx, K, rows, qps_levels, slots=4, budget=32 << 30, forms=("served", "plain")):
    """The serving form: a device-resident fact table (the 100M-row group-by
    table plus a 12-value column g) and the 1M-row dimension table, a request
    mix of filtered COUNT(*), GROUP BY g / key, INNER / SEMI joins on
    predicate-selected probe ranges and ORDER BY v2 DESC LIMIT 10, Poisson
    arrivals at each offered QPS.  The same arrival streams are also served by
    the plain per-query form: one slot, no shared scans.  The two forms run in
    separate phases, each on its own warmed-up server."""
    import falcondb_amd as fa
    from falcondb_amd import datagen as dg
    from falcondb_amd.operators import DictColumn
    from falcondb_amd.serving import Server, mix_request, run_alone
    ctx.set_tuning("keep_on_device", 0)  # run_alone reads its results back
    cols = dg.c5_fact(rows, K)
    fact = fa.Table(ctx, {"key": cols["key"], "v1": cols["v1"], "v2": cols["v2"], "f": cols["f"],
                          "g": DictColumn(cols["g"], dg.G_NAMES)})
    del cols
    dim = fa.Table(ctx, dg.c3_build(K))
    idx = fa.JoinIndex(ctx, dim, "bk")
    # the same request streams per offered QPS for both forms
    streams = []
    for i, q in enumerate(qps_levels):
        n = int(min(max(q * 1.5, 60), 600))
        mix = dg.c5_mix(n, seed_=dg.seed(5, 32 + i))
        streams.append(([mix_request(sp, fact, dim, idx) for sp in mix], [sp["kind"] for sp in mix]))
    mix0 = dg.c5_mix(48, seed_=dg.seed(5, 1))
    reqs0 = [mix_request(sp, fact, dim, idx) for sp in mix0]
    same = True
    results = {}
    # one form at a time: a server is created, warmed up (scratch arenas
    # allocated, request shapes learned), checked against the requests run
    # alone, swept over every offered QPS and closed before the other starts
    for name in forms:
        srv_slots, share, always = SERVE_FORMS[name]
        srv_slots = slots if srv_slots > 1 else 1
        srv = Server(0, slots=srv_slots, hbm_budget=budget, share_counts=share, overlap_always=always)
        ts = [srv.submit(r) for r in reqs0]
        for sp, r, t in zip(mix0, reqs0, ts):
            got, _ = srv.wait(t)
            exp = run_alone(ctx, r)
            if sp["kind"] == "groupby":
                got, exp = got.sort_by(sp["key"]), exp.sort_by(sp["key"])
            same &= got.num_rows == exp.num_rows and all(
                np.array_equal(a, b) for a, b in zip(got.columns, exp.columns))
        results[name] = [_serve_stream(srv, reqs, kinds, q, 100 + i)
                         for i, (q, (reqs, kinds)) in enumerate(zip(qps_levels, streams))]
        srv.close()
    for i, q in enumerate(qps_levels):
        log(f"[bench] C5 qps {q}: " + " | ".join(
            f"{name} p50 {r[i]['p50_ms']:.2f} p99 {r[i]['p99_ms']:.2f} ms {r[i]['qps_served']:.0f} q/s "
            f"{r[i]['rows_per_s'] / 1e9:.1f} G rows/s" for name, r in results.items()))
    idx.free()
    dim.free()
    fact.free()
    return {"rows": rows, "dim_rows": int(K.size), "slots": slots, "budget_GB": budget / 1e9,
            "mix": "40% COUNT(*) WHERE f<t; 30% GROUP BY g|key WHERE f<t; 20% INNER|SEMI join WHERE f<t; "
                   "10% ORDER BY v2 DESC LIMIT 10",
            "served_gate": "a request starts beside a running one only while >= 2 wait",
            "parity_vs_alone": bool(same),
            **{("plain_one_slot" if n == "plain" else n): r for n, r in results.items()}}


def cpu_info():
    """CPU model, logical CPUs and the oracle's thread count on this host."""
    import oracle as ora
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "threads": ora.threads()}


def timed_runs(fn, runs):
    """Median and min wall time (s) of `runs` calls after one warm-up call."""
    fn()
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), float(np.min(ts))


def cpu_baselines(c2, K, gpu_c2, aggs, runs=5, c3_probe_rows=20_000_000):
    """The reference's CPU operators (the oracle's restatement, on this host's
    cores): C1 at 1M rows and C2 at full size, median (and min) of `runs`
    runs each; C3 INNER and SEMI over the first `c3_probe_rows` probe rows
    against the full 1M-row build side.  Also checks every output column of
    this run's GPU C2 result against the oracle's exact result."""
    import oracle as ora
    from falcondb_amd import datagen as dg
    from falcondb_amd._abi import AVG, COUNT, INNER, SEMI, SUM
    from falcondb_amd.operators import AggExpr, DictColumn, col
    info = cpu_info()
    out = {"host": info, "runs": runs}
    c1 = dg.c1_table(1_000_000)
    c1cols = {"id": c1["id"], "dept": DictColumn(c1["dept"], dg.DEPTS), "score": c1["score"],
              "salary": c1["salary"]}
    c1aggs = [AggExpr(COUNT), AggExpr(SUM, "score"), AggExpr(AVG, "salary")]
    med, mn = timed_runs(lambda: ora.groupby(c1cols, col("score") > 50, "dept", c1aggs, exact=False), runs)
    out["c1"] = {"rows": 1_000_000, "median_s": med, "min_s": mn, "rows_per_s": 1_000_000 / med}
    n = len(c2["key"])
    med, mn = timed_runs(lambda: ora.groupby(c2, col("f") < 40, "key", aggs, exact=False), runs)
    out["c2"] = {"rows": n, "median_s": med, "min_s": mn, "rows_per_s": n / med}
    # parity of this run: every column of the GPU result against the exact oracle
    e = ora.groupby(c2, col("f") < 40, "key", aggs).exact.sort_by("key")
    g = gpu_c2.sort_by("key")
    cols_ok = {}
    for i, name in enumerate(e.names):
        a, b = g.columns[i], e.columns[i]
        va = g.validity[i] if g.validity[i] is not None else np.ones(a.size, bool)
        vb = e.validity[i] if e.validity[i] is not None else np.ones(b.size, bool)
        same = g.num_rows == e.num_rows and np.array_equal(va, vb)
        if same and e.dtypes[i] == "Float64":
            same = np.array_equal(a[va].astype(np.float64).view(np.int64), b[vb].view(np.int64))
        elif same:
            same = np.array_equal(a[va], b[vb].astype(a.dtype))
        cols_ok[name] = bool(same)
    out["c2_parity_by_column"] = cols_ok
    Kb = dg.keys(C3_BUILD, 40, seed_=dg.seed(3, 15))
    bc = dg.c3_build(Kb)
    pc = dg.c3_probe(c3_probe_rows, Kb)
    for name, jt, jaggs in (("inner", INNER, [AggExpr(SUM, "bpay", side="build"), AggExpr(COUNT)]),
                            ("semi", SEMI, [AggExpr(SUM, "ppay"), AggExpr(COUNT)])):
        med, mn = timed_runs(lambda: ora.join(pc, bc, "pk", "bk", jt, jaggs, exact=False), runs)
        out[f"c3_{name}"] = {"probe_rows": c3_probe_rows, "build_rows": C3_BUILD, "median_s": med, "min_s": mn,
                             "rows_per_s": (c3_probe_rows + C3_BUILD) / med}
    log(f"[bench] CPU ({info['threads']} threads, {info['cpu_model']}): C1 {out['c1']['rows_per_s'] / 1e6:.0f} M rows/s, "
        f"C2 {out['c2']['rows_per_s'] / 1e6:.0f} M rows/s, C3 inner {out['c3_inner']['rows_per_s'] / 1e6:.0f} "
        f"semi {out['c3_semi']['rows_per_s'] / 1e6:.0f} M rows/s; C2 parity {cols_ok}")
    return out


def bench_c1(ctx, reps=50):
    """C1 on the GPU: SELECT dept, COUNT(*), SUM(score), AVG(salary) FROM t
    WHERE score > 50 GROUP BY dept over 1M rows.  Latency is the whole call
    with the 12 result rows copied back; the kernel time comes from HIP events
    on the context's stream."""
    import falcondb_amd as fa
    from falcondb_amd import datagen as dg
    from falcondb_amd._abi import AVG, COUNT, SUM
    from falcondb_amd.operators import AggExpr, DictColumn, col
    c1 = dg.c1_table(1_000_000)
    t = fa.Table(ctx, {"id": c1["id"], "dept": DictColumn(c1["dept"], dg.DEPTS), "score": c1["score"],
                       "salary": c1["salary"]})
    aggs = [AggExpr(COUNT), AggExpr(SUM, "score"), AggExpr(AVG, "salary")]
    gb = fa.GroupByExec(["dept"], aggs)
    ctx.set_tuning("keep_on_device", 0)  # the caller restores it
    for _ in range(5):
        r = gb.execute(ctx, t, col("score") > 50)
    lat, kern = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        r = gb.execute(ctx, t, col("score") > 50)
        lat.append((time.perf_counter() - t0) * 1e3)
        kern.append(ctx.last_timing()["total_ms"])
    med = float(np.median(lat))
    out = {"rows": 1_000_000, "groups": int(r.num_rows), "latency_ms_p50": med,
           "latency_ms_min": float(np.min(lat)), "kernels_ms": float(np.median(kern)),
           "rows_per_s": 1_000_000 / (med * 1e-3), "plan": ctx.last_plan()}
    log(f"[bench] C1: {med:.3f} ms per query with copy-back ({out['kernels_ms']:.3f} ms of kernels), "
        f"{out['rows_per_s'] / 1e9:.2f} G rows/s")
    t.free()
    return out


TRAFFIC_FILE = ROOT / "profiles" / "traffic_c2_groupby_100M_and_c3_join.txt"


def profiled_traffic(kernel: str):
    """HBM bytes per launch of `kernel` (FETCH_SIZE + WRITE_SIZE, each scaled by
    what the counter reads for a plain copy of known size), from the committed
    rocprofv3 counter summary of this same bench command; None if absent."""
    try:
        for line in TRAFFIC_FILE.read_text().splitlines():
            f = line.split()
            if len(f) > 2 and f[0] == "traffic" and f[1] == kernel:
                kv = dict(x.split("=", 1) for x in f[2:])
                return int(kv["bytes_per_launch"])
    except (OSError, ValueError, KeyError):
        pass
    return None


def calibrate_copy(n_bytes=1 << 30, reps=3):
    """A plain device-to-device copy of known size (the HIP runtime's copy
    kernel), for scaling FETCH_SIZE / WRITE_SIZE counter readings: profiled
    beside the operator kernels, it says what the counters read per byte moved."""
    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    for ptr in (src, dst):
        if hip.hipMalloc(C.byref(ptr), C.c_size_t(n_bytes)):
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(src, 1, C.c_size_t(n_bytes))
    for _ in range(reps):
        if hip.hipMemcpy(dst, src, C.c_size_t(n_bytes), 3):  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"calibration_copy_bytes": n_bytes, "reps": reps}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rows", type=int, default=C2_ROWS)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-join", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--batch-trials", type=int, default=18,
                    help="configurations the C4 batch's autotuner tries (1: the first candidate only)")
    ap.add_argument("--qps", default="100,300,1000,3000")
    ap.add_argument("--serve-forms", default="served,plain",
                    help="serving forms swept, from: " + ", ".join(SERVE_FORMS))
    ap.add_argument("--serve-only", action="store_true",
                    help="run only the serving leg (no JSON bench line)")
    ap.add_argument("--gb-strategy", choices=["auto", "hash"], default="auto",
                    help="C2 GROUP BY strategy: the planner's choice, or the single hash table "
                         "(the radix-partitioned path switched off), to measure what partitioning buys")
    ap.add_argument("--ceiling", action="store_true",
                    help="only run the STREAM-like read probe (the scan's four columns, its tiling)")
    ap.add_argument("--calibrate", action="store_true",
                    help="only run a 1 GiB device-to-device copy (counter calibration)")
    args = ap.parse_args()
    if args.calibrate:
        calibrate_copy()
        return
    if args.ceiling:  # a child process: this one has not touched the GPU
        import subprocess
        exe = ROOT / "falcondb_amd" / "fgpu_stream_read"
        if not exe.exists():
            sys.exit(f"{exe} is not built (python -m falcondb_amd.build)")
        sys.exit(subprocess.run([str(exe)]).returncode)
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU runs are not part of this build", file=sys.stderr)
        sys.exit(2)

    import falcondb_amd as fa
    from falcondb_amd import _abi
    from falcondb_amd import datagen as dg
    from falcondb_amd._abi import ANTI, COUNT, INNER, LEFT, MAX, MIN, SEMI, SUM
    from falcondb_amd.operators import AggExpr, _aggs_struct, col

    lib = _abi.load()
    forms = tuple(f for f in args.serve_forms.split(",") if f)
    if any(f not in SERVE_FORMS for f in forms):
        sys.exit(f"--serve-forms: unknown form in {args.serve_forms!r}")
    if args.serve_only:
        res = bench_serving(fa.Context(0), dg.keys(C2_KEYS, 40), args.rows,
                            [int(q) for q in args.qps.split(",")], forms=forms)
        print(json.dumps({"c5_serving": res}), flush=True)
        return
    t0 = time.time()
    K = dg.keys(C2_KEYS, 40)
    c2 = dg.c2_table(args.rows, K)
    log(f"[bench] generated C2 {args.rows} rows in {time.time() - t0:.1f}s")
    ctx = fa.Context(0)
    if args.gb_strategy == "hash":
        ctx.set_tuning("gb_partition", 0)
    t0 = time.time()
    table = fa.Table(ctx, c2)
    log(f"[bench] uploaded {table.device_bytes() / 1e9:.2f} GB in {time.time() - t0:.1f}s")

    aggs = [AggExpr(COUNT), AggExpr(SUM, "v1"), AggExpr(MIN, "v1"), AggExpr(MAX, "v1"),
            AggExpr(SUM, "v2")]
    pred, keep = (col("f") < 40)._compile(table)
    arr, keep2 = _aggs_struct(aggs, lambda side, name: table.column_index(name))
    spec = _abi.GroupBySpec(table.column_index("key"), len(aggs), C.cast(arr, C.POINTER(_abi.Agg)))

    # correctness of this run against the oracle is checked on the result copied to host
    gpu_res = fa.GroupByExec(["key"], aggs).execute(ctx, table, col("f") < 40)
    n_groups = gpu_res.num_rows
    ctx.set_tuning("keep_on_device", 1)
    ctx.set_timing(True)
    for _ in range(args.warmup):
        groupby_call(lib, ctx, table, pred, spec)
    main_ms, fin_ms, total_ms = [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        groupby_call(lib, ctx, table, pred, spec)
        tm = ctx.last_timing()
        main_ms.append(tm["main_ms"])
        fin_ms.append(tm["finalize_ms"])
        total_ms.append(tm["total_ms"])
    wall = time.perf_counter() - t0
    plan = ctx.last_plan()
    ms_per_step = wall * 1e3 / args.steps
    rows_per_s = args.rows * args.steps / wall
    # HIP events on the context's stream bracket each kernel: main = the fused
    # scan (pass 1), finalize = pass 2 + emission (or the hash table emission)
    kern_ms = float(np.mean(main_ms))
    pass2_ms = float(np.mean(fin_ms))
    partitioned = plan.startswith("groupby:partitioned") and "fallback" not in plan
    kernel, pass2 = ("k_gbp_scan", "k_gbp_agg") if partitioned else ("k_gb_hash", "k_gb_emit")
    scan_bytes = args.rows * (8 + 8 + 8 + 4)          # key, v1, v2, f once
    alg_bytes = scan_bytes + n_groups * 48            # + the group rows written
    achieved = scan_bytes / (kern_ms * 1e-3) / 1e9
    query_gbs = alg_bytes / ((kern_ms + pass2_ms) * 1e-3) / 1e9
    log(f"[bench] C2: {ms_per_step:.3f} ms/step, {kernel} {kern_ms:.3f} ms ({achieved:.0f} GB/s), "
        f"{pass2} {pass2_ms:.3f} ms, query {query_gbs:.0f} GB/s, groups {n_groups}; plan {plan}")

    # host-inclusive rate (results copied back over PCIe): reported, never `value`
    ctx.set_tuning("keep_on_device", 0)
    t0 = time.perf_counter()
    for _ in range(3):
        groupby_call(lib, ctx, table, pred, spec)
    host_ms = (time.perf_counter() - t0) * 1e3 / 3
    # the same query as SQL text (Database.sql) against the operator call, both
    # reading the result back: the difference is the SQL layer's parse and plan
    from falcondb_amd.sql import Database
    db = Database(ctx)
    db.register("t", table)
    c2_sql = "SELECT key, COUNT(*), SUM(v1), MIN(v1), MAX(v1), SUM(v2) FROM t WHERE f < 40 GROUP BY key"
    db.sql(c2_sql)
    t0 = time.perf_counter()
    for _ in range(3):
        db.sql(c2_sql)
    sql_ms = (time.perf_counter() - t0) * 1e3 / 3
    t0 = time.perf_counter()
    for _ in range(3):
        fa.GroupByExec(["key"], aggs).execute(ctx, table, col("f") < 40)
    op_ms = (time.perf_counter() - t0) * 1e3 / 3
    ctx.set_tuning("keep_on_device", 1)
    log(f"[bench] C2 as SQL: {sql_ms:.3f} ms per query, operator call {op_ms:.3f} ms (both read the result back)")

    secondary = {"c2_sql": {"statement": c2_sql, "sql_ms": sql_ms, "operator_ms": op_ms,
                            "parse_plan_ms": sql_ms - op_ms}}
    secondary["c1"] = bench_c1(ctx)
    ctx.set_tuning("keep_on_device", 1)
    if not args.no_join:
        del c2
        Kb = dg.keys(C3_BUILD, 40, seed_=dg.seed(3, 15))
        bc = dg.c3_build(Kb)
        pc = dg.c3_probe(C3_PROBE, Kb)
        bt, pt = fa.Table(ctx, bc), fa.Table(ctx, pc)
        idx = fa.JoinIndex(ctx, bt, "bk")
        for name, jt, jaggs, byts in [
            ("inner", INNER, [AggExpr(SUM, "bpay", side="build"), AggExpr(COUNT)], 16),
            ("left", LEFT, [AggExpr(SUM, "bpay", side="build"), AggExpr(COUNT)], 16),
            ("semi", SEMI, [AggExpr(SUM, "ppay"), AggExpr(COUNT)], 8),
            ("anti", ANTI, [AggExpr(SUM, "ppay"), AggExpr(COUNT)], 8),
        ]:
            def resolve(side, nm):
                return (bt if side == 1 else pt).column_index(nm)
            jarr, jkeep = _aggs_struct(jaggs, resolve)
            jspec = _abi.JoinSpec(jt, pt.column_index("pk"), bt.column_index("bk"), len(jaggs),
                                  C.cast(jarr, C.POINTER(_abi.Agg)))
            for _ in range(2):
                join_call(lib, ctx, pt, bt, idx, jspec)
            jms, t0 = [], time.perf_counter()
            steps = max(3, args.steps // 2)
            for _ in range(steps):
                join_call(lib, ctx, pt, bt, idx, jspec)
                jms.append(ctx.last_timing()["main_ms"])
            jwall = (time.perf_counter() - t0) / steps
            # algorithmic bytes: probe keys (+ probe payload for SEMI/ANTI) + build side
            jb = C3_PROBE * (8 if byts == 16 else 16) + C3_BUILD * (16 if byts == 16 else 8)
            km = float(np.mean(jms))
            secondary[f"join_{name}"] = {
                "rows_per_s": (C3_PROBE + C3_BUILD) / jwall, "ms_per_step": jwall * 1e3,
                "kernels_ms": km, "achieved_GBs": jb / (km * 1e-3) / 1e9,
                "frac": jb / (km * 1e-3) / 1e9 / HBM_PEAK_GBS}
            log(f"[bench] join {name}: {jwall * 1e3:.3f} ms/step, partition + probe kernels {km:.3f} ms")
        # index build time, and the cold join: no prebuilt index, so the call
        # builds its index and partitioned sub-tables per query, as the
        # reference builds its hash table per query (src/query/join.rs:607-645)
        t0 = time.perf_counter()
        idx2 = fa.JoinIndex(ctx, bt, "bk")
        secondary["join_build_ms"] = (time.perf_counter() - t0) * 1e3
        idx2.free()
        jaggs = [AggExpr(SUM, "bpay", side="build"), AggExpr(COUNT)]
        jarr, jkeep = _aggs_struct(jaggs, lambda side, nm: (bt if side == 1 else pt).column_index(nm))
        jspec = _abi.JoinSpec(INNER, pt.column_index("pk"), bt.column_index("bk"), len(jaggs),
                              C.cast(jarr, C.POINTER(_abi.Agg)))

        class _NoIndex:
            handle = None
        for _ in range(2):
            join_call(lib, ctx, pt, bt, _NoIndex, jspec)
        t0 = time.perf_counter()
        for _ in range(5):
            join_call(lib, ctx, pt, bt, _NoIndex, jspec)
        cold = (time.perf_counter() - t0) / 5
        secondary["join_inner_cold"] = {"ms_per_step": cold * 1e3, "rows_per_s": (C3_PROBE + C3_BUILD) / cold,
                                        "plan": ctx.last_plan()}
        log(f"[bench] join inner, cold (index and sub-tables built per query): {cold * 1e3:.3f} ms/step")
        pt.free()
        bt.free()
        del pc, bc

    if not args.no_join:
        secondary["zone_maps"] = bench_zone_maps(ctx, args.rows)

    if not args.no_batch:
        secondary["c4_batched"] = bench_batched(ctx, lib, max(3, args.steps // 4), args.batch_trials)

    if not args.no_serve:
        table.free()
        secondary["c5_serving"] = bench_serving(ctx, K, args.rows, [int(q) for q in args.qps.split(",")],
                                                forms=forms)

    cpu = None
    if not args.no_cpu:
        c2 = dg.c2_table(args.rows, K) if "c2" not in locals() else c2
        base = cpu_baselines(c2, K, gpu_res, aggs)
        secondary["cpu_baselines"] = base
        info = base["host"]
        cpu = {"value": base["c2"]["rows_per_s"], "unit": "rows/s", "cores": info["threads"], "kind": "port",
               "sample": f"the full {args.rows}-row C2 query, median of {base['runs']} runs of the oracle's "
                         f"restatement of the reference operators ({base['c2']['median_s']:.2f} s; min "
                         f"{base['c2']['min_s']:.2f} s) on {info['threads']} threads of {info['nproc']} "
                         f"({info['cpu_model']})",
               "median_s": base["c2"]["median_s"], "min_s": base["c2"]["min_s"],
               "cpu_model": info["cpu_model"], "nproc": info["nproc"]}
        secondary["parity_vs_cpu"] = all(base["c2_parity_by_column"].values())

    out = {
        "metric": METRIC, "value": rows_per_s, "unit": "rows/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int64",
        "data": "synthetic",
        "config": {"workload": "c2_filter_groupby_100M", "rows": args.rows,
                   "distinct_keys": C2_KEYS, "zipf_s": 1.1, "predicate": "f < 40",
                   "aggregates": "COUNT(*), SUM(v1), MIN(v1), MAX(v1), SUM(v2)",
                   "groups_out": int(n_groups), "parallelism": "1 GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": profiled_traffic(kernel),
                     "traffic_source": TRAFFIC_FILE.name,
                     "kernel": kernel, "kernel_ms": kern_ms, "algorithmic_bytes": scan_bytes,
                     "pass2_kernel": pass2, "pass2_ms": pass2_ms,
                     "query_achieved": query_gbs, "query_frac": query_gbs / HBM_PEAK_GBS,
                     "query_algorithmic_bytes": alg_bytes, "plan": plan},
        "cpu_baseline": cpu,
        "host_inclusive_ms_per_step": host_ms,
        "secondary": secondary,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
