"""Zone-map pruning in the GROUP BY, join and ORDER BY ... LIMIT scans (the
reference prunes every execute_select scan by batch statistics:
BatchStats::can_skip, src/storage/stats.rs:57-113, used at
src/query/select.rs:119-135).  A batch whose bounds prove that no row passes
the WHERE is skipped; results must equal the unpruned scan and the oracle for
every strategy (dense, hash, partitioned and scalar GROUP BY; partitioned and
direct joins of every type; top-k), and the plan reports the live batches."""
from __future__ import annotations

import numpy as np
import pytest

import falcondb_amd as fa
import oracle as ora
from falcondb_amd._abi import ANTI, AVG, COUNT, INNER, LEFT, MAX, MIN, SEMI, SUM
from falcondb_amd.operators import AggExpr, NullableColumn, col, execute_sort_limit

from test_gpu_parity import assert_batches_equal

pytestmark = pytest.mark.gpu

BR, NB = 4096, 41
N = (NB - 1) * BR + 123


def _clustered():
    rng = np.random.default_rng(29)
    bi = np.arange(N) // BR
    drift = bi.astype(np.float64) + rng.random(N) * 0.5
    drift[5 * BR: 5 * BR + 300] = np.nan
    gvalid = ~((bi >= 10) & (bi <= 12) & (rng.random(N) < 0.5))
    return {"ts": np.arange(N, dtype=np.int64) * 3, "drift": drift, "k": bi.astype(np.int32),
            "g": NullableColumn(rng.integers(-50, 50, N), gvalid),
            "key": rng.integers(0, 60_000, N), "v": rng.integers(-10**6, 10**6, N), "x": rng.normal(size=N)}


# (predicate, live batches when pruned)
CASES = [(col("ts") < 3 * BR * 2, 2), ((col("ts") > 1000) & (col("ts") < 50_000), 5),
         (col("drift") > 39.2, 2), (col("k") == 3, 1), (col("k") < 3.5, 0),
         ((col("k") == 3) | col("g").is_null(), 4), (col("drift") != 5.25, NB)]


def _contexts(**tune):
    on, off = fa.Context(0), fa.Context(0)
    for c in (on, off):
        for k, v in tune.items():
            c.set_tuning(k, v)
    off.set_tuning("zone_maps", 0)
    return on, off


@pytest.mark.parametrize("strategy", ["auto", "partitioned", "hash", "dense", "scalar"])
def test_groupby_scans_prune(strategy):
    cols = _clustered()
    tune = {"partitioned": {"gb_part_min_rows": 1}, "hash": {"gb_partition": 0}}.get(strategy, {})
    on, off = _contexts(**tune)
    key = {"dense": "k", "scalar": None}.get(strategy, "key")
    aggs = [AggExpr(COUNT), AggExpr(SUM, "v"), AggExpr(MIN, "v"), AggExpr(MAX, "x"), AggExpr(AVG, "x")]
    t_on, t_off = fa.Table(on, cols, batch_rows=BR), fa.Table(off, cols, batch_rows=BR)
    for p, live in CASES:
        if key is None:
            g_on, g_off = fa.execute_agg_only(on, t_on, aggs, p), fa.execute_agg_only(off, t_off, aggs, p)
        else:
            g_on = fa.GroupByExec([key], aggs).execute(on, t_on, p)
            g_off = fa.GroupByExec([key], aggs).execute(off, t_off, p)
        plan = on.last_plan()
        assert f"zone maps: {live} of {NB} batches live" in plan, (plan, p.nodes)
        assert_batches_equal(g_on, g_off, key)
        assert_batches_equal(g_on, ora.groupby(cols, p, key, aggs, batch_rows=BR).exact, key)
    t_on.free()
    t_off.free()
    on.close()
    off.close()


@pytest.mark.parametrize("path", ["partitioned", "direct"])
def test_join_probe_scans_prune(path):
    cols = _clustered()
    rng = np.random.default_rng(31)
    bk = rng.choice(60_000, 20_000, replace=False).astype(np.int64)
    bcols = {"bk": bk, "bpay": rng.integers(-1000, 1000, bk.size)}
    tune = {"join_part_min_rows": 1} if path == "partitioned" else {"join_partition": 0}
    on, off = _contexts(**tune)
    pt_on, pt_off = fa.Table(on, cols, batch_rows=BR), fa.Table(off, cols, batch_rows=BR)
    bt_on, bt_off = fa.Table(on, bcols), fa.Table(off, bcols)
    for jt in (INNER, LEFT, SEMI, ANTI):
        aggs = [AggExpr(COUNT), AggExpr(SUM, "v")] + ([AggExpr(SUM, "bpay", side="build")] if jt in (INNER, LEFT) else [])
        for p, live in CASES[:4] + CASES[5:6]:
            g_on = fa.execute_hash_join(on, pt_on, bt_on, "key", "bk", jt, aggs, filter=p)
            plan = on.last_plan()
            g_off = fa.execute_hash_join(off, pt_off, bt_off, "key", "bk", jt, aggs, filter=p)
            assert f"zone maps: {live} of {NB} batches live" in plan, (plan, jt, p.nodes)
            assert path == "direct" or plan.startswith("join:partitioned"), plan
            assert_batches_equal(g_on, g_off, None)
            assert_batches_equal(g_on, ora.join(cols, bcols, "key", "bk", jt, aggs, expr=p, batch_rows=BR).exact, None)
    for t in (pt_on, pt_off, bt_on, bt_off):
        t.free()
    on.close()
    off.close()


def test_topk_scan_prunes():
    cols = _clustered()
    on, off = _contexts()
    t_on, t_off = fa.Table(on, cols, batch_rows=BR), fa.Table(off, cols, batch_rows=BR)
    for p, live in CASES:
        for desc in (True, False):
            r_on, v_on = execute_sort_limit(on, t_on, "x", 10, desc=desc, filter=p)
            plan = on.last_plan()
            r_off, v_off = execute_sort_limit(off, t_off, "x", 10, desc=desc, filter=p)
            assert f"zone maps: {live} of {NB} batches live" in plan, (plan, p.nodes)
            assert np.array_equal(r_on, r_off) and np.array_equal(v_on, v_off)
            assert np.array_equal(r_on, ora.topk(cols, p, "x", 10, desc=desc, batch_rows=BR))
    t_on.free()
    t_off.free()
    on.close()
    off.close()
