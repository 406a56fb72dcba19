"""Serving form (fgpu_server_*): every request served in a concurrent mix gives
the result it gives alone; admission respects the HBM budget; COUNT(*)
requests on one column share a scan."""
import numpy as np
import pytest

import falcondb_amd as fa
from falcondb_amd import datagen as dg
from falcondb_amd.operators import AggExpr, DictColumn, FalconError, col
from falcondb_amd.serving import Server, count_request, groupby_request, mix_request, run_alone, topk_request
from falcondb_amd._abi import COUNT, MAX, MIN, SUM

pytestmark = pytest.mark.gpu

FLOAT_RTOL = 1e-12  # float SUM / AVG tolerance of the contract (the sums are exact, so equal in fact)


@pytest.fixture(scope="module")
def c5(ctx):
    K = dg.keys(20_000, 40, seed_=dg.seed(5, 8))
    cols = dg.c5_fact(2_000_000, K)
    fact = fa.Table(ctx, {"key": cols["key"], "v1": cols["v1"], "v2": cols["v2"], "f": cols["f"],
                          "g": DictColumn(cols["g"], dg.G_NAMES)})
    dim = fa.Table(ctx, dg.c3_build(K))
    idx = fa.JoinIndex(ctx, dim, "bk")
    yield cols, fact, dim, idx
    idx.free()
    dim.free()
    fact.free()


def _same(a: fa.RecordBatch, b: fa.RecordBatch, sort_key: str | None):
    assert a.names == b.names and a.dtypes == b.dtypes
    assert a.num_rows == b.num_rows
    if sort_key is not None:
        a, b = a.sort_by(sort_key), b.sort_by(sort_key)
    for i, name in enumerate(a.names):
        x, y = a.columns[i], b.columns[i]
        vx = a.validity[i] if a.validity[i] is not None else np.ones(x.size, bool)
        vy = b.validity[i] if b.validity[i] is not None else np.ones(y.size, bool)
        assert np.array_equal(vx, vy), name
        if x.dtype == object:
            assert list(x[vx]) == list(y[vy]), name
        elif a.dtypes[i] == "Float64":
            xs, ys = x[vx], y[vy]
            assert np.all((xs == ys) | np.isclose(xs, ys, rtol=FLOAT_RTOL, atol=0)), name
        else:
            assert np.array_equal(x[vx], y[vy]), name


def _sort_key(spec):
    return spec["key"] if spec["kind"] == "groupby" else None


def test_served_mix_equals_alone(ctx, c5):
    cols, fact, dim, idx = c5
    mix = dg.c5_mix(96, seed_=dg.seed(5, 9))
    reqs = [mix_request(s, fact, dim, idx) for s in mix]
    alone = [run_alone(ctx, r) for r in reqs]
    budget = 1 << 30  # a 2M-row high-cardinality GROUP BY needs ~340 MB: at most two run at once
    with Server(0, slots=4, hbm_budget=budget) as srv:
        tickets = [srv.submit(r) for r in reqs]
        served = [srv.wait(t) for t in tickets]
        st = srv.stats()
    for spec, (got, rs), exp in zip(mix, served, alone):
        _same(got, exp, _sort_key(spec))
        assert rs["latency_ms"] >= rs["run_ms"] >= 0 and rs["reserved_bytes"] <= budget
    assert st["completed"] == len(mix) and st["failed"] == 0
    assert 0 < st["peak_held_bytes"] <= budget
    # the plain results themselves: COUNT(*) and ORDER BY against numpy
    for spec, exp in zip(mix, alone):
        if spec["kind"] == "count":
            assert exp.columns[0][0] == int(np.count_nonzero(cols["f"] < spec["t"]))
        if spec["kind"] == "topk":
            order = np.lexsort((np.arange(cols["v2"].size), -cols["v2"]))[: spec["k"]]
            assert np.array_equal(exp.columns[0], order)
            assert np.array_equal(exp.columns[1], cols["v2"][order])


def test_shared_count_scan(c5):
    cols, fact, dim, idx = c5
    ts = list(range(-5, 106, 4))
    with Server(0, slots=1, hbm_budget=64 << 20) as srv:
        srv.pause(True)  # queue everything, then let one slot take it
        tickets = [srv.submit(count_request(fact, col("f") < t)) for t in ts]
        t_v2 = srv.submit(count_request(fact, col("v2") >= 0.25, alias="n"))
        t_key = srv.submit(count_request(fact, col("key") > 2**39))
        t_and = srv.submit(count_request(fact, (col("f") < 50) & (col("v2") < 0)))
        srv.pause(False)
        got = [srv.wait(t) for t in tickets]
        g_v2, g_key, g_and = srv.wait(t_v2), srv.wait(t_key), srv.wait(t_and)
        st = srv.stats()
    for t, (res, rs) in zip(ts, got):
        assert res.columns[0][0] == int(np.count_nonzero(cols["f"] < t)), t
        assert res.names == ["col0"]
    assert all(rs["shared"] == len(ts) for _, rs in got)  # one scan answered all of them
    assert g_v2[0].names == ["n"] and g_v2[0].columns[0][0] == int(np.count_nonzero(cols["v2"] >= 0.25))
    assert g_key[0].columns[0][0] == int(np.count_nonzero(cols["key"] > 2**39))
    assert g_and[0].columns[0][0] == int(np.count_nonzero((cols["f"] < 50) & (cols["v2"] < 0)))
    assert g_and[1]["shared"] == 1  # two columns: runs alone
    assert st["shared_launches"] >= 1 and st["shared_requests"] >= len(ts)


def test_budget_bounds_admission(c5):
    cols, fact, dim, idx = c5
    aggs = [AggExpr(COUNT), AggExpr(SUM, "v1"), AggExpr(MIN, "v1"), AggExpr(MAX, "v1"), AggExpr(SUM, "v2")]
    with Server(0, slots=2, hbm_budget=8 << 20) as srv:
        big = srv.submit(groupby_request(fact, "key", aggs, col("f") < 40))
        small = srv.submit(count_request(fact, col("f") < 40))
        with pytest.raises(FalconError) as e:
            srv.wait(big)
        assert e.value.code == 101  # OutOfMemory: more scratch than the budget
        res, rs = srv.wait(small)
        assert res.columns[0][0] == int(np.count_nonzero(cols["f"] < 40))
        st = srv.stats()
    assert st["failed"] == 1 and st["retries"] >= 1 and st["peak_held_bytes"] <= 8 << 20


def test_learned_need_and_trim(ctx, c5):
    """A large request admitted after smaller ones have left arenas behind: idle
    arenas are released so the large one fits, and results stay equal."""
    cols, fact, dim, idx = c5
    g_aggs = [AggExpr(COUNT), AggExpr(SUM, "v1"), AggExpr(SUM, "v2")]
    k_aggs = [AggExpr(COUNT), AggExpr(SUM, "v1"), AggExpr(MIN, "v1"), AggExpr(MAX, "v1"), AggExpr(SUM, "v2")]
    r_key = groupby_request(fact, "key", k_aggs, col("f") < 70)
    exp_key = run_alone(ctx, r_key)
    exp_g = run_alone(ctx, groupby_request(fact, "g", g_aggs, col("f") < 30))
    need = ctx.scratch_bytes()
    with Server(0, slots=3, hbm_budget=max(need, 64 << 20) * 3 // 2) as srv:
        first = [srv.submit(groupby_request(fact, "key", k_aggs, col("f") < 70)) for _ in range(3)]
        for t in first:
            _same(srv.wait(t)[0], exp_key, "key")
        ts = [srv.submit(groupby_request(fact, "g", g_aggs, col("f") < 30)) for _ in range(6)]
        ts.append(srv.submit(r_key))
        for t in ts[:-1]:
            _same(srv.wait(t)[0], exp_g, "g")
        _same(srv.wait(ts[-1])[0], exp_key, "key")
        st = srv.stats()
    assert st["failed"] == 0 and st["peak_held_bytes"] <= st["budget_bytes"]


def test_topk_served(c5):
    cols, fact, dim, idx = c5
    with Server(0, slots=2, hbm_budget=64 << 20) as srv:
        t1 = srv.submit(topk_request(fact, "v1", 5))
        t2 = srv.submit(topk_request(fact, "v2", 7, desc=True, filter=col("f") < 3))
        (a, _), (b, _) = srv.wait(t1), srv.wait(t2)
    o1 = np.lexsort((np.arange(cols["v1"].size), cols["v1"]))[:5]
    assert np.array_equal(a.columns[0], o1) and np.array_equal(a.columns[1], cols["v1"][o1])
    sel = np.nonzero(cols["f"] < 3)[0]
    o2 = sel[np.lexsort((sel, -cols["v2"][sel]))][:7]
    assert np.array_equal(b.columns[0], o2) and np.array_equal(b.columns[1], cols["v2"][o2])


def test_backlog_gate_runs_light_load_one_at_a_time(ctx, c5):
    """Without overlap_always a request starts beside a running one only while
    at least two wait: two queued requests run one after the other.  With
    overlap_always the same requests give the same results."""
    cols, fact, dim, idx = c5
    reqs = [groupby_request(fact, "key", [AggExpr(COUNT), AggExpr(SUM, "v1")], filter=col("f") < 50),
            groupby_request(fact, "g", [AggExpr(MIN, "v2"), AggExpr(MAX, "v2")], filter=col("f") < 20)]
    alone = [run_alone(ctx, r) for r in reqs]
    for always in (False, True):
        with Server(0, slots=4, hbm_budget=1 << 30, overlap_always=always) as srv:
            srv.pause(True)
            tickets = [srv.submit(r) for r in reqs]
            srv.pause(False)
            served = [srv.wait(t)[0] for t in tickets]
            st = srv.stats()
        for got, exp, key in zip(served, alone, ("key", "g")):
            _same(got, exp, key)
        assert st["completed"] == 2 and st["failed"] == 0
        if not always:
            assert st["peak_running"] == 1
