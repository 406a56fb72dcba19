"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on the
same seeded inputs.  Integer results and group keys must match bit for bit;
float SUM / AVG are exactly rounded on both sides, so they must match bit for
bit too, and they must lie within 1e-12 relative of the reference-order
float sum (the tolerance the reference itself uses,
src/query/aggregate.rs:3662-3667)."""
import math

import numpy as np
import pytest

import falcondb_amd as fa
import oracle as ora
from falcondb_amd import datagen as dg
from falcondb_amd._abi import ANTI, AVG, COUNT, INNER, LEFT, MAX, MIN, NO_ROW, SEMI, SUM
from falcondb_amd.operators import AggExpr, DictColumn, NullableColumn, col

pytestmark = pytest.mark.gpu

BOUNDARY_ROWS = [0, 1, 63, 64, 65, 127, 128, 129, 1000, 4095, 10000, 65536 + 7]
FLOAT_RTOL = 1e-12


def _table(ctx, cols, batch_rows=65536):
    return fa.Table(ctx, cols, batch_rows=batch_rows)


def assert_batches_equal(gpu: fa.RecordBatch, exp: fa.RecordBatch, key: str | None,
                         ref: fa.RecordBatch | None = None):
    assert gpu.names == exp.names
    assert gpu.num_rows == exp.num_rows, (gpu.num_rows, exp.num_rows)
    if key is not None:
        gpu = gpu.sort_by(key)
        exp = exp.sort_by(key)
        if ref is not None:
            ref = ref.sort_by(key)
    for i, name in enumerate(exp.names):
        g, e = gpu.columns[i], exp.columns[i]
        gv = gpu.validity[i] if gpu.validity[i] is not None else np.ones(g.size, bool)
        ev = exp.validity[i] if exp.validity[i] is not None else np.ones(e.size, bool)
        assert np.array_equal(gv, ev), f"validity of {name}"
        if g.dtype == object:
            assert list(g[gv]) == list(e[ev]), name
            continue
        if gpu.dtypes[i] == "Float64":
            ge, ee = g[gv], e[ev].astype(np.float64)
            assert np.array_equal(ge.view(np.int64), ee.view(np.int64)) or np.array_equal(ge, ee, equal_nan=True), \
                f"{name}: float results differ: max |d| {np.max(np.abs(ge - ee)) if ge.size else 0}"
            if ref is not None and ref.dtypes[i] == "Float64":
                r = ref.columns[i][ev]
                ok = np.isclose(ge, r, rtol=FLOAT_RTOL, atol=0) | (ge == r)
                # the reference-order sum may itself be off for cancelling groups
                frac = 1.0 - ok.mean() if ok.size else 0.0
                assert frac < 1e-3, f"{name}: {frac:.2%} beyond 1e-12 of reference order"
        else:
            assert np.array_equal(g[gv], e[ev].astype(g.dtype)), name


# ─── WHERE ───────────────────────────────────────────────────────────────────
PREDS = [
    lambda: col("score") > 50,
    lambda: col("score") >= 50.5,
    lambda: col("score") < 10,
    lambda: col("score") <= 0.5,
    lambda: col("score") == 50.5,
    lambda: col("score") != 50.5,
    lambda: col("id") > 500,
    lambda: col("id") == 7,
    lambda: col("k32") < 40,
    lambda: col("k32") >= 4294967336,  # Int64 literal narrowed `as i32` -> 40
    lambda: col("dept") == "dept03",
    lambda: col("dept") > "dept05",
    lambda: (col("score") > 20) & ((col("id") < 300) | (col("dept") != "dept01")),
    lambda: col("id") > 12.5,  # Int64 column widened to f64
    lambda: col("score") > "x",  # unsupported pairing -> no rows
]


def _filter_cols(n, seed=7):
    i = np.arange(n, dtype=np.int64)
    score = (i % 100) + 0.5
    return {
        "id": i,
        "score": score,
        "k32": (i % 100).astype(np.int32),
        "dept": DictColumn((i % 12).astype(np.uint32), dg.DEPTS),
    }


@pytest.mark.parametrize("n", BOUNDARY_ROWS)
def test_filter_count_and_mask_boundaries(ctx, n):
    cols = _filter_cols(n)
    t = _table(ctx, cols)
    for mk in PREDS:
        p = mk()
        assert fa.count_star(ctx, t, p) == ora.filter_count(cols, p)
        assert np.array_equal(p.eval_mask_bits(ctx, t), ora.filter_mask(cols, p))
    # closed form: (i % 100) + 0.5 > 50 keeps exactly half of every 100 rows
    if n % 100 == 0:
        assert fa.count_star(ctx, t, col("score") > 50) == n // 2


def test_filter_nulls_and_nan(ctx):
    n = 5000
    rng = np.random.default_rng(3)
    v = rng.normal(size=n)
    v[::17] = np.nan
    valid = rng.random(n) > 0.2
    iv = rng.integers(-100, 100, n)
    ivalid = rng.random(n) > 0.3
    codes = rng.integers(0, 12, n).astype(np.uint32)
    codes[::11] = 0xFFFFFFFF
    cols = {"v": NullableColumn(v, valid), "i": NullableColumn(iv, ivalid),
            "d": DictColumn(codes, dg.DEPTS)}
    t = _table(ctx, cols)
    for p in [col("v") > 0, col("v") != 0.25, col("i") <= 3, col("d") == "dept02",
              col("d") != "dept02", col("v").is_null(), col("i").is_not_null(),
              col("d").is_null(), (col("v") > -1) | col("i").is_null()]:
        assert np.array_equal(p.eval_mask_bits(ctx, t), ora.filter_mask(cols, p))
        assert fa.count_star(ctx, t, p) == ora.filter_count(cols, p)


def test_zone_maps_prune_exactly(ctx):
    # Clustered columns: a sorted timestamp, a float drifting by batch with NaN
    # in one batch, an int32 constant per batch, NULLs in three batches.  Every
    # predicate gives the same mask, count and selection with zone maps on and
    # off, equal to the oracle; batches are skipped where the bounds allow.
    br, nb = 4096, 41
    n = (nb - 1) * br + 123
    rng = np.random.default_rng(17)
    bi = np.arange(n) // br
    drift = bi.astype(np.float64) + rng.random(n) * 0.5
    drift[5 * br: 5 * br + 300] = np.nan
    gvalid = ~((bi >= 10) & (bi <= 12) & (rng.random(n) < 0.5))
    cols = {"ts": np.arange(n, dtype=np.int64) * 3, "drift": drift, "k": bi.astype(np.int32),
            "g": NullableColumn(rng.integers(-50, 50, n), gvalid)}
    t = _table(ctx, cols, batch_rows=br)
    off = fa.Context(0)
    off.set_tuning("zone_maps", 0)
    cases = [  # (predicate, live batches expected when pruned, or None for all live)
        (col("ts") < 3 * br * 2, 2), (col("ts") >= 3 * (n - 100), 1),
        ((col("ts") > 1000) & (col("ts") < 50_000), 5), ((col("ts") < 5000) | (col("ts") > 3 * (n - 5000)), 4),
        (col("drift") < 2.0, 2), (col("drift") > 39.2, 2), (col("drift") != 5.25, None),
        (col("drift") == 7.3, 1), (col("k") != 7, nb - 1), (col("k") == 3, 1), (col("k") >= 38, 3),
        (col("g").is_null(), 3), (col("g").is_not_null(), None), (col("ts") < 1.5e5, 13),
        (col("k") < 3.5, 0), ((col("k") == 3) | col("g").is_null(), 4),
    ]
    for p, expect_live in cases:
        m_on = p.eval_mask_bits(ctx, t)
        plan = ctx.last_plan()
        m_off = p.eval_mask_bits(off, t)
        assert np.array_equal(m_on, m_off), (p.nodes, plan)
        assert np.array_equal(m_on, ora.filter_mask(cols, p, br)), p.nodes
        c_on = fa.count_star(ctx, t, p)
        assert c_on == fa.count_star(off, t, p) == ora.filter_count(cols, p, br)
        live = nb if expect_live is None else expect_live
        assert f"zone maps: {live} of {nb} batches live" in plan, (plan, p.nodes)
        g_on = fa.execute_select(ctx, t, p, ["ts", "g"])
        g_off = fa.execute_select(off, t, p, ["ts", "g"])
        assert_batches_equal(g_on, g_off, None)
    off.close()
    t.free()


def test_filter_select_materialises_in_row_order(ctx):
    n = 200_003
    t1 = dg.c1_table(n)
    cols = {"id": t1["id"], "dept": DictColumn(t1["dept"], dg.DEPTS), "score": t1["score"],
            "salary": t1["salary"]}
    t = _table(ctx, cols)
    p = col("score") > 50
    g = fa.execute_select(ctx, t, p, ["id", "score", "dept", "salary"])
    e = ora.filter_select(cols, p, ["id", "score", "dept", "salary"])
    assert_batches_equal(g, e, None)


# ─── GROUP BY ────────────────────────────────────────────────────────────────
C1_AGGS = [AggExpr(COUNT), AggExpr(SUM, "score"), AggExpr(AVG, "salary")]


def test_groupby_c1_one_million_rows(ctx):
    t1 = dg.c1_table(1_000_000)
    cols = {"id": t1["id"], "dept": DictColumn(t1["dept"], dg.DEPTS), "score": t1["score"],
            "salary": t1["salary"]}
    t = _table(ctx, cols)
    p = col("score") > 50
    g = fa.GroupByExec(["dept"], C1_AGGS).execute(ctx, t, p)
    o = ora.groupby(cols, p, "dept", C1_AGGS)
    assert g.names == ["dept", "col1", "col2", "col3"]
    assert g.dtypes == ["Utf8", "Int64", "Float64", "Float64"]
    assert_batches_equal(g, o.exact, "dept", o.reference)


C2_AGGS = [AggExpr(COUNT), AggExpr(SUM, "v1"), AggExpr(MIN, "v1"), AggExpr(MAX, "v1"),
           AggExpr(SUM, "v2")]


@pytest.fixture(params=["partitioned", "partitioned_small_cache", "hash"])
def gctx(request, ctx_gbp, ctx_gbp_small, ctx_gbh):
    return {"partitioned": ctx_gbp, "partitioned_small_cache": ctx_gbp_small, "hash": ctx_gbh}[request.param]


@pytest.mark.parametrize("n,nk", [(1000, 100), (300_000, 50_000), (2_000_000, 200_000)])
def test_groupby_high_cardinality_zipf(gctx, n, nk):
    ctx = gctx
    K = dg.keys(nk, 40)
    cols = dg.c2_table(n, K)
    t = _table(ctx, cols)
    p = col("f") < 40
    g = fa.GroupByExec(["key"], C2_AGGS).execute(ctx, t, p)
    plan = ctx.last_plan()
    if ctx.mode == "partitioned":
        assert plan.startswith("groupby:partitioned") and "fallback" not in plan, plan
    else:
        assert plan.split(" zone maps:")[0] == "groupby:hash", plan
    o = ora.groupby(cols, p, "key", C2_AGGS)
    assert_batches_equal(g, o.exact, "key", o.reference)
    # reference output typing: SUM/MIN/MAX of Int64 come back as Float64
    assert g.to_reference_types().dtypes == ["Int64", "Int64", "Float64", "Float64", "Float64", "Float64"]


@pytest.mark.parametrize("pred,tiles", [
    (col("f") < 40, "fast4"), (col("f") != 7, "fast4"), (col("f") >= -3, "fast4"),
    (col("v2") < 0.5, "fast8"), (col("v2") < 1, "fast8"), (col("v1") >= 0, "fast8"),
    (col("v1") < 0.5, "fast8"), (col("key") > 2**39, "fast8"),
    (col("f") < 40.5, "generic"),  # int32 column vs float literal: no row passes
    ((col("f") < 40) & (col("v1") > 0), "generic"),
])
def test_groupby_partitioned_tile_paths(ctx_gbp, pred, tiles):
    # One comparison on a column without NULLs, an INT64 key and 64-bit inputs
    # take the scan's specialised tile loads; INT64_MIN keys still fold into
    # the NULL group there.
    K = dg.keys(40_000, 40)
    cols = dg.c2_table(300_000, K)
    cols["key"][::101] = np.iinfo(np.int64).min
    t = _table(ctx_gbp, cols)
    g = fa.GroupByExec(["key"], C2_AGGS).execute(ctx_gbp, t, pred)
    plan = ctx_gbp.last_plan()
    assert plan.startswith("groupby:partitioned") and f"tiles={tiles}" in plan, plan
    o = ora.groupby(cols, pred, "key", C2_AGGS)
    imin = np.iinfo(np.int64).min
    for r in (g, o.exact, o.reference):
        if r.validity[0] is not None:
            r.columns[0] = np.where(r.validity[0], r.columns[0], imin)
            r.validity[0] = None
    assert_batches_equal(g, o.exact, "key", o.reference)


@pytest.mark.parametrize("path", ["dense", "partitioned", "hash", "scalar", "join"])
def test_float_minmax_nan_semantics(ctx, ctx_gbp, ctx_gbh, ctx_part, path):
    # MIN / MAX fold like f64::min / f64::max: NaN skipped, a group of NaN
    # values alone gives NaN, a group without values NULL; equal to the oracle
    rng = np.random.default_rng(91)
    n = 200_000
    wide = path in ("partitioned", "hash")
    key = rng.integers(0, 50_000 if wide else 40, n)
    x = rng.normal(size=n)
    x[rng.random(n) < 0.3] = np.nan
    x[key < 5] = np.nan               # groups of NaN alone
    x[key == 7] = np.where(np.arange(n)[key == 7] % 2 == 0, np.nan, 2.5)  # NaN first in row order
    aggs = [AggExpr(MIN, "x"), AggExpr(MAX, "x"), AggExpr(COUNT)]
    cols = {"key": key, "x": x}
    if path == "join":
        build = {"bk": np.arange(0, 60, dtype=np.int64), "bx": rng.random(60)}
        pt, bt = _table(ctx_part, cols), _table(ctx_part, build)
        for jt in (INNER, LEFT, SEMI, ANTI):
            g = fa.execute_hash_join(ctx_part, pt, bt, "key", "bk", jt, aggs)
            assert_batches_equal(g, ora.join(cols, build, "key", "bk", jt, aggs).exact, None)
        return
    c = {"dense": ctx, "partitioned": ctx_gbp, "hash": ctx_gbh, "scalar": ctx}[path]
    t = _table(c, cols)
    keyname = None if path == "scalar" else "key"
    for p in (None, col("key") < 6):
        g = (fa.GroupByExec([keyname], aggs).execute(c, t, p) if keyname
             else fa.execute_agg_only(c, t, aggs, p))
        o = ora.groupby(cols, p, keyname, aggs)
        assert_batches_equal(g, o.exact, keyname)
    if path == "scalar":  # NaN alone, and no rows at all
        for p, expect_nan in ((col("key") < 5, True), (col("key") > 10**9, False)):
            g = fa.execute_agg_only(c, t, aggs[:2], p)
            if expect_nan:
                assert np.isnan(g.columns[0][0]) and np.isnan(g.columns[1][0])
            else:
                assert not g.validity[0][0] and not g.validity[1][0]


@pytest.mark.parametrize("path", ["scalar", "dense", "partitioned", "hash", "join"])
def test_float_sum_wide_exponents(ctx, ctx_gbp, ctx_gbh, ctx_part, path):
    # Exponents spanning ~330 binades (1e-50 .. 1e50) fit the 15-word exact
    # accumulator: every path equals the oracle's exact sums.  Spanning ~1660
    # binades (1e-250 .. 1e250): aggregates with one output row (no GROUP BY,
    # joins) round values to the window's LSB and stay within 1e-12 of the
    # exact sum when it does not cancel; GROUP BY refuses the column, since a
    # group far below the column's largest values would lose precision.
    rng = np.random.default_rng(77)
    n = 100_000
    key = rng.integers(0, 30_000 if path in ("partitioned", "hash") else 50, n)
    x = rng.choice([-1.0, 1.0], n) * 10.0 ** rng.uniform(-50, 50, n)
    w = np.abs(rng.choice([-1.0, 1.0], n) * 10.0 ** rng.uniform(-250, 250, n))
    cols = {"key": key, "x": x, "w": w}
    aggs = [AggExpr(SUM, "x"), AggExpr(AVG, "x"), AggExpr(COUNT)]

    def close(a, v):
        e = math.fsum(v)
        return abs(a - e) <= 1e-12 * abs(e)
    if path == "join":
        build = {"bk": np.arange(0, 50, dtype=np.int64), "bx": rng.random(50)}
        pt, bt = _table(ctx_part, cols), _table(ctx_part, build)
        for jt in (INNER, SEMI):
            g = fa.execute_hash_join(ctx_part, pt, bt, "key", "bk", jt, aggs)
            assert_batches_equal(g, ora.join(cols, build, "key", "bk", jt, aggs).exact, None)
            g = fa.execute_hash_join(ctx_part, pt, bt, "key", "bk", jt, [AggExpr(SUM, "w")])
            assert close(g.columns[0][0], w[key < 50])
        return
    c = {"scalar": ctx, "dense": ctx, "partitioned": ctx_gbp, "hash": ctx_gbh}[path]
    t = _table(c, cols)
    if path == "scalar":
        assert_batches_equal(fa.execute_agg_only(c, t, aggs), ora.groupby(cols, None, None, aggs).exact, None)
        assert close(fa.execute_agg_only(c, t, [AggExpr(SUM, "w")]).columns[0][0], w)
        return
    g = fa.GroupByExec(["key"], aggs).execute(c, t)
    assert_batches_equal(g, ora.groupby(cols, None, "key", aggs).exact, "key")
    with pytest.raises(fa.FalconError) as e:
        fa.GroupByExec(["key"], [AggExpr(SUM, "w")]).execute(c, t)
    assert e.value.code == 102 and "too wide" in str(e.value)


@pytest.mark.parametrize("case", ["c2_like", "negative_ranges", "int32_input", "fits_64_bits", "no_fit"])
def test_groupby_packed_records(ctx_gbp, case):
    # The partitioned GROUP BY packs a row record's key and integer input 0
    # into one word when their ranges fit 64 bits together: identical results
    # with the packing on and off, and equal to the oracle.
    rng = np.random.default_rng(sum(map(ord, case)))
    n = 1_500_000
    if case == "c2_like":
        K = dg.keys(40_000, 40, seed_=dg.seed(2, 11))
        cols = dg.c2_table(n, K)
        key, v, expect = "key", "v1", True
    else:
        if case == "negative_ranges":
            k = -(1 << 45) + rng.integers(0, 300_000, n) * 1009  # sparse: not a dense range
            v1 = rng.integers(-(1 << 15), 1 << 15, n)
        elif case == "int32_input":
            k = rng.integers(0, 200_000, n) * 7919
            v1 = rng.integers(-(1 << 31), (1 << 31) - 1, n).astype(np.int32)
        elif case == "fits_64_bits":
            k = rng.integers(0, 1 << 50, 200_000)[rng.integers(0, 200_000, n)]
            v1 = rng.integers(-(1 << 13), 1 << 13, n)
        else:
            k = rng.integers(0, 1 << 55, 200_000)[rng.integers(0, 200_000, n)]
            v1 = rng.integers(-(1 << 20), 1 << 20, n)
        cols = {"key": k, "v1": v1, "v2": rng.normal(size=n), "f": rng.integers(0, 100, n).astype(np.int32)}
        key, v, expect = "key", "v1", case != "no_fit"
    aggs = [AggExpr(SUM, v), AggExpr(MIN, v), AggExpr(MAX, v), AggExpr(SUM, "v2"), AggExpr(COUNT)]
    t = _table(ctx_gbp, cols)
    res = []
    for pack in (1, 0):
        ctx_gbp.set_tuning("gbp_pack", pack)
        res.append(fa.GroupByExec([key], aggs).execute(ctx_gbp, t, col("f") < 60))
        plan = ctx_gbp.last_plan()
        assert plan.startswith("groupby:partitioned") and ("packed" in plan) == (pack == 1 and expect), plan
    ctx_gbp.set_tuning("gbp_pack", 1)
    assert_batches_equal(res[0], res[1], key)
    assert_batches_equal(res[0], ora.groupby(cols, col("f") < 60, key, aggs).exact, key)


def _with_specials(x, rng, pinf=0, ninf=0, nan=0):
    """x with `pinf` +inf, `ninf` -inf and `nan` NaN values at distinct random rows."""
    x = x.copy()
    rows = rng.choice(x.size, pinf + ninf + nan, replace=False)
    x[rows[:pinf]] = np.inf
    x[rows[pinf:pinf + ninf]] = -np.inf
    x[rows[pinf + ninf:]] = np.nan
    return x


@pytest.mark.parametrize("path", ["hash", "partitioned"])
def test_groupby_float_sums_with_inf_and_nan(ctx, ctx_gbp, path):
    # a group whose summands hold +inf sums to +inf, one with both infinities or
    # a NaN to NaN (IEEE addition, in any order); the other groups keep their
    # exact sums
    c = ctx if path == "hash" else ctx_gbp
    K = dg.keys(40_000, 40)
    cols = dg.c2_table(300_000, K)
    v2 = cols["v2"].copy()
    key = cols["key"]
    groups, counts = np.unique(key, return_counts=True)
    g_pinf, g_both, g_nan = groups[counts >= 2][:3]
    v2[np.flatnonzero(key == g_pinf)[:2]] = np.inf
    both = np.flatnonzero(key == g_both)
    v2[both[0]], v2[both[-1]] = np.inf, -np.inf
    v2[np.flatnonzero(key == g_nan)[0]] = np.nan
    cols["v2"] = v2
    t = _table(c, cols)
    aggs = [AggExpr(COUNT), AggExpr(SUM, "v2"), AggExpr(SUM, "v1")]
    g = fa.GroupByExec(["key"], aggs).execute(c, t, None)
    if path == "partitioned":
        assert c.last_plan().startswith("groupby:partitioned"), c.last_plan()
    o = ora.groupby(cols, None, "key", aggs)
    assert_batches_equal(g, o.exact, "key")
    res = dict(zip(g.columns[0].tolist(), g.columns[2].tolist()))
    assert res[int(g_pinf)] == np.inf and np.isnan(res[int(g_both)]) and np.isnan(res[int(g_nan)])
    assert sum(np.isfinite(x) for x in res.values()) == len(res) - 3


@pytest.mark.parametrize("card", [8, 1000, 70_000, 1_000_000])
def test_groupby_int32_dense_keys(ctx, card):
    cols = dg.c4_batch(400_000, card, 1234 + card)
    t = _table(ctx, cols)
    p = col("score") < 37.5
    aggs = [AggExpr(COUNT), AggExpr(SUM, "val"), AggExpr(MIN, "val"), AggExpr(MAX, "val"),
            AggExpr(SUM, "score")]
    g = fa.GroupByExec(["key"], aggs).execute(ctx, t, p)
    o = ora.groupby(cols, p, "key", aggs)
    assert_batches_equal(g, o.exact, "key", o.reference)


def test_groupby_nulls_and_float_minmax(gctx):
    ctx = gctx
    n = 50_000
    rng = np.random.default_rng(11)
    key = rng.integers(-50, 50, n) * 1_000_003  # sparse: hash path, not dense
    kvalid = rng.random(n) > 0.1
    key[::97] = np.iinfo(np.int64).min  # folds into the NULL group like the reference
    x = rng.normal(size=n) * 1e3
    xvalid = rng.random(n) > 0.25
    y = rng.integers(-10**12, 10**12, n)
    cols = {"k": NullableColumn(key, kvalid), "x": NullableColumn(x, xvalid), "y": y}
    t = _table(ctx, cols)
    aggs = [AggExpr(COUNT), AggExpr(COUNT, "x"), AggExpr(SUM, "x"), AggExpr(AVG, "x"),
            AggExpr(MIN, "x"), AggExpr(MAX, "x"), AggExpr(AVG, "y"), AggExpr(SUM, "y")]
    g = fa.GroupByExec(["k"], aggs).execute(ctx, t, None)
    o = ora.groupby(cols, None, "k", aggs)
    # NULL group sorts by its placeholder key; compare keyed by (valid, key)
    gk = np.where(g.validity[0], g.columns[0], np.iinfo(np.int64).min)
    ok = np.where(o.exact.validity[0], o.exact.columns[0], np.iinfo(np.int64).min)
    g.columns[0], o.exact.columns[0] = gk, ok
    assert_batches_equal(g, o.exact, "k")


def test_upload_batches_merges_dictionaries(ctx):
    # A loader-shaped text column: one dictionary per 65,536-row batch.  Upload
    # merges them like concat_all (remapping the codes on the device); every
    # query then equals the same strings uploaded with one table-wide dictionary.
    rng = np.random.default_rng(21)
    vocab = [f"dept{i:02d}" for i in range(20)] + [""]
    rows = [65536] * 30 + [12345]
    dicts, parts = [], []
    for n in rows:
        d = [str(s) for s in rng.permutation(vocab)[: int(rng.integers(1, len(vocab) + 1))]]
        c = rng.integers(0, len(d), n).astype(np.uint32)
        c[rng.random(n) < 0.05] = 0xFFFFFFFF
        dicts.append(d)
        parts.append(c)
    codes = np.concatenate(parts)
    x = rng.normal(size=codes.size)
    m = ora.concat_dicts(codes, rows, dicts)
    tb = _table(ctx, {"d": fa.BatchDictColumn(codes, rows, dicts), "x": x})
    tm = _table(ctx, {"d": DictColumn(m.codes, m.dictionary), "x": x})
    # row by row: the strings the batches' own dictionaries give
    sel = fa.execute_select(ctx, tb, None, ["d"]).columns[0]
    expect = [None if c == 0xFFFFFFFF else m.dictionary[c] for c in m.codes.tolist()]
    assert list(sel) == expect
    # the device codes are the merged numbering itself
    assert list(fa.execute_select(ctx, tm, None, ["d"]).columns[0]) == expect
    aggs = [AggExpr(COUNT), AggExpr(SUM, "x")]
    gb = fa.GroupByExec(["d"], aggs).execute(ctx, tb, col("x") > 0)
    gm = fa.GroupByExec(["d"], aggs).execute(ctx, tm, col("x") > 0)
    assert_batches_equal(gb, gm, "d")
    for lit in ("dept03", "", "zzz"):
        assert fa.count_star(ctx, tb, col("d") == lit) == fa.count_star(ctx, tm, col("d") == lit)
    assert fa.count_star(ctx, tb, col("d") < "dept10") == fa.count_star(ctx, tm, col("d") < "dept10")
    # batches sharing one dictionary keep their codes
    same = _table(ctx, {"d": fa.BatchDictColumn(parts[0][:300], [100, 200], [dicts[0], list(dicts[0])])})
    assert list(fa.execute_select(ctx, same, None, ["d"]).columns[0]) == \
        [None if c == 0xFFFFFFFF else dicts[0][c] for c in parts[0][:300].tolist()]
    # a code past its batch's dictionary is an error, not a read past the map
    bad = codes.copy()
    bad[65536 + 7] = len(dicts[1])
    with pytest.raises(fa.FalconError) as e:
        _table(ctx, {"d": fa.BatchDictColumn(bad, rows, dicts)})
    assert e.value.code == 3  # IndexOutOfBounds
    for t in (tb, tm, same):
        t.free()


def test_groupby_utf8_with_null_keys(ctx):
    n = 30_000
    rng = np.random.default_rng(5)
    codes = rng.integers(0, 5, n).astype(np.uint32)
    codes[::7] = 0xFFFFFFFF
    d = ["a", "b", "", "c", "d"]  # NULL joins the "" group
    cols = {"g": DictColumn(codes, d), "v": rng.integers(0, 100, n)}
    t = _table(ctx, cols)
    aggs = [AggExpr(COUNT), AggExpr(SUM, "v")]
    g = fa.GroupByExec(["g"], aggs).execute(ctx, t, col("v") > 10)
    o = ora.groupby(cols, col("v") > 10, "g", aggs)
    assert_batches_equal(g, o.exact, "g")


def test_scalar_aggregates(ctx):
    n = 1_234_567
    rng = np.random.default_rng(8)
    cols = {"a": rng.normal(size=n), "b": rng.integers(-10**6, 10**6, n),
            "c": rng.integers(0, 100, n).astype(np.int32)}
    t = _table(ctx, cols)
    aggs = [AggExpr(COUNT), AggExpr(SUM, "a"), AggExpr(MIN, "b"), AggExpr(MAX, "a"),
            AggExpr(AVG, "b"), AggExpr(SUM, "c")]
    for p in [None, col("c") < 30, col("c") > 1000]:
        g = fa.execute_agg_only(ctx, t, aggs, p)
        o = ora.groupby(cols, p, None, aggs)
        assert_batches_equal(g, o.exact, None)


def test_closed_form_sums(ctx):
    # diag_agg: score = (i % 100) + 0.5 over 100k rows -> SUM 5,000,000, MIN 0.5,
    # MAX 99.5, AVG 50, SUM(id) = n(n-1)/2 (tests/diag_agg.rs:51-111;
    # benches/query_bench.rs:697)
    n = 100_000
    i = np.arange(n, dtype=np.int64)
    t = _table(ctx, {"id": i, "score": (i % 100) + 0.5}, batch_rows=1024)
    r = fa.execute_agg_only(ctx, t, [AggExpr(SUM, "score"), AggExpr(COUNT), AggExpr(MIN, "score"),
                                     AggExpr(MAX, "score"), AggExpr(AVG, "score"),
                                     AggExpr(SUM, "id")])
    assert [c[0] for c in r.columns] == [5_000_000.0, 100_000, 0.5, 99.5, 50.0, n * (n - 1) // 2]


# ─── JOIN ────────────────────────────────────────────────────────────────────
def _c3(nb, np_):
    K = dg.keys(nb, 40, seed_=99)
    b = dg.c3_build(K)
    p = dg.c3_probe(np_, K)
    return p, b


@pytest.fixture(params=["partitioned", "partitioned_compact", "partitioned_hash64", "direct"])
def jctx(request, ctx_part, ctx_part_compact, ctx_part_hash64, ctx_direct):
    return {"partitioned": ctx_part, "partitioned_compact": ctx_part_compact,
            "partitioned_hash64": ctx_part_hash64, "direct": ctx_direct}[request.param]


@pytest.mark.parametrize("jt", [INNER, LEFT, SEMI, ANTI])
def test_join_aggregates_match_oracle(jctx, jt):
    ctx = jctx
    pcols, bcols = _c3(20_000, 400_000)
    pt, bt = _table(ctx, pcols), _table(ctx, bcols)
    if jt in (INNER, LEFT):
        aggs = [AggExpr(SUM, "bpay", side="build"), AggExpr(COUNT)]
    else:
        aggs = [AggExpr(SUM, "ppay"), AggExpr(COUNT)]
    g = fa.execute_hash_join(ctx, pt, bt, "pk", "bk", jt, aggs)
    o = ora.join(pcols, bcols, "pk", "bk", jt, aggs)
    assert_batches_equal(g, o.exact, None, o.reference)
    idx = fa.JoinIndex(ctx, bt, "bk")
    g2 = fa.execute_hash_join(ctx, pt, bt, "pk", "bk", jt, aggs, index=idx)
    assert_batches_equal(g2, o.exact, None)
    # pairs as multisets
    gl, gr = fa.join_pairs(ctx, pt, bt, "pk", "bk", jt)
    _, ol, orr = ora.join(pcols, bcols, "pk", "bk", jt, [], pairs=True)
    gp = np.sort(gl.astype(np.uint64) << 32 | gr)
    op = np.sort(ol.astype(np.uint64) << 32 | orr)
    assert np.array_equal(gp, op)


def test_join_stride_keys_closed_form(jctx):
    ctx = jctx
    # src/query/join.rs:1747-1848: 70,000 left rows, right = every 5th id ->
    # INNER 14,000, LEFT 70,000, SEMI 14,000, ANTI 56,000
    left = {"id": np.arange(70_000, dtype=np.int64), "lv": np.arange(70_000, dtype=np.int64) * 2}
    right = {"rid": np.arange(0, 70_000, 5, dtype=np.int64),
             "rv": np.arange(0, 70_000, 5, dtype=np.float64)}
    lt, rt = _table(ctx, left), _table(ctx, right)
    expect = {INNER: 14_000, LEFT: 70_000, SEMI: 14_000, ANTI: 56_000}
    for jt, n in expect.items():
        r = fa.execute_hash_join(ctx, lt, rt, "id", "rid", jt, [AggExpr(COUNT)])
        assert r.columns[0][0] == n


# The key whose 64-bit finaliser hash (MurmurHash3 fmix64) is 0x8000000000000000,
# the empty-slot marker of the partitioned join's hash-keyed sub-tables.
HASH_SENTINEL_KEY = 1401494638771588894


def test_hash_sentinel_key_is_the_marker_preimage():
    m = (1 << 64) - 1
    h = HASH_SENTINEL_KEY
    h ^= h >> 33
    h = h * 0xff51afd7ed558ccd & m
    h ^= h >> 33
    h = h * 0xc4ceb9fe1a85ec53 & m
    h ^= h >> 33
    assert h == 1 << 63


def test_join_duplicates_nulls_and_sentinel_keys(jctx):
    ctx = jctx
    rng = np.random.default_rng(21)
    nb, np_ = 3000, 20000
    bk = rng.integers(0, 500, nb)
    bk[::101] = np.iinfo(np.int64).min  # a real key equal to the direct index's empty-slot marker
    bk[7::211] = HASH_SENTINEL_KEY       # a real key whose hash is the partitioned tables' marker
    bvalid = rng.random(nb) > 0.05
    pk = rng.integers(0, 700, np_)
    pk[::203] = np.iinfo(np.int64).min
    pk[5::307] = HASH_SENTINEL_KEY
    pvalid = rng.random(np_) > 0.05
    bcols = {"bk": NullableColumn(bk, bvalid), "bx": rng.normal(size=nb)}
    pcols = {"pk": NullableColumn(pk, pvalid), "px": rng.integers(-50, 50, np_)}
    pt, bt = _table(ctx, pcols), _table(ctx, bcols)
    for jt in (INNER, LEFT, SEMI, ANTI):
        gl, gr = fa.join_pairs(ctx, pt, bt, "pk", "bk", jt)
        _, ol, orr = ora.join(pcols, bcols, "pk", "bk", jt, [], pairs=True)
        assert np.array_equal(np.sort(gl.astype(np.uint64) << 32 | gr),
                              np.sort(ol.astype(np.uint64) << 32 | orr)), jt
        aggs = [AggExpr(COUNT), AggExpr(SUM, "px"), AggExpr(MAX, "px")]
        if jt in (INNER, LEFT):
            aggs += [AggExpr(SUM, "bx", side="build")]
        g = fa.execute_hash_join(ctx, pt, bt, "pk", "bk", jt, aggs)
        o = ora.join(pcols, bcols, "pk", "bk", jt, aggs)
        assert_batches_equal(g, o.exact, None)


def test_join_with_probe_predicate(jctx):
    ctx = jctx
    pcols, bcols = _c3(5000, 100_000)
    pcols["f"] = (np.arange(100_000) % 100).astype(np.int32)
    pt, bt = _table(ctx, pcols), _table(ctx, bcols)
    p = col("f") < 30
    for jt in (INNER, SEMI):
        aggs = [AggExpr(COUNT), AggExpr(SUM, "ppay")]
        g = fa.execute_hash_join(ctx, pt, bt, "pk", "bk", jt, aggs, filter=p)
        o = ora.join(pcols, bcols, "pk", "bk", jt, aggs, expr=p)
        assert_batches_equal(g, o.exact, None)


@pytest.mark.parametrize("jt", [INNER, LEFT, SEMI, ANTI])
def test_join_float_sums_with_inf_and_nan(jctx, jt):
    # non-finite summands on both sides: +inf alone gives +inf, both infinities
    # or a NaN give NaN (IEEE addition, in any order); columns without them
    # keep their exact sums
    ctx = jctx
    rng = np.random.default_rng(41)
    pcols, bcols = _c3(20_000, 400_000)
    bcols["b_pinf"] = _with_specials(bcols["bpay"], rng, pinf=3)
    bcols["b_both"] = _with_specials(bcols["bpay"], rng, pinf=2, ninf=2)
    bcols["b_nan"] = _with_specials(bcols["bpay"], rng, nan=1)
    pcols["pf"] = _with_specials(rng.normal(size=pcols["pk"].size), rng, ninf=40)
    pt, bt = _table(ctx, pcols), _table(ctx, bcols)
    cases = [("pf", [AggExpr(SUM, "pf"), AggExpr(COUNT)])]
    if jt in (INNER, LEFT):
        cases += [(c, [AggExpr(SUM, c, side="build"), AggExpr(COUNT)]) for c in ("bpay", "b_pinf", "b_both", "b_nan")]
    for name, aggs in cases:
        g = fa.execute_hash_join(ctx, pt, bt, "pk", "bk", jt, aggs)
        o = ora.join(pcols, bcols, "pk", "bk", jt, aggs)
        assert_batches_equal(g, o.exact, None)
        s = g.columns[0][0]
        if jt == INNER:  # every build row has ~18 probe matches, so each special is summed
            expect = {"bpay": np.isfinite, "b_pinf": lambda x: x == np.inf, "b_both": np.isnan,
                      "b_nan": np.isnan, "pf": lambda x: x == -np.inf}[name]
            assert expect(s), (name, s)


@pytest.mark.parametrize("fmt", ["compact", "hash64"])
def test_join_skewed_probe_keys_exact_partitions(ctx_part_compact, ctx_part_hash64, fmt):
    # Zipf(1.1) probe keys: the hottest key alone fills its partition's cells,
    # so the partitioned join redoes the partitioning exactly (count, offsets,
    # chunked pass 2); results must still equal the oracle's
    ctx_part = ctx_part_compact if fmt == "compact" else ctx_part_hash64
    K = dg.keys(20_000, 40, seed_=dg.seed(3, 9))
    probe = dg.c2_table(3_000_000, K)
    pcols = {"pk": probe["key"], "ppay": probe["v1"], "f": probe["f"]}
    bcols = dg.c3_build(K)
    pt, bt = _table(ctx_part, pcols), _table(ctx_part, bcols)
    idx = fa.JoinIndex(ctx_part, bt, "bk")
    # with a predicate column and without one (the larger staging tiles)
    for p in (col("f") < 60, None):
        for jt in (INNER, LEFT, SEMI, ANTI):
            aggs = [AggExpr(COUNT), AggExpr(SUM, "ppay"), AggExpr(MAX, "ppay")]
            if jt in (INNER, LEFT):
                aggs += [AggExpr(SUM, "bpay", side="build")]
            g = fa.execute_hash_join(ctx_part, pt, bt, "pk", "bk", jt, aggs, filter=p, index=idx)
            assert "exact partitions" in ctx_part.last_plan(), ctx_part.last_plan()
            assert f"records={fmt}" in ctx_part.last_plan(), ctx_part.last_plan()
            o = ora.join(pcols, bcols, "pk", "bk", jt, aggs, expr=p)
            assert_batches_equal(g, o.exact, None)
    idx.free()


def test_join_partitioned_matches_direct_at_scale(ctx_part, ctx_direct):
    # 4M probe rows over a 200k build: both paths, several aggregates per side
    pcols, bcols = _c3(200_000, 4_000_000)
    pcols["f"] = (np.arange(4_000_000) % 100).astype(np.int32)
    for jt in (INNER, LEFT, SEMI, ANTI):
        aggs = [AggExpr(COUNT), AggExpr(SUM, "ppay"), AggExpr(MIN, "ppay")]
        if jt in (INNER, LEFT):
            aggs += [AggExpr(SUM, "bpay", side="build")]
        res = []
        for c in (ctx_part, ctx_direct):
            pt, bt = _table(c, pcols), _table(c, bcols)
            res.append(fa.execute_hash_join(c, pt, bt, "pk", "bk", jt, aggs, filter=col("f") >= 25))
        assert_batches_equal(res[0], res[1], None)


def test_join_index_keeps_partitioned_sub_tables(ctx_part, ctx_direct):
    # The first partitioned join through an index builds the build side's
    # sub-tables into it; joins holding another build column in the slots add
    # their own; repeats reuse them and give the same results as the direct path.
    pcols, bcols = _c3(150_000, 3_000_000)
    pcols["f"] = (np.arange(3_000_000) % 100).astype(np.int32)
    bcols["bpay2"] = np.arange(bcols["bk"].size, dtype=np.float64) * 0.5
    pt, bt = _table(ctx_part, pcols), _table(ctx_part, bcols)
    idx = fa.JoinIndex(ctx_part, bt, "bk")
    a1 = [AggExpr(COUNT), AggExpr(SUM, "bpay", side="build")]
    a2 = [AggExpr(COUNT), AggExpr(MAX, "bpay2", side="build"), AggExpr(SUM, "ppay")]
    sizes, got = [idx.device_bytes()], []
    for jt, aggs in ((INNER, a1), (LEFT, a1), (INNER, a2), (INNER, a1), (SEMI, a2[:1] + a2[2:]), (LEFT, a2)):
        got.append(fa.execute_hash_join(ctx_part, pt, bt, "pk", "bk", jt, aggs, filter=col("f") < 70, index=idx))
        assert ctx_part.last_plan().startswith("join:partitioned"), ctx_part.last_plan()
        sizes.append(idx.device_bytes())
    # new entries: bpay slots (1st join), bpay2 slots (3rd), build-row slots (5th)
    assert sizes[1] > sizes[0] and sizes[2] == sizes[1] and sizes[3] > sizes[2]
    assert sizes[4] == sizes[3] and sizes[5] > sizes[4] and sizes[6] == sizes[5]
    pd, bd = _table(ctx_direct, pcols), _table(ctx_direct, bcols)
    for (jt, aggs), g in zip(((INNER, a1), (LEFT, a1), (INNER, a2), (INNER, a1), (SEMI, a2[:1] + a2[2:]), (LEFT, a2)),
                             got):
        e = fa.execute_hash_join(ctx_direct, pd, bd, "pk", "bk", jt, aggs, filter=col("f") < 70)
        assert_batches_equal(g, e, None)
    idx.free()


def test_groupby_partitioned_overflow_falls_back(ctx_gbp):
    # A fan-out and pass-2 table far too small for the keys: the bound is
    # detected, and the query is answered by the single-table hash path.
    K = dg.keys(50_000, 40)
    cols = dg.c2_table(400_000, K)
    c = fa.Context(0)
    c.set_tuning("gb_part_min_rows", 1)
    c.set_tuning("gbp_table_slots", 64)
    c.set_tuning("gbp_partitions", 8)
    t = _table(c, cols)
    g = fa.GroupByExec(["key"], C2_AGGS).execute(c, t, col("f") < 40)
    assert "fallback" in c.last_plan(), c.last_plan()
    o = ora.groupby(cols, col("f") < 40, "key", C2_AGGS)
    assert_batches_equal(g, o.exact, "key")
    c.close()


def test_groupby_partitioned_cell_overflow_resizes(ctx_gbp):
    # No hot-key cache and a fine fan-out: the Zipf head overflows its cells in
    # pass 1; the retry sizes cells from the counts it kept and is exact.
    K = dg.keys(50_000, 40)
    cols = dg.c2_table(400_000, K)
    c = fa.Context(0)
    c.set_tuning("gb_part_min_rows", 1)
    c.set_tuning("gbp_cache_slots", 0)
    c.set_tuning("gbp_partitions", 256)
    t = _table(c, cols)
    g = fa.GroupByExec(["key"], C2_AGGS).execute(c, t, col("f") < 40)
    plan = c.last_plan()
    assert "resized" in plan and "fallback" not in plan, plan
    o = ora.groupby(cols, col("f") < 40, "key", C2_AGGS)
    assert_batches_equal(g, o.exact, "key")
    c.close()


def test_groupby_partitioned_at_scale_matches_hash(ctx_gbp, ctx_gbh):
    K = dg.keys(1_000_000, 40)
    cols = dg.c2_table(20_000_000, K)
    res, plans = [], []
    for c in (ctx_gbp, ctx_gbh):
        t = _table(c, cols)
        res.append(fa.GroupByExec(["key"], C2_AGGS).execute(c, t, col("f") < 40))
        plans.append(c.last_plan())
        t.free()
    assert "partitioned" in plans[0] and "fallback" not in plans[0], plans[0]
    assert plans[1].split(" zone maps:")[0] == "groupby:hash", plans[1]
    assert res[0].num_rows > 500_000
    assert_batches_equal(res[0], res[1], "key")


# ─── register float sums near the top of the double range ───────────────────
def _fraction_sum(x):
    """The correctly rounded sum by exact rational arithmetic."""
    from fractions import Fraction
    v, c = np.unique(x, return_counts=True)
    return float(sum(Fraction(float(a)) * int(n) for a, n in zip(v, c)))


def _huge_column(n):
    # +/-1.7e308 alternating: every thread of a strided loop sees one sign, so
    # its partial sum would overflow without the power-of-two prescale; every
    # 1000th pair is 1e300, and the column is balanced so that the exact total
    # is one 1.7e308 plus the 1e300 terms
    x = np.where(np.arange(n) % 2 == 0, 1.7e308, -1.7e308)
    x[::1000] = 1e300
    x[1::1000] = 1e300
    extra = int((x == 1.7e308).sum() - (x == -1.7e308).sum()) - 1
    idx = np.flatnonzero(x == (1.7e308 if extra > 0 else -1.7e308))[:abs(extra)]
    x[idx] = 0.0
    return x


@pytest.mark.parametrize("n", [10_000, 1_000_003])
def test_scalar_float_sum_near_dbl_max(ctx, n):
    x = _huge_column(n)
    t = _table(ctx, {"x": x})
    g = fa.execute_agg_only(ctx, t, [AggExpr(SUM, "x"), AggExpr(AVG, "x")])
    exact = _fraction_sum(x)  # math.fsum raises on the overflowing partials
    assert math.isfinite(exact)
    assert g.columns[0][0] == exact == ora.exact_sum(x)
    assert g.columns[1][0] == exact / n


@pytest.mark.parametrize("jt", [INNER, SEMI])
def test_join_float_sum_near_dbl_max(jctx, jt):
    n = 300_000
    x = _huge_column(n)
    probe = {"pk": np.arange(n, dtype=np.int64), "px": x}
    build = {"bk": np.arange(n, dtype=np.int64), "bx": x}
    side = "build" if jt == INNER else "probe"
    aggs = [AggExpr(SUM, "bx" if jt == INNER else "px", side=side), AggExpr(COUNT)]
    g = fa.execute_hash_join(jctx, _table(jctx, probe), _table(jctx, build), "pk", "bk", jt, aggs)
    assert g.columns[0][0] == _fraction_sum(x)
    assert g.columns[1][0] == n


# ─── MIN / MAX over NaN ──────────────────────────────────────────────────────
def test_float_minmax_skip_nan_even_when_first(ctx, ctx_gbp, ctx_gbh):
    # A group whose first value is NaN: MIN / MAX skip NaN on every path, the
    # way the reference's execute_agg_only does (f64::min / f64::max); the
    # reference's GROUP BY fast path would keep the leading NaN (its
    # `v < state.sum` never replaces it), which the oracle's reference-order
    # columns restate.
    k = np.array([0, 0, 0, 1, 1, 2, 2], np.int64)
    x = np.array([np.nan, 1.0, -2.0, 3.0, np.nan, np.nan, np.nan])
    cols = {"k": k, "x": x}
    aggs = [AggExpr(MIN, "x"), AggExpr(MAX, "x"), AggExpr(COUNT, "x")]
    o = ora.groupby(cols, None, "k", aggs)
    exp = o.exact.sort_by("k")
    assert list(exp.columns[1][:2]) == [-2.0, 3.0] and list(exp.columns[2][:2]) == [1.0, 3.0]
    # a group of NaN values alone: NaN, as f64::min / f64::max fold it
    assert exp.validity[1][2] and exp.validity[2][2]
    assert math.isnan(exp.columns[1][2]) and math.isnan(exp.columns[2][2])
    ref = o.reference.sort_by("k")
    assert math.isnan(ref.columns[1][0]) and math.isnan(ref.columns[2][0])
    for c in (ctx, ctx_gbp, ctx_gbh):
        t = _table(c, cols)
        assert_batches_equal(fa.GroupByExec(["k"], aggs).execute(c, t), o.exact, "k")
        s = fa.execute_agg_only(c, t, aggs)
        assert s.columns[0][0] == -2.0 and s.columns[1][0] == 3.0


@pytest.mark.parametrize("jt", [INNER, LEFT, SEMI, ANTI])
def test_join_many_aggregates_run_in_parts(jctx, jt):
    # more aggregates than the probe kernels' four register accumulators: the
    # call runs them in parts (one join per part) and joins the columns in order
    ctx = jctx
    pcols, bcols = _c3(20_000, 300_000)
    rng = np.random.default_rng(61)
    pcols["w"] = rng.integers(-1000, 1000, 300_000)
    pcols["z"] = rng.normal(size=300_000)
    pt, bt = _table(ctx, pcols), _table(ctx, bcols)
    aggs = [AggExpr(COUNT), AggExpr(SUM, "ppay"), AggExpr(MIN, "ppay"), AggExpr(MAX, "ppay"), AggExpr(SUM, "w"),
            AggExpr(MIN, "z"), AggExpr(MAX, "z"), AggExpr(AVG, "w"), AggExpr(COUNT, "z")]
    if jt in (INNER, LEFT):
        aggs += [AggExpr(SUM, "bpay", side="build"), AggExpr(MAX, "bpay", side="build")]
    g = fa.execute_hash_join(ctx, pt, bt, "pk", "bk", jt, aggs, filter=col("w") > -800)
    assert "aggregates in" in ctx.last_plan(), ctx.last_plan()
    o = ora.join(pcols, bcols, "pk", "bk", jt, aggs, expr=col("w") > -800)
    assert_batches_equal(g, o.exact, None)
    pt.free()
    bt.free()
