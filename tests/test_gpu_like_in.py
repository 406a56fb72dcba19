"""LIKE / ILIKE / IN / NOT IN predicates on the device against the oracle and
the reference's row counts: the filter kernels, GROUP BY (dense, partitioned
and hash paths), the join's probe filter, ORDER BY and the SQL layer."""
import numpy as np
import pytest

import falcondb_amd as fa
import oracle as ora
from falcondb_amd import datagen as dg
from falcondb_amd._abi import COUNT, INNER, MAX, MIN, SEMI, SUM
from falcondb_amd.operators import AggExpr, DictColumn, NullableColumn, col, execute_sort
from falcondb_amd.sql import Database

from known_answers import check_like_in
from test_gpu_parity import BOUNDARY_ROWS, C2_AGGS, _c3, _table, assert_batches_equal, jctx  # noqa: F401

pytestmark = pytest.mark.gpu

WORDS = ["", "a", "ab", "abc", "Abc", "ba", "b%a", "x_y", "xay", "é", "éa", "aaa", "banana", "BANANA"]


class Engine:
    def __init__(self, ctx):
        self.ctx = ctx

    def filter_count(self, cols, pred):
        return fa.count_star(self.ctx, fa.Table(self.ctx, cols), pred)


def test_like_in_reference_known_answers(ctx):
    assert check_like_in(Engine(ctx)) >= 35


def _cols(n, seed):
    rng = np.random.default_rng(seed)
    codes = rng.integers(0, len(WORDS), n).astype(np.uint32)
    codes[rng.random(n) < 0.1] = 0xFFFFFFFF
    f = rng.choice([-1.5, -0.0, 0.0, 1.0, 2.0, 2.5, np.nan, np.inf], n)
    return {
        "s": DictColumn(codes, WORDS),
        "i": NullableColumn(rng.integers(-5, 6, n), rng.random(n) > 0.1),
        "f": NullableColumn(f, rng.random(n) > 0.05),
        "j": rng.integers(-3, 4, n).astype(np.int32),
        "k": rng.integers(0, 50, n),
    }


def _preds():
    big = list(range(-1000, 1000, 7))  # a long list: binary search over 286 values
    return [
        col("s").like("a%"), col("s").not_like("%a%"), col("s").ilike("a%"), col("s").like("b\\%a"),
        col("s").like("_"), col("s").like("%é%"), col("s").not_like("%"), col("s").like("x_y"),
        col("i").isin([1, 3, -5]), col("i").not_in([0, 2]), col("i").isin([2.0, 7]), col("i").isin([]),
        col("i").not_in([]), col("f").isin([0, 2.5]), col("f").not_in([1.0, np.inf]), col("f").isin([np.nan]),
        col("j").isin([-3, 3, 2**32 + 1]), col("j").not_in([0]), col("j").isin([1.5]), col("j").not_in(["a"]),
        col("s").isin(["a", "é", "zzz"]), col("s").not_in(["", "banana"]), col("s").isin([1]),
        col("k").isin(big), col("k").not_in(big),
        col("s").like("%a") & col("i").isin([1, 2, 3]), col("s").like("b%") | col("f").not_in([0.0]),
        col("s").isin(["ab", "abc"]) | col("j").isin([0]),
    ]


@pytest.mark.parametrize("n", BOUNDARY_ROWS)
def test_filter_like_in_boundaries(ctx, n):
    cols = _cols(n, n)
    t = _table(ctx, cols)
    for p in _preds():
        assert fa.count_star(ctx, t, p) == ora.filter_count(cols, p)
        assert np.array_equal(p.eval_mask_bits(ctx, t), ora.filter_mask(cols, p))
    t.free()


def test_select_and_sort_with_like_in(ctx):
    cols = _cols(200_000, 3)
    t = _table(ctx, cols)
    for p in (col("s").like("%a%") & col("i").not_in([0]), col("k").isin(list(range(0, 50, 3)))):
        res = fa.execute_select(ctx, t, p, ["k", "i"])
        exp = ora.filter_select(cols, p, ["k", "i"])
        assert_batches_equal(res, exp, None)
        rows, _ = execute_sort(ctx, t, [("k", True), ("s", False)], 500, p)
        assert rows.tolist() == ora.sort_limit(cols, p, [("k", True), ("s", False)], 500).tolist()
    t.free()


@pytest.mark.parametrize("mode", ["dense", "partitioned", "hash"])
def test_groupby_with_like_in(ctx, ctx_gbp, ctx_gbh, mode):
    if mode == "dense":  # a Utf8 key: the dense LDS path
        cols = _cols(300_000, 5)
        c, key = ctx, "s"
        aggs = [AggExpr(COUNT), AggExpr(SUM, "i"), AggExpr(MIN, "f"), AggExpr(MAX, "j")]
        preds = [col("i").isin([1, 2, 3]), col("s").not_like("%a%") | col("j").isin([0, 1])]
    else:
        K = dg.keys(50_000, 40)
        cols = dg.c2_table(1_000_000, K)
        c, key, aggs = (ctx_gbp if mode == "partitioned" else ctx_gbh), "key", C2_AGGS
        preds = [col("f").isin(list(range(0, 100, 3))), col("f").not_in([1, 2, 3]) & (col("v1") > 0)]
    t = _table(c, cols)
    for p in preds:
        g = fa.GroupByExec([key], aggs).execute(c, t, p)
        o = ora.groupby(cols, p, key, aggs)
        # f holds NaN: MIN skips it, where the reference order keeps a leading NaN,
        # so only the exact columns are compared there
        assert_batches_equal(g, o.exact, key, None if mode == "dense" else o.reference)
        s = fa.execute_agg_only(c, t, aggs, p)
        assert_batches_equal(s, ora.groupby(cols, p, None, aggs).exact, None)
    t.free()


def test_join_probe_filter_like_in(jctx):
    ctx = jctx
    pcols, bcols = _c3(5000, 100_000)
    rng = np.random.default_rng(8)
    pcols["f"] = (np.arange(100_000) % 100).astype(np.int32)
    pcols["tag"] = DictColumn(rng.integers(0, len(WORDS), 100_000).astype(np.uint32), WORDS)
    pt, bt = _table(ctx, pcols), _table(ctx, bcols)
    for p in (col("f").isin([1, 5, 9, 50]), col("tag").like("%a%") & col("f").not_in([3])):
        for jt in (INNER, SEMI):
            aggs = [AggExpr(COUNT), AggExpr(SUM, "ppay")]
            g = fa.execute_hash_join(ctx, pt, bt, "pk", "bk", jt, aggs, filter=p)
            o = ora.join(pcols, bcols, "pk", "bk", jt, aggs, expr=p)
            assert_batches_equal(g, o.exact, None)


def test_sql_like_in(ctx):
    d = Database(ctx)
    cols = _cols(50_000, 9)
    d.create_table("t", cols)
    from falcondb_amd.sql import where_expr
    for w in ("s LIKE 'a%' AND i NOT IN (1, 2)", "s NOT ILIKE 'ab%'", "f IN (2.5, 0) OR s IN ('a', 'ba')",
              "j NOT IN (0, 1) AND s NOT LIKE '%a'"):
        n = d.sql(f"SELECT COUNT(*) FROM t WHERE {w}").columns[0][0]
        assert n == ora.filter_count(cols, where_expr(w)), w
    for nm in list(d.tables):
        d.drop(nm)
