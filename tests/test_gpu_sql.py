"""The SQL subset on the device: each statement equals the operator call it
plans to (same names, types and values), and C1 equals the oracle."""
import numpy as np
import pytest

import falcondb_amd as fa
import oracle as ora
from falcondb_amd import datagen as dg
from falcondb_amd._abi import ANTI, AVG, COUNT, INNER, LEFT, MAX, MIN, SEMI, SUM
from falcondb_amd.operators import AggExpr, DictColumn, NullableColumn, col
from falcondb_amd.sql import Database

from test_gpu_parity import assert_batches_equal

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def db(ctx):
    d = Database(ctx)
    t1 = dg.c1_table(300_000)
    c1 = {"id": t1["id"], "dept": DictColumn(t1["dept"], dg.DEPTS), "score": t1["score"], "salary": t1["salary"]}
    d.create_table("t", c1)
    K = dg.keys(50_000, 40, seed_=dg.seed(2, 9))
    c2 = dg.c2_table(3_000_000, K)
    rng = np.random.default_rng(4)
    c2["v3"] = NullableColumn(rng.normal(size=c2["f"].size), rng.random(c2["f"].size) > 0.1)
    d.create_table("fact", c2)
    d.create_table("build", dg.c3_build(K))
    d.create_table("probe", dg.c3_probe(3_000_000, K) | {"f": (np.arange(3_000_000) % 100).astype(np.int32)})
    yield d, c1, c2
    for n in list(d.tables):
        d.drop(n)


def test_c1_groupby(db, ctx):
    d, c1, _ = db
    g = d.sql("SELECT dept, COUNT(*), SUM(score), AVG(salary) FROM t WHERE score > 50 GROUP BY dept")
    assert g.names == ["dept", "col1", "col2", "col3"]
    aggs = [AggExpr(COUNT), AggExpr(SUM, "score"), AggExpr(AVG, "salary")]
    assert_batches_equal(g, fa.GroupByExec(["dept"], aggs).execute(ctx, d.table("t"), col("score") > 50), "dept")
    o = ora.groupby(c1, col("score") > 50, "dept", aggs)
    assert_batches_equal(g, o.exact, "dept", o.reference)


def test_groupby_several_keys(db, ctx):
    # GROUP BY two keys (the reference's slow path): columns in SELECT order,
    # the groups equal to the operator call's and to the oracle's restatement
    from test_gpu_groupby_keys import _assert_same_groups
    d, c1, _ = db
    g = d.sql("SELECT COUNT(*) AS n, salary, dept, SUM(score) FROM t WHERE score > 90 GROUP BY dept, salary")
    assert g.names == ["n", "salary", "dept", "col3"]
    aggs = [AggExpr(COUNT, None, "n"), AggExpr(SUM, "score", "col3")]
    o = ora.groupby_keys(c1, col("score") > 90, ["dept", "salary"], aggs)
    # the SQL result in the oracle's column order: the keys, then the aggregates
    idx = [2, 1, 0, 3]
    g2 = fa.RecordBatch([g.names[i] for i in idx], [g.dtypes[i] for i in idx], [g.columns[i] for i in idx],
                        [g.validity[i] for i in idx], [g.agg_funcs[i] for i in idx])
    _assert_same_groups(g2, o, 2)


def test_select_order_and_aliases(db, ctx):
    d, _, _ = db
    g = d.sql("SELECT COUNT(*) AS n, dept, MAX(salary) FROM t GROUP BY dept")
    assert g.names == ["n", "dept", "col2"]
    e = fa.GroupByExec(["dept"], [AggExpr(COUNT), AggExpr(MAX, "salary")]).execute(ctx, d.table("t"))
    order = np.argsort(np.array(g.columns[1], dtype=object), kind="stable")
    eorder = np.argsort(np.array(e.columns[0], dtype=object), kind="stable")
    assert np.array_equal(g.columns[0][order], e.columns[1][eorder])
    assert np.array_equal(g.columns[2][order], e.columns[2][eorder])


def test_c2_groupby_and_aggregates(db, ctx):
    d, _, _ = db
    t = d.table("fact")
    g = d.sql("SELECT key, COUNT(*), SUM(v1), MIN(v1), MAX(v1), SUM(v2) FROM fact WHERE f < 40 GROUP BY key")
    aggs = [AggExpr(COUNT), AggExpr(SUM, "v1"), AggExpr(MIN, "v1"), AggExpr(MAX, "v1"), AggExpr(SUM, "v2")]
    assert_batches_equal(g, fa.GroupByExec(["key"], aggs).execute(ctx, t, col("f") < 40), "key")
    n = d.sql("SELECT COUNT(*) FROM fact WHERE f < 40")
    assert n.names == ["col0"] and n.columns[0][0] == fa.count_star(ctx, t, col("f") < 40)
    a = d.sql("SELECT SUM(v1), MIN(v2) AS lo, COUNT(v3) FROM fact WHERE f >= 10 AND v3 IS NOT NULL")
    e = fa.execute_agg_only(ctx, t, [AggExpr(SUM, "v1", "col0"), AggExpr(MIN, "v2", "lo"), AggExpr(COUNT, "v3", "col2")],
                            (col("f") >= 10) & col("v3").is_not_null())
    assert_batches_equal(a, e, None)


@pytest.mark.parametrize("text,jt", [
    ("SELECT SUM(build.bpay), COUNT(*) FROM probe JOIN build ON probe.pk = build.bk", INNER),
    ("SELECT SUM(b.bpay), COUNT(*) FROM probe p LEFT JOIN build b ON b.bk = p.pk WHERE p.f < 30", LEFT),
])
def test_joins(db, ctx, text, jt):
    d, _, _ = db
    g = d.sql(text)
    f = col("f") < 30 if "WHERE" in text else None
    e = fa.execute_hash_join(ctx, d.table("probe"), d.table("build"), "pk", "bk", jt,
                             [AggExpr(SUM, "bpay", "col0", side="build"), AggExpr(COUNT, None, "col1")], filter=f)
    assert g.names == ["col0", "col1"]
    assert_batches_equal(g, e, None)


@pytest.mark.parametrize("neg", [False, True])
def test_exists_is_semi_or_anti(db, ctx, neg):
    d, _, _ = db
    g = d.sql(f"SELECT SUM(probe.ppay), COUNT(*) FROM probe WHERE {'NOT ' if neg else ''}EXISTS "
              "(SELECT 1 FROM build WHERE build.bk = probe.pk) AND probe.f < 50")
    e = fa.execute_hash_join(ctx, d.table("probe"), d.table("build"), "pk", "bk", ANTI if neg else SEMI,
                             [AggExpr(SUM, "ppay", "col0"), AggExpr(COUNT, None, "col1")], filter=col("f") < 50)
    assert_batches_equal(g, e, None)


def test_order_by_limit(db, ctx):
    d, _, c2 = db
    t = d.table("fact")
    g = d.sql("SELECT v2 FROM fact ORDER BY v2 DESC LIMIT 10")
    rows, vals = fa.operators.execute_sort_limit(ctx, t, "v2", 10, desc=True)
    assert g.names == ["v2"] and np.array_equal(g.columns[0], vals)
    # a column with NULLs: they lead for ASC and trail for DESC (the reference's placement)
    for thr, desc in ((-995_000, False), (-999_000, True)):
        g = d.sql(f"SELECT v3 FROM fact WHERE f < 1 AND v1 < {thr} ORDER BY v3 {'DESC' if desc else 'ASC'} LIMIT 30")
        sel = (c2["f"] < 1) & (c2["v1"] < thr)
        x = c2["v3"].values[sel & c2["v3"].valid]
        n_null = int(np.count_nonzero(sel & ~c2["v3"].valid))
        k = min(30, int(np.count_nonzero(sel)))
        assert g.columns[0].size == k and n_null > 0
        i = np.arange(k)
        if desc:
            n_val = min(k, x.size)
            assert np.array_equal(g.validity[0], i < n_val)
            assert np.array_equal(g.columns[0][:n_val], np.sort(x)[::-1][:n_val])
        else:
            n_lead = min(k, n_null)
            assert np.array_equal(g.validity[0], i >= n_lead)
            assert np.array_equal(g.columns[0][n_lead:], np.sort(x)[: k - n_lead])
    # past the fused top-k's 32 rows the general device sort takes over: same order
    g = d.sql("SELECT v2 FROM fact ORDER BY v2 LIMIT 33")
    rows = ora.sort_limit({"v2": c2["v2"]}, None, [("v2", False)], 33)
    assert np.array_equal(g.columns[0], c2["v2"][rows])


def test_select_materialises(db, ctx):
    d, _, _ = db
    g = d.sql("SELECT id, score FROM t WHERE score > 99.5")
    e = fa.execute_select(ctx, d.table("t"), col("score") > 99.5, ["id", "score"])
    assert g.names == ["id", "score"]
    assert_batches_equal(g, e, None)


@pytest.mark.parametrize("text,code", [
    ("SELECT COUNT(*) FROM nosuch", 6),
    ("SELECT COUNT(*) FROM probe JOIN build ON probe.pk = build.bk WHERE build.bpay > 0", 102),
    ("SELECT dept, score FROM t GROUP BY dept", 6),
    ("SELECT COUNT(*) FROM t WHERE nosuch > 1", 1),
])
def test_errors(db, text, code):
    d, _, _ = db
    with pytest.raises(fa.FalconError) as e:
        d.sql(text)
    assert e.value.code == code
