"""The join's output as a device table (fgpu_join_materialize): rows and
names against the oracle's join pairs, GROUP BY / WHERE / SELECT * over a join
through SQL, and the reference's join_projection known answers
(tests/join_projection.rs:15-180)."""
import numpy as np
import pytest

import falcondb_amd as fa
import oracle as ora
from falcondb_amd._abi import ANTI, COUNT, INNER, LEFT, NO_ROW, SEMI, SUM
from falcondb_amd.operators import AggExpr, DictColumn, NullableColumn, col
from falcondb_amd.sql import Database

from known_answers import GOLDEN
from test_gpu_parity import assert_batches_equal

pytestmark = pytest.mark.gpu


def _host_rows(cols: dict, rows: np.ndarray):
    """Column values at `rows` (NO_ROW -> None), as Python objects."""
    out = {}
    for name, c in cols.items():
        if isinstance(c, DictColumn):
            vals = [None if r == NO_ROW or c.codes[r] == 0xFFFFFFFF else c.dictionary[c.codes[r]] for r in rows]
        elif isinstance(c, NullableColumn):
            vals = [None if r == NO_ROW or not c.valid[r] else c.values[r].item() for r in rows]
        else:
            vals = [None if r == NO_ROW else c[r].item() for r in rows]
        out[name] = vals
    return out


def _sorted_rows(d: dict, names):
    rows = list(zip(*[d[n] for n in names])) if names else []
    return sorted(rows, key=lambda t: tuple((x is None, x if x is not None else 0) for x in t))


def _batch_rows(b: fa.RecordBatch) -> dict:
    out = {}
    for i, n in enumerate(b.names):
        v = b.validity[i]
        c = b.columns[i]
        out[n] = [None if (v is not None and not v[j]) or c[j] is None else
                  (c[j].item() if hasattr(c[j], "item") else c[j]) for j in range(b.num_rows)]
    return out


@pytest.mark.parametrize("jt", [INNER, LEFT, SEMI, ANTI])
def test_materialised_join_matches_oracle_pairs(ctx, jt):
    rng = np.random.default_rng(100 + jt)
    n_p, n_b = 20_000, 3_000
    pk = rng.integers(0, 4_000, n_p)
    bk = rng.integers(0, 4_000, n_b)  # duplicate build keys
    probe = {"k": NullableColumn(pk, rng.random(n_p) > 0.05),
             "v": rng.integers(-100, 100, n_p),
             "tag": DictColumn(rng.integers(0, 3, n_p).astype(np.uint32), ["a", "b", "c"])}
    build = {"k": NullableColumn(bk, rng.random(n_b) > 0.05),
             "v": rng.normal(size=n_b),
             "w": NullableColumn(rng.integers(0, 9, n_b).astype(np.int32), rng.random(n_b) > 0.2),
             "name": DictColumn(rng.integers(0, 4, n_b).astype(np.uint32), ["x", "y", "z", ""])}
    pt, bt = fa.Table(ctx, probe), fa.Table(ctx, build)
    semi = jt in (SEMI, ANTI)
    t = fa.execute_hash_join_projected(ctx, pt, bt, "k", "k", jt, ["tag", "v", "k"],
                                       None if not semi else [], filter=col("v") > -50)
    names = ["tag", "v", "k"] + ([] if semi else ["right.k", "right.v", "w", "name"])
    assert t.names == names
    got = _batch_rows(t.read())
    _, l, r = ora.join(probe, build, "k", "k", jt, [AggExpr(COUNT)], expr=col("v") > -50, pairs=True)
    exp = _host_rows({"tag": probe["tag"], "v": probe["v"], "k": probe["k"]}, l)
    if not semi:
        exp |= {f"right.{k}" if k in ("k", "v") else k: v
                for k, v in _host_rows(build, r).items()}
    assert t.n_rows == l.size
    assert _sorted_rows(got, names) == _sorted_rows(exp, names)
    # GROUP BY over the join output, against the oracle on the same rows
    if not semi:
        g = fa.GroupByExec(["tag"], [AggExpr(COUNT), AggExpr(SUM, "right.v"), AggExpr(SUM, "w")]).execute(ctx, t)
        host = {"tag": DictColumn(np.array([["a", "b", "c"].index(x) for x in exp["tag"]], np.uint32), ["a", "b", "c"]),
                "rv": NullableColumn(np.array([0.0 if x is None else x for x in exp["right.v"]]),
                                     np.array([x is not None for x in exp["right.v"]])),
                "w": NullableColumn(np.array([0 if x is None else x for x in exp["w"]], np.int32),
                                    np.array([x is not None for x in exp["w"]]))}
        o = ora.groupby(host, None, "tag", [AggExpr(COUNT), AggExpr(SUM, "rv"), AggExpr(SUM, "w")])
        e = o.exact
        assert_batches_equal(g, e, "tag")
    t.free()


def _join_db(ctx, shared=False):
    d = Database(ctx)
    i = np.arange(40, dtype=np.int64)
    j = np.arange(0, 40, 4, dtype=np.int64)
    if shared:
        d.create_table("l", {"k": i, "v": i * 10})
        d.create_table("r", {"k": j, "v": j * 100})
    else:
        tags = [f"L{x}" for x in range(3)]
        d.create_table("l", {"k": i, "lv": i * 10, "tag": DictColumn((i % 3).astype(np.uint32), tags)})
        d.create_table("r", {"rk": j, "rv": j * 100})
    return d


def _one(res):
    return res.columns[0][0].item() if hasattr(res.columns[0][0], "item") else res.columns[0][0]


def test_join_projection_known_answers(ctx):
    g = GOLDEN["join_projection"]
    d = _join_db(ctx)
    assert _one(d.sql("SELECT COUNT(*) AS n FROM l JOIN r ON l.k = r.rk")) == g["count"]
    assert _one(d.sql("SELECT SUM(lv) AS n FROM l JOIN r ON l.k = r.rk")) == g["sum_lv"]
    assert _one(d.sql("SELECT SUM(rv) AS n FROM l JOIN r ON l.k = r.rk")) == g["sum_rv"]
    assert _one(d.sql("SELECT COUNT(*) AS n FROM l JOIN r ON l.k = r.rk WHERE lv > 100")) == g["count_where_lv_gt_100"]
    with pytest.raises(fa.FalconError):  # a WHERE on a right-side column is a documented gap
        d.sql("SELECT COUNT(*) AS n FROM l JOIN r ON l.k = r.rk WHERE rv > 1000")
    res = d.sql("SELECT tag, COUNT(*) AS n FROM l JOIN r ON l.k = r.rk GROUP BY tag")
    assert res.num_rows == g["group_by_tag_groups"] and int(res.columns[1].sum()) == g["group_by_tag_total"]
    res = d.sql("SELECT lv FROM l JOIN r ON l.k = r.rk ORDER BY rv DESC")
    assert res.num_rows == 10 and res.columns[0][0] == g["order_by_rv_desc_first_lv"]
    assert res.columns[0].tolist() == list(range(360, -1, -40))
    res = d.sql("SELECT * FROM l JOIN r ON l.k = r.rk")
    assert res.num_rows == 10 and len(res.names) == g["select_star_columns"]
    sh = g["shared_names"]
    d2 = _join_db(ctx, shared=True)
    assert _one(d2.sql("SELECT SUM(v) AS n FROM l JOIN r ON l.k = r.k")) == sh["sum_left_v"]
    star = d2.sql("SELECT * FROM l JOIN r ON l.k = r.k")
    for n in sh["star_names_contain"]:
        assert n in star.names, star.names
    for db in (d, d2):
        for n in list(db.tables):
            db.drop(n)


def test_join_empty_build_side(ctx):
    # SEMI against an empty right side returns nothing, ANTI every left row
    # (src/query/join.rs:60-67)
    n = GOLDEN["join_empty_build"]["left_rows"]
    pt = fa.Table(ctx, {"k": np.arange(n, dtype=np.int64), "v": np.ones(n, np.int64)})
    bt = fa.Table(ctx, {"k": np.zeros(0, np.int64)})
    for jt, exp in ((SEMI, 0), (ANTI, n), (INNER, 0), (LEFT, n)):
        r = fa.execute_hash_join(ctx, pt, bt, "k", "k", jt, [AggExpr(COUNT), AggExpr(SUM, "v")])
        assert r.columns[0][0] == exp, jt
        t = fa.execute_hash_join_projected(ctx, pt, bt, "k", "k", jt, ["v"], [])
        assert t.n_rows == exp
        t.free()
