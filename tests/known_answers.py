"""The reference's own known-answer tests (tests/golden/reference_known_answers.json),
restated as inputs for any engine that exposes filter_count / groupby / join.
Shared by the CPU oracle tests and the GPU tests."""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np

from falcondb_amd._abi import ANTI, AVG, COUNT, INNER, LEFT, MAX, MIN, SEMI, SUM
from falcondb_amd.operators import AggExpr, DictColumn, NullableColumn, col

GOLDEN = json.loads((Path(__file__).parent / "golden" / "reference_known_answers.json").read_text())

# The count_star_filter predicates as Expr builders.
PRED = {
    "n >= 0": lambda: col("n") >= 0,
    "n < 0": lambda: col("n") < 0,
    "n >= 0 AND n < 100000": lambda: (col("n") >= 0) & (col("n") < 100000),
    "n > 10": lambda: col("n") > 10,
    "maybe IS NULL": lambda: col("maybe").is_null(),
    "maybe IS NOT NULL": lambda: col("maybe").is_not_null(),
    "s = 'a'": lambda: col("s") == "a",
    "s <> 'a'": lambda: col("s") != "a",
    "s LIKE 'b%'": lambda: col("s").like("b%"),
    "n IN (1, 2, 3)": lambda: col("n").isin([1, 2, 3]),
    "n NOT IN (1, 2, 3)": lambda: col("n").not_in([1, 2, 3]),
    "n > 5 AND n < 20": lambda: (col("n") > 5) & (col("n") < 20),
    "n < 2 OR n > 9995": lambda: (col("n") < 2) | (col("n") > 9995),
}


def count_star_table(rows: int) -> dict:
    i = np.arange(rows, dtype=np.int64)
    return {
        "n": i,
        "maybe": NullableColumn(i.copy(), i % 3 != 0),
        "s": DictColumn((i % 3).astype(np.uint32), ["a", "b", "c"]),
    }


def closed_form_count(pred: str, rows: int) -> int:
    """The number of rows each predicate selects, from the table's closed form."""
    n = np.arange(rows)
    m = {
        "n >= 0": n >= 0,
        "n < 0": n < 0,
        "n >= 0 AND n < 100000": (n >= 0) & (n < 100000),
        "n > 10": n > 10,
        "maybe IS NULL": n % 3 == 0,
        "maybe IS NOT NULL": n % 3 != 0,
        "s = 'a'": n % 3 == 0,
        "s <> 'a'": n % 3 != 0,
        "s LIKE 'b%'": n % 3 == 1,
        "n IN (1, 2, 3)": (n >= 1) & (n <= 3),
        "n NOT IN (1, 2, 3)": ~((n >= 1) & (n <= 3)),
        "n > 5 AND n < 20": (n > 5) & (n < 20),
        "n < 2 OR n > 9995": (n < 2) | (n > 9995),
    }[pred]
    return int(m.sum())


def scalar(batch, i=0):
    c = batch.columns[i]
    v = batch.validity[i]
    if v is not None and not v[0]:
        return None
    return c[0].item() if hasattr(c[0], "item") else c[0]


def check_all(engine) -> int:
    """Runs every known answer through `engine`; returns the number of checks."""
    checks = 0
    g = GOLDEN["count_star_filter"]
    for rows in g["rows"]:
        cols = count_star_table(rows)
        for p in g["predicates"]:
            got = engine.filter_count(cols, PRED[p]())
            assert got == closed_form_count(p, rows), (rows, p, got)
            checks += 1

    g = GOLDEN["count_star_declines"]
    cols = count_star_table(g["rows"])
    r = engine.groupby(cols, col("n") >= 0, None, [AggExpr(COUNT, "maybe")])
    assert scalar(r) == g["count_maybe_where_n_ge_0"]
    r = engine.groupby(cols, col("n") < 10, None, [AggExpr(COUNT), AggExpr(SUM, "n")])
    assert [scalar(r, 0), float(scalar(r, 1))] == g["count_and_sum_where_n_lt_10"]
    r = engine.groupby(cols, col("n") >= 0, "s", [AggExpr(COUNT)])
    assert r.num_rows == g["groups_by_s"]
    checks += 3

    g = GOLDEN["diag_agg"]
    i = np.arange(g["rows"], dtype=np.int64)
    cols = {"id": i, "score": (i % 100) + 0.5}
    r = engine.groupby(cols, None, None, [AggExpr(SUM, "score"), AggExpr(COUNT), AggExpr(MIN, "score"),
                                          AggExpr(MAX, "score"), AggExpr(AVG, "score")],
                       batch_rows=g["batch_rows"])
    assert [scalar(r, k) for k in range(5)] == [g["sum"], g["count"], g["min"], g["max"], g["avg"]]
    vals = [x for b in g["nulls_batches"] for x in b]
    v = np.array([0.0 if x is None else x for x in vals])
    ok = np.array([x is not None for x in vals])
    cols = {"v": NullableColumn(v, ok)}
    r = engine.groupby(cols, None, None, [AggExpr(SUM, "v"), AggExpr(MIN, "v"), AggExpr(MAX, "v"),
                                          AggExpr(AVG, "v"), AggExpr(COUNT, "v"), AggExpr(COUNT)])
    assert [scalar(r, k) for k in range(6)] == [g["nulls_sum"], g["nulls_min"], g["nulls_max"],
                                               g["nulls_avg"], g["nulls_count_v"], g["nulls_count_star"]]
    cols = {"v": NullableColumn(np.zeros(2), np.zeros(2, bool))}
    r = engine.groupby(cols, None, None, [AggExpr(SUM, "v"), AggExpr(MIN, "v")])
    assert scalar(r, 0) is None and scalar(r, 1) is None
    checks += 3

    g = GOLDEN["sql_group_by"]

    def grouped(d, vals, agg):
        names = sorted(set(d))
        cols = {"dept": DictColumn(np.array([names.index(x) for x in d], np.uint32), names)}
        aggs = [AggExpr(COUNT)] if vals is None else [AggExpr(agg, "v")]
        if vals is not None:
            cols["v"] = np.array(vals, dtype=np.int64)
        r = engine.groupby(cols, None, "dept", aggs)
        return {k: (r.columns[1][j].item() if r.validity[1] is None or r.validity[1][j] else None)
                for j, k in enumerate(r.columns[0])}
    assert grouped(g["count"]["dept"], None, None) == g["count"]["expect"]
    assert grouped(g["sum"]["dept"], g["sum"]["salary"], SUM) == g["sum"]["expect"]
    assert grouped(g["avg"]["dept"], g["avg"]["score"], AVG) == g["avg"]["expect"]
    assert grouped(g["min_max"]["dept"], g["min_max"]["score"], MIN) == {"eng": g["min_max"]["expect_min"]}
    assert grouped(g["min_max"]["dept"], g["min_max"]["score"], MAX) == {"eng": g["min_max"]["expect_max"]}
    assert grouped(g["single_group"]["dept"], g["single_group"]["n"], SUM) == {"eng": g["single_group"]["expect"]}
    assert len(grouped(g["all_unique"]["dept"], g["all_unique"]["n"], SUM)) == g["all_unique"]["expect_groups"]
    checks += 7

    g = GOLDEN["join_stride"]
    left = {"k": np.arange(g["left_rows"], dtype=np.int64)}
    right = {"rk": np.arange(0, g["left_rows"], g["stride"], dtype=np.int64)}
    for jt, name in [(INNER, "inner"), (LEFT, "left"), (SEMI, "semi"), (ANTI, "anti")]:
        r = engine.join(left, right, "k", "rk", jt, [AggExpr(COUNT)])
        assert scalar(r) == g[name], (name, scalar(r))
        checks += 1

    g = GOLDEN["join_null_keys"]

    def side(name, vals):
        return {name: NullableColumn(np.array([0 if x is None else x for x in vals], np.int64),
                                     np.array([x is not None for x in vals]))}
    left, right = side("k", g["left"]), side("rk", g["right"])
    for jt, name in [(INNER, "inner"), (LEFT, "left_join"), (SEMI, "semi"), (ANTI, "anti")]:
        r = engine.join(left, right, "k", "rk", jt, [AggExpr(COUNT)])
        assert scalar(r) == g[name], (name, scalar(r))
        checks += 1
    sl, sr = side("k", g["self_join_keys"]), side("rk", g["self_join_keys"])
    assert scalar(engine.join(sl, sr, "k", "rk", INNER, [AggExpr(COUNT)])) == g["self_inner"]
    checks += 1

    # an empty right side: SEMI keeps nothing, ANTI everything (src/query/join.rs:60-67)
    n = GOLDEN["join_empty_build"]["left_rows"]
    left = {"k": np.arange(n, dtype=np.int64)}
    right = {"rk": np.zeros(0, np.int64)}
    for jt, exp in ((SEMI, 0), (ANTI, n), (INNER, 0), (LEFT, n)):
        assert scalar(engine.join(left, right, "k", "rk", jt, [AggExpr(COUNT)])) == exp, jt
        checks += 1
    return checks


# ─── integer SUM past 2^53 and past int64 (src/query/aggregate.rs:1936-1993) ──
I64_MAX = 2**63 - 1


def _int_sum_cases():
    """(name, values, exact integer total) for the reference's int_sum tests."""
    g = GOLDEN["int_sum"]
    out = []
    for n in g["ordinary"]["rows"]:
        v = [(i % 997) - 498 for i in range(n)]
        out.append((f"ordinary_{n}", v, sum(v)))
    v = [(1 << 53) + i for i in range(8)]
    assert sum(v) == g["past_float_limit"]["exact"]
    out.append(("past_float_limit", v, sum(v)))
    v = list(g["overflow"]["data"])
    out.append(("overflow", v, sum(v)))
    out.append(("overflow_negative", [-x for x in v], -sum(v)))
    # partial sums leave int64 but the total fits: the exact int64 comes back
    v = [I64_MAX, I64_MAX, -I64_MAX, -I64_MAX, 5, -(2**63), 2**62, 2**62]
    out.append(("wide_but_fits", v, sum(v)))
    for n in g["ragged_tail"]["n"]:
        v = list(range(1, n + 1))
        out.append((f"ragged_{n}", v, n * (n + 1) // 2))
    return out


def _as_total(x):
    """A returned SUM as an exact Python number (int for Int64, float for Float64)."""
    return x.item() if hasattr(x, "item") else x


def _check_total(name, got, exact):
    """The exact total when it fits int64, else its correctly rounded double
    (Python's int -> float conversion rounds half to even)."""
    if -2**63 <= exact <= I64_MAX:
        assert isinstance(got, int) and got == exact, (name, got, exact)
        # the reference's Float64 result is the total as f64 (sum_i64)
        assert float(got) == float(exact)
    else:
        assert isinstance(got, float) and got == float(exact), (name, got, exact)
    if name.startswith("overflow"):
        assert abs(got) > GOLDEN["int_sum"]["overflow"]["greater_than"], (name, got)
        assert (got > 0) == (exact > 0), (name, got)


def check_int_sum(engine, key_rows: int = 3) -> int:
    """The int_sum known answers through every aggregate path `engine` offers:
    scalar SUM / AVG, GROUP BY where one group holds the test vector and other
    groups hold small values, and SUM over a join (probe side via SEMI, build
    side via INNER).  Returns the number of checks."""
    checks = 0
    for name, vals, exact in _int_sum_cases():
        v = np.array(vals, dtype=np.int64)
        n = v.size
        # scalar
        r = engine.groupby({"v": v}, None, None, [AggExpr(SUM, "v"), AggExpr(COUNT)])
        if n:
            _check_total(name, _as_total(r.columns[0][0]), exact)
        else:
            assert r.validity[0] is not None and not r.validity[0][0], name  # SUM of nothing is NULL
        checks += 1
        if n:
            r = engine.groupby({"v": v}, None, None, [AggExpr(AVG, "v")])
            assert r.columns[0][0] == float(exact) / n, (name, r.columns[0][0])
            checks += 1
        # GROUP BY: key 0 holds the vector, keys 1..key_rows one small value each
        k = np.concatenate([np.zeros(n, np.int64), np.arange(1, key_rows + 1, dtype=np.int64)])
        w = np.concatenate([v, np.arange(1, key_rows + 1, dtype=np.int64) * 7])
        for kname, kcol in (("int64 key", k), ("wide key", k * 1_000_003 + 5)):
            r = engine.groupby({"k": kcol, "w": w}, None, "k", [AggExpr(SUM, "w"), AggExpr(COUNT)])
            r = r.sort_by("k")
            sums = [_as_total(x) for x in r.columns[1]]
            if n:
                _check_total(f"{name} / {kname}", sums[0], exact)
                rest = sums[1:]
            else:
                rest = sums
            overflowed = not (-2**63 <= exact <= I64_MAX)
            for j, s in enumerate(rest):
                # the column is Float64 when any group left int64
                assert s == (float(7 * (j + 1)) if overflowed else 7 * (j + 1)), (name, kname, j, s)
                assert isinstance(s, float) == overflowed, (name, kname, s)
            checks += 1
        if not n:
            continue
        # joins: SEMI sums the probe column; INNER sums the build column
        probe = {"pk": np.arange(n, dtype=np.int64), "pv": v}
        build = {"bk": np.arange(n, dtype=np.int64), "bv": v}
        r = engine.join(probe, build, "pk", "bk", SEMI, [AggExpr(SUM, "pv")])
        _check_total(f"{name} / semi", _as_total(r.columns[0][0]), exact)
        r = engine.join(probe, build, "pk", "bk", INNER, [AggExpr(SUM, "bv", side="build")])
        _check_total(f"{name} / inner", _as_total(r.columns[0][0]), exact)
        checks += 2
    return checks


# ─── ORDER BY [LIMIT [OFFSET]] ────────────────────────────────────────────────
def sort_column(spec: dict):
    """A golden column spec ({"dtype", "values"} with None for NULL) as an
    upload-ready column; Utf8 dictionaries are in first-seen order, which
    is not sorted, so a sort by code instead of by string shows."""
    vals, dt = spec["values"], spec["dtype"]
    if dt == "utf8":
        dic = list(dict.fromkeys(v for v in vals if v is not None))
        codes = np.array([0xFFFFFFFF if v is None else dic.index(v) for v in vals], np.uint32)
        return DictColumn(codes, dic)
    np_dt = np.int64 if dt == "int64" else np.float64
    arr = np.array([0 if v is None else v for v in vals], np_dt)
    if any(v is None for v in vals):
        return NullableColumn(arr, np.array([v is not None for v in vals]))
    return arr


def column_values(c, rows) -> list:
    """Python values (None for NULL) of column c at the given rows."""
    rows = np.asarray(rows, np.int64)
    if isinstance(c, DictColumn):
        return [None if c.codes[r] == 0xFFFFFFFF else c.dictionary[c.codes[r]] for r in rows]
    if isinstance(c, NullableColumn):
        return [c.values[r].item() if c.valid[r] else None for r in rows]
    return [c[r].item() for r in rows]


def check_sort(engine) -> int:
    """The reference's ORDER BY known answers through engine.sort(cols, pred,
    keys, limit) -> row numbers in output order; OFFSET drops the first rows
    of a sort kept to limit + offset (tests/sort_and_dedup_paths.rs:116-121)."""
    g = GOLDEN["sort"]
    n = 0
    for case in g["cases"]:
        cols = {k: sort_column(v) for k, v in case["cols"].items()}
        keys = [tuple(k) for k in case["keys"]]
        lim, off = case.get("limit"), case.get("offset", 0)
        rows = engine.sort(cols, None, keys, None if lim is None else lim + off)[off:]
        for name, want in case.get("out", {}).items():
            assert column_values(cols[name], rows) == want, case["src"]
        for name, at in case.get("out_at", {}).items():
            got = column_values(cols[name], rows)
            for i, want in at.items():
                assert got[int(i)] == want, case["src"]
        n += 1
    cf = g["closed_forms"]
    v = np.arange(80000, dtype=np.int64)[::-1] % 1000
    got = column_values(v, engine.sort({"v": v}, None, [("v", False)], None))
    assert len(got) == cf["large_sort"]["rows"] and got[0] == cf["large_sort"]["first"]
    assert got[-1] == cf["large_sort"]["last"] and got == sorted(got)
    v = (np.arange(5000, dtype=np.int64) * 7919) % 1000
    full = engine.sort({"v": v}, None, [("v", False)], None)
    top = engine.sort({"v": v}, None, [("v", False)], cf["top_n_matches_full"]["limit"])
    assert top.tolist() == full[:25].tolist()
    s = [f"{(i * 37) % 64:03d}" for i in range(64)]
    sc = sort_column({"dtype": "utf8", "values": s})
    assert column_values(sc, engine.sort({"s": sc}, None, [("s", False)], None)) == sorted(s)
    m = 2 * 8192
    f = ((np.arange(m) * 7919) % m).astype(np.float64) - m / 2
    got = column_values(f, engine.sort({"score": f}, None, [("score", False)], None))
    assert all(a <= b for a, b in zip(got, got[1:]))
    i = np.arange(1, 21, dtype=np.int64)
    wide = {"a": i, "b": 2 * i, "c": 3 * i}
    got = column_values(i, engine.sort(wide, col("b") > 20, [("a", False)], None))
    assert len(got) == cf["wide_filter"]["rows"] and got[0] == cf["wide_filter"]["first"]
    got = column_values(i, engine.sort(wide, None, [("c", True)], None))
    assert len(got) == cf["wide_order_by_unselected"]["rows"] and got[0] == cf["wide_order_by_unselected"]["first"]
    return n + 6


# ─── LIKE / ILIKE / IN / NOT IN ───────────────────────────────────────────────
def check_like_in(engine) -> int:
    """The reference's LIKE and IN row counts through engine.filter_count."""
    from falcondb_amd.sql import where_expr
    n = 0
    for case in GOLDEN["like_in"]["cases"]:
        cols = {k: sort_column(v) for k, v in case["cols"].items()}
        for text, want in case["where"].items():
            assert engine.filter_count(cols, where_expr(text)) == want, (case["src"], text)
            n += 1
    return n
