"""ORDER BY k1 [DESC] [NULLS FIRST | LAST], k2 ... [LIMIT n [OFFSET m]] on
the device (fgpu_sort_limit) against the oracle's stable multi-key sort and
the reference's ORDER BY known answers."""
import numpy as np
import pytest

import falcondb_amd as fa
import oracle as ora
from falcondb_amd.operators import DictColumn, NullableColumn, col, execute_sort
from falcondb_amd.sql import Database

from known_answers import check_sort, column_values

pytestmark = pytest.mark.gpu


class SortEngine:
    def __init__(self, ctx):
        self.ctx = ctx

    def sort(self, cols, pred, keys, limit):
        rows, _ = execute_sort(self.ctx, fa.Table(self.ctx, cols), keys, limit, pred)
        return rows


def test_sort_reference_known_answers(ctx):
    assert check_sort(SortEngine(ctx)) >= 25


WORDS = ["pear", "apple", "", "fig", "Zebra", "apple pie", "é", "a\x7f", "fig"]  # first-seen, unsorted, a repeat


def _cols(n, seed):
    rng = np.random.default_rng(seed)
    f = rng.choice([-np.inf, -2.5, -0.0, 0.0, 1.0, 3.25, np.inf, np.nan], n)
    f[rng.random(n) < 0.01] = np.frombuffer(np.array([0xfff8000000000001], np.uint64).tobytes(), np.float64)[0]
    i32 = rng.integers(-3, 3, n).astype(np.int32)
    i32[rng.random(n) < 0.01] = np.iinfo(np.int32).min
    i32[rng.random(n) < 0.01] = np.iinfo(np.int32).max
    i64 = rng.integers(-4, 4, n) * (1 << 40)
    i64[rng.random(n) < 0.01] = np.iinfo(np.int64).min
    codes = rng.integers(0, len(WORDS), n).astype(np.uint32)
    codes[rng.random(n) < 0.1] = 0xFFFFFFFF
    return {
        "i": NullableColumn(i64, rng.random(n) > 0.15),
        "j": i32,
        "f": NullableColumn(f, rng.random(n) > 0.1),
        "s": DictColumn(codes, WORDS),
        "w": rng.integers(0, 100, n).astype(np.int32),
        "u": np.arange(n, dtype=np.int64) * 3,
    }


def _bits(vals):
    """Python values with doubles as their bit patterns, so NaNs compare."""
    return [np.array([v]).view(np.int64)[0].item() if isinstance(v, float) else v for v in vals]


def _check(res, cols, names, rows):
    assert res.num_rows == len(rows)
    for k, nm in enumerate(names):
        got = res.columns[k].tolist() if res.columns[k].dtype != object else list(res.columns[k])
        if res.validity[k] is not None:
            got = [g if ok else None for g, ok in zip(got, res.validity[k].tolist())]
        want = column_values(cols[nm], rows)
        assert _bits(got) == _bits(want), nm


@pytest.mark.parametrize("n,seed", [(1, 0), (63, 1), (65, 2), (5000, 3), (200_000, 4), (3_000_000, 5)])
def test_sort_matches_oracle(ctx, n, seed):
    cols = _cols(n, seed)
    t = fa.Table(ctx, cols)
    rng = np.random.default_rng(100 + seed)
    names = list(cols)
    for trial in range(10):
        kn = list(rng.permutation(["i", "j", "f", "s", "w"])[: int(rng.integers(1, 4))])
        keys = [(c, bool(rng.integers(0, 2)), bool(rng.integers(0, 2))) for c in kn]
        limit = [None, 0, 1, 10, 100, n // 2, n + 3][int(rng.integers(0, 7))]
        filt = (col("w") < 60) if trial % 2 else None
        rows, res = execute_sort(ctx, t, keys, limit, filt, columns=names)
        exp = ora.sort_limit(cols, filt, keys, limit)
        assert rows.tolist() == exp.tolist(), (keys, limit, filt is not None)
        _check(res, cols, names, exp)
    t.free()


def test_sort_edge_cases(ctx):
    # nothing passes, an empty table, all-NULL keys, one distinct value, LIMIT 0
    cols = _cols(1000, 9)
    t = fa.Table(ctx, cols)
    rows, res = execute_sort(ctx, t, [("i", False)], None, col("w") > 1000)
    assert rows.size == 0 and res.num_rows == 0
    rows, res = execute_sort(ctx, t, [("u", True)], 0)
    assert rows.size == 0
    t.free()
    empty = {"a": np.zeros(0, np.int64), "s": DictColumn(np.zeros(0, np.uint32), [])}
    te = fa.Table(ctx, empty)
    rows, res = execute_sort(ctx, te, [("s", False), ("a", True)], None, columns=["a", "s"])
    assert rows.size == 0 and res.num_rows == 0
    te.free()
    n = 70_000
    alln = {"x": NullableColumn(np.zeros(n), np.zeros(n, bool)), "c": np.full(n, 7, np.int64),
            "s": DictColumn(np.full(n, 0xFFFFFFFF, np.uint32), ["a"])}
    ta = fa.Table(ctx, alln)
    for keys in ([("x", False)], [("x", True), ("c", False)], [("s", True)], [("c", True), ("s", False, True)]):
        rows, _ = execute_sort(ctx, ta, keys, None)
        assert rows.tolist() == list(range(n)), keys  # every row ties: row order
    ta.free()


def test_sort_large_two_keys(ctx):
    # 50M rows, two 8-byte keys: the full-size shape of an unlimited ORDER BY
    n = 50_000_000
    rng = np.random.default_rng(77)
    cols = {"a": rng.integers(0, 1000, n), "b": rng.normal(size=n)}
    t = fa.Table(ctx, cols)
    rows, _ = execute_sort(ctx, t, [("a", True), ("b", False)], None, columns=[])
    assert rows.size == n
    # properties that fix the order without a CPU sort of 50M rows: a
    # permutation, keys non-increasing / non-decreasing within equal a
    assert np.array_equal(np.bincount(rows, minlength=n), np.ones(n, np.int64))
    a, b = cols["a"][rows], cols["b"][rows]
    assert np.all(a[:-1] >= a[1:])
    same = a[:-1] == a[1:]
    assert np.all(b[:-1][same] <= b[1:][same])
    top, _ = execute_sort(ctx, t, [("a", True), ("b", False)], 1000, columns=[])
    assert top.tolist() == rows[:1000].tolist()
    t.free()


def test_sql_order_by(ctx):
    d = Database(ctx)
    cols = _cols(20_000, 11)
    d.create_table("t", cols)
    q = d.sql("SELECT u, s, f FROM t WHERE w < 70 ORDER BY s DESC NULLS FIRST, f, i DESC LIMIT 500 OFFSET 7")
    keys = [("s", True, False), ("f", False, False), ("i", True, True)]
    exp = ora.sort_limit(cols, col("w") < 70, keys, 507)[7:]
    assert q.names == ["u", "s", "f"]
    _check(q, cols, ["u", "s", "f"], exp)
    # ORDER BY a column that is not selected, and no LIMIT (tests/projection_pushdown.rs:50-56)
    q = d.sql("SELECT u FROM t ORDER BY j DESC, w")
    _check(q, cols, ["u"], ora.sort_limit(cols, None, [("j", True), ("w", False)]))
    # one key, LIMIT <= 32, the key selected alone: the fused top-k, same rows
    q = d.sql("SELECT f FROM t ORDER BY f DESC LIMIT 20")
    _check(q, cols, ["f"], ora.sort_limit(cols, None, [("f", True)], 20))
    q = d.sql("SELECT f FROM t ORDER BY f DESC NULLS FIRST LIMIT 20")
    _check(q, cols, ["f"], ora.sort_limit(cols, None, [("f", True, False)], 20))
    for n in list(d.tables):
        d.drop(n)
