"""The oracle's restatement of the GROUP BY slow path (oracle.groupby_keys),
pinned before the GPU is compared with it.

The reference has no known answers for several GROUP BY keys, so the
restatement is pinned by a closed form (keys i mod 3, i mod 5: fifteen groups
given by the Chinese remainder theorem), by an independent brute force over
value tuples, by the display-string special cases of GroupKey::from_row
(src/query/aggregate.rs:79-83: NULL prints "NULL", every NaN prints "NaN",
-0.0 and 0.0 print differently), and by the single-key fast path where the two
must agree.  A planted mistake (-0.0 printed as 0.0, merging their groups)
must be caught.
"""
from __future__ import annotations

import math
import struct
from collections import OrderedDict

import numpy as np
import pytest

import oracle as ora
from falcondb_amd._abi import COUNT, MAX, MIN, SUM
from falcondb_amd.operators import AggExpr, DictColumn, NullableColumn, col


def _rows(b):
    return [tuple(c[i] if (b.validity[j] is None or b.validity[j][i]) else None
                  for j, c in enumerate(b.columns)) for i in range(b.num_rows)]


def test_closed_form_crt_groups():
    n = 30_000
    i = np.arange(n, dtype=np.int64)
    cols = {"a": i % 3, "b": i % 5, "v": i}
    r = ora.groupby_keys(cols, None, ["a", "b"], [AggExpr(COUNT), AggExpr(SUM, "v")])
    assert r.names[:2] == ["a", "b"] and r.num_rows == 15
    for a, b, cnt, tot in zip(r.columns[0], r.columns[1], r.columns[2], r.columns[3]):
        c = next(x for x in range(15) if x % 3 == a and x % 5 == b)
        assert cnt == n // 15
        assert tot == n // 15 * c + 15 * (n // 15) * (n // 15 - 1) // 2


def _canon(v):
    """A hashable value with the reference's grouping: NaN one value, -0.0
    apart from 0.0 (their bit patterns), NULL apart from everything."""
    if v is None:
        return ("null",)
    if isinstance(v, float):
        return ("nan",) if math.isnan(v) else ("f", struct.pack("<d", v))
    return ("v", v)


def _brute(cols, keys, aggs_cols, mask):
    groups = OrderedDict()

    def value(c, r):
        if isinstance(c, DictColumn):
            x = int(c.codes[r])
            return None if x == 0xFFFFFFFF else c.dictionary[x]
        if isinstance(c, NullableColumn):
            return (c.values[r].item() if c.valid[r] else None)
        return c[r].item()
    nrows = mask.size
    for r in range(nrows):
        if not mask[r]:
            continue
        kv = tuple(value(cols[k], r) for k in keys)
        key = tuple(_canon(x) for x in kv)
        g = groups.setdefault(key, {"key": kv, "n": 0, "s": 0})
        g["n"] += 1
        g["s"] += int(cols[aggs_cols][r])
    return groups


@pytest.mark.parametrize("seed", range(4))
def test_brute_force_tuples(seed):
    rng = np.random.default_rng(900 + seed)
    n = 5000
    f = rng.choice([0.0, -0.0, 1.5, np.nan, 2.0], n)
    cols = {
        "i": NullableColumn(rng.integers(-3, 4, n), rng.random(n) > 0.1),
        "s": DictColumn(np.where(rng.random(n) < 0.1, 0xFFFFFFFF, rng.integers(0, 3, n)).astype(np.uint32),
                        ["x", "y", "z"]),
        "j": rng.integers(0, 2, n).astype(np.int32),
        "f": NullableColumn(f, rng.random(n) > 0.05),
        "v": rng.integers(-100, 100, n),
    }
    for keys in (["i", "s"], ["s", "j", "i"], ["f"], ["i"], ["j", "i", "s"]):
        pred = col("v") > -50
        mask = ora._unpack_mask(ora.filter_mask(cols, pred), n)
        r = ora.groupby_keys(cols, pred, keys, [AggExpr(COUNT), AggExpr(SUM, "v")])
        b = _brute(cols, keys, "v", mask)
        assert r.num_rows == len(b), keys
        got = {}
        for row in _rows(r):
            kv = row[:len(keys)]
            got[tuple(_canon(x.item() if hasattr(x, "item") else x) for x in kv)] = row[len(keys):]
        for key, g in b.items():
            cnt, tot = got[key]
            assert cnt == g["n"] and tot == g["s"], (keys, key)


def _check_special_cases():
    f = np.array([0.0, -0.0, np.nan, -np.nan, float("nan"), 1.0, 0.0, np.inf])
    valid = np.array([1, 1, 1, 1, 1, 1, 1, 0], bool)
    r = ora.groupby_keys({"f": NullableColumn(f, valid)}, None, ["f"], [AggExpr(COUNT)])
    got = {}
    for v, ok, c in zip(r.columns[0], r.validity[0], r.columns[1]):
        k = "NULL" if not ok else "nan" if math.isnan(v) else repr(float(v))
        got[k] = int(c)
    assert got == {"0.0": 2, "-0.0": 1, "nan": 3, "1.0": 1, "NULL": 1}


def test_display_string_special_cases():
    _check_special_cases()
    # a NULL text key and the text "NULL" print alike, so they form one group
    s = DictColumn(np.array([0, 1, 0xFFFFFFFF, 1], np.uint32), ["a", "NULL"])
    r = ora.groupby_keys({"s": s, "k": np.zeros(4, np.int64)}, None, ["s", "k"], [AggExpr(COUNT)])
    assert sorted(zip([x for x in r.columns[0]], r.columns[2].tolist())) == [("NULL", 3), ("a", 1)]


def test_agrees_with_single_key_fast_path():
    rng = np.random.default_rng(5)
    n = 20_000
    cols = {"k": rng.integers(0, 50, n), "c": np.full(n, 7, np.int64), "v": rng.normal(size=n)}
    aggs = [AggExpr(COUNT), AggExpr(SUM, "v"), AggExpr(MIN, "v"), AggExpr(MAX, "v")]
    slow = ora.groupby_keys(cols, col("v") < 1.0, ["k", "c"], aggs)
    fast = ora.groupby(cols, col("v") < 1.0, "k", aggs).exact.sort_by("k")
    o = np.argsort(slow.columns[0])
    assert np.array_equal(slow.columns[0][o], fast.columns[0])
    assert np.all(slow.columns[1] == 7)
    for i in range(1, 5):
        assert np.array_equal(slow.columns[i + 1][o], fast.columns[i])


def test_planted_zero_merge_is_caught(monkeypatch):
    # a restatement that prints a float after adding 0.0 (so -0.0 prints as
    # 0.0 and joins its group) must fail the special-case check
    good = ora._display_keys

    def planted(c, n):
        if not isinstance(c, (DictColumn, NullableColumn)) or np.asarray(
                c.values if isinstance(c, NullableColumn) else c).dtype.kind != "f":
            return good(c, n)
        vals = np.asarray(c.values, np.float64) + 0.0
        return good(NullableColumn(vals, c.valid), n)

    monkeypatch.setattr(ora, "_display_keys", planted)
    with pytest.raises(AssertionError):
        _check_special_cases()
