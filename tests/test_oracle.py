"""Pins the CPU oracle (oracle/liboracle.so) before it is trusted as the checker.

Parity status of the oracle's exported functions:
  ora_filter_mask / ora_filter_count   pinned: reference known answers
      (tests/count_star_filter.rs:51-90 via closed forms) + brute force below
  ora_filter_select                    pinned: brute force (row order, values)
  ora_groupby (exact columns)          pinned: reference known answers
      (tests/diag_agg.rs, tests/sql_integration.rs:237-306,
      tests/count_star_filter.rs:95-117) + brute force + math.fsum
  ora_groupby (reference columns)      pinned to within 1e-12 of the exact sum;
      their last bits follow the restated lane / batch order, which nothing
      outside the Rust binary can confirm
  ora_join                             pinned: reference known answers
      (src/query/join.rs:1747-1848, :2607-2676) + nested-loop brute force
  ora_exact_sum                        pinned: math.fsum, bit for bit, and exact
      rational sums (fractions.Fraction) over subnormals, ties and partial sums
      beyond the double range (where math.fsum raises)
  ora_sum_i64 / ora_checked_sum_i64    pinned: the reference's int_sum known
      answers (src/query/aggregate.rs:1936-1993)
  exact integer SUM columns            pinned: int_sum known answers through
      scalar, GROUP BY and join (exact int64, or the correctly rounded double
      past int64, checked against Python's exact integers)
  ora_sort_limit                       pinned: the reference's ORDER BY known
      answers (tests/sort_and_dedup_paths.rs, src/query/sort.rs unit tests,
      tests/projection_pushdown.rs:42-56), Python's stable sorted() with the
      same keys on random multi-key inputs, and the float total order
  LIKE / ILIKE / IN / NOT IN leaves     pinned: the reference's LIKE and IN
      row counts (src/query/expr.rs:2050-2082, :2415-2462, :2700-2739;
      tests/sql_integration.rs:1189-1238; tests/count_star_filter.rs:67-69),
      and brute force against Python's regular expressions (LIKE) and
      membership tests (IN); ILIKE folds ASCII letters only
  concat_dicts (Python)                pinned: closed form of the merged dictionary
      (ordered union of the batches' dictionaries), every row decoding to the
      same string, the shared-dictionary fast path, and a planted row-order
      numbering that the checks reject.  The reference's own concat_all tests
      cover primitive columns only (src/storage/column.rs:549-590).
"""
import math

import numpy as np
import pytest

import oracle as ora
from falcondb_amd._abi import ANTI, AVG, COUNT, INNER, LEFT, MAX, MIN, NO_ROW, SEMI, SUM
from falcondb_amd.operators import AggExpr, DictColumn, NullableColumn, col

from known_answers import GOLDEN, check_all, check_int_sum, check_like_in, check_sort, column_values


class OracleEngine:
    def filter_count(self, cols, pred):
        return ora.filter_count(cols, pred)

    def groupby(self, cols, pred, key, aggs, batch_rows=65536):
        return ora.groupby(cols, pred, key, aggs, batch_rows=batch_rows).exact

    def join(self, pcols, bcols, pk, bk, jt, aggs):
        return ora.join(pcols, bcols, pk, bk, jt, aggs).exact

    def sort(self, cols, pred, keys, limit):
        return ora.sort_limit(cols, pred, keys, limit)


def test_reference_known_answers():
    assert check_all(OracleEngine()) > 100


def test_sum_i64_restatement_known_answers():
    # aggregate.rs:1936-1993, the reference's int_sum tests, on the oracle's
    # restatement of sum_i64 / checked_sum_i64 / sum_lanes
    g = GOLDEN["int_sum"]
    for n in g["ordinary"]["rows"]:
        d = np.array([(i % 997) - 498 for i in range(n)], np.int64)
        assert ora.sum_i64(d) == ora.sum_lanes(d.astype(np.float64)) == float(sum(d.tolist()))
    d = np.array([(1 << 53) + i for i in range(8)], np.int64)
    exact = g["past_float_limit"]["exact"]
    new, old = ora.sum_i64(d), ora.sum_lanes(d.astype(np.float64))
    assert int(new) == int(float(exact))  # the correctly rounded total
    assert abs(int(new) - exact) < abs(int(old) - exact)
    d = np.array(g["overflow"]["data"], np.int64)
    assert ora.checked_sum_i64(d) is None
    got = ora.sum_i64(d)
    assert got == ora.sum_lanes(d.astype(np.float64)) and got > g["overflow"]["greater_than"]
    for n in g["ragged_tail"]["n"]:
        assert ora.sum_i64(np.arange(1, n + 1, dtype=np.int64)) == n * (n + 1) // 2
    # a planted wrapping sum is caught: it gives the wrapped (negative) total
    wrapped = int(np.sum(d, dtype=np.int64))
    assert not wrapped > g["overflow"]["greater_than"]


def test_int_sum_known_answers_every_path():
    assert check_int_sum(OracleEngine()) > 100


@pytest.mark.parametrize("seed", range(6))
def test_exact_sum_is_fsum(seed):
    rng = np.random.default_rng(seed)
    for _ in range(100):
        n = int(rng.integers(0, 60))
        x = rng.normal(size=n) * 10.0 ** rng.integers(-30, 30, size=n)
        if n and seed % 2:
            x = np.concatenate([x, -x[: n // 2] * (1 + 1e-15)])  # heavy cancellation
        assert ora.exact_sum(x) == math.fsum(x)
    assert ora.exact_sum(np.array([1e308, -1e308, 1e292])) == math.fsum([1e308, -1e308, 1e292])
    assert math.isnan(ora.exact_sum(np.array([1.0, np.inf, -np.inf])))
    assert ora.exact_sum(np.array([5e-324, 5e-324])) == 1e-323


def _brute_mask(vals, valid, op, lit):
    f = {"<": lambda a: a < lit, "<=": lambda a: a <= lit, ">": lambda a: a > lit,
         ">=": lambda a: a >= lit, "=": lambda a: a == lit, "!=": lambda a: a != lit}[op]
    return [bool(ok and f(v)) for v, ok in zip(vals, valid)]


@pytest.mark.parametrize("seed", range(4))
def test_filter_brute_force(seed):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(1, 400))
    iv = rng.integers(-20, 20, n)
    ivalid = rng.random(n) > 0.2
    fv = rng.normal(size=n)
    fv[rng.random(n) < 0.05] = np.nan
    i32 = rng.integers(-5, 5, n).astype(np.int32)
    cols = {"i": NullableColumn(iv, ivalid), "f": fv, "k": i32}
    ops = {"<": lambda c, v: c < v, "<=": lambda c, v: c <= v, ">": lambda c, v: c > v,
           ">=": lambda c, v: c >= v, "=": lambda c, v: c == v, "!=": lambda c, v: c != v}
    for op, mk in ops.items():
        for name, vals, valid, lit in [("i", iv, ivalid, 3), ("f", fv, np.ones(n, bool), 0.25),
                                       ("k", i32, np.ones(n, bool), -1)]:
            expect = _brute_mask(vals.tolist(), valid.tolist(), op, lit)
            words = ora.filter_mask(cols, mk(col(name), lit))
            bits = np.unpackbits(words.view(np.uint8), bitorder="little")[:n].astype(bool)
            assert bits.tolist() == expect, (op, name)
            assert not np.unpackbits(words.view(np.uint8), bitorder="little")[n:].any()
            assert ora.filter_count(cols, mk(col(name), lit)) == sum(expect)
    sel = ora.filter_select(cols, col("i") > 0, ["i", "f"])
    rows = [r for r in range(n) if ivalid[r] and iv[r] > 0]
    assert sel.columns[0].tolist() == iv[rows].tolist()
    assert np.array_equal(sel.columns[1], fv[rows], equal_nan=True)


@pytest.mark.parametrize("seed", range(4))
def test_groupby_brute_force(seed):
    rng = np.random.default_rng(200 + seed)
    n = int(rng.integers(1, 500))
    key = rng.integers(0, 15, n)
    kvalid = rng.random(n) > 0.1
    x = rng.normal(size=n) * 100
    xvalid = rng.random(n) > 0.2
    y = rng.integers(-1000, 1000, n)
    f = rng.integers(0, 10, n).astype(np.int32)
    cols = {"k": NullableColumn(key, kvalid), "x": NullableColumn(x, xvalid), "y": y, "f": f}
    aggs = [AggExpr(COUNT), AggExpr(COUNT, "x"), AggExpr(SUM, "x"), AggExpr(AVG, "x"),
            AggExpr(MIN, "y"), AggExpr(MAX, "y"), AggExpr(SUM, "y"), AggExpr(AVG, "y")]
    r = ora.groupby(cols, col("f") < 7, "k", aggs).exact
    groups = {}
    for i in range(n):
        if f[i] >= 7:
            continue
        g = key[i] if kvalid[i] else None
        groups.setdefault(g, []).append(i)
    assert r.num_rows == len(groups)
    for j in range(r.num_rows):
        g = r.columns[0][j] if r.validity[0][j] else None
        idx = groups[g]
        xs = [x[i] for i in idx if xvalid[i]]
        ys = [int(y[i]) for i in idx]
        assert r.columns[1][j] == len(idx)
        assert r.columns[2][j] == len(xs)
        if xs:
            assert r.columns[3][j] == math.fsum(xs)
            assert r.columns[4][j] == math.fsum(xs) / len(xs)
        else:
            assert not r.validity[3][j]
        assert r.columns[5][j] == min(ys) and r.columns[6][j] == max(ys)
        assert r.columns[7][j] == sum(ys) and r.columns[8][j] == sum(ys) / len(ys)


def test_groupby_reference_order_within_tolerance():
    rng = np.random.default_rng(7)
    n = 300_000
    cols = {"k": rng.integers(0, 50, n), "v": rng.normal(size=n)}
    r = ora.groupby(cols, None, "k", [AggExpr(SUM, "v")])
    e, ref = r.exact.columns[1], r.reference.columns[1]
    assert np.all(np.abs(e - ref) <= 1e-12 * np.maximum(np.abs(e), 1e-300) + 1e-300) or \
        np.mean(np.abs(e - ref) <= 1e-12 * np.abs(e)) > 0.98


@pytest.mark.parametrize("seed", range(4))
def test_join_brute_force(seed):
    rng = np.random.default_rng(300 + seed)
    nb, np_ = int(rng.integers(1, 60)), int(rng.integers(1, 200))
    bk = rng.integers(0, 30, nb)
    bvalid = rng.random(nb) > 0.1
    pk = rng.integers(0, 40, np_)
    pvalid = rng.random(np_) > 0.1
    bx = rng.normal(size=nb)
    px = rng.integers(-9, 9, np_)
    bcols = {"bk": NullableColumn(bk, bvalid), "bx": bx}
    pcols = {"pk": NullableColumn(pk, pvalid), "px": px}
    for jt in (INNER, LEFT, SEMI, ANTI):
        pairs = []
        for i in range(np_):
            ms = [j for j in range(nb) if pvalid[i] and bvalid[j] and pk[i] == bk[j]]
            if jt == INNER:
                pairs += [(i, j) for j in ms]
            elif jt == LEFT:
                pairs += [(i, j) for j in ms] or [(i, NO_ROW)]
            elif jt == SEMI and ms:
                pairs.append((i, NO_ROW))
            elif jt == ANTI and not ms:
                pairs.append((i, NO_ROW))
        aggs = [AggExpr(COUNT), AggExpr(SUM, "px")] + ([AggExpr(SUM, "bx", side="build")] if jt in (INNER, LEFT) else [])
        res, l, r = ora.join(pcols, bcols, "pk", "bk", jt, aggs, pairs=True)
        assert sorted(zip(l.tolist(), r.tolist())) == sorted(pairs), jt
        e = res.exact
        assert e.columns[0][0] == len(pairs)
        assert e.columns[1][0] == sum(int(px[i]) for i, _ in pairs)
        if jt in (INNER, LEFT):
            bs = [bx[j] for _, j in pairs if j != NO_ROW]
            if bs:
                assert e.columns[2][0] == math.fsum(bs)
            else:
                assert not e.validity[2][0]


def test_planted_mistake_is_caught():
    """A mistake planted in the restatement (here: GT evaluated as GE) changes a
    known answer, so the checks above would reject such an oracle."""
    from known_answers import closed_form_count, count_star_table
    cols = count_star_table(1000)
    good = ora.filter_count(cols, col("n") > 10)
    wrong = ora.filter_count(cols, col("n") >= 10)  # what a GT->GE slip computes
    assert good == closed_form_count("n > 10", 1000) and wrong != good


# ─── ORDER BY ... LIMIT k ─────────────────────────────────────────────────────
def _topk_values(cols, name, rows):
    v = cols[name]
    vals = v.values if isinstance(v, NullableColumn) else v
    valid = v.valid if isinstance(v, NullableColumn) else np.ones(len(vals), bool)
    return [vals[r] if valid[r] else None for r in rows]


def test_topk_reference_known_answers():
    # tests/sort_and_dedup_paths.rs:107-125: v = 9, 3, 7, 1, 5
    cols = {"v": np.array([9, 3, 7, 1, 5], np.int64)}
    assert _topk_values(cols, "v", ora.topk(cols, None, "v", 3, desc=False)) == [1, 3, 5]
    assert _topk_values(cols, "v", ora.topk(cols, None, "v", 2, desc=True)) == [9, 7]
    assert sorted(_topk_values(cols, "v", ora.topk(cols, None, "v", 10))) == [1, 3, 5, 7, 9]
    # tests/top_k_scan.rs:60-95: the limited sort is the head of the full sort
    n = 1000
    s = np.array([((i * 7919) % n) + 0.5 for i in range(n)])
    ok = np.array([i % 97 != 0 for i in range(n)])
    cols = {"score": NullableColumn(s, ok)}
    keyed = sorted(s[ok].tolist())
    for desc in (False, True):
        full = ([None] * int((~ok).sum()) + keyed) if not desc else (keyed[::-1] + [None] * int((~ok).sum()))
        for k in (1, 2, 10):
            assert _topk_values(cols, "score", ora.topk(cols, None, "score", k, desc)) == full[:k]
    # tests/top_k_scan.rs:124-131: more NULLs than the limit, placed first
    s = np.arange(n, dtype=np.float64)
    cols = {"score": NullableColumn(s, np.arange(n) >= n - 3)}
    assert _topk_values(cols, "score", ora.topk(cols, None, "score", 5, desc=False)) == [None] * 5


def test_topk_float_total_order():
    # f64_ord_key (src/query/sort.rs:452-454) is IEEE 754 totalOrder:
    # -NaN < -inf < ... < -0.0 < +0.0 < ... < +inf < +NaN
    neg_nan = np.frombuffer(np.array([0xfff8000000000000], np.uint64).tobytes(), np.float64)[0]
    v = np.array([1.0, np.nan, -np.inf, np.inf, -0.0, 0.0, neg_nan])
    order = ora.topk({"v": v}, None, "v", 7, desc=True)
    assert order.tolist() == [1, 3, 0, 5, 4, 2, 6]
    assert ora.topk({"v": v}, None, "v", 7, desc=False).tolist() == [6, 2, 4, 5, 0, 3, 1]


@pytest.mark.parametrize("seed", range(3))
def test_topk_brute_force(seed):
    rng = np.random.default_rng(400 + seed)
    n = int(rng.integers(1, 3000))
    v = rng.integers(-50, 50, n)  # many ties: broken by row
    f = rng.integers(0, 10, n)
    cols = {"v": v, "f": f}
    for desc in (False, True):
        for k in (1, 7, 32, 5000):
            rows = [i for i in range(n) if f[i] < 6]
            exp = sorted(rows, key=lambda i: (-v[i] if desc else v[i], i))[:k]
            assert ora.topk(cols, col("f") < 6, "v", k, desc).tolist() == exp


def test_like_in_reference_known_answers():
    assert check_like_in(OracleEngine()) >= 35


def _like_re(pattern: str, fold: bool):
    import re
    pat = pattern.encode()
    if fold:
        pat = pat.lower()  # bytes.lower folds ASCII letters only
    body = b"".join(b".*" if ch == ord("%") else b"." if ch == ord("_") else re.escape(bytes([ch])) for ch in pat)
    return re.compile(body, re.DOTALL)


@pytest.mark.parametrize("seed", range(3))
def test_like_in_brute_force(seed):
    rng = np.random.default_rng(700 + seed)
    words = ["", "a", "ab", "abc", "Abc", "ba", "b%a", "x_y", "xay", "é", "éa", "aaa", "banana", "BANANA"]
    n = 3000
    codes = rng.integers(0, len(words), n).astype(np.uint32)
    codes[rng.random(n) < 0.1] = 0xFFFFFFFF
    iv = rng.integers(-5, 6, n)
    ok = rng.random(n) > 0.1
    fv = rng.choice([-1.5, -0.0, 0.0, 1.0, 2.0, 2.5, np.nan, np.inf], n)
    cols = {"s": DictColumn(codes, words), "i": NullableColumn(iv, ok), "f": fv,
            "j": rng.integers(-3, 4, n).astype(np.int32)}
    s_vals = [None if c == 0xFFFFFFFF else words[c] for c in codes]
    for pat in ["%", "", "a%", "%a", "_", "__", "%a%a%", "b%a", "_a%", "x_y", "%é%", "A%", "%%", "a_c"]:
        for neg in (False, True):
            for fold in (False, True):
                r = _like_re(pat, fold)
                exp = sum(1 for v in s_vals if (neg if v is None else
                          (r.fullmatch(v.encode().lower() if fold else v.encode()) is not None) != neg))
                got = ora.filter_count(cols, col("s").like(pat, negated=neg, case_insensitive=fold))
                assert got == exp, (pat, neg, fold)
    for lst in ([1, 3], [-5, 5, 0], [2.0, 7], [2.5], [], ["a"], [1, 2.5]):
        for neg in (False, True):
            for name in ("i", "f", "j"):
                c = cols[name]
                vals = c.values if isinstance(c, NullableColumn) else c
                valid = c.valid if isinstance(c, NullableColumn) else np.ones(n, bool)
                def member(x):
                    for v in lst:
                        if isinstance(v, str):
                            continue  # a text literal matches no number (an unsupported pairing)
                        if isinstance(v, float) and name == "j":
                            continue  # Int32 against a float: unsupported pairing
                        if isinstance(v, int) and name == "j":
                            if np.int32(x) == np.int32(np.int64(v).astype(np.int32)):
                                return True
                        elif float(x) == float(v):
                            return True
                    return False
                exp = sum(1 for x, okr in zip(vals.tolist(), valid.tolist()) if okr and member(x) != neg)
                got = ora.filter_count(cols, col(name).isin(lst, negated=neg))
                assert got == exp, (name, lst, neg)


def test_sort_reference_known_answers():
    assert check_sort(OracleEngine()) >= 25


def _sort_key_fn(cols, name, desc, nulls_last):
    """Python's sort key for one ORDER BY key: NULLs placed by nulls_last,
    values by their total order (f64_ord_key for doubles, byte order for
    strings), negated or inverted for DESC so that one ascending stable
    sorted() reproduces the reference's comparator."""
    c = cols[name]
    vals = column_values(c, np.arange(len(c.codes) if isinstance(c, DictColumn) else
                                      len(c.values) if isinstance(c, NullableColumn) else len(c)))

    def ordkey(v):
        if isinstance(v, float):
            b = int(np.array([v]).view(np.uint64)[0])
            return (b ^ (1 << 63)) if b >> 63 == 0 else (~b & ((1 << 64) - 1))
        if isinstance(v, str):
            return v.encode()
        return v

    def key(r):
        v = vals[r]
        if v is None:
            return (0 if not nulls_last else 2, 0)
        k = ordkey(v)
        if desc:
            k = bytes(255 - x for x in k) + b"\xff" if isinstance(k, bytes) else -k
        return (1, k)
    return key


@pytest.mark.parametrize("seed", range(4))
def test_sort_brute_force(seed):
    rng = np.random.default_rng(900 + seed)
    n = int(rng.integers(1, 2500))
    words = ["", "a", "ab", "b", "ba", "zz", "Z", "é", "a\x7f"]
    cols = {
        "i": NullableColumn(rng.integers(-5, 5, n), rng.random(n) > 0.2),
        "f": NullableColumn(rng.choice([-np.inf, -1.5, -0.0, 0.0, 2.0, np.inf, np.nan], n), rng.random(n) > 0.1),
        "s": DictColumn(np.where(rng.random(n) < 0.15, 0xFFFFFFFF, rng.integers(0, len(words), n)).astype(np.uint32),
                        words),
        "w": rng.integers(0, 10, n).astype(np.int32),
    }
    for trial in range(12):
        names = list(rng.permutation(["i", "f", "s", "w"])[: int(rng.integers(1, 4))])
        keys = [(c, bool(rng.integers(0, 2)), bool(rng.integers(0, 2))) for c in names]
        limit = [None, 0, 1, 17, n + 5][int(rng.integers(0, 5))]
        pred = col("w") < 7 if trial % 2 else None
        rows = [r for r in range(n) if pred is None or cols["w"][r] < 7]
        fns = [_sort_key_fn(cols, *k) for k in keys]
        exp = sorted(rows, key=lambda r: tuple(f(r) for f in fns))
        exp = exp if limit is None else exp[:limit]
        assert ora.sort_limit(cols, pred, keys, limit).tolist() == exp, (keys, limit)


def test_sort_second_key_breaks_ties():
    # a comparator that ignored the second key would still pass one-key checks
    cols = {"d": sort_column_utf8(["eng", "hr", "eng"]), "s": np.array([7.0, 9.0, 8.0])}
    assert ora.sort_limit(cols, None, [("d", False), ("s", True)]).tolist() == [2, 0, 1]
    assert ora.sort_limit(cols, None, [("d", False)]).tolist() == [0, 2, 1]


def sort_column_utf8(vals):
    dic = list(dict.fromkeys(vals))
    return DictColumn(np.array([dic.index(v) for v in vals], np.uint32), dic)


def _batched_dicts(seed):
    rng = np.random.default_rng(seed)
    vocab = [f"w{i:02d}" for i in range(40)]
    rows, dicts, codes = [], [], []
    for b in range(7):
        d = list(rng.permutation(vocab)[: int(rng.integers(0, 15))])
        n = int(rng.integers(0, 300))
        c = rng.integers(0, max(len(d), 1), n).astype(np.uint32)
        if not d:
            c[:] = 0xFFFFFFFF
        c[rng.random(n) < 0.1] = 0xFFFFFFFF
        # rows reach the dictionary's entries in another order than the dictionary's
        if len(d) > 1 and n > 1:
            c[0] = len(d) - 1
        rows.append(n)
        dicts.append(d)
        codes.append(c)
    return np.concatenate(codes) if codes else np.zeros(0, np.uint32), rows, dicts


def _decode(codes, rows, dicts):
    out, r0 = [], 0
    for n, d in zip(rows, dicts):
        out += [None if c == 0xFFFFFFFF else d[c] for c in codes[r0:r0 + n].tolist()]
        r0 += n
    return out


@pytest.mark.parametrize("seed", range(6))
def test_concat_dicts_merge(seed):
    codes, rows, dicts = _batched_dicts(seed)
    m = ora.concat_dicts(codes, rows, dicts)
    # closed form: the ordered union of the dictionaries, batch after batch
    assert m.dictionary == list(dict.fromkeys(s for d in dicts for s in d))
    # every row decodes to the string it had
    got = [None if c == 0xFFFFFFFF else m.dictionary[c] for c in m.codes.tolist()]
    assert got == _decode(codes, rows, dicts)


def test_concat_dicts_shared_dictionary_keeps_codes():
    d = ["x", "y", "z"]
    codes = np.array([2, 0, 0xFFFFFFFF, 1, 2], np.uint32)
    m = ora.concat_dicts(codes, [2, 3], [d, list(d)])
    assert m.dictionary == d and np.array_equal(m.codes, codes)


def test_concat_dicts_planted_row_order_numbering_is_caught():
    # Numbering entries in the order rows first use them (instead of the
    # dictionaries' order) decodes every row correctly but breaks the closed form.
    codes, rows, dicts = _batched_dicts(3)
    used = []
    for s in _decode(codes, rows, dicts):
        if s is not None and s not in used:
            used.append(s)
    merged = used + [s for d in dicts for s in d if s not in used]
    merged = list(dict.fromkeys(merged))
    assert merged != list(dict.fromkeys(s for d in dicts for s in d))
    assert ora.concat_dicts(codes, rows, dicts).dictionary != merged


def _fraction_sum(x):
    """The correctly rounded sum by exact rational arithmetic."""
    from fractions import Fraction
    v, c = np.unique(x, return_counts=True)
    return float(sum(Fraction(float(a)) * int(n) for a, n in zip(v, c)))


def _huge_column(n):
    """+/-1.7e308 alternating (a strided per-thread partial overflows), every
    1000th pair 1e300, balanced so that the total is 1.7e308 plus the 1e300s."""
    x = np.where(np.arange(n) % 2 == 0, 1.7e308, -1.7e308)
    x[::1000] = 1e300
    x[1::1000] = 1e300
    extra = int((x == 1.7e308).sum() - (x == -1.7e308).sum()) - 1
    idx = np.flatnonzero(x == (1.7e308 if extra > 0 else -1.7e308))[:abs(extra)]
    x[idx] = 0.0
    return x


def _fold_min_max(vals):
    """f64::min / f64::max folded left to right (the NaN-skipping IEEE minNum /
    maxNum): NaN unless some value is not NaN, then the min / max of those."""
    clean = [v for v in vals if not math.isnan(v)]
    return (min(clean), max(clean)) if clean else (math.nan, math.nan)


def test_float_minmax_nan_semantics():
    # Exact columns fold MIN / MAX like f64::min / f64::max, as the reference's
    # execute_agg_only (src/query/aggregate.rs:3296-3330) and its GROUP BY slow
    # path (Accumulator::update, :283-302) do: NaN values are skipped and a
    # group of NaN values alone yields NaN.  The reference-order columns
    # restate the GROUP BY fast path, whose `!has_val || v < sum` keeps a NaN
    # that comes first (:1093-1101).
    k = np.array([0, 0, 0, 1, 1, 2, 2], np.int64)
    x = np.array([np.nan, 1.0, -2.0, 3.0, np.nan, np.nan, np.nan])
    o = ora.groupby({"k": k, "x": x}, None, "k", [AggExpr(MIN, "x"), AggExpr(MAX, "x")])
    e, r = o.exact.sort_by("k"), o.reference.sort_by("k")
    for g in range(3):
        lo, hi = _fold_min_max(list(x[k == g]))
        assert (math.isnan(lo) and math.isnan(e.columns[1][g])) or e.columns[1][g] == lo
        assert (math.isnan(hi) and math.isnan(e.columns[2][g])) or e.columns[2][g] == hi
        assert e.validity[1] is None or e.validity[1][g]
    assert list(e.columns[1][:2]) == [-2.0, 3.0] and list(e.columns[2][:2]) == [1.0, 3.0]
    assert math.isnan(e.columns[1][2]) and math.isnan(e.columns[2][2])
    assert math.isnan(r.columns[1][0]) and math.isnan(r.columns[2][0])
    assert list(r.columns[1][1:2]) == [3.0] and list(r.columns[2][1:2]) == [3.0]
    s = ora.groupby({"x": x}, None, None, [AggExpr(MIN, "x"), AggExpr(MAX, "x")]).exact
    assert [s.columns[0][0], s.columns[1][0]] == [-2.0, 3.0]
    # no GROUP BY over NaN values alone: NaN (a value exists); over no rows: NULL
    s = ora.groupby({"x": np.array([np.nan, np.nan])}, None, None, [AggExpr(MIN, "x"), AggExpr(MAX, "x")]).exact
    assert math.isnan(s.columns[0][0]) and math.isnan(s.columns[1][0])
    s = ora.groupby({"x": np.array([1.0])}, col("x") > 5, None, [AggExpr(MIN, "x")]).exact
    assert s.validity[0] is not None and not s.validity[0][0]


@pytest.mark.parametrize("seed", range(3))
def test_float_minmax_brute_force(seed):
    # random groups with NaN sprinkled in, some groups NaN only
    rng = np.random.default_rng(600 + seed)
    n = 3000
    k = rng.integers(0, 60, n)
    x = rng.normal(size=n)
    x[rng.random(n) < 0.2] = np.nan
    x[k < 3] = np.nan
    e = ora.groupby({"k": k, "x": x}, None, "k", [AggExpr(MIN, "x"), AggExpr(MAX, "x")]).exact.sort_by("k")
    for i, g in enumerate(e.columns[0]):
        lo, hi = _fold_min_max(list(x[k == g]))
        got_lo, got_hi = e.columns[1][i], e.columns[2][i]
        assert (math.isnan(lo) and math.isnan(got_lo)) or got_lo == lo
        assert (math.isnan(hi) and math.isnan(got_hi)) or got_hi == hi


@pytest.mark.parametrize("seed", range(3))
def test_exact_sum_is_the_rational_sum(seed):
    from fractions import Fraction
    rng = np.random.default_rng(50 + seed)
    for t in range(400):
        n = int(rng.integers(1, 40))
        x = rng.normal(size=n) * 10.0 ** rng.integers(-320, 300, size=n)
        if t % 3 == 0:
            x = np.concatenate([x, -x[: n // 2]])
        if t % 7 == 0:
            x = np.concatenate([x, [5e-324, -1e-310, 2.2250738585072014e-308]])
        assert ora.exact_sum(x) == float(sum(Fraction(float(v)) for v in x))
    # round half to even at the last place, and a sticky bit far below it
    assert ora.exact_sum(np.array([1.0, 2.0 ** -53])) == 1.0
    assert ora.exact_sum(np.array([1.0, 2.0 ** -53, 2.0 ** -1000])) == 1.0 + 2 ** -52
    assert ora.exact_sum(np.array([1.0 + 2 ** -52, 2.0 ** -53])) == 1.0 + 2 ** -51
    assert ora.exact_sum(np.array([1.7976931348623157e308, 1e292])) == math.inf


def test_exact_sum_near_dbl_max():
    x = _huge_column(10_000)
    exact = _fraction_sum(x)  # math.fsum raises on the overflowing partials
    assert math.isfinite(exact) and ora.exact_sum(x) == exact
    r = ora.groupby({"x": x}, None, None, [AggExpr(SUM, "x")]).exact
    assert r.columns[0][0] == exact
