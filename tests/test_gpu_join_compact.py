"""Partitioned join with compact 4-byte record keys, against the oracle.

The partitioned join keys its records on cmix(key - min) over the build key
range when that range spans at most 31 + log2(partitions) bits (4-byte record
keys), else on hash64(key) (8-byte records).  These cases drive the compact
form through its edges: probe keys below and above the build range, negative
ranges, dense ranges, ranges wide enough to need more partitions, duplicate
and NULL build keys (the non-unique probe), Int32 keys, and a range too wide
for compact records.  Each join type's aggregates must equal the oracle's (src/query/join.rs semantics: NULL keys
match nothing, LEFT and ANTI keep unmatched probe rows).
"""
from __future__ import annotations

import numpy as np
import pytest

import falcondb_amd as fa
import oracle as ora
from falcondb_amd._abi import ANTI, COUNT, INNER, LEFT, MIN, SEMI, SUM
from falcondb_amd.operators import AggExpr, NullableColumn, col

pytestmark = pytest.mark.gpu


def _assert_same(g: fa.RecordBatch, o: fa.RecordBatch):
    assert g.names == o.names
    for a, b, va, vb in zip(g.columns, o.columns, g.validity, o.validity):
        if va is not None or vb is not None:
            assert np.array_equal(va if va is not None else np.ones(a.size, bool),
                                  vb if vb is not None else np.ones(b.size, bool))
        assert np.array_equal(a, b.astype(a.dtype)) or np.array_equal(a, b, equal_nan=True), (a, b)


def _aggs(jt):
    aggs = [AggExpr(COUNT), AggExpr(SUM, "ppay"), AggExpr(MIN, "ppay")]
    if jt in (INNER, LEFT):
        aggs += [AggExpr(SUM, "bpay", side="build")]
    return aggs


def _check_all_types(ctx, pcols, bcols, expect_records, pred=None):
    pt, bt = fa.Table(ctx, pcols), fa.Table(ctx, bcols)
    idx = fa.JoinIndex(ctx, bt, "bk")
    for jt in (INNER, LEFT, SEMI, ANTI):
        o = ora.join(pcols, bcols, "pk", "bk", jt, _aggs(jt), expr=pred)
        for index in (None, idx):
            g = fa.execute_hash_join(ctx, pt, bt, "pk", "bk", jt, _aggs(jt), filter=pred, index=index)
            plan = ctx.last_plan()
            assert plan.startswith("join:partitioned") and f"records={expect_records}" in plan, plan
            _assert_same(g, o.exact)
    idx.free()
    pt.free()
    bt.free()


def _sample_keys(rng, lo, hi, n):
    """n distinct int64 keys drawn from [lo, hi)."""
    k = np.unique(rng.integers(lo, hi, int(n * 1.3), dtype=np.int64))
    rng.shuffle(k)
    assert k.size >= n
    return k[:n]


# (name, build key range [lo, hi), probe key range, expected record form)
RANGES = [
    ("negative", -(1 << 30), 1 << 30, (-(1 << 31), 1 << 31), "compact"),
    ("c3_like", 0, 1 << 40, None, "compact"),
    ("wide_42_bits", 0, 1 << 42, None, "compact"),        # 2,048 partitions
    ("too_wide_50_bits", 0, 1 << 50, None, "hash64"),     # would need 2^19 partitions
    ("dense", 0, 150_000, (-5_000, 160_000), "compact"),
]


@pytest.mark.parametrize("name,lo,hi,prange,records", RANGES, ids=[r[0] for r in RANGES])
def test_compact_records_key_ranges(ctx_part_compact, name, lo, hi, prange, records):
    rng = np.random.default_rng(sum(name.encode()))
    nb, npr = 100_000, 1_500_000
    bk = np.arange(lo, hi, dtype=np.int64) if hi - lo <= nb * 2 else _sample_keys(rng, lo, hi, nb)
    rng.shuffle(bk)
    bk = bk[:nb]
    if prange is None:  # 90 % of the probe keys from the build keys, the rest anywhere in [lo, hi)
        pk = np.where(rng.random(npr) < 0.9, bk[rng.integers(0, bk.size, npr)],
                      rng.integers(lo, hi, npr, dtype=np.int64))
    else:
        pk = rng.integers(prange[0], prange[1], npr, dtype=np.int64)
    bcols = {"bk": bk, "bpay": rng.random(bk.size)}
    pcols = {"pk": pk, "ppay": rng.integers(-10**6, 10**6, npr, dtype=np.int64),
             "f": rng.integers(0, 100, npr).astype(np.int32)}
    _check_all_types(ctx_part_compact, pcols, bcols, records)
    _check_all_types(ctx_part_compact, pcols, bcols, records, pred=col("f") < 45)


def test_compact_records_duplicates_and_nulls(ctx_part_compact):
    # duplicate build keys take the general probe (every match walked); NULL
    # keys on either side match nothing; probe keys outside the build range
    # on both sides
    rng = np.random.default_rng(7)
    nb, npr = 20_000, 600_000
    bk = rng.integers(-300, 500, nb, dtype=np.int64)
    pk = rng.integers(-600, 900, npr, dtype=np.int64)
    bcols = {"bk": NullableColumn(bk, rng.random(nb) > 0.03), "bpay": rng.normal(size=nb)}
    pcols = {"pk": NullableColumn(pk, rng.random(npr) > 0.05),
             "ppay": rng.integers(-50, 50, npr, dtype=np.int64)}
    _check_all_types(ctx_part_compact, pcols, bcols, "compact")


def test_compact_records_int32_keys(ctx_part_compact):
    # Int32 keys on both sides (the reference's typed path needs one key type,
    # src/query/join.rs:1394-1415), probe keys over the whole Int32 range
    rng = np.random.default_rng(11)
    nb, npr = 50_000, 800_000
    bk = _sample_keys(rng, -(1 << 20), 1 << 24, nb).astype(np.int32)
    pk = np.where(rng.random(npr) < 0.8, bk[rng.integers(0, nb, npr)],
                  rng.integers(-(1 << 31), (1 << 31) - 1, npr)).astype(np.int32)
    bcols = {"bk": bk, "bpay": rng.random(nb)}
    pcols = {"pk": pk, "ppay": rng.integers(-10**6, 10**6, npr, dtype=np.int64)}
    _check_all_types(ctx_part_compact, pcols, bcols, "compact")


@pytest.mark.parametrize("pay", ["narrow_int64", "narrow_int32", "wide_int64", "float64"])
def test_compact_records_payload_widths(ctx_part, ctx_part_compact, pay):
    # a probe payload rides as a 4-byte offset from its minimum when it is an
    # integer column whose range fits 32 bits, else as 8 bytes; the default
    # policy makes records compact for key-only and 4-byte-payload records
    from falcondb_amd import datagen as dg
    rng = np.random.default_rng(sum(map(ord, pay)))
    K = dg.keys(100_000, 40, seed_=dg.seed(3, 13))
    bcols, pcols = dg.c3_build(K), dg.c3_probe(2_000_000, K)
    n = pcols["pk"].size
    pcols["ppay"] = {"narrow_int64": rng.integers(-(1 << 31) + 5, (1 << 31) - 5, n),
                     "narrow_int32": rng.integers(-(1 << 31), (1 << 31) - 1, n).astype(np.int32),
                     "wide_int64": rng.integers(-(1 << 40), 1 << 40, n),
                     "float64": rng.normal(size=n)}[pay]
    narrow = pay.startswith("narrow")
    for c, forced in ((ctx_part, False), (ctx_part_compact, True)):
        pt, bt = fa.Table(c, pcols), fa.Table(c, bcols)
        for jt in (INNER, LEFT, SEMI, ANTI):
            aggs = [AggExpr(COUNT), AggExpr(SUM, "ppay"), AggExpr(MIN, "ppay")]
            g = fa.execute_hash_join(c, pt, bt, "pk", "bk", jt, aggs)
            plan = c.last_plan()
            if narrow:
                assert "records=compact pay4" in plan, plan
            elif forced:
                assert "records=compact" in plan and "pay4" not in plan, plan
            else:
                assert "records=hash64" in plan, plan
            _assert_same(g, ora.join(pcols, bcols, "pk", "bk", jt, aggs).exact)
        pt.free()
        bt.free()


def test_compact_and_hash64_records_agree(ctx_part, ctx_part_compact, ctx_part_hash64):
    # the same C3-shaped join in every record form: identical results; by
    # default only key-only records (INNER / LEFT here) are compact
    from falcondb_amd import datagen as dg
    K = dg.keys(200_000, 40, seed_=dg.seed(3, 12))
    bcols, pcols = dg.c3_build(K), dg.c3_probe(5_000_000, K)
    for jt in (INNER, LEFT, SEMI, ANTI):
        res = []
        default = "compact"  # key-only records, and SEMI / ANTI's 21-bit payload as 4 bytes
        for c, rec in ((ctx_part, default), (ctx_part_compact, "compact"), (ctx_part_hash64, "hash64")):
            pt, bt = fa.Table(c, pcols), fa.Table(c, bcols)
            # INNER / LEFT carry no probe payload (key-only records), SEMI / ANTI one
            aggs = ([AggExpr(COUNT), AggExpr(SUM, "bpay", side="build")] if jt in (INNER, LEFT)
                    else [AggExpr(COUNT), AggExpr(SUM, "ppay")])
            res.append(fa.execute_hash_join(c, pt, bt, "pk", "bk", jt, aggs))
            assert f"records={rec}" in c.last_plan(), c.last_plan()
            pt.free()
            bt.free()
        _assert_same(res[0], res[1])
        _assert_same(res[0], res[2])
