"""Batched GROUP BY (fgpu_batch_*): every query of a batch must give exactly
what the oracle and the plain per-query fgpu_groupby give, on every replay of
the captured graph and after autotuning."""
import numpy as np
import pytest

import falcondb_amd as fa
import oracle as ora
from falcondb_amd import datagen as dg
from falcondb_amd._abi import AVG, COUNT, MAX, MIN, SUM
from falcondb_amd.operators import AggExpr, DictColumn, NullableColumn, col

from test_gpu_parity import _table, assert_batches_equal

pytestmark = pytest.mark.gpu

C4_AGGS = [AggExpr(COUNT), AggExpr(SUM, "val"), AggExpr(MIN, "val"), AggExpr(MAX, "val"),
           AggExpr(SUM, "score")]


def _mixed_batch(ctx):
    """(table, key, aggs, filter, columns) for a batch that exercises both
    pass-1 modes, NULL keys and inputs, a skewed key, and a fallback query."""
    out = []
    shapes = [(10_000, 8), (50_000, 300), (120_000, 5_000), (300_000, 60_000), (80_000, 250_000),
              (1_000_000, 900_000), (20_000, 1)]
    for i, (n, card) in enumerate(shapes):
        cols = dg.c4_batch(n, card, dg.seed(4, 100 + i))
        theta = [37.5, 90.0, 12.0, 50.0, 99.9, 63.0, 100.0][i]
        out.append((cols, "key", C4_AGGS, col("score") < theta))
    # int64 keys with an offset range, NULL keys and nullable inputs
    rng = np.random.default_rng(3)
    n = 60_000
    k = rng.integers(10**12, 10**12 + 40_000, n)
    kv = rng.random(n) > 0.05
    x = rng.normal(size=n)
    xv = rng.random(n) > 0.2
    y = rng.integers(-10**9, 10**9, n)
    out.append(({"k": NullableColumn(k, kv), "x": NullableColumn(x, xv), "y": y}, "k",
                [AggExpr(COUNT), AggExpr(COUNT, "x"), AggExpr(SUM, "x"), AggExpr(AVG, "x"),
                 AggExpr(MIN, "y"), AggExpr(MAX, "y"), AggExpr(AVG, "y")], col("y") > -5 * 10**8))
    # a skewed key: most rows in one partition, so pass 1 overflows a cell first
    n = 400_000
    ks = np.where(rng.random(n) < 0.9, 7, rng.integers(0, 200_000, n)).astype(np.int32)
    out.append(({"key": ks, "val": rng.integers(-10**6, 10**6, n), "score": rng.random(n) * 100}, "key",
                C4_AGGS, None))
    # Utf8 key: not a batch-kernel shape, runs as a plain call inside the batch
    codes = rng.integers(0, 5, 30_000).astype(np.uint32)
    out.append(({"g": DictColumn(codes, ["a", "b", "", "c", "d"]), "val": rng.integers(0, 100, 30_000)}, "g",
                [AggExpr(COUNT), AggExpr(SUM, "val")], col("val") > 10))
    return [(_table(ctx, cols), key, aggs, filt, cols) for cols, key, aggs, filt in out]


def _null_safe_key(rb):
    """NULL keys sort as INT64_MIN (their stored value is arbitrary)."""
    if rb.validity[0] is not None and rb.dtypes[0] != "Utf8":
        rb.columns[0] = np.where(rb.validity[0], rb.columns[0], np.iinfo(np.int64).min)
    return rb


def test_batch_matches_oracle_and_plain_calls(ctx):
    qs = _mixed_batch(ctx)
    batch = fa.BatchGroupBy(ctx, [(t, k, a, f) for t, k, a, f, _ in qs])
    assert batch.fallbacks() == 1, batch.info()
    for rep in range(3):  # graph replays, with eager plain calls in between
        res = batch.run()
        for (t, key, aggs, filt, cols), got in zip(qs, res):
            plain = fa.GroupByExec([key], aggs).execute(ctx, t, filt)
            o = ora.groupby(cols, filt, key, aggs)
            exp = o.exact
            for rb in (got, plain, exp):
                _null_safe_key(rb)
            assert_batches_equal(got, exp, key)
            assert_batches_equal(got, plain, key)
    batch.free()


def test_batch_autotune_keeps_results(ctx):
    qs = []
    for i, (n, card) in enumerate([(200_000, 1_000), (500_000, 400_000), (30_000, 20), (700_000, 70_000)]):
        cols = dg.c4_batch(n, card, dg.seed(4, 200 + i))
        qs.append((_table(ctx, cols), "key", C4_AGGS, col("score") < 25.0 + 20 * i, cols))
    batch = fa.BatchGroupBy(ctx, [(t, k, a, f) for t, k, a, f, _ in qs])
    before = batch.run()
    info = batch.tune(6)
    assert "tuned=" in info, info
    after = batch.run()
    for (t, key, aggs, filt, cols), b0, b1 in zip(qs, before, after):
        assert_batches_equal(b1, b0, key)
        assert_batches_equal(b1, ora.groupby(cols, filt, key, aggs).exact, key)
    batch.free()


def test_batch_of_c4_shapes_small(ctx):
    """The C4 recipe at a hundredth of its row counts: 64 queries."""
    qdesc = dg.c4_queries(64, min_rows=1_000, max_rows=100_000)
    qs = []
    for q in qdesc:
        cols = dg.c4_batch(q["rows"], q["card"], q["seed"])
        qs.append((_table(ctx, cols), "key", C4_AGGS, col("score") < q["theta"], cols))
    batch = fa.BatchGroupBy(ctx, [(t, k, a, f) for t, k, a, f, _ in qs])
    assert batch.fallbacks() == 0
    res = batch.run()
    for (t, key, aggs, filt, cols), got in zip(qs, res):
        assert_batches_equal(got, ora.groupby(cols, filt, key, aggs).exact, key)
    batch.free()


def test_batch_packed_records_match_unpacked(ctx):
    """BQ_PART records with the key slot and an integer input 0 in one word
    (ranges from the upload statistics fit 64 bits), against unpacked records
    and the oracle: Int64 and Int32 inputs, a negative input minimum, and a
    query whose ranges are too wide to pack."""
    rng = np.random.default_rng(41)
    n = 400_000
    shapes = [
        dg.c4_batch(n, 300_000, dg.seed(4, 300)),
        {"key": rng.integers(-50_000, 150_000, n).astype(np.int32),
         "v32": rng.integers(-(1 << 31), (1 << 31) - 1, n).astype(np.int32), "score": rng.random(n) * 100},
        # 18 key bits + 61 value bits: left unpacked
        {"key": rng.integers(0, 250_000, n), "val": rng.integers(-(1 << 60), 1 << 60, n),
         "score": rng.random(n)},
    ]
    aggs = [C4_AGGS, [AggExpr(COUNT), AggExpr(SUM, "v32"), AggExpr(MIN, "v32"), AggExpr(MAX, "v32"),
                      AggExpr(AVG, "score")], [AggExpr(COUNT), AggExpr(MIN, "val"), AggExpr(MAX, "val")]]
    filters = [col("score") < 40.0, col("score") > 10.0, None]
    tabs = [_table(ctx, c) for c in shapes]
    results = {}
    try:
        for pack in (1, 0):
            ctx.set_tuning("gbp_pack", pack)
            batch = fa.BatchGroupBy(ctx, [(t, "key", a, f) for t, a, f in zip(tabs, aggs, filters)])
            assert f"packed={2 if pack else 0}" in batch.info(), batch.info()
            results[pack] = batch.run()
            batch.free()
    finally:
        ctx.set_tuning("gbp_pack", 1)
    for cols, a, f, p1, p0 in zip(shapes, aggs, filters, results[1], results[0]):
        assert_batches_equal(p1, p0, "key")
        assert_batches_equal(p1, ora.groupby(cols, f, "key", a).exact, "key")


def test_batch_six_nullable_inputs(ctx):
    """The widest BQ_PART record: key slot, six inputs (some with NULLs) and
    the validity word."""
    rng = np.random.default_rng(43)
    n = 300_000
    cols = {"key": rng.integers(0, 120_000, n)}
    aggs = [AggExpr(COUNT)]
    for j in range(6):
        vals = rng.integers(-10**9, 10**9, n) if j % 2 == 0 else rng.normal(size=n) * 1e3
        cols[f"c{j}"] = NullableColumn(vals, rng.random(n) > 0.15 * (j % 3)) if j % 3 else vals
        aggs.append(AggExpr([SUM, MIN, MAX, COUNT, AVG, SUM][j], f"c{j}"))
    t = _table(ctx, cols)
    batch = fa.BatchGroupBy(ctx, [(t, "key", aggs, col("c0") > -5 * 10**8)])
    got = batch.run()[0]
    assert_batches_equal(got, ora.groupby(cols, col("c0") > -5 * 10**8, "key", aggs).exact, "key")
    batch.free()


def test_batch_specialised_tile_shapes(ctx):
    """The specialised pass-1 tile (NULL-free key, inputs and predicate column,
    one comparison or none, records of one or two words) against the oracle:
    Int32 and Int64 keys, Int32 / Int64 / Float64 inputs, predicates of each
    comparison class and none; a nullable input keeps a query on the generic
    tile."""
    rng = np.random.default_rng(47)
    n = 250_000
    k32 = rng.integers(-70_000, 90_000, n).astype(np.int32)
    k64 = rng.integers(10**12, 10**12 + 150_000, n)
    i32 = rng.integers(-(1 << 31), (1 << 31) - 1, n).astype(np.int32)
    i64 = rng.integers(-10**6, 10**6, n)
    f64 = rng.normal(size=n) * 50
    shapes = [
        ({"key": k32, "v": i64, "x": f64}, [AggExpr(COUNT), AggExpr(SUM, "v"), AggExpr(MAX, "x")], col("x") < 10.0),
        ({"key": k64, "v": i32, "w": i32}, [AggExpr(SUM, "v"), AggExpr(MIN, "v"), AggExpr(MAX, "v")], col("w") >= 0),
        ({"key": k32, "x": f64}, [AggExpr(COUNT), AggExpr(SUM, "x"), AggExpr(MIN, "x")], col("key") != 5),
        ({"key": k64, "v": i64, "w": i64}, [AggExpr(COUNT), AggExpr(SUM, "v"), AggExpr(AVG, "w")], None),
        ({"key": k32, "v": i64, "w": i32}, [AggExpr(MAX, "v"), AggExpr(SUM, "w")], col("v") < 5.5e5),
        ({"key": k32, "v": NullableColumn(i64, rng.random(n) > 0.1)}, [AggExpr(SUM, "v")], col("key") > 0),
    ]
    tabs = [_table(ctx, c) for c, _, _ in shapes]
    batch = fa.BatchGroupBy(ctx, [(t, "key", a, f) for t, (_, a, f) in zip(tabs, shapes)])
    assert "fast=5" in batch.info(), batch.info()
    for rep in range(2):
        for (cols, aggs, f), got in zip(shapes, batch.run()):
            assert_batches_equal(got, ora.groupby(cols, f, "key", aggs).exact, "key")
    batch.free()


def test_batch_freed_when_its_context_closes():
    # closing a context frees the batches prepared on it (their stream and
    # memory belong to the context); a later free is a no-op
    c = fa.Context(0)
    cols = dg.c4_batch(20_000, 3_000, dg.seed(4, 400))
    t = _table(c, cols)
    b = fa.BatchGroupBy(c, [(t, "key", C4_AGGS, col("score") < 50.0)])
    assert_batches_equal(b.run()[0], ora.groupby(cols, col("score") < 50.0, "key", C4_AGGS).exact, "key")
    t.free()
    c.close()
    assert not b.handle
    b.free()


@pytest.mark.parametrize("unit_rows", [4096, 1 << 16])
def test_batch_many_units_per_query(ctx, unit_rows):
    """Queries cut into more than 64 pass-1 units (the merge then finds a
    record's cell by binary search over the units' counts) and into fewer."""
    rng = np.random.default_rng(53)
    qs = []
    for n, card in ((600_000, 200_000), (300_000, 5_000), (90_000, 40_000)):
        cols = {"key": rng.integers(0, card, n).astype(np.int32), "val": rng.integers(-10**6, 10**6, n),
                "score": rng.random(n) * 100}
        qs.append((_table(ctx, cols), cols))
    try:
        ctx.set_tuning("batch_unit_rows", unit_rows)
        batch = fa.BatchGroupBy(ctx, [(t, "key", C4_AGGS, col("score") < 70.0) for t, _ in qs])
    finally:
        ctx.set_tuning("batch_unit_rows", 0)
    assert f"unit_rows={unit_rows}" in batch.info(), batch.info()
    for (t, cols), got in zip(qs, batch.run()):
        assert_batches_equal(got, ora.groupby(cols, col("score") < 70.0, "key", C4_AGGS).exact, "key")
    batch.free()
