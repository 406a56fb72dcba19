"""Full-size parity, C2: SELECT key, COUNT(*), SUM(v1), MIN(v1), MAX(v1), SUM(v2)
FROM t WHERE f < 40 GROUP BY key over 100M rows (1M Zipf(1.1) keys), every
output column against the oracle: keys, counts and integer aggregates bit for
bit, float sums exactly rounded on both sides (so equal bit for bit) and within
1e-12 relative of the reference-order sum."""
import numpy as np
import pytest

import falcondb_amd as fa
import oracle as ora
from falcondb_amd import datagen as dg
from falcondb_amd._abi import AVG, COUNT, MAX, MIN, SUM
from falcondb_amd.operators import AggExpr, col

from test_gpu_parity import assert_batches_equal

pytestmark = pytest.mark.gpu

ROWS, KEYS = 100_000_000, 1_000_000
AGGS = [AggExpr(COUNT), AggExpr(SUM, "v1"), AggExpr(MIN, "v1"), AggExpr(MAX, "v1"), AggExpr(SUM, "v2")]


@pytest.fixture(scope="module")
def c2():
    print(f"[fullsize] generating C2: {ROWS} rows, {KEYS} keys", flush=True)
    K = dg.keys(KEYS, 40)
    cols = dg.c2_table(ROWS, K)
    print("[fullsize] oracle C2", flush=True)
    o = ora.groupby(cols, col("f") < 40, "key", AGGS)
    return cols, o


def test_c2_full_size_every_column(ctx, c2):
    cols, o = c2
    t = fa.Table(ctx, cols)
    g = fa.GroupByExec(["key"], AGGS).execute(ctx, t, col("f") < 40)
    assert ctx.last_plan().startswith("groupby:partitioned"), ctx.last_plan()
    assert g.num_rows == o.exact.num_rows > 900_000
    assert_batches_equal(g, o.exact, "key", ref=o.reference)
    t.free()


def test_c2_full_size_avg_and_no_predicate(ctx, c2):
    cols, _ = c2
    aggs = [AggExpr(COUNT), AggExpr(AVG, "v1"), AggExpr(AVG, "v2")]
    o = ora.groupby(cols, None, "key", aggs)
    t = fa.Table(ctx, cols)
    g = fa.GroupByExec(["key"], aggs).execute(ctx, t)
    assert_batches_equal(g, o.exact, "key")
    t.free()
