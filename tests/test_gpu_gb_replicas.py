"""State replicas of the dense-LDS and scalar GROUP BY kernels: every
workgroup merges its LDS states into replica (workgroup index mod R) of the
HBM state table and k_gb_fold folds replicas 1..R-1 into replica 0 before
emission.  The result must not depend on R: each replica count (one, a count
that does not divide the grid, the default, more than the fold's usual) gives
the oracle's exact aggregates, for a key of 40 values and for no key."""
from __future__ import annotations

import numpy as np
import pytest

import falcondb_amd as fa
import oracle as ora
from falcondb_amd._abi import AVG, COUNT, MAX, MIN, SUM
from falcondb_amd.operators import AggExpr, NullableColumn, col

from test_gpu_parity import assert_batches_equal

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("replicas", [1, 3, 4, 16])
def test_replicas_match_oracle(ctx, replicas):
    rng = np.random.default_rng(500 + replicas)
    n = 3_000_000  # about 700 workgroups: every replica takes many merges
    cols = {"key": rng.integers(-5, 35, n), "v": NullableColumn(rng.integers(-10**12, 10**12, n), rng.random(n) > 0.2),
            "x": rng.normal(size=n) * 10.0 ** rng.integers(-20, 20, n), "f": rng.integers(0, 100, n).astype(np.int32)}
    aggs = [AggExpr(COUNT), AggExpr(SUM, "v"), AggExpr(MIN, "v"), AggExpr(MAX, "x"), AggExpr(SUM, "x"),
            AggExpr(AVG, "x"), AggExpr(COUNT, "v")]
    t = fa.Table(ctx, cols)
    ctx.set_tuning("gb_replicas", replicas)
    try:
        for p in (None, col("f") < 37):
            g = fa.GroupByExec(["key"], aggs).execute(ctx, t, p)
            assert ctx.last_plan().startswith("groupby:dense_lds")
            assert_batches_equal(g, ora.groupby(cols, p, "key", aggs).exact, "key")
            s = fa.execute_agg_only(ctx, t, aggs, p)
            assert_batches_equal(s, ora.groupby(cols, p, None, aggs).exact, None)
        # a second call on the same context starts from fresh replicas
        g = fa.GroupByExec(["key"], aggs).execute(ctx, t, col("f") < 3)
        assert_batches_equal(g, ora.groupby(cols, col("f") < 3, "key", aggs).exact, "key")
    finally:
        ctx.set_tuning("gb_replicas", 4)
        t.free()
