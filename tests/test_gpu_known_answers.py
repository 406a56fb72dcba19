"""The reference's known answers (tests/golden/reference_known_answers.json)
replayed on the GPU path through the C ABI, for every aggregation strategy:
dense LDS / dense HBM / hash / radix-partitioned GROUP BY, scalar aggregates,
and the direct and partitioned join probes."""
import numpy as np
import pytest

import falcondb_amd as fa
from falcondb_amd._abi import AVG, COUNT, SUM
from falcondb_amd.operators import AggExpr

from known_answers import GOLDEN, check_all, check_int_sum

pytestmark = pytest.mark.gpu


class GpuEngine:
    def __init__(self, ctx, jctx=None):
        self.ctx = ctx
        self.jctx = jctx or ctx

    def filter_count(self, cols, pred):
        return fa.count_star(self.ctx, fa.Table(self.ctx, cols), pred)

    def groupby(self, cols, pred, key, aggs, batch_rows=65536):
        t = fa.Table(self.ctx, cols, batch_rows=batch_rows)
        if key is None:
            return fa.execute_agg_only(self.ctx, t, aggs, pred)
        return fa.GroupByExec([key], aggs).execute(self.ctx, t, pred)

    def join(self, pcols, bcols, pk, bk, jt, aggs):
        pt, bt = fa.Table(self.jctx, pcols), fa.Table(self.jctx, bcols)
        return fa.execute_hash_join(self.jctx, pt, bt, pk, bk, jt, aggs)


def test_reference_known_answers_on_gpu(ctx):
    assert check_all(GpuEngine(ctx)) > 100


@pytest.mark.parametrize("gb", ["default", "partitioned", "hash"])
@pytest.mark.parametrize("key_rows", [3, 20_000])
def test_int_sum_known_answers_every_path(gb, key_rows, ctx, ctx_gbp, ctx_gbh, ctx_part, ctx_direct):
    # aggregate.rs:1936-1993: exact below 2^63, correctly rounded past it, never wrapped
    g = {"default": ctx, "partitioned": ctx_gbp, "hash": ctx_gbh}[gb]
    j = ctx_part if gb == "partitioned" else ctx_direct
    assert check_int_sum(GpuEngine(g, j), key_rows=key_rows) > 50


def test_int_sum_batched_matches_plain(ctx):
    # the batch kernels (dense key ranges) on groups whose sums leave int64
    v = np.array(GOLDEN["int_sum"]["overflow"]["data"], np.int64)
    qs, tables = [], []
    for s in (1, -1):
        k = np.concatenate([np.zeros(v.size, np.int64), np.arange(1, 50, dtype=np.int64)])
        w = np.concatenate([s * v, np.arange(1, 50, dtype=np.int64)])
        t = fa.Table(ctx, {"k": k, "w": w})
        tables.append(t)
        qs.append((t, "k", [AggExpr(SUM, "w"), AggExpr(COUNT), AggExpr(AVG, "w")], None))
    b = fa.BatchGroupBy(ctx, qs)
    outs = b.run()
    for (t, key, aggs, _), got in zip(qs, outs):
        exp = fa.GroupByExec([key], aggs).execute(ctx, t).sort_by("k")
        got = got.sort_by("k")
        assert got.dtypes == exp.dtypes and got.dtypes[1] == "Float64"
        for a, e in zip(got.columns, exp.columns):
            assert np.array_equal(a, e)
        assert abs(got.columns[1][0]) > 3.6e19
    b.free()
