"""A Parquet file loaded onto the device (per-batch dictionaries merged on
upload) answers queries exactly as the same data uploaded with one dictionary."""
import numpy as np
import pytest

import falcondb_amd as fa
import oracle as ora
from falcondb_amd._abi import COUNT, MAX, SUM
from falcondb_amd.loaders import read_parquet, read_parquet_columns
from falcondb_amd.operators import AggExpr, DictColumn, col

from test_gpu_parity import assert_batches_equal
from test_loaders import write_sample

pytestmark = pytest.mark.gpu


def test_parquet_table_equals_single_dictionary_upload(ctx, tmp_path):
    path = tmp_path / "t.parquet"
    write_sample(path, with_f32=False)
    t = read_parquet(ctx, path, batch_rows=512)
    host = read_parquet_columns(path, batch_rows=512)
    sc = host["s"]
    m = ora.concat_dicts(sc.codes, sc.batch_rows, sc.dictionaries)
    ref_cols = dict(host)
    ref_cols["s"] = DictColumn(m.codes, m.dictionary)
    r = fa.Table(ctx, ref_cols, batch_rows=512)
    aggs = [AggExpr(COUNT), AggExpr(SUM, "I64"), AggExpr(MAX, "f")]
    for p in (None, col("i8") > 0, col("I64").is_not_null() & (col("s") != "w3")):
        g = fa.GroupByExec(["s"], aggs).execute(ctx, t, p)
        e = fa.GroupByExec(["s"], aggs).execute(ctx, r, p)
        assert_batches_equal(g, e, "s")
        assert fa.count_star(ctx, t, p) == fa.count_star(ctx, r, p)
    t.free()
    r.free()
