"""Parquet loader (read_parquet_projected restated over pyarrow): row groups cut
into batches, per-batch text dictionaries, projection rules and type mapping."""
import numpy as np
import pyarrow as pa
import pyarrow.parquet as pq
import pytest

import oracle as ora
from falcondb_amd.loaders import read_parquet_columns
from falcondb_amd.operators import BatchDictColumn, FalconError, NullableColumn

N = 5500


def write_sample(path, with_f32=True):
    rng = np.random.default_rng(8)
    words = [f"w{i}" for i in range(30)]
    s = [None if rng.random() < 0.05 else words[int(rng.integers(0, 6 + i // 500))] for i in range(N)]
    i64 = rng.integers(-10**12, 10**12, N)
    i64_mask = rng.random(N) < 0.1
    f = rng.normal(size=N)
    f[::97] = np.nan
    f_mask = rng.random(N) < 0.05
    cols = {
        "i8": pa.array(rng.integers(-100, 100, N).astype(np.int8)),
        "u32": pa.array(rng.integers(0, 2**32, N, dtype=np.uint64).astype(np.uint32)),
        "I64": pa.array(i64, mask=i64_mask),
        "f": pa.array(f, mask=f_mask),
        "s": pa.array(s, type=pa.string()),
    }
    if with_f32:
        cols["f32"] = pa.array(rng.random(N).astype(np.float32))
    pq.write_table(pa.table(cols), str(path), row_group_size=1000)
    return cols, s, i64, i64_mask, f, f_mask


def test_columns_batches_and_types(tmp_path):
    path = tmp_path / "t.parquet"
    cols, s, i64, i64_mask, f, f_mask = write_sample(path)
    got = read_parquet_columns(path, ["i64", "F", "s", "U32", "i8", "nosuch"], batch_rows=512)
    assert list(got) == ["i8", "u32", "I64", "f", "s"]  # file order, case-insensitive
    assert got["i8"].dtype == np.int32 and np.array_equal(got["i8"], cols["i8"].to_numpy())
    u = cols["u32"].to_numpy()
    assert got["u32"].dtype == np.int32 and np.array_equal(got["u32"], u.view(np.int32))  # `as i32`
    assert isinstance(got["I64"], NullableColumn)
    assert np.array_equal(got["I64"].valid, ~i64_mask)
    assert np.array_equal(got["I64"].values[~i64_mask], i64[~i64_mask])
    fv = got["f"]
    assert np.array_equal(fv.valid, ~f_mask)
    a, b = fv.values[~f_mask], f[~f_mask]
    assert np.array_equal(np.isnan(a), np.isnan(b)) and np.array_equal(a[~np.isnan(a)], b[~np.isnan(b)])
    sc = got["s"]
    assert isinstance(sc, BatchDictColumn)
    assert list(sc.batch_rows) == [512, 488] * 5 + [500]  # row groups of 1000 (the last 500) cut at 512
    for d in sc.dictionaries:  # each batch's own dictionary, first-appearance order
        assert len(d) == len(set(d))
    m = ora.concat_dicts(sc.codes, sc.batch_rows, sc.dictionaries)
    assert [None if c == 0xFFFFFFFF else m.dictionary[c] for c in m.codes.tolist()] == s


def test_unsupported_type_and_fallback_projection(tmp_path):
    path = tmp_path / "t.parquet"
    write_sample(path)
    with pytest.raises(FalconError) as e:
        read_parquet_columns(path)  # f32 is not a device type
    assert e.value.code == 102
    with pytest.raises(FalconError):
        read_parquet_columns(path, ["nosuch"])  # nothing matched: every column, f32 included
    path2 = tmp_path / "u.parquet"
    write_sample(path2, with_f32=False)
    assert list(read_parquet_columns(path2, ["nosuch"])) == ["i8", "u32", "I64", "f", "s"]
