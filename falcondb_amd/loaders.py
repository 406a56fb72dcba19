"""Parquet into device-resident tables: read_parquet_projected
(src/storage/parquet_io.rs:403-501) and batch_from_arrow (:230-256) as host
code in front of fgpu_table_upload_batches.

Row groups are read in file order and cut into batches of at most `batch_rows`
rows.  A text column keeps one dictionary per batch, in first-appearance order
as Utf8Column::from_strs builds it, and the upload merges them on the device
(concat_all's numbering).  The projection matches column names
case-insensitively; names the file lacks are ignored, and if none matches,
every column is read (:420-437).  Types follow arrow_array_to_falcon
(:150-228) on the device's types:
- Int8 / Int16 / Int32 / UInt8 / UInt16 / UInt32 become Int32 (UInt32 wraps as
  `as i32`);
- Int64 / UInt64 become Int64;
- Float64 stays Float64;
- String / LargeString become Utf8.
Other types fail with FalconError Unsupported.
"""
from __future__ import annotations

from collections.abc import Iterable

import numpy as np

from .operators import BatchDictColumn, Context, FalconError, NullableColumn, Table

UNSUPPORTED = 102


def _arrow():
    try:
        import pyarrow as pa
        import pyarrow.parquet as pq
    except ImportError as e:  # pragma: no cover - pyarrow ships with the image
        raise FalconError(UNSUPPORTED, f"reading Parquet needs pyarrow: {e}") from None
    return pa, pq


def _numeric(arr, np_type):
    """An arrow array as (values, validity or None); NULL slots hold 0."""
    pa, _ = _arrow()
    valid = None
    if arr.null_count:
        valid = np.asarray(arr.is_valid().to_numpy(zero_copy_only=False), dtype=bool)
        arr = arr.fill_null(0)
    vals = np.asarray(arr.to_numpy(zero_copy_only=False))
    if vals.dtype.kind in "iu" and np_type == np.int32:
        vals = vals.astype(np.int64).astype(np.uint32).view(np.int32) if vals.dtype == np.uint32 \
            else vals.astype(np.int32)
    elif np_type == np.int64:
        vals = vals.view(np.int64) if vals.dtype == np.uint64 else vals.astype(np.int64)
    else:
        vals = vals.astype(np.float64)
    return vals, valid


def _strings(arr):
    """An arrow string array as (u32 codes, dictionary): first-appearance
    order, NULL = 0xFFFFFFFF."""
    pa, _ = _arrow()
    enc = arr.dictionary_encode()
    idx = enc.indices
    codes = np.asarray(idx.fill_null(0).to_numpy(zero_copy_only=False)).astype(np.uint32)
    if idx.null_count:
        codes[~np.asarray(idx.is_valid().to_numpy(zero_copy_only=False), dtype=bool)] = 0xFFFFFFFF
    return codes, [str(s) for s in enc.dictionary.to_pylist()]


def _kind(t) -> str:
    pa, _ = _arrow()
    if pa.types.is_int8(t) or pa.types.is_int16(t) or pa.types.is_int32(t) or pa.types.is_uint8(t) \
            or pa.types.is_uint16(t) or pa.types.is_uint32(t):
        return "i32"
    if pa.types.is_int64(t) or pa.types.is_uint64(t):
        return "i64"
    if pa.types.is_float64(t):
        return "f64"
    if pa.types.is_string(t) or pa.types.is_large_string(t):
        return "utf8"
    return ""


def read_parquet_columns(path, columns: Iterable[str] | None = None, batch_rows: int = 65536) -> dict:
    """The file's columns as host columns for falcondb_amd.Table: numpy arrays,
    NullableColumn where NULLs occur, BatchDictColumn for text."""
    pa, pq = _arrow()
    pf = pq.ParquetFile(str(path))
    names = pf.schema_arrow.names
    if columns is not None:
        want = {c.lower() for c in columns}
        pick = [n for n in names if n.lower() in want]
        names = pick or names  # nothing matched: read everything
    kinds = {}
    for n in names:
        k = _kind(pf.schema_arrow.field(n).type)
        if not k:
            raise FalconError(UNSUPPORTED, f"column '{n}': type {pf.schema_arrow.field(n).type} is not on the device path")
        kinds[n] = k
    parts = {n: [] for n in names}
    for g in range(pf.num_row_groups):
        tbl = pf.read_row_group(g, columns=names)
        for b in tbl.to_batches(max_chunksize=batch_rows):
            for n in names:
                arr = b.column(b.schema.get_field_index(n))
                k = kinds[n]
                if k == "utf8":
                    parts[n].append(_strings(arr))
                else:
                    parts[n].append(_numeric(arr, {"i32": np.int32, "i64": np.int64, "f64": np.float64}[k]))
    out = {}
    for n in names:
        if kinds[n] == "utf8":
            codes = [c for c, _ in parts[n]]
            out[n] = BatchDictColumn(np.concatenate(codes) if codes else np.zeros(0, np.uint32),
                                     [c.size for c in codes], [d for _, d in parts[n]])
            continue
        dt = {"i32": np.int32, "i64": np.int64, "f64": np.float64}[kinds[n]]
        vals = [v for v, _ in parts[n]]
        values = np.concatenate(vals) if vals else np.zeros(0, dt)
        if any(v is not None for _, v in parts[n]):
            valid = np.concatenate([np.ones(v.size, bool) if m is None else m for v, m in parts[n]])
            out[n] = NullableColumn(values, valid)
        else:
            out[n] = values
    return out


def read_parquet(ctx: Context, path, columns: Iterable[str] | None = None, batch_rows: int = 65536) -> Table:
    """A Parquet file as a device table (text dictionaries merged on upload)."""
    return Table(ctx, read_parquet_columns(path, columns, batch_rows), batch_rows=batch_rows)
