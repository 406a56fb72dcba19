"""Python mirror of FalconDB's physical-operator interface on the MI355X backend.

The names follow the reference's query layer so code written against it reads
the same here:

  Expr.eval_mask_bits        src/query/expr.rs:635        packed LSB-first mask
  execute_select             src/query/select.rs:105      WHERE + projection
  count_star                 src/db.rs:1439-1476          COUNT(*) ... WHERE
  GroupByExec.execute        src/query/aggregate.rs:535   GROUP BY (fused with WHERE)
  execute_agg_only           src/query/aggregate.rs:1394  aggregates without GROUP BY
  execute_hash_join          src/query/join.rs:184        INNER / LEFT join + aggregates
  execute_hash_semi_anti_join src/query/join.rs:49        EXISTS / NOT EXISTS + aggregates

Every call goes through libfalcon_gpu.so (include/falcon_gpu.h); errors come
back as FalconError with the reference's variant name.  Results are
RecordBatch objects of numpy columns; `to_reference_types()` applies the
reference's output typing (integer SUM/MIN/MAX as Float64,
src/query/aggregate.rs:3390-3420).
"""
from __future__ import annotations

import atexit
import ctypes as C
import weakref
from dataclasses import dataclass, field
from typing import Iterable, Sequence

import numpy as np

from . import _abi
from ._abi import (ANTI, AVG, COUNT, EQ, FLOAT64, GE, GT, INNER, INT32, INT64, LE, LEFT, LT,
                   MAX, MIN, NE, NO_ROW, P_AND, P_CMP, P_IN, P_IS_NOT_NULL, P_IS_NULL, P_LIKE, P_OR, SEMI,
                   SUM, UTF8)


class FalconError(Exception):
    """An error returned through the C ABI.  `kind` is the FalconError variant."""

    def __init__(self, code: int, message: str):
        self.code = code
        self.kind = _abi.ERROR_NAMES.get(code, f"Error{code}")
        super().__init__(f"{self.kind}: {message}")


def _lib():
    return _abi.load()


def _check(rc: int) -> None:
    if rc != _abi.OK:
        raise FalconError(rc, _lib().fgpu_last_error().decode(errors="replace"))


# ─── context ─────────────────────────────────────────────────────────────────
# Handles still open at interpreter exit are released by an atexit hook, while
# the library is still loaded (module teardown may already have dropped it by
# the time __del__ runs): batches first, then contexts.
_LIVE_BATCHES: "weakref.WeakSet" = weakref.WeakSet()
_LIVE_CONTEXTS: "weakref.WeakSet" = weakref.WeakSet()


@atexit.register
def _release_at_exit():
    for b in list(_LIVE_BATCHES):
        b.free()
    for c in list(_LIVE_CONTEXTS):
        c.close()


class Context:
    """A device, a HIP stream and a scratch arena (fgpu_open)."""

    def __init__(self, device: int = 0, timing: bool = False):
        h = C.c_void_p()
        _check(_lib().fgpu_open(device, C.byref(h)))
        _LIVE_CONTEXTS.add(self)
        self._h = h
        self.device = device
        if timing:
            self.set_timing(True)

    @property
    def handle(self):
        return self._h

    def set_timing(self, on: bool) -> None:
        _check(_lib().fgpu_set_timing(self._h, 1 if on else 0))

    def last_timing(self) -> dict:
        t = _abi.Timing()
        _check(_lib().fgpu_last_timing(self._h, C.byref(t)))
        return {"total_ms": t.total_ms, "main_ms": t.main_ms, "build_ms": t.build_ms,
                "finalize_ms": t.finalize_ms, "bytes_main": t.bytes_main}

    def set_tuning(self, name: str, value: int) -> None:
        _check(_lib().fgpu_set_tuning(self._h, name.encode(), int(value)))

    def last_plan(self) -> str:
        """The strategy the last call used, e.g. "groupby:partitioned P=2048 ..."."""
        return _lib().fgpu_last_plan(self._h).decode()

    def scratch_bytes(self) -> int:
        return int(_lib().fgpu_scratch_bytes(self._h))

    def close(self) -> None:
        if self._h:
            # a batch prepared on this context holds its stream and memory
            for b in list(_LIVE_BATCHES):
                if b.ctx is self:
                    b.free()
            _lib().fgpu_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ─── columns and tables ──────────────────────────────────────────────────────
@dataclass
class DictColumn:
    """Dictionary-encoded text column: u32 codes (0xFFFFFFFF = NULL) + strings
    (Utf8Column{dict, codes}, src/storage/column.rs:1272-1277)."""
    codes: np.ndarray
    dictionary: Sequence[str]


@dataclass
class BatchDictColumn:
    """A text column as record batches with one dictionary each (what the CSV /
    Parquet loaders build): batch-local u32 codes of every batch in row order,
    the batches' row counts and their dictionaries.  Upload merges the
    dictionaries like Utf8Column::concat_all (src/storage/column.rs:1539-1595)."""
    codes: np.ndarray
    batch_rows: Sequence[int]
    dictionaries: Sequence[Sequence[str]]


@dataclass
class NullableColumn:
    """A primitive column with a validity mask (True = valid)."""
    values: np.ndarray
    valid: np.ndarray


def pack_validity(valid: np.ndarray) -> np.ndarray:
    """bool per row -> LSB-first u64 words (src/storage/column.rs:132-133)."""
    valid = np.asarray(valid, dtype=bool)
    n = valid.size
    pad = (-n) % 64
    bits = np.packbits(np.concatenate([valid, np.zeros(pad, bool)]), bitorder="little")
    return bits.view(np.uint64).copy() if n else np.zeros(0, np.uint64)


def _np_dtype_code(a: np.ndarray) -> int:
    if a.dtype == np.int32:
        return INT32
    if a.dtype == np.int64:
        return INT64
    if a.dtype == np.float64:
        return FLOAT64
    raise FalconError(2, f"unsupported column dtype {a.dtype}")


class Table:
    """A device-resident table (fgpu_table_upload).  Column names resolve
    case-insensitively (Schema::column_index, src/storage/schema.rs:81-86)."""

    def __init__(self, ctx: Context, columns: dict, batch_rows: int = 65536):
        self.ctx = ctx
        self.names = list(columns.keys())
        self.dtypes: list[int] = []
        self.dicts: list[list[str] | None] = []
        keep = []  # host buffers alive for the duration of the upload
        descs = (_abi.ColumnDesc * max(len(columns), 1))()
        utf8 = (_abi.Utf8Batches * max(len(columns), 1))()
        batched = False
        n_rows = None
        for i, (name, col) in enumerate(columns.items()):
            d = descs[i]
            d.name = name.encode()
            valid_words = None
            if isinstance(col, BatchDictColumn):
                codes = np.ascontiguousarray(col.codes, dtype=np.uint32)
                rows = np.ascontiguousarray(col.batch_rows, dtype=np.int64)
                if len(col.dictionaries) != rows.size:
                    raise FalconError(6, f"column '{name}': {rows.size} batches, {len(col.dictionaries)} dictionaries")
                lens = np.array([len(dd) for dd in col.dictionaries], dtype=np.int32)
                arrs = [(C.c_char_p * max(len(dd), 1))(*[x.encode() for x in dd]) for dd in col.dictionaries]
                parr = (C.POINTER(C.c_char_p) * max(len(arrs), 1))(*[C.cast(a, C.POINTER(C.c_char_p)) for a in arrs])
                keep += [codes, rows, lens, arrs, parr]
                u = utf8[i]
                u.n_batches = rows.size
                u.batch_rows = rows.ctypes.data_as(C.POINTER(C.c_int64))
                u.dict_len = lens.ctypes.data_as(C.POINTER(C.c_int32))
                u.dicts = C.cast(parr, C.POINTER(C.POINTER(C.c_char_p)))
                d.dtype, d.values = UTF8, codes.ctypes.data
                n = codes.size
                batched = True
                self.dicts.append(None)  # merged on upload
            elif isinstance(col, DictColumn):
                codes = np.ascontiguousarray(col.codes, dtype=np.uint32)
                strs = [s.encode() for s in col.dictionary]
                arr = (C.c_char_p * max(len(strs), 1))(*strs)
                keep += [codes, strs, arr]
                d.dtype, d.values = UTF8, codes.ctypes.data
                d.dict_len, d.dict = len(strs), C.cast(arr, C.POINTER(C.c_char_p))
                n = codes.size
                self.dicts.append(list(col.dictionary))
            else:
                if isinstance(col, NullableColumn):
                    vals = np.ascontiguousarray(col.values)
                    valid_words = pack_validity(col.valid)
                    if np.asarray(col.valid).size != vals.size:
                        raise FalconError(6, f"column '{name}': validity length mismatch")
                else:
                    vals = np.ascontiguousarray(col)
                d.dtype = _np_dtype_code(vals)
                d.values = vals.ctypes.data
                keep.append(vals)
                n = vals.size
                self.dicts.append(None)
            if valid_words is not None:
                keep.append(valid_words)
                d.validity = valid_words.ctypes.data
            self.dtypes.append(d.dtype)
            if n_rows is None:
                n_rows = n
            elif n != n_rows:
                raise FalconError(6, f"column '{name}' has {n} rows, expected {n_rows}")
        h = C.c_void_p()
        if batched:
            _check(_lib().fgpu_table_upload_batches(ctx.handle, descs, utf8, len(columns), int(n_rows or 0),
                                                    batch_rows, C.byref(h)))
        else:
            _check(_lib().fgpu_table_upload(ctx.handle, descs, len(columns), int(n_rows or 0),
                                            batch_rows, C.byref(h)))
        self._h = h
        self.n_rows = int(n_rows or 0)
        self.batch_rows = batch_rows

    @classmethod
    def from_handle(cls, ctx: Context, h, batch_rows: int = 65536) -> "Table":
        """Wraps a device table the library built (a materialised join); the
        schema comes from the library."""
        t = cls.__new__(cls)
        t.ctx = ctx
        t._h = h
        lib = _lib()
        t.n_rows = int(lib.fgpu_table_rows(h))
        t.batch_rows = batch_rows
        t.names, t.dtypes, t.dicts = [], [], []
        for i in range(int(lib.fgpu_table_cols(h))):
            v = _abi.ColumnView()
            _check(lib.fgpu_table_column_info(h, i, C.byref(v)))
            t.names.append(v.name.decode())
            t.dtypes.append(int(v.dtype))
            t.dicts.append([v.dict[k].decode() for k in range(v.dict_len)] if v.dtype == UTF8 else None)
        return t

    @property
    def handle(self):
        return self._h

    def column_index(self, name: str) -> int:
        low = name.lower()
        for i, n in enumerate(self.names):  # an exact name first ("right.v" is a column name)
            if n.lower() == low:
                return i
        if "." in low:  # qualified t.col: resolve by the bare name
            low = low.split(".", 1)[1]
        for i, n in enumerate(self.names):
            if n.lower() == low:
                return i
        raise FalconError(1, f"column '{name}' not found")

    def read(self, ctx: Context | None = None) -> "RecordBatch":
        """Every row and column, copied back to the host."""
        ctx = ctx or self.ctx
        idx = list(range(len(self.names)))
        arr = (C.c_int32 * max(len(idx), 1))(*idx)
        h = C.c_void_p()
        _check(_lib().fgpu_filter_select(ctx.handle, self._h, None, arr, len(idx), C.byref(h)))
        return _read_result(h, [None] * len(idx))

    def device_bytes(self) -> int:
        return int(_lib().fgpu_table_device_bytes(self._h))

    def column_stats(self, name: str) -> dict:
        mn, mx, nulls, nd = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        _check(_lib().fgpu_table_column_stats(self._h, self.column_index(name), C.byref(mn),
                                              C.byref(mx), C.byref(nulls), C.byref(nd)))
        return {"min": mn.value, "max": mx.value, "nulls": nulls.value, "distinct": nd.value}

    def free(self) -> None:
        if getattr(self, "_h", None):
            _lib().fgpu_table_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ─── expressions ─────────────────────────────────────────────────────────────
_OPS = {"=": EQ, "!=": NE, "<": LT, "<=": LE, ">": GT, ">=": GE}
_SWAP = {EQ: EQ, NE: NE, LT: GT, LE: GE, GT: LT, GE: LE}


@dataclass
class _Node:
    kind: int
    col: str | None = None
    op: int = 0
    lit: object = None       # comparison literal, LIKE pattern, or the IN list
    lit_type: int = INT64


class Expr:
    """A WHERE predicate as a postfix program of comparisons and AND/OR
    (the subset Expr::try_fast_mask evaluates, src/query/expr.rs:656-808)."""

    def __init__(self, nodes: list[_Node]):
        self.nodes = nodes

    def __and__(self, other: "Expr") -> "Expr":
        return Expr(self.nodes + other.nodes + [_Node(P_AND)])

    def __or__(self, other: "Expr") -> "Expr":
        return Expr(self.nodes + other.nodes + [_Node(P_OR)])

    def _compile(self, table: Table):
        nodes = (_abi.PredNode * len(self.nodes))()
        keep = []
        for i, n in enumerate(self.nodes):
            d = nodes[i]
            d.kind = n.kind
            if n.kind in (P_AND, P_OR):
                continue
            d.col = table.column_index(n.col)
            d.op = n.op
            d.lit_type = n.lit_type
            if n.kind == P_IN:
                vals = list(n.lit)
                if n.lit_type == UTF8:
                    bs = [str(v).encode() for v in vals]
                    arr = (C.c_char_p * max(len(bs), 1))(*bs)
                    keep.append(bs)
                elif n.lit_type == FLOAT64:
                    arr = (C.c_double * max(len(vals), 1))(*[float(v) for v in vals])
                else:
                    arr = (C.c_int64 * max(len(vals), 1))(*[int(v) for v in vals])
                keep.append(arr)
                d.n_list = len(vals)
                d.list = C.cast(arr, C.c_void_p)
            elif n.lit_type == UTF8:
                b = str(n.lit).encode()
                keep.append(b)
                d.lit_str = b
            elif n.lit_type == FLOAT64:
                d.lit_f64 = float(n.lit)
            elif n.kind == P_CMP:
                d.lit_i64 = int(n.lit)
        p = _abi.Pred(len(self.nodes), C.cast(nodes, C.POINTER(_abi.PredNode)))
        return p, (nodes, keep)

    def eval_mask_bits(self, ctx: Context, table: Table) -> np.ndarray:
        """Packed mask of the rows that pass: ceil(n/64) LSB-first u64 words."""
        p, keep = self._compile(table)
        out = np.zeros((table.n_rows + 63) // 64, dtype=np.uint64)
        _check(_lib().fgpu_filter_mask(ctx.handle, table.handle, C.byref(p), out.ctypes.data))
        return out


def _lit_type(v) -> int:
    if isinstance(v, str):
        return UTF8
    if isinstance(v, (float, np.floating)):
        return FLOAT64
    if isinstance(v, np.int32):
        return INT32
    return INT64


class ColumnRef:
    """`col("name")`; comparing it with a literal yields an Expr."""

    def __init__(self, name: str):
        self.name = name

    def _cmp(self, op: int, v) -> Expr:
        return Expr([_Node(P_CMP, self.name, op, v, _lit_type(v))])

    def __gt__(self, v): return self._cmp(GT, v)
    def __ge__(self, v): return self._cmp(GE, v)
    def __lt__(self, v): return self._cmp(LT, v)
    def __le__(self, v): return self._cmp(LE, v)
    def __eq__(self, v): return self._cmp(EQ, v)  # type: ignore[override]
    def __ne__(self, v): return self._cmp(NE, v)  # type: ignore[override]
    __hash__ = object.__hash__

    def is_null(self) -> Expr:
        return Expr([_Node(P_IS_NULL, self.name)])

    def like(self, pattern: str, negated: bool = False, case_insensitive: bool = False) -> Expr:
        """`col [NOT] [I]LIKE pattern` on a Utf8 column: '%' any bytes, '_'
        one byte, no escape (Expr::Like, src/query/expr.rs:697-754)."""
        return Expr([_Node(P_LIKE, self.name, int(negated) | 2 * int(case_insensitive), str(pattern), UTF8)])

    def not_like(self, pattern: str) -> Expr:
        return self.like(pattern, negated=True)

    def ilike(self, pattern: str) -> Expr:
        return self.like(pattern, case_insensitive=True)

    def isin(self, values, negated: bool = False) -> Expr:
        """`col [NOT] IN (values)`: the union of the values' `=` comparisons;
        NOT IN drops the NULL rows (Expr::InList, src/query/expr.rs:756-803).
        Literals of different types become one list node each."""
        groups: dict[int, list] = {}
        for v in values:
            lt = _lit_type(v)
            groups.setdefault(INT64 if lt == INT32 else lt, []).append(v)
        if not groups:
            groups[INT64] = []
        nodes: list[_Node] = []
        for i, (lt, vals) in enumerate(groups.items()):
            nodes.append(_Node(P_IN, self.name, int(negated), vals, lt))
            if i:
                nodes.append(_Node(P_AND if negated else P_OR))
        return Expr(nodes)

    def not_in(self, values) -> Expr:
        return self.isin(values, negated=True)

    def is_not_null(self) -> Expr:
        return Expr([_Node(P_IS_NOT_NULL, self.name)])


def col(name: str) -> ColumnRef:
    return ColumnRef(name)


def cmp(column: str, op: str, literal) -> Expr:
    """`cmp("score", ">", 50)` == `col("score") > 50`."""
    return col(column)._cmp(_OPS[op], literal)


def lit_cmp(literal, op: str, column: str) -> Expr:
    """`literal op column`, swapped like the reference (src/query/expr.rs:679-684)."""
    return col(column)._cmp(_SWAP[_OPS[op]], literal)


# ─── aggregates ──────────────────────────────────────────────────────────────
class AggFunc:
    Count = COUNT
    Sum = SUM
    Avg = AVG
    Min = MIN
    Max = MAX


@dataclass
class AggExpr:
    """One aggregate.  input=None is COUNT(*); `side` picks the join input
    ("probe"/"left" or "build"/"right")."""
    func: int
    input: str | None = None
    alias: str | None = None
    side: str = "probe"


def _aggs_struct(aggs: Sequence[AggExpr], resolve):
    arr = (_abi.Agg * max(len(aggs), 1))()
    keep = []
    for i, a in enumerate(aggs):
        arr[i].func = a.func
        side = 1 if a.side in ("build", "right") else 0
        arr[i].side = side
        arr[i].col = -1 if a.input is None else resolve(side, a.input)
        if a.alias:
            b = a.alias.encode()
            keep.append(b)
            arr[i].alias = b
    return arr, keep


# ─── results ─────────────────────────────────────────────────────────────────
_DT_NAME = {INT64: "Int64", FLOAT64: "Float64", UTF8: "Utf8", INT32: "Int32"}


@dataclass
class RecordBatch:
    """Schema + columns (RecordBatch, src/storage/table.rs:45-49): `columns[i]`
    is a numpy array; `validity[i]` a bool array or None; Utf8 columns hold
    python strings (NULL = None)."""
    names: list[str]
    dtypes: list[str]
    columns: list[np.ndarray]
    validity: list[np.ndarray | None]
    agg_funcs: list[int | None] = field(default_factory=list)

    @property
    def num_rows(self) -> int:
        return int(self.columns[0].size) if self.columns else 0

    def column(self, name: str) -> np.ndarray:
        return self.columns[self.names.index(name)]

    def to_reference_types(self) -> "RecordBatch":
        """Integer SUM / MIN / MAX as Float64, like the reference's build_agg_col."""
        cols, dts = [], []
        for c, dt, f in zip(self.columns, self.dtypes, self.agg_funcs or [None] * len(self.columns)):
            if f in (SUM, MIN, MAX) and dt == "Int64":
                cols.append(c.astype(np.float64))
                dts.append("Float64")
            else:
                cols.append(c)
                dts.append(dt)
        return RecordBatch(list(self.names), dts, cols, list(self.validity), list(self.agg_funcs))

    def sort_by(self, name: str) -> "RecordBatch":
        key = self.column(name)
        order = np.argsort(np.array(key, dtype=object) if key.dtype == object else key, kind="stable")
        return RecordBatch(list(self.names), list(self.dtypes), [c[order] for c in self.columns],
                           [None if v is None else v[order] for v in self.validity],
                           list(self.agg_funcs))


def _read_result(h, agg_funcs: list[int | None]) -> RecordBatch:
    lib = _lib()
    try:
        n = int(lib.fgpu_result_rows(h))
        nc = int(lib.fgpu_result_cols(h))
        names, dts, cols, valid = [], [], [], []
        for i in range(nc):
            v = _abi.ColumnView()
            _check(lib.fgpu_result_column(h, i, C.byref(v)))
            names.append(v.name.decode())
            if v.dtype == UTF8:
                codes = np.ctypeslib.as_array(C.cast(v.values, C.POINTER(C.c_uint32)), (n,)).copy() if n else np.zeros(0, np.uint32)
                d = [v.dict[k].decode() for k in range(v.dict_len)]
                strs = np.array([d[c] if c < len(d) else ("" if c == len(d) else None) for c in codes.tolist()], dtype=object)
                cols.append(strs)
                dts.append("Utf8")
            else:
                ct = C.c_double if v.dtype == FLOAT64 else C.c_int64
                arr = np.ctypeslib.as_array(C.cast(v.values, C.POINTER(ct)), (n,)).copy() if n else np.zeros(0, np.float64 if v.dtype == FLOAT64 else np.int64)
                cols.append(arr)
                dts.append(_DT_NAME[v.dtype])
            if v.valid and n:
                valid.append(np.ctypeslib.as_array(v.valid, (n,)).astype(bool))
            else:
                valid.append(None)
        return RecordBatch(names, dts, cols, valid, agg_funcs)
    finally:
        lib.fgpu_result_free(h)


# ─── operators ───────────────────────────────────────────────────────────────
def _pred(filter: Expr | None, table: Table):
    if filter is None:
        return None, None
    p, keep = filter._compile(table)
    return C.byref(p), (p, keep)


def count_star(ctx: Context, table: Table, filter: Expr | None = None) -> int:
    """SELECT COUNT(*) FROM t WHERE filter (src/db.rs:1439-1476)."""
    p, keep = _pred(filter, table)
    out = C.c_int64()
    _check(_lib().fgpu_filter_count(ctx.handle, table.handle, p, C.byref(out)))
    return out.value


def execute_select(ctx: Context, table: Table, filter: Expr | None,
                   columns: Iterable[str]) -> RecordBatch:
    """Rows passing `filter`, columns `columns`, in row order
    (execute_select + filter_by_bitmask_cols)."""
    p, keep = _pred(filter, table)
    idx = [table.column_index(c) for c in columns]
    arr = (C.c_int32 * max(len(idx), 1))(*idx)
    h = C.c_void_p()
    _check(_lib().fgpu_filter_select(ctx.handle, table.handle, p, arr, len(idx), C.byref(h)))
    return _read_result(h, [None] * len(idx))


class GroupByExec:
    """GROUP BY with COUNT/SUM/AVG/MIN/MAX aggregates (GroupByExec,
    src/query/aggregate.rs:535).  One Int32 / Int64 / Utf8 key takes the
    reference's fast path (fgpu_groupby); several keys, or one Float64 key,
    its slow path (execute_slow, :1232-1346; fgpu_groupby_keys), which groups
    on the keys' display strings.  The WHERE filter is fused into the same
    device pass instead of materialising filtered batches."""

    def __init__(self, group_by: Sequence[str | ColumnRef], aggs: Sequence[AggExpr]):
        if not 1 <= len(group_by) <= _abi.MAX_GROUP_KEYS:
            raise FalconError(6, f"GROUP BY takes 1 to {_abi.MAX_GROUP_KEYS} key columns on this backend")
        self.keys = [g.name if isinstance(g, ColumnRef) else g for g in group_by]
        self.key = self.keys[0]
        self.aggs = list(aggs)

    def execute(self, ctx: Context, table: Table, filter: Expr | None = None) -> RecordBatch:
        if len(self.keys) == 1 and table.dtypes[table.column_index(self.key)] != _abi.FLOAT64:
            return _groupby(ctx, table, filter, table.column_index(self.key), self.aggs)
        return _groupby_keys(ctx, table, filter, [table.column_index(k) for k in self.keys], self.aggs)


class BatchGroupBy:
    """Many GROUP BY queries run together (fgpu_batch_*): prepared once, then
    each run() is one replay of a captured graph of two persistent kernels.
    Query i is (table, key column, aggregates, WHERE); its result is the one
    GroupByExec gives (src/query/aggregate.rs:535).  The prepare / run split
    follows the reference's streaming operator shape (pipeline::Operator
    process / flush, src/query/pipeline.rs:46-72)."""

    def __init__(self, ctx: Context, queries: Sequence[tuple]):
        self.ctx = ctx
        self._keep = []
        self._funcs = []
        qs = (_abi.BatchQuery * max(1, len(queries)))()
        for i, (table, key, aggs, filt) in enumerate(queries):
            p, keep = _pred(filt, table)
            arr, keep2 = _aggs_struct(list(aggs), lambda side, name, t=table: t.column_index(name))
            spec = _abi.GroupBySpec(table.column_index(key), len(aggs), C.cast(arr, C.POINTER(_abi.Agg)))
            self._keep += [p, keep, arr, keep2, spec]
            qs[i].table = table.handle
            qs[i].pred = C.addressof(keep[0]) if keep else None
            qs[i].spec = C.cast(C.pointer(spec), C.c_void_p)
            self._funcs.append([None] + [a.func for a in aggs])
        self.n = len(queries)
        self.handle = C.c_void_p()
        _check(_lib().fgpu_batch_prepare(ctx.handle, qs, self.n, C.byref(self.handle)))
        _LIVE_BATCHES.add(self)

    def tune(self, max_trials: int = 0) -> str:
        _check(_lib().fgpu_batch_tune(self.handle, max_trials))
        return self.info()

    def info(self) -> str:
        return _lib().fgpu_batch_info(self.handle).decode()

    def fallbacks(self) -> int:
        return _lib().fgpu_batch_fallbacks(self.handle)

    def run_raw(self) -> list:
        """Runs the batch and returns the raw result handles (caller frees)."""
        outs = (C.c_void_p * max(1, self.n))()
        _check(_lib().fgpu_batch_run(self.handle, outs))
        return [outs[i] for i in range(self.n)]

    def run(self) -> list[RecordBatch]:
        return [_read_result(C.c_void_p(h), f) for h, f in zip(self.run_raw(), self._funcs)]

    def free(self):
        if self.handle:
            _lib().fgpu_batch_free(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def execute_sort_limit(ctx: Context, table: Table, column: str, limit: int, desc: bool = False,
                       filter: Expr | None = None) -> tuple[np.ndarray, np.ndarray]:
    """ORDER BY column [DESC] LIMIT limit over the rows that pass `filter`
    (execute_sort_limit, src/query/sort.rs:59-115).  Returns (row indices,
    values of the column) in output order; values of NULL rows are NaN for
    Float64 and 0 for integers, and their row indices say which they are."""
    p, keep = _pred(filter, table)
    rows = np.zeros(max(limit, 1), np.int64)
    vals = np.zeros(max(limit, 1), np.int64)
    n = C.c_int64()
    ci = table.column_index(column)
    _check(_lib().fgpu_topk(ctx.handle, table.handle, p, ci, int(desc), limit, rows.ctypes.data,
                            vals.ctypes.data, C.byref(n)))
    rows, vals = rows[: n.value], vals[: n.value]
    if table.dtypes[ci] == FLOAT64:
        vals = vals.view(np.float64).copy()
    return rows, vals


def execute_sort(ctx: Context, table: Table, keys: Sequence, limit: int | None = None,
                 filter: Expr | None = None, columns: Sequence[str] | None = None
                 ) -> tuple[np.ndarray, RecordBatch]:
    """ORDER BY keys [LIMIT limit] for any limit and number of keys
    (execute_sort_limit, src/query/sort.rs:59-115).  keys are (column, desc)
    or (column, desc, nulls_last); nulls_last defaults to desc (SortExpr::desc
    / ::asc).  Returns the row numbers in output order and the rows' `columns`
    (default: the sort columns)."""
    p, keep = _pred(filter, table)
    ks = (_abi.SortKey * max(len(keys), 1))()
    for i, k in enumerate(keys):
        ks[i].col = table.column_index(k[0])
        ks[i].desc = int(k[1])
        ks[i].nulls_last = int(k[2]) if len(k) > 2 else int(k[1])
    if columns is None:
        columns = list(dict.fromkeys(k[0] for k in keys))
    ci = [table.column_index(c) for c in columns]
    arr = (C.c_int32 * max(len(ci), 1))(*ci)
    lim = -1 if limit is None else int(limit)
    cap = table.n_rows if lim < 0 else min(lim, table.n_rows)
    rows = np.zeros(max(cap, 1), np.int64)
    h = C.c_void_p()
    _check(_lib().fgpu_sort_limit(ctx.handle, table.handle, p, ks, len(keys), lim, arr, len(ci),
                                  rows.ctypes.data, C.byref(h)))
    m = int(_lib().fgpu_result_rows(h))
    return rows[:m], _read_result(h, [None] * len(ci))


def execute_agg_only(ctx: Context, table: Table, aggs: Sequence[AggExpr],
                     filter: Expr | None = None) -> RecordBatch:
    """Aggregates without GROUP BY: one row (src/query/aggregate.rs:1394)."""
    return _groupby(ctx, table, filter, -1, list(aggs))


def _groupby(ctx, table, filter, key_col, aggs) -> RecordBatch:
    p, keep = _pred(filter, table)
    arr, keep2 = _aggs_struct(aggs, lambda side, name: table.column_index(name))
    spec = _abi.GroupBySpec(key_col, len(aggs), C.cast(arr, C.POINTER(_abi.Agg)))
    h = C.c_void_p()
    _check(_lib().fgpu_groupby(ctx.handle, table.handle, p, C.byref(spec), C.byref(h)))
    funcs = ([None] if key_col >= 0 else []) + [a.func for a in aggs]
    return _read_result(h, funcs)


def _groupby_keys(ctx, table, filter, key_cols, aggs) -> RecordBatch:
    p, keep = _pred(filter, table)
    arr, keep2 = _aggs_struct(aggs, lambda side, name: table.column_index(name))
    spec = _abi.GroupBySpec(-1, len(aggs), C.cast(arr, C.POINTER(_abi.Agg)))
    kc = (C.c_int32 * len(key_cols))(*key_cols)
    h = C.c_void_p()
    _check(_lib().fgpu_groupby_keys(ctx.handle, table.handle, p, kc, len(key_cols), C.byref(spec), C.byref(h)))
    return _read_result(h, [None] * len(key_cols) + [a.func for a in aggs])


class JoinIndex:
    """Device hash index over a build table's key (fgpu_join_build); reusable."""

    def __init__(self, ctx: Context, build: Table, key: str):
        self.build = build
        self.key_col = build.column_index(key)
        h = C.c_void_p()
        _check(_lib().fgpu_join_build(ctx.handle, build.handle, self.key_col, C.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def device_bytes(self) -> int:
        return int(_lib().fgpu_join_index_device_bytes(self._h))

    def free(self):
        if getattr(self, "_h", None):
            _lib().fgpu_join_index_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class JoinType:
    Inner = INNER
    Left = LEFT
    Semi = SEMI
    Anti = ANTI


def _join_spec(probe, build, probe_key, build_key, join_type, aggs):
    def resolve(side, name):
        return (build if side == 1 else probe).column_index(name)
    arr, keep = _aggs_struct(aggs, resolve)
    spec = _abi.JoinSpec(join_type, probe.column_index(probe_key), build.column_index(build_key),
                         len(aggs), C.cast(arr, C.POINTER(_abi.Agg)))
    return spec, (arr, keep)


def execute_hash_join(ctx: Context, probe: Table, build: Table, probe_key: str, build_key: str,
                      join_type: int, aggs: Sequence[AggExpr], filter: Expr | None = None,
                      index: JoinIndex | None = None) -> RecordBatch:
    """SELECT aggs FROM probe [LEFT] JOIN build ON probe.k = build.k [WHERE filter on
    probe] -- the join fused with the aggregate that consumes it
    (execute_hash_join_projected + execute_agg_only).  SEMI/ANTI also accepted."""
    spec, keep = _join_spec(probe, build, probe_key, build_key, join_type, aggs)
    p, keep2 = _pred(filter, probe)
    h = C.c_void_p()
    _check(_lib().fgpu_join_agg(ctx.handle, probe.handle, p, build.handle,
                                index.handle if index else None, C.byref(spec), C.byref(h)))
    return _read_result(h, [a.func for a in aggs])


def execute_hash_join_projected(ctx: Context, probe: Table, build: Table, probe_key: str, build_key: str,
                                join_type: int, probe_columns: Sequence[str] | None = None,
                                build_columns: Sequence[str] | None = None, filter: Expr | None = None,
                                index: JoinIndex | None = None) -> Table:
    """The join's output as a new device table: the kept probe columns, then
    the kept build columns (a build column whose name the probe side has is
    named right.<name>; NULL where a LEFT row found no match), rows in probe
    order (execute_hash_join_projected, src/query/join.rs:184-282).  None keeps
    every column of that side; SEMI / ANTI keep probe columns only."""
    spec, keep = _join_spec(probe, build, probe_key, build_key, join_type, [])
    p, keep2 = _pred(filter, probe)
    if probe_columns is None:
        probe_columns = list(probe.names)
    if build_columns is None:
        build_columns = [] if join_type in (SEMI, ANTI) else list(build.names)
    pi = [probe.column_index(c) for c in probe_columns]
    bi = [build.column_index(c) for c in build_columns]
    pa = (C.c_int32 * max(len(pi), 1))(*pi)
    ba = (C.c_int32 * max(len(bi), 1))(*bi)
    h = C.c_void_p()
    _check(_lib().fgpu_join_materialize(ctx.handle, probe.handle, p, build.handle, index.handle if index else None,
                                        C.byref(spec), pa, len(pi), ba, len(bi), C.byref(h)))
    return Table.from_handle(ctx, h, probe.batch_rows)


def execute_hash_semi_anti_join(ctx: Context, probe: Table, build: Table, probe_key: str,
                                build_key: str, anti: bool, aggs: Sequence[AggExpr],
                                filter: Expr | None = None,
                                index: JoinIndex | None = None) -> RecordBatch:
    """WHERE [NOT] EXISTS (SELECT 1 FROM build WHERE build.k = probe.k), aggregated
    (execute_hash_semi_anti_join, src/query/join.rs:49-132)."""
    return execute_hash_join(ctx, probe, build, probe_key, build_key, ANTI if anti else SEMI,
                             aggs, filter, index)


def join_pairs(ctx: Context, probe: Table, build: Table, probe_key: str, build_key: str,
               join_type: int, filter: Expr | None = None,
               index: JoinIndex | None = None) -> tuple[np.ndarray, np.ndarray]:
    """(probe row, build row) pairs of the join, NO_ROW = 0xFFFFFFFF."""
    spec, keep = _join_spec(probe, build, probe_key, build_key, join_type, [])
    p, keep2 = _pred(filter, probe)
    n = C.c_int64()
    lib = _lib()
    rc = lib.fgpu_join_pairs(ctx.handle, probe.handle, p, build.handle,
                             index.handle if index else None, C.byref(spec), None, None, 0,
                             C.byref(n))
    if rc not in (_abi.OK, 3):
        _check(rc)
    cap = max(n.value, 1)
    l = np.zeros(cap, np.uint32)
    r = np.zeros(cap, np.uint32)
    _check(lib.fgpu_join_pairs(ctx.handle, probe.handle, p, build.handle,
                               index.handle if index else None, C.byref(spec), l.ctypes.data,
                               r.ctypes.data, cap, C.byref(n)))
    return l[:n.value], r[:n.value]


def host_exact_sum(x: np.ndarray) -> float:
    """The kernels' exact fixed-point float summation, run on the host."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = C.c_double()
    _check(_lib().fgpu_host_exact_sum(x.ctypes.data, x.size, C.byref(out)))
    return out.value
