"""Serving form: a request server with continuous batching over device-resident
tables (fgpu_server_* in include/falcon_gpu.h).

The reference serves concurrent `Database::sql` calls from REST worker threads
over one `Sync` database (src/db.rs:80-91, :242; src/bin/server.rs:181-203).
Here a `Server` owns a few slots (one HIP stream and worker thread each) and an
HBM budget for their scratch memory.  Requests queue in arrival order and start
as soon as a slot is free and their scratch fits the budget, so requests that
arrive while others run overlap with them on the device.  COUNT(*) requests
with one comparison on the same numeric column ride one shared scan.  Each
request's result is the one the plain operator gives for it.

    srv = Server(slots=4, hbm_budget=8 << 30)
    t = srv.submit(count_request(fact, col("f") < 40))
    batch, stats = srv.wait(t)       # RecordBatch, per-request timing
"""
from __future__ import annotations

import atexit
import ctypes as C
import weakref
from dataclasses import dataclass
from typing import Sequence

import numpy as np

from . import _abi
from ._abi import COUNT, FLOAT64, INNER, MAX, MIN, SEMI, SUM
from .operators import (AggExpr, Expr, FalconError, JoinIndex, RecordBatch, Table, _aggs_struct,
                        _check, _lib, _read_result)


@dataclass
class Request:
    """A request ready to submit: the C struct plus what keeps its pointers alive
    until fgpu_server_submit has copied them."""
    kind: int
    struct: _abi.Request
    agg_funcs: list
    keep: tuple
    topk_f64: bool = False


def _pred_ptr(filter: Expr | None, table: Table):
    if filter is None:
        return None, None
    p, keep = filter._compile(table)
    return C.cast(C.pointer(p), C.c_void_p), (p, keep)


def count_request(table: Table, filter: Expr | None = None, alias: str | None = None) -> Request:
    """SELECT COUNT(*) FROM table WHERE filter (src/db.rs:1439-1476)."""
    pp, keep = _pred_ptr(filter, table)
    s = _abi.Request(kind=_abi.REQ_COUNT, table=table.handle, pred=pp, alias=alias.encode() if alias else None)
    return Request(_abi.REQ_COUNT, s, [COUNT], (keep, table))


def groupby_request(table: Table, key: str | None, aggs: Sequence[AggExpr],
                    filter: Expr | None = None) -> Request:
    """SELECT key, aggs FROM table WHERE filter GROUP BY key (key None = no
    GROUP BY), as GroupByExec::execute (src/query/aggregate.rs:535)."""
    pp, keep = _pred_ptr(filter, table)
    arr, keep2 = _aggs_struct(list(aggs), lambda side, name: table.column_index(name))
    kc = table.column_index(key) if key is not None else -1
    spec = _abi.GroupBySpec(kc, len(aggs), C.cast(arr, C.POINTER(_abi.Agg)))
    s = _abi.Request(kind=_abi.REQ_GROUPBY, table=table.handle, pred=pp,
                     groupby=C.cast(C.pointer(spec), C.c_void_p))
    funcs = ([None] if kc >= 0 else []) + [a.func for a in aggs]
    return Request(_abi.REQ_GROUPBY, s, funcs, (keep, arr, keep2, spec, table))


def join_request(probe: Table, build: Table, index: JoinIndex, probe_key: str, build_key: str,
                 join_type: int, aggs: Sequence[AggExpr], filter: Expr | None = None) -> Request:
    """SELECT aggs FROM probe [LEFT|SEMI|ANTI] JOIN build ON probe.k = build.k
    WHERE filter-on-probe, as execute_hash_join (fgpu_join_agg)."""
    if index is None:
        raise FalconError(6, "a served join needs a prebuilt JoinIndex")

    def resolve(side, name):
        return (build if side == 1 else probe).column_index(name)
    arr, keep2 = _aggs_struct(list(aggs), resolve)
    spec = _abi.JoinSpec(join_type, probe.column_index(probe_key), build.column_index(build_key),
                         len(aggs), C.cast(arr, C.POINTER(_abi.Agg)))
    pp, keep = _pred_ptr(filter, probe)
    s = _abi.Request(kind=_abi.REQ_JOIN, table=probe.handle, pred=pp, build=build.handle,
                     index=index.handle, join=C.cast(C.pointer(spec), C.c_void_p))
    return Request(_abi.REQ_JOIN, s, [a.func for a in aggs], (keep, arr, keep2, spec, probe, build, index))


def topk_request(table: Table, column: str, limit: int, desc: bool = False,
                 filter: Expr | None = None) -> Request:
    """SELECT ... ORDER BY column [DESC] LIMIT limit (execute_sort_limit,
    src/query/sort.rs:59-115): result columns `row` and `value`."""
    pp, keep = _pred_ptr(filter, table)
    ci = table.column_index(column)
    s = _abi.Request(kind=_abi.REQ_TOPK, table=table.handle, pred=pp, topk_col=ci,
                     topk_desc=int(desc), topk_k=limit)
    return Request(_abi.REQ_TOPK, s, [None, None], (keep, table), table.dtypes[ci] == FLOAT64)


def run_alone(ctx, req: Request) -> RecordBatch:
    """Runs `req` by itself through the plain entry points on `ctx`: the result
    a served request must equal."""
    lib, r = _lib(), req.struct
    pred = C.cast(r.pred, C.POINTER(_abi.Pred)) if r.pred else None
    if req.kind == _abi.REQ_COUNT:
        n = C.c_int64()
        _check(lib.fgpu_filter_count(ctx.handle, r.table, pred, C.byref(n)))
        name = r.alias.decode() if r.alias else "col0"
        return RecordBatch([name], ["Int64"], [np.array([n.value], np.int64)], [None], [COUNT])
    h = C.c_void_p()
    if req.kind == _abi.REQ_GROUPBY:
        _check(lib.fgpu_groupby(ctx.handle, r.table, pred, C.cast(r.groupby, C.POINTER(_abi.GroupBySpec)),
                                C.byref(h)))
        return _read_result(h, req.agg_funcs)
    if req.kind == _abi.REQ_JOIN:
        _check(lib.fgpu_join_agg(ctx.handle, r.table, pred, r.build, r.index,
                                 C.cast(r.join, C.POINTER(_abi.JoinSpec)), C.byref(h)))
        return _read_result(h, req.agg_funcs)
    k = max(r.topk_k, 1)
    rows, vals, n = np.zeros(k, np.int64), np.zeros(k, np.int64), C.c_int64()
    _check(lib.fgpu_topk(ctx.handle, r.table, pred, r.topk_col, r.topk_desc, r.topk_k, rows.ctypes.data,
                         vals.ctypes.data, C.byref(n)))
    rows, vals = rows[: n.value], vals[: n.value]
    if req.topk_f64:
        return RecordBatch(["row", "value"], ["Int64", "Float64"], [rows, vals.view(np.float64)], [None, None],
                           [None, None])
    return RecordBatch(["row", "value"], ["Int64", "Int64"], [rows, vals], [None, None], [None, None])


# The serving workload's request shapes (datagen.c5_mix gives the mix).
MIX_GROUPBY_AGGS = {
    "g": [AggExpr(COUNT), AggExpr(SUM, "v1"), AggExpr(SUM, "v2")],
    "key": [AggExpr(COUNT), AggExpr(SUM, "v1"), AggExpr(MIN, "v1"), AggExpr(MAX, "v1"), AggExpr(SUM, "v2")],
}
MIX_JOIN_AGGS = {
    "inner": [AggExpr(SUM, "bpay", side="build"), AggExpr(COUNT)],
    "semi": [AggExpr(SUM, "v1"), AggExpr(COUNT)],
}


def mix_request(spec: dict, fact: Table, dim: Table, index: JoinIndex) -> Request:
    """The request for one entry of datagen.c5_mix over the fact table (key,
    v1, v2, f, g) and the dimension table (bk, bpay) with its join index."""
    from .operators import col
    kind = spec["kind"]
    if kind == "count":
        return count_request(fact, col("f") < spec["t"])
    if kind == "groupby":
        return groupby_request(fact, spec["key"], MIX_GROUPBY_AGGS[spec["key"]], col("f") < spec["t"])
    if kind == "join":
        jt = INNER if spec["type"] == "inner" else SEMI
        return join_request(fact, dim, index, "key", "bk", jt, MIX_JOIN_AGGS[spec["type"]], col("f") < spec["t"])
    return topk_request(fact, "v2", spec["k"], desc=True)


_LIVE_SERVERS: "weakref.WeakSet" = weakref.WeakSet()


@atexit.register
def _close_servers_at_exit():
    # registered after the operator layer's hook, so it runs first: servers
    # (their worker threads and slot contexts) go before the caller's contexts
    for s in list(_LIVE_SERVERS):
        s.close()


class Server:
    """Continuous-batching request server (fgpu_server_*).  Close it (or use it
    as a context manager) before freeing the tables its requests read."""

    def __init__(self, device: int = 0, slots: int = 4, hbm_budget: int = 8 << 30,
                 share_counts: bool = True, overlap_always: bool = False):
        # overlap_always=False: beside a running request a new one starts only
        # while at least two wait (one after the other under light load)
        h = C.c_void_p()
        flags = (1 if share_counts else 0) | (2 if overlap_always else 0)
        _check(_lib().fgpu_server_open(device, slots, hbm_budget, flags, C.byref(h)))
        self._h = h
        self._funcs: dict[int, list] = {}
        _LIVE_SERVERS.add(self)

    def submit(self, req: Request) -> int:
        t = C.c_uint64()
        _check(_lib().fgpu_server_submit(self._h, C.byref(req.struct), C.byref(t)))
        self._funcs[t.value] = req.agg_funcs
        return t.value

    def wait(self, ticket: int) -> tuple[RecordBatch, dict]:
        h = C.c_void_p()
        st = _abi.RequestStats()
        rc = _lib().fgpu_server_wait(self._h, ticket, C.byref(h), C.byref(st))
        funcs = self._funcs.pop(ticket, [])
        _check(rc)
        stats = {f: getattr(st, f) for f, _ in _abi.RequestStats._fields_}
        return _read_result(h, funcs), stats

    def pause(self, on: bool = True) -> None:
        """Holds (True) or resumes (False) admissions; queued requests wait."""
        _check(_lib().fgpu_server_pause(self._h, int(on)))

    def stats(self) -> dict:
        st = _abi.ServerStats()
        _check(_lib().fgpu_server_get_stats(self._h, C.byref(st)))
        return {f: getattr(st, f) for f, _ in _abi.ServerStats._fields_}

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib().fgpu_server_close(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
