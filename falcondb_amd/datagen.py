"""Synthetic tables for the filter / GROUP BY / join workloads (numpy arrays
filled by libfgpu_datagen.so; bit-identical for any thread count).

  c1_table(n)        id, dept (12-value dictionary), score, salary
  keys(n, bits)      n distinct int64 keys in [0, 2^bits), random order
  c2_table(n, K)     key ~ Zipf(1.1) over K, v1, v2 ~ N(0,1), f in [0, 100)
  c3_build(K)        bk = K shuffled, bpay = U[0,1)
  c3_probe(n, K)     pk: 90 % members of K, 10 % non-members; ppay
  c4_batch(n, card)  key in [0, card), val, score
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libfgpu_datagen.so"
DEPTS = [f"dept{i:02d}" for i in range(12)]
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} missing: run `python -m falcondb_amd.build`")
        L = C.CDLL(str(LIB_PATH))
        vp, i64, u64, dbl, i32 = C.c_void_p, C.c_int64, C.c_uint64, C.c_double, C.c_int32
        L.fgen_seed.restype = u64
        L.fgen_seed.argtypes = [C.c_int, C.c_int]
        L.fgen_keys.argtypes = [i64, C.c_int, u64, vp]
        L.fgen_c1.argtypes = [i64, vp, vp, vp, vp]
        L.fgen_c2.argtypes = [i64, vp, i64, dbl, vp, vp, vp, vp]
        L.fgen_c3_build.argtypes = [i64, vp, vp, vp]
        L.fgen_c3_probe.argtypes = [i64, vp, i64, C.c_int, dbl, vp, vp]
        L.fgen_c4.argtypes = [i64, u64, i32, vp, vp, vp]
        L.fgen_uniform_f64.argtypes = [i64, u64, dbl, dbl, vp]
        L.fgen_uniform_i64.argtypes = [i64, u64, i64, i64, vp]
        for f in ("fgen_keys", "fgen_c1", "fgen_c2", "fgen_c3_build", "fgen_c3_probe", "fgen_c4",
                  "fgen_uniform_f64", "fgen_uniform_i64"):
            getattr(L, f).restype = C.c_int
        _lib = L
    return _lib


def _ok(rc):
    if rc != 0:
        raise ValueError("generator rejected its arguments")


def seed(config: int, column: int) -> int:
    return int(lib().fgen_seed(config, column))


def c1_table(n: int) -> dict:
    id_ = np.empty(n, np.int64)
    dept = np.empty(n, np.uint32)
    score = np.empty(n, np.float64)
    salary = np.empty(n, np.int64)
    _ok(lib().fgen_c1(n, id_.ctypes.data, dept.ctypes.data, score.ctypes.data, salary.ctypes.data))
    return {"id": id_, "dept": dept, "score": score, "salary": salary}


def keys(n: int, bits: int = 40, seed_: int | None = None) -> np.ndarray:
    out = np.empty(n, np.int64)
    _ok(lib().fgen_keys(n, bits, seed(2, 15) if seed_ is None else seed_, out.ctypes.data))
    return out


def c2_table(n: int, K: np.ndarray, s: float = 1.1) -> dict:
    K = np.ascontiguousarray(K, np.int64)
    key = np.empty(n, np.int64)
    v1 = np.empty(n, np.int64)
    v2 = np.empty(n, np.float64)
    f = np.empty(n, np.int32)
    _ok(lib().fgen_c2(n, K.ctypes.data, K.size, s, key.ctypes.data, v1.ctypes.data,
                      v2.ctypes.data, f.ctypes.data))
    return {"key": key, "v1": v1, "v2": v2, "f": f}


def c3_build(K: np.ndarray) -> dict:
    K = np.ascontiguousarray(K, np.int64)
    bk = np.empty(K.size, np.int64)
    bpay = np.empty(K.size, np.float64)
    _ok(lib().fgen_c3_build(K.size, K.ctypes.data, bk.ctypes.data, bpay.ctypes.data))
    return {"bk": bk, "bpay": bpay}


def c3_probe(n: int, K: np.ndarray, bits: int = 40, match_frac: float = 0.9) -> dict:
    K = np.ascontiguousarray(K, np.int64)
    pk = np.empty(n, np.int64)
    ppay = np.empty(n, np.int64)
    _ok(lib().fgen_c3_probe(n, K.ctypes.data, K.size, bits, match_frac, pk.ctypes.data,
                            ppay.ctypes.data))
    return {"pk": pk, "ppay": ppay}


def c4_batch(n: int, card: int, seed_: int) -> dict:
    key = np.empty(n, np.int32)
    val = np.empty(n, np.int64)
    score = np.empty(n, np.float64)
    _ok(lib().fgen_c4(n, seed_, card, key.ctypes.data, val.ctypes.data, score.ctypes.data))
    return {"key": key, "val": val, "score": score}


def uniform_f64(n: int, seed_: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
    out = np.empty(n, np.float64)
    _ok(lib().fgen_uniform_f64(n, seed_, lo, hi, out.ctypes.data))
    return out


def uniform_i64(n: int, seed_: int, lo: int, hi: int) -> np.ndarray:
    out = np.empty(n, np.int64)
    _ok(lib().fgen_uniform_i64(n, seed_, lo, hi, out.ctypes.data))
    return out


def c4_queries(n_queries: int = 256, min_rows: int = 10_000, max_rows: int = 10_000_000,
               min_card: int = 8, max_card: int = 1_000_000) -> list[dict]:
    """Per-query shape of the batched workload: rows log-uniform in
    [min_rows, max_rows], key cardinality uniform in [min_card, max_card] and
    the predicate threshold `score < theta` with theta uniform in [0, 100).
    Batch i's columns are c4_batch(rows, card, seed(4, 16 + i))."""
    u_rows = uniform_f64(n_queries, seed(4, 0))
    u_card = uniform_f64(n_queries, seed(4, 1))
    u_theta = uniform_f64(n_queries, seed(4, 2))
    lo, hi = np.log(min_rows), np.log(max_rows)
    out = []
    for i in range(n_queries):
        rows = int(np.floor(np.exp(lo + (hi - lo) * u_rows[i])))
        card = int(min_card + np.floor(u_card[i] * (max_card - min_card + 1)))
        out.append({"rows": max(rows, 1), "card": min(card, max_card), "theta": float(100.0 * u_theta[i]),
                    "seed": seed(4, 16 + i)})
    return out


G_NAMES = [f"g{i:02d}" for i in range(12)]


def c5_fact(n: int, K: np.ndarray) -> dict:
    """Serving fact table: the 100M-row group-by schema (key, v1, v2, f) plus
    `g`, dictionary codes of 12 values (G_NAMES) drawn uniformly."""
    cols = c2_table(n, K)
    cols["g"] = np.minimum(np.floor(uniform_f64(n, seed(5, 4)) * 12), 11).astype(np.uint32)
    return cols


def c5_mix(n_requests: int, seed_: int | None = None) -> list[dict]:
    """Serving request mix: 40 % `COUNT(*) WHERE f < t`; 30 % GROUP BY g (12
    groups) or GROUP BY key (high cardinality), each WHERE f < t; 20 % INNER or
    SEMI join of the fact table (probe, WHERE f < t) with the dimension table;
    10 % `ORDER BY v2 DESC LIMIT 10`.  t is uniform in [1, 100], so a join
    probes 1 % .. 100 % of the fact rows."""
    u = uniform_f64(3 * n_requests, seed(5, 0) if seed_ is None else seed_)
    out = []
    for i in range(n_requests):
        a, b, c = u[3 * i], u[3 * i + 1], u[3 * i + 2]
        t = 1 + int(np.floor(c * 100))
        if a < 0.4:
            out.append({"kind": "count", "t": t})
        elif a < 0.7:
            out.append({"kind": "groupby", "key": "g" if b < 0.5 else "key", "t": t})
        elif a < 0.9:
            out.append({"kind": "join", "type": "inner" if b < 0.5 else "semi", "t": t})
        else:
            out.append({"kind": "topk", "k": 10})
    return out
