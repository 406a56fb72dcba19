// fgpu_join.hip — hash join build / probe for gfx950, fused with the
// aggregates that follow the join in the reference plan.
//
// Reference path replaced: build_hash_table + IntHashTable (chained table,
// serial insert; src/query/join.rs:607-645, :1054-1107, :1318-1350), probe_one /
// collect_matches (:657-723) and, for SEMI / ANTI, execute_hash_semi_anti_join
// (:49-132); the INNER / LEFT output is not materialised, it is folded straight
// into the aggregate accumulators (execute_agg_only, src/query/aggregate.rs:1394).
//
// Index: open addressing with linear probing, one slot per build row (duplicate
// keys occupy several slots; a probe walks until an empty slot).  Slot = int64
// key (EMPTY = INT64_MIN, claimed with a 64-bit CAS) + u32 build row.  Build
// rows whose key equals the EMPTY sentinel are kept in a side list.  A NULL key
// never matches (src/query/join.rs:1009-1019).
#include "fgpu_state.h"

namespace fgpu {

__global__ void k_join_build(JoinBuildArgs a) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < a.n; r += stride) {
        if (a.key_valid && !((a.key_valid[r >> 6] >> (r & 63)) & 1)) continue;  // NULL joins nothing
        const int64_t key = a.key_dtype == FGPU_INT32 ? (int64_t)((const int32_t *)a.keys_col)[r]
                                                      : ((const int64_t *)a.keys_col)[r];
        if (key == EMPTY_KEY) {
            const unsigned long long k = atomicAdd(a.counters + 2, 1ull);
            if ((int64_t)k < a.min_key_cap) a.min_key_rows[k] = (uint32_t)r;
            continue;
        }
        uint64_t i = hash64((uint64_t)key) & a.cap_mask;
        bool done = false;
        for (uint64_t p = 0; p <= a.cap_mask; ++p) {
            int64_t expect = EMPTY_KEY;
            if (__hip_atomic_compare_exchange_strong(a.keys + i, &expect, key, __ATOMIC_RELAXED,
                                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                a.rows[i] = (uint32_t)r;
                done = true;
                break;
            }
            i = (i + 1) & a.cap_mask;
        }
        if (!done) atomicExch(a.counters + 1, 1ull);
        else atomicAdd(a.counters, 1ull);
    }
}

hipError_t launch_join_build(hipStream_t s, const JoinBuildArgs &a, int grid) {
    hipLaunchKernelGGL(k_join_build, dim3(grid), dim3(BLOCK), 0, s, a);
    return hipGetLastError();
}

FGPU_DEV bool probe_key(const JoinProbeArgs &a, int64_t r, int64_t &key) {
    const int k = a.pkey;
    if (a.pcols.valid[k] && !((a.pcols.valid[k][r >> 6] >> (r & 63)) & 1)) return false;
    key = load_raw(a.pcols, k, r);
    return true;
}

// Calls f(build_row) for every build row whose key equals `key`, in slot
// order; stops early when f returns false.  Returns the number of matches seen.
template <class F>
FGPU_DEV int64_t for_matches(const JoinProbeArgs &a, int64_t key, F f) {
    int64_t m = 0;
    if (key == EMPTY_KEY) {
        for (int64_t j = 0; j < a.n_min_key_rows; ++j) {
            ++m;
            if (!f(a.min_key_rows[j])) return m;
        }
        return m;
    }
    uint64_t i = hash64((uint64_t)key) & a.cap_mask;
    for (;;) {
        const int64_t k = a.keys[i];
        if (k == EMPTY_KEY) break;
        if (k == key) {
            ++m;
            if (!f(a.rows[i])) return m;
        }
        i = (i + 1) & a.cap_mask;
    }
    return m;
}

FGPU_DEV bool side_value(const JoinProbeArgs &a, const JoinOp &o, int64_t pr, uint32_t br,
                         int64_t &v) {
    if (o.side == 0) {
        if (!load_valid(a.pcols, o.col, pr)) return false;
        v = load_raw(a.pcols, o.col, pr);
        return true;
    }
    if (br == NO_ROW || !load_valid(a.bcols, o.col, br)) return false;  // LEFT: NULL build side
    v = load_raw(a.bcols, o.col, br);
    return true;
}

constexpr int JOIN_UNROLL = 4;
constexpr int JOIN_MAX_OPS = 4;  // register accumulators per thread

struct JoinAcc {
    int64_t rows;
    int64_t acc[JOIN_MAX_OPS];
    double hi[JOIN_MAX_OPS], lo[JOIN_MAX_OPS];
};

// One output row of the join: probe row pr joined with build row br (NO_ROW =
// NULL-extended / not part of the output).
FGPU_DEV void join_emit(const JoinProbeArgs &a, JoinAcc &j, int64_t pr, uint32_t br) {
    ++j.rows;
#pragma unroll
    for (int i = 0; i < JOIN_MAX_OPS; ++i) {
        if (i < a.n_ops) {
            const JoinOp o = a.ops[i];
            int64_t v;
            if (o.col >= 0 && side_value(a, o, pr, br, v)) reg_add(o.kind, v, j.acc[i], j.hi[i], j.lo[i]);
        }
    }
}

// Probe fused with the aggregates: every output row of the join (a match for
// INNER; a match or the NULL-extended row for LEFT; the kept probe row for
// SEMI / ANTI) is folded into per-thread registers, then once per workgroup
// into the state row.
template <int TYPE>
__global__ void k_join_probe_agg(JoinProbeArgs a) {
    __shared__ uint64_t lrow[64];
    for (int i = threadIdx.x; i < a.S; i += blockDim.x) lrow[i] = 0;
    __syncthreads();
    if (threadIdx.x < JOIN_MAX_OPS && (int)threadIdx.x < a.n_ops) {
        const JoinOp o = a.ops[threadIdx.x];
        if (o.col >= 0 && o.kind != OP_FSUM) lrow[o.word] = (uint64_t)reg_ident(o.kind);
    }
    __syncthreads();
    JoinAcc j;
    j.rows = 0;
#pragma unroll
    for (int i = 0; i < JOIN_MAX_OPS; ++i) {
        j.acc[i] = i < a.n_ops ? reg_ident(a.ops[i].kind) : 0;
        j.hi[i] = 0.0;
        j.lo[i] = 0.0;
    }
    const int64_t tile = (int64_t)blockDim.x * JOIN_UNROLL;
    for (int64_t base = (int64_t)blockIdx.x * tile; base < a.n_probe; base += (int64_t)gridDim.x * tile) {
        bool sel[JOIN_UNROLL], has[JOIN_UNROLL];
        int64_t key[JOIN_UNROLL];
#pragma unroll
        for (int u = 0; u < JOIN_UNROLL; ++u) {
            const int64_t r = base + u * blockDim.x + threadIdx.x;
            sel[u] = r < a.n_probe && eval_pred(a.pred, a.pcols, r);
            key[u] = 0;
            has[u] = sel[u] && probe_key(a, r, key[u]);
        }
#pragma unroll
        for (int u = 0; u < JOIN_UNROLL; ++u) {
            if (!sel[u]) continue;
            const int64_t r = base + u * blockDim.x + threadIdx.x;
            int64_t m = 0;
            if (has[u]) {
                if (key[u] == EMPTY_KEY) {
                    for (int64_t q = 0; q < a.n_min_key_rows; ++q) {
                        ++m;
                        if (TYPE == FGPU_INNER || TYPE == FGPU_LEFT) join_emit(a, j, r, a.min_key_rows[q]);
                        else break;
                    }
                } else {
                    uint64_t i = hash64((uint64_t)key[u]) & a.cap_mask;
                    for (;;) {
                        const int64_t k = a.keys[i];
                        if (k == EMPTY_KEY) break;
                        if (k == key[u]) {
                            ++m;
                            if (TYPE == FGPU_INNER || TYPE == FGPU_LEFT) join_emit(a, j, r, a.rows[i]);
                            else break;
                        }
                        i = (i + 1) & a.cap_mask;
                    }
                }
            }
            if ((TYPE == FGPU_LEFT && m == 0) || (TYPE == FGPU_SEMI && m > 0) ||
                (TYPE == FGPU_ANTI && m == 0))
                join_emit(a, j, r, NO_ROW);
        }
    }
    if (j.rows) __hip_atomic_fetch_add(lrow, (uint64_t)j.rows, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
#pragma unroll
    for (int i = 0; i < JOIN_MAX_OPS; ++i) {
        if (i < a.n_ops) {
            const JoinOp jo = a.ops[i];
            if (jo.col >= 0) {
                DOp o;
                o.kind = jo.kind;
                o.col = 0;
                o.W = jo.W;
                o.has_flags = jo.has_flags;
                o.word = jo.word;
                o.L = jo.L;
                fold_reg(lrow, o, j.acc[i], j.hi[i], j.lo[i]);
            }
        }
    }
    __syncthreads();
    // merge the workgroup row into the HBM state row, one word per thread
    for (int w = threadIdx.x; w < a.S; w += blockDim.x) {
        int kind = OP_ISUM;  // word 0 (rows), counts, sums and fixed-point words add
        for (int i = 0; i < a.n_ops; ++i) {
            const JoinOp o = a.ops[i];
            if (o.col < 0) continue;
            if (w == o.word && o.kind != OP_FSUM) kind = o.kind;
            if (o.kind == OP_FSUM && o.has_flags && w == o.word + o.W) kind = -1;  // flags: OR
        }
        const uint64_t v = lrow[w];
        uint64_t *d = a.state + w;
        switch (kind) {
            case OP_IMIN: a_smin<DEV_SCOPE>(d, (int64_t)v); break;
            case OP_IMAX: a_smax<DEV_SCOPE>(d, (int64_t)v); break;
            case OP_FMIN: a_umin<DEV_SCOPE>(d, v); break;
            case OP_FMAX: a_umax<DEV_SCOPE>(d, v); break;
            case -1:
                if (v) a_or<DEV_SCOPE>(d, v);
                break;
            default:
                if (v) a_add<DEV_SCOPE>(d, v);
                break;
        }
    }
}

hipError_t launch_join_probe_agg(hipStream_t s, const JoinProbeArgs &a, int grid, int block) {
    if (a.n_ops > JOIN_MAX_OPS) return hipErrorInvalidValue;
    switch (a.type) {
        case FGPU_INNER: hipLaunchKernelGGL(k_join_probe_agg<FGPU_INNER>, dim3(grid), dim3(block), 0, s, a); break;
        case FGPU_LEFT: hipLaunchKernelGGL(k_join_probe_agg<FGPU_LEFT>, dim3(grid), dim3(block), 0, s, a); break;
        case FGPU_SEMI: hipLaunchKernelGGL(k_join_probe_agg<FGPU_SEMI>, dim3(grid), dim3(block), 0, s, a); break;
        default: hipLaunchKernelGGL(k_join_probe_agg<FGPU_ANTI>, dim3(grid), dim3(block), 0, s, a); break;
    }
    return hipGetLastError();
}

// Pair mode.  Pass 0 counts the output rows of each probe row; pass 1 writes
// (probe row, build row) pairs at the scanned offsets.
__global__ void k_join_pairs(JoinProbeArgs a, int pass) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < a.n_probe; r += stride) {
        uint32_t outn = 0;
        int64_t o = pass ? a.offs[r] : 0;
        const bool sel = eval_pred(a.pred, a.pcols, r);
        int64_t key;
        const bool has_key = sel && probe_key(a, r, key);
        int64_t m = 0;
        if (has_key) {
            if (a.type == FGPU_INNER || a.type == FGPU_LEFT) {
                m = for_matches(a, key, [&](uint32_t br) {
                    if (pass) {
                        a.out_l[o] = (uint32_t)r;
                        a.out_r[o] = br;
                        ++o;
                    }
                    ++outn;
                    return true;
                });
            } else {
                m = for_matches(a, key, [&](uint32_t) { return false; });
            }
        }
        if (sel && ((a.type == FGPU_LEFT && m == 0) || (a.type == FGPU_SEMI && m > 0) ||
                    (a.type == FGPU_ANTI && m == 0))) {
            if (pass) {
                a.out_l[o] = (uint32_t)r;
                a.out_r[o] = NO_ROW;
            }
            ++outn;
        }
        if (!pass) a.counts[r] = outn;
    }
}

hipError_t launch_join_pairs(hipStream_t s, const JoinProbeArgs &a, int pass, int grid) {
    hipLaunchKernelGGL(k_join_pairs, dim3(grid), dim3(BLOCK), 0, s, a, pass);
    return hipGetLastError();
}

}  // namespace fgpu
