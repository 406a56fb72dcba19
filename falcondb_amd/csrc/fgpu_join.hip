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
// keys occupy several slots; a probe walks until an empty slot).  A slot is 16
// bytes, {int64 key (EMPTY = INT64_MIN, claimed with a 64-bit CAS), build row},
// so one 16-byte load returns both.  Capacity >= 2x the build rows keeps the
// walks short and guarantees an empty slot.  Build rows whose key equals the
// EMPTY sentinel are kept in a side list.  A NULL key never matches
// (src/query/join.rs:1009-1019).
//
// Probe: each thread takes JOIN_UNROLL probe rows of a tile; the predicate
// columns, then the keys, then the first slot of every row are loaded before
// any of them is examined, so each wave has JOIN_UNROLL independent random
// slot loads in flight.
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
            if (__hip_atomic_compare_exchange_strong(a.slots + 2 * i, &expect, key, __ATOMIC_RELAXED,
                                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                a.slots[2 * i + 1] = r;
                done = true;
                break;
            }
            if (expect == key) atomicExch(a.counters + 3, 1ull);  // duplicate build key
            i = (i + 1) & a.cap_mask;
        }
        if (!done) atomicExch(a.counters + 1, 1ull);
    }
}

hipError_t launch_join_build(hipStream_t s, const JoinBuildArgs &a, int grid) {
    hipLaunchKernelGGL(k_join_build, dim3(grid), dim3(BLOCK), 0, s, a);
    return hipGetLastError();
}

FGPU_DEV longlong2 load_slot(const int64_t *slots, uint64_t i) {
    return ((const longlong2 *)slots)[i];
}

FGPU_DEV bool side_value(const JoinProbeArgs &a, const JoinOp &o, int64_t pr, uint32_t br,
                         int64_t &v) {
    bool ok;
    if (o.side == 0) {
        load_cell(a.pcols, o.col, pr, v, ok);
        return ok;
    }
    if (br == NO_ROW) return false;  // LEFT: NULL build side
    load_cell(a.bcols, o.col, br, v, ok);
    return ok;
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

// Loads, for every row of a tile, whether it is selected, its key and the
// first slot of its probe sequence.
template <int U, int NPC>
FGPU_DEV void probe_tile(const JoinProbeArgs &a, const int64_t (&rows)[U], bool (&sel)[U],
                         bool (&has)[U], int64_t (&key)[U], uint64_t (&slot)[U],
                         longlong2 (&first)[U]) {
    pred_tile<U, NPC>(a.pred, a.pcols, rows, a.n_probe, sel);
#pragma unroll
    for (int u = 0; u < U; ++u) {
        key[u] = 0;
        bool ok = false;
        if (sel[u]) load_cell(a.pcols, a.pkey, rows[u], key[u], ok);
        has[u] = sel[u] && ok;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        slot[u] = hash64((uint64_t)key[u]) & a.cap_mask;
        first[u] = make_longlong2(EMPTY_KEY, 0);
        if (has[u] && key[u] != EMPTY_KEY) first[u] = load_slot(a.slots, slot[u]);
    }
}

// Walks the probe sequence of `key` starting with the already loaded slot
// `cur` at index i; f(build_row) is called per match and returns false to stop.
template <class F>
FGPU_DEV int64_t walk(const JoinProbeArgs &a, int64_t key, uint64_t i, longlong2 cur, F f) {
    int64_t m = 0;
    if (key == EMPTY_KEY) {
        for (int64_t q = 0; q < a.n_min_key_rows; ++q) {
            ++m;
            if (!f(a.min_key_rows[q])) break;
        }
        return m;
    }
    while (cur.x != EMPTY_KEY) {
        if (cur.x == key) {
            ++m;
            if (!f((uint32_t)cur.y)) break;
        }
        i = (i + 1) & a.cap_mask;
        cur = load_slot(a.slots, i);
    }
    return m;
}

// Folds the thread's accumulators into the workgroup's LDS row, then merges
// the row into the HBM state row (one word per thread).  Every thread of the
// workgroup must call it.
FGPU_DEV void acc_finish(const JoinProbeArgs &a, JoinAcc &j, uint64_t *lrow) {
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

// Zero / identity-initialises the workgroup's LDS state row.
FGPU_DEV void acc_init(const JoinProbeArgs &a, JoinAcc &j, uint64_t *lrow) {
    for (int i = threadIdx.x; i < a.S; i += blockDim.x) lrow[i] = 0;
    __syncthreads();
    if (threadIdx.x < JOIN_MAX_OPS && (int)threadIdx.x < a.n_ops) {
        const JoinOp o = a.ops[threadIdx.x];
        if (o.col >= 0 && o.kind != OP_FSUM) lrow[o.word] = (uint64_t)reg_ident(o.kind);
    }
    __syncthreads();
    j.rows = 0;
#pragma unroll
    for (int i = 0; i < JOIN_MAX_OPS; ++i) {
        j.acc[i] = i < a.n_ops ? reg_ident(a.ops[i].kind) : 0;
        j.hi[i] = 0.0;
        j.lo[i] = 0.0;
    }
}

// Probe fused with the aggregates: every output row of the join (a match for
// INNER; a match or the NULL-extended row for LEFT; the kept probe row for
// SEMI / ANTI) is folded into per-thread registers, then once per workgroup
// into the state row.
template <int TYPE, int NPC>
__global__ void k_join_probe_agg(JoinProbeArgs a) {
    __shared__ uint64_t lrow[64];
    JoinAcc j;
    acc_init(a, j, lrow);
    constexpr bool kPairs = TYPE == FGPU_INNER || TYPE == FGPU_LEFT;
    const int64_t tile = (int64_t)blockDim.x * JOIN_UNROLL;
    for (int64_t base = (int64_t)blockIdx.x * tile; base < a.n_probe; base += (int64_t)gridDim.x * tile) {
        int64_t rows[JOIN_UNROLL], key[JOIN_UNROLL];
        bool sel[JOIN_UNROLL], has[JOIN_UNROLL];
        uint64_t slot[JOIN_UNROLL];
        longlong2 first[JOIN_UNROLL];
#pragma unroll
        for (int u = 0; u < JOIN_UNROLL; ++u) rows[u] = base + u * blockDim.x + threadIdx.x;
        probe_tile<JOIN_UNROLL, NPC>(a, rows, sel, has, key, slot, first);
#pragma unroll
        for (int u = 0; u < JOIN_UNROLL; ++u) {
            if (!sel[u]) continue;
            int64_t m = 0;
            if (has[u]) {
                if (kPairs) {
                    m = walk(a, key[u], slot[u], first[u], [&](uint32_t br) {
                        join_emit(a, j, rows[u], br);
                        return !a.unique_keys;
                    });
                } else {
                    m = walk(a, key[u], slot[u], first[u], [&](uint32_t) { return false; });
                }
            }
            if ((TYPE == FGPU_LEFT && m == 0) || (TYPE == FGPU_SEMI && m > 0) ||
                (TYPE == FGPU_ANTI && m == 0))
                join_emit(a, j, rows[u], NO_ROW);
        }
    }
    acc_finish(a, j, lrow);
}

template <int NPC>
static void launch_probe_npc(hipStream_t s, const JoinProbeArgs &a, int grid, int block) {
    switch (a.type) {
        case FGPU_INNER: hipLaunchKernelGGL((k_join_probe_agg<FGPU_INNER, NPC>), dim3(grid), dim3(block), 0, s, a); break;
        case FGPU_LEFT: hipLaunchKernelGGL((k_join_probe_agg<FGPU_LEFT, NPC>), dim3(grid), dim3(block), 0, s, a); break;
        case FGPU_SEMI: hipLaunchKernelGGL((k_join_probe_agg<FGPU_SEMI, NPC>), dim3(grid), dim3(block), 0, s, a); break;
        default: hipLaunchKernelGGL((k_join_probe_agg<FGPU_ANTI, NPC>), dim3(grid), dim3(block), 0, s, a); break;
    }
}

hipError_t launch_join_probe_agg(hipStream_t s, const JoinProbeArgs &a, int grid, int block) {
    if (a.n_ops > JOIN_MAX_OPS) return hipErrorInvalidValue;
    if (a.pred.n_pcols <= 2) launch_probe_npc<2>(s, a, grid, block);
    else launch_probe_npc<MAX_PCOLS>(s, a, grid, block);
    return hipGetLastError();
}

// ─── radix-partitioned join ──────────────────────────────────────────────────
FGPU_DEV int part_of(uint64_t h, int log_p) { return log_p ? (int)(h >> (64 - log_p)) : 0; }

__global__ void k_jp_build(JoinPartArgs a, JoinBuildArgs b) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const uint64_t tmask = (uint64_t)a.T - 1;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < b.n; r += stride) {
        if (b.key_valid && !((b.key_valid[r >> 6] >> (r & 63)) & 1)) continue;  // NULL joins nothing
        const int64_t key = b.key_dtype == FGPU_INT32 ? (int64_t)((const int32_t *)b.keys_col)[r]
                                                      : ((const int64_t *)b.keys_col)[r];
        if (key == EMPTY_KEY) {
            const unsigned long long k = atomicAdd(a.flags + 2, 1ull);
            if ((int64_t)k < a.min_key_cap) a.min_key_rows[k] = (uint32_t)r;
            continue;
        }
        const uint64_t h = hash64((uint64_t)key);
        const int p = part_of(h, a.log_p);  // k_jp_count checked every sub-table has room
        const int64_t pay = a.build_pay_col >= 0 ? load_raw(a.pa.bcols, a.build_pay_col, r) : r;
        int64_t *sub = a.sub + (int64_t)p * a.T * 2;
        uint64_t slot = h & tmask;
        bool done = false;
        for (int q = 0; q < a.T && !done; ++q) {
            int64_t expect = EMPTY_KEY;
            if (__hip_atomic_compare_exchange_strong(sub + 2 * slot, &expect, key, __ATOMIC_RELAXED,
                                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                sub[2 * slot + 1] = pay;
                done = true;
            } else if (expect == key) {
                atomicExch(a.flags + 3, 1ull);  // duplicate build key
            }
            slot = (slot + 1) & tmask;
        }
        if (!done) atomicExch(a.flags + 1, 1ull);
    }
}

// Keys per sub-table: per-workgroup LDS histogram, one atomic per partition.
__global__ void k_jp_count(JoinPartArgs a, JoinBuildArgs b) {
    extern __shared__ unsigned int hist[];
    const int P = 1 << a.log_p;
    for (int i = threadIdx.x; i < P; i += blockDim.x) hist[i] = 0;
    __syncthreads();
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < b.n; r += stride) {
        if (b.key_valid && !((b.key_valid[r >> 6] >> (r & 63)) & 1)) continue;
        const int64_t key = b.key_dtype == FGPU_INT32 ? (int64_t)((const int32_t *)b.keys_col)[r]
                                                      : ((const int64_t *)b.keys_col)[r];
        if (key == EMPTY_KEY) continue;
        atomicAdd(hist + part_of(hash64((uint64_t)key), a.log_p), 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < P; i += blockDim.x)
        if (hist[i]) atomicAdd(a.sub_cnt + i, hist[i]);
}

hipError_t launch_join_part_count(hipStream_t s, const JoinPartArgs &a, const JoinBuildArgs &b, int grid) {
    hipLaunchKernelGGL(k_jp_count, dim3(grid), dim3(BLOCK), (size_t)(1 << a.log_p) * 4, s, a, b);
    return hipGetLastError();
}

hipError_t launch_join_part_build(hipStream_t s, const JoinPartArgs &a, const JoinBuildArgs &b, int grid) {
    hipLaunchKernelGGL(k_jp_build, dim3(grid), dim3(BLOCK), 0, s, a, b);
    return hipGetLastError();
}

// Pass 1: rows with a NULL key or the sentinel key are joined here directly;
// every other selected row becomes a record in cell (workgroup, partition).
template <int TYPE, int NPC>
__global__ void k_jp_scatter(JoinPartArgs a) {
    extern __shared__ unsigned int cur[];  // one cursor per partition
    __shared__ uint64_t lrow[64];
    const JoinProbeArgs &pa = a.pa;
    const int P = 1 << a.log_p;
    for (int i = threadIdx.x; i < P; i += blockDim.x) cur[i] = 0;
    JoinAcc j;
    acc_init(pa, j, lrow);  // synchronises the workgroup
    const int64_t r0 = (int64_t)blockIdx.x * a.rows_per_wg;
    const int64_t r1 = r0 + a.rows_per_wg < pa.n_probe ? r0 + a.rows_per_wg : pa.n_probe;
    int64_t *cells = a.cells + (int64_t)blockIdx.x * P * a.Cw * a.RW;
    for (int64_t base = r0; base < r1; base += (int64_t)blockDim.x * JOIN_UNROLL) {
        int64_t rows[JOIN_UNROLL], key[JOIN_UNROLL], f1[JOIN_UNROLL], f2[JOIN_UNROLL];
        bool sel[JOIN_UNROLL], has[JOIN_UNROLL];
#pragma unroll
        for (int u = 0; u < JOIN_UNROLL; ++u) rows[u] = base + u * blockDim.x + threadIdx.x;
        pred_tile<JOIN_UNROLL, NPC>(pa.pred, pa.pcols, rows, r1, sel);
#pragma unroll
        for (int u = 0; u < JOIN_UNROLL; ++u) {
            key[u] = f1[u] = f2[u] = 0;
            bool ok = false;
            if (sel[u]) {
                load_cell(pa.pcols, pa.pkey, rows[u], key[u], ok);
                if (a.n_pay > 0) f1[u] = load_raw(pa.pcols, a.pay_col[0], rows[u]);
                if (a.n_pay > 1) f2[u] = load_raw(pa.pcols, a.pay_col[1], rows[u]);
            }
            has[u] = sel[u] && ok;
        }
#pragma unroll
        for (int u = 0; u < JOIN_UNROLL; ++u) {
            if (!sel[u]) continue;
            const int64_t r = rows[u];
            if (!has[u]) {  // NULL key: matches nothing
                if (TYPE == FGPU_LEFT || TYPE == FGPU_ANTI) join_emit(pa, j, r, NO_ROW);
                continue;
            }
            if (key[u] == EMPTY_KEY) {  // sentinel-valued key: the build side list
                int64_t m = 0;
                for (int64_t q = 0; q < pa.n_min_key_rows; ++q) {
                    ++m;
                    if (TYPE == FGPU_INNER || TYPE == FGPU_LEFT) join_emit(pa, j, r, pa.min_key_rows[q]);
                    else break;
                }
                if ((TYPE == FGPU_LEFT && m == 0) || (TYPE == FGPU_SEMI && m > 0) ||
                    (TYPE == FGPU_ANTI && m == 0))
                    join_emit(pa, j, r, NO_ROW);
                continue;
            }
            const int p = part_of(hash64((uint64_t)key[u]), a.log_p);
            const unsigned pos = atomicAdd(cur + p, 1u);
            if (pos < (unsigned)a.Cw) {
                int64_t *rec = cells + ((int64_t)p * a.Cw + pos) * a.RW;
                rec[0] = key[u];
                if (a.n_pay > 0) rec[1] = f1[u];
                if (a.n_pay > 1) rec[2] = f2[u];
            } else {
                atomicExch(a.flags, 1ull);  // cell full: the host reruns on the direct path
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < P; i += blockDim.x) a.cell_cnt[(int64_t)blockIdx.x * P + i] = cur[i];
    acc_finish(pa, j, lrow);
}

// One output row of the partitioned join: probe payloads f1, f2 come from the
// record, the build value from the sub-table slot (`present` = a match).
FGPU_DEV void part_emit(const JoinPartArgs &a, JoinAcc &j, int64_t f1, int64_t f2, bool present,
                        int64_t pay) {
    ++j.rows;
#pragma unroll
    for (int i = 0; i < JOIN_MAX_OPS; ++i) {
        if (i < a.pa.n_ops) {
            const JoinOp o = a.pa.ops[i];
            if (o.col < 0) continue;
            int64_t v;
            bool ok = true;
            if (o.side == 0) {
                v = a.field_of_op[i] == 2 ? f2 : f1;
            } else {
                if (!present) continue;
                if (a.build_pay_col >= 0) v = pay;
                else load_cell(a.pa.bcols, o.col, (int64_t)(uint32_t)pay, v, ok);
            }
            if (ok) reg_add(o.kind, v, j.acc[i], j.hi[i], j.lo[i]);
        }
    }
}

// Pass 2: XCD-aware.  Workgroups are dispatched round-robin over the 8 XCDs
// (workgroup b runs on XCD b % 8), so XCD x owns partitions x, x+8, x+16, ...
// and all of its workgroups walk those partitions in the same order: while
// they probe partition p, p's table (T 16-byte slots, about 1 MB) is the hot
// set of that XCD's L2.  Workgroup b covers cells [l*G/K, (l+1)*G/K) of each
// partition, l = b / 8, K = gridDim.x / 8; one cell per wave at a time.
template <int TYPE>
__global__ void k_jp_probe(JoinPartArgs a) {
    __shared__ uint64_t lrow[64];
    JoinAcc j;
    acc_init(a.pa, j, lrow);  // synchronises the workgroup
    const int P = 1 << a.log_p;
    const int x = blockIdx.x & 7, l = blockIdx.x >> 3, K = gridDim.x >> 3;
    const uint64_t tmask = (uint64_t)a.T - 1;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int w0 = (int)((int64_t)a.G * l / K), w1 = (int)((int64_t)a.G * (l + 1) / K);
    for (int p = x; p < P; p += 8) {
        const int64_t *tab = a.sub + (int64_t)p * a.T * 2;
        for (int w = w0 + wave; w < w1; w += nw) {
            unsigned n = a.cell_cnt[(int64_t)w * P + p];
            n = n < (unsigned)a.Cw ? n : (unsigned)a.Cw;
            const int64_t *cell = a.cells + ((int64_t)w * P + p) * a.Cw * a.RW;
            constexpr int R = 4;  // records per lane in flight
            for (unsigned i0 = lane; i0 < n; i0 += 64 * R) {
                int64_t key[R], f1[R], f2[R];
                uint64_t slot[R];
                longlong2 sl[R];
#pragma unroll
                for (int q = 0; q < R; ++q) {
                    const unsigned i = i0 + 64 * q;
                    key[q] = f1[q] = f2[q] = 0;
                    if (i < n) {
                        const int64_t *rec = cell + (int64_t)i * a.RW;
                        key[q] = rec[0];
                        if (a.RW > 1) f1[q] = rec[1];
                        if (a.RW > 2) f2[q] = rec[2];
                    }
                }
#pragma unroll
                for (int q = 0; q < R; ++q) {
                    slot[q] = hash64((uint64_t)key[q]) & tmask;
                    sl[q] = make_longlong2(EMPTY_KEY, 0);
                    if (i0 + 64 * q < n) sl[q] = load_slot(tab, slot[q]);
                }
#pragma unroll
                for (int q = 0; q < R; ++q) {
                    if (i0 + 64 * q >= n) continue;
                    int64_t m = 0;
                    longlong2 cur = sl[q];
                    uint64_t si = slot[q];
                    while (cur.x != EMPTY_KEY) {
                        if (cur.x == key[q]) {
                            ++m;
                            if (TYPE == FGPU_INNER || TYPE == FGPU_LEFT) {
                                part_emit(a, j, f1[q], f2[q], true, cur.y);
                                if (a.pa.unique_keys) break;
                            } else {
                                break;
                            }
                        }
                        si = (si + 1) & tmask;
                        cur = load_slot(tab, si);
                    }
                    if ((TYPE == FGPU_LEFT && m == 0) || (TYPE == FGPU_SEMI && m > 0) ||
                        (TYPE == FGPU_ANTI && m == 0))
                        part_emit(a, j, f1[q], f2[q], false, 0);
                }
            }
        }
    }
    acc_finish(a.pa, j, lrow);
}

template <int NPC>
static void launch_scatter_npc(hipStream_t s, const JoinPartArgs &a, int block, size_t lds) {
    switch (a.pa.type) {
        case FGPU_INNER: hipLaunchKernelGGL((k_jp_scatter<FGPU_INNER, NPC>), dim3(a.G), dim3(block), lds, s, a); break;
        case FGPU_LEFT: hipLaunchKernelGGL((k_jp_scatter<FGPU_LEFT, NPC>), dim3(a.G), dim3(block), lds, s, a); break;
        case FGPU_SEMI: hipLaunchKernelGGL((k_jp_scatter<FGPU_SEMI, NPC>), dim3(a.G), dim3(block), lds, s, a); break;
        default: hipLaunchKernelGGL((k_jp_scatter<FGPU_ANTI, NPC>), dim3(a.G), dim3(block), lds, s, a); break;
    }
}

hipError_t launch_join_part_scatter(hipStream_t s, const JoinPartArgs &a, int block) {
    const size_t lds = (size_t)(1 << a.log_p) * 4;
    if (a.pa.pred.n_pcols <= 2) launch_scatter_npc<2>(s, a, block, lds);
    else launch_scatter_npc<MAX_PCOLS>(s, a, block, lds);
    return hipGetLastError();
}

hipError_t launch_join_part_probe(hipStream_t s, const JoinPartArgs &a, int block) {
    const dim3 grid((unsigned)(8 * a.split));  // a.split = workgroups per XCD
    switch (a.pa.type) {
        case FGPU_INNER: hipLaunchKernelGGL(k_jp_probe<FGPU_INNER>, grid, dim3(block), 0, s, a); break;
        case FGPU_LEFT: hipLaunchKernelGGL(k_jp_probe<FGPU_LEFT>, grid, dim3(block), 0, s, a); break;
        case FGPU_SEMI: hipLaunchKernelGGL(k_jp_probe<FGPU_SEMI>, grid, dim3(block), 0, s, a); break;
        default: hipLaunchKernelGGL(k_jp_probe<FGPU_ANTI>, grid, dim3(block), 0, s, a); break;
    }
    return hipGetLastError();
}

// Pair mode.  Pass 0 counts the output rows of each probe row; pass 1 writes
// (probe row, build row) pairs at the scanned offsets.
__global__ void k_join_pairs(JoinProbeArgs a, int pass) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < a.n_probe; base += stride) {
        int64_t rows[1] = {base + threadIdx.x}, key[1];
        bool sel[1], has[1];
        uint64_t slot[1];
        longlong2 first[1];
        probe_tile<1, MAX_PCOLS>(a, rows, sel, has, key, slot, first);
        if (!sel[0]) {
            if (!pass && rows[0] < a.n_probe) a.counts[rows[0]] = 0;
            continue;
        }
        const int64_t r = rows[0];
        uint32_t outn = 0;
        int64_t o = pass ? a.offs[r] : 0;
        int64_t m = 0;
        if (has[0]) {
            if (a.type == FGPU_INNER || a.type == FGPU_LEFT) {
                m = walk(a, key[0], slot[0], first[0], [&](uint32_t br) {
                    if (pass) {
                        a.out_l[o] = (uint32_t)r;
                        a.out_r[o] = br;
                        ++o;
                    }
                    ++outn;
                    return true;
                });
            } else {
                m = walk(a, key[0], slot[0], first[0], [&](uint32_t) { return false; });
            }
        }
        if ((a.type == FGPU_LEFT && m == 0) || (a.type == FGPU_SEMI && m > 0) ||
            (a.type == FGPU_ANTI && m == 0)) {
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
