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
#include <mutex>
#include <vector>

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
            if (o.col >= 0 && side_value(a, o, pr, br, v)) reg_add(o.kind, o.W, o.rscale, v, j.acc[i], j.hi[i], j.lo[i]);
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
                o.sc = jo.sc;
                o.rscale = jo.rscale;
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

// The partitioned join works on h = hash64(key) instead of the key: hash64 is
// a bijection of 64-bit values, so equal hashes are equal keys.  Sub-table
// slots and pass-1 records hold h; the partition is h's top bits.  A sub-table
// is T / 4 buckets of 4 slots; a key's home bucket is given by h's low bits and
// it is stored in the first free slot from the start of that bucket on
// (linear probing that spills into the next buckets only when its own is
// full).  Slots are never freed, so a bucket with a free slot proves a key
// that is not in it is absent.  The one key whose hash is the empty-slot marker EMPTY_KEY
// (the "sentinel key") is kept out of the sub-tables: its build rows go to a
// side list and its probe rows to a side partition.
FGPU_DEV int64_t build_key_at(const JoinBuildArgs &b, int64_t r) {
    return b.key_dtype == FGPU_INT32 ? (int64_t)((const int32_t *)b.keys_col)[r] : ((const int64_t *)b.keys_col)[r];
}

__global__ void k_jp_build(JoinPartArgs a, JoinBuildArgs b) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const uint64_t tmask = (uint64_t)a.T - 1;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < b.n; r += stride) {
        if (b.key_valid && !((b.key_valid[r >> 6] >> (r & 63)) & 1)) continue;  // NULL joins nothing
        const int64_t h = (int64_t)hash64((uint64_t)build_key_at(b, r));
        if (h == EMPTY_KEY) {
            const unsigned long long k = atomicAdd(a.flags + 2, 1ull);
            if ((int64_t)k < a.min_key_cap) a.min_key_rows[k] = (uint32_t)r;
            continue;
        }
        const int p = part_of((uint64_t)h, a.log_p);  // k_jp_count checked every sub-table has room
        const int64_t pay = a.build_pay_col >= 0 ? load_raw(a.pa.bcols, a.build_pay_col, r) : r;
        int64_t *sub = a.sub + (int64_t)p * a.T * 2;
        uint64_t slot = ((uint64_t)h << 2) & tmask;  // first slot of the home bucket (4 slots)
        bool done = false;
        for (int q = 0; q < a.T && !done; ++q) {
            int64_t expect = EMPTY_KEY;
            if (__hip_atomic_compare_exchange_strong(sub + 2 * slot, &expect, h, __ATOMIC_RELAXED,
                                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                sub[2 * slot + 1] = pay;
                done = true;
            } else if (expect == h) {
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
        const uint64_t h = hash64((uint64_t)build_key_at(b, r));
        if ((int64_t)h == EMPTY_KEY) continue;
        atomicAdd(hist + part_of(h, a.log_p), 1u);
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

// Block-wide exclusive scan of hist[0..n) into off[0..n); returns the total.
// Thread t sums a run of ceil(n / blockDim) entries, the runs are scanned
// across the block, then written back.  `tmp` holds blockDim / 64 words of
// LDS.  Every thread must call it.
FGPU_DEV unsigned block_scan_hist(const unsigned *hist, unsigned *off, int n, unsigned *tmp) {
    const int nt = blockDim.x, t = threadIdx.x;
    const int per = (n + nt - 1) / nt;
    const int first = t * per, last = first + per < n ? first + per : n;
    unsigned run = 0;
    for (int i = first; i < last; ++i) run += hist[i];
    const int lane = t & 63, wave = t >> 6, nw = nt >> 6;
    unsigned incl = run;
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned y = __shfl_up(incl, o);
        if (lane >= o) incl += y;
    }
    if (lane == 63) tmp[wave] = incl;
    __syncthreads();
    unsigned before = 0, total = 0;
    for (int i = 0; i < nw; ++i) {
        before += i < wave ? tmp[i] : 0u;
        total += tmp[i];
    }
    unsigned x = before + incl - run;
    for (int i = first; i < last; ++i) {
        off[i] = x;
        x += hist[i];
    }
    __syncthreads();
    return total;
}

// Pass 1: pure partitioning.  Every selected probe row becomes a record (h =
// hash64(key), then the probe payloads) of partition part_of(h); rows with a
// NULL key go to partition P and rows with the sentinel key to partition
// P + 1 (pass 2 gives them their join semantics).  A tile of JS_TPB * U rows is
// counted per partition in LDS, staged in LDS in partition order and written
// to the workgroup's cells in runs: consecutive threads write consecutive
// records of one partition, so stores leave in whole cache lines.  The loads
// of tile t + 1 are issued before tile t is staged (every load is issued, a
// row past the range re-reads the range's last row), so they overlap the
// LDS phases and the stores of tile t.
template <int NPC, int U>
struct ScatterRaw {
    RawCell3 pc[U][NPC > 0 ? NPC : 1];
    RawCell3 key[U];
    RawCell3 f1[U], f2[U];
};

// MODE selects where records go.  JS_CELLS: cell (w, p) of fixed capacity Cw,
// partition-major; a full cell sets flags[0] (once per workgroup) and the host
// redoes the partitioning exactly.  JS_COUNT: nothing is written; the number
// of records of (w, p) goes to cell_cnt[p * G + w].  JS_EXACT: records of
// (w, p) go to a flat array at xoff[p * G + w] (the exclusive prefix sum of
// the JS_COUNT pass), so each partition's records are contiguous.
enum { JS_CELLS = 0, JS_COUNT = 1, JS_EXACT = 2 };

template <int NPC, int RW, int U, int MODE>
__global__ __launch_bounds__(JS_TPB) void k_jp_scatter(JoinPartArgs a) {
    extern __shared__ uint64_t js_lds[];
    __shared__ int overflow;
    constexpr int TILE = JS_TPB * U;
    constexpr bool STAGE = MODE != JS_COUNT;
    const JoinProbeArgs &pa = a.pa;
    const int P2 = (1 << a.log_p) + 2;
    const int Cw = a.Cw, log_p = a.log_p;
    int64_t *stage = (int64_t *)js_lds;                                     // TILE * RW words (staging modes)
    uint32_t *dest = (uint32_t *)(stage + (STAGE ? (size_t)TILE * RW : 0));  // TILE: partition << 17 | position
    unsigned *hist = dest + (STAGE ? TILE : 0);                             // P2
    unsigned *off = hist + P2;                                              // P2
    unsigned *cur = off + P2;                                               // P2: records of this workgroup so far
    unsigned *tmp = cur + P2;                                               // 16
    const ColRef kc = col_ref(pa.pcols, pa.pkey);
    ColRef pc[NPC > 0 ? NPC : 1];
#pragma unroll
    for (int j = 0; j < (NPC > 0 ? NPC : 1); ++j)
        pc[j] = col_ref(pa.pcols, pa.pred.pcol[j < pa.pred.n_pcols ? j : 0]);
    // payload columns have no NULLs (the host checks), so no validity word is read
    const ColRef pay0 = col_ref(pa.pcols, a.pay_col[0]), pay1 = col_ref(pa.pcols, a.pay_col[1]);
    for (int i = threadIdx.x; i < P2; i += blockDim.x) hist[i] = cur[i] = 0;
    if (threadIdx.x == 0) overflow = 0;
    __syncthreads();
    const int64_t r0 = (int64_t)blockIdx.x * a.rows_per_wg;
    const int64_t r1 = r0 + a.rows_per_wg < pa.n_probe ? r0 + a.rows_per_wg : pa.n_probe;
    const int64_t G = a.G;
    auto issue = [&](int64_t base, ScatterRaw<NPC, U> &t) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            int64_t r = base + u * JS_TPB + threadIdx.x;
            r = r < r1 ? r : r1 - 1;
#pragma unroll
            for (int j = 0; j < NPC; ++j) cell_issue3(pc[j], r, t.pc[u][j]);
            cell_issue3(kc, r, t.key[u]);
            if (STAGE && RW > 1) cell_issue3(pay0, r, t.f1[u]);
            if (STAGE && RW > 2) cell_issue3(pay1, r, t.f2[u]);
        }
    };
    ScatterRaw<NPC, U> raw;
    if (r0 < r1) issue(r0, raw);
    for (int64_t base = r0; base < r1; base += TILE) {
        ScatterRaw<NPC, U> nxt;
        issue(base + TILE, nxt);
        int part[U];
        int64_t h[U], f1[U], f2[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t r = base + u * JS_TPB + threadIdx.x;
            f1[u] = f2[u] = 0;
            bool sel = r < r1;
            if (NPC > 0 && pa.pred.n > 0) {
                int64_t v[NPC > 0 ? NPC : 1];
                uint32_t ok = 0;
#pragma unroll
                for (int j = 0; j < NPC; ++j) {
                    bool okj;
                    cell_finish3(pc[j], r, raw.pc[u][j], v[j], okj);
                    ok |= (okj ? 1u : 0u) << j;
                }
                sel = sel && eval_reg<(NPC > 0 ? NPC : 1)>(pa.pred, v, ok);
            }
            int64_t key;
            bool ok, pok;
            cell_finish3(kc, r, raw.key[u], key, ok);
            if (STAGE && RW > 1) cell_finish3(pay0, r, raw.f1[u], f1[u], pok);
            if (STAGE && RW > 2) cell_finish3(pay1, r, raw.f2[u], f2[u], pok);
            h[u] = (int64_t)hash64((uint64_t)key);
            part[u] = !sel ? -1 : !ok ? P2 - 2 : h[u] == EMPTY_KEY ? P2 - 1 : part_of((uint64_t)h[u], log_p);
        }
        if (!STAGE) {
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (part[u] >= 0) atomicAdd(cur + part[u], 1u);
            raw = nxt;
            continue;
        }
        unsigned rank[U];
#pragma unroll
        for (int u = 0; u < U; ++u) rank[u] = part[u] >= 0 ? atomicAdd(hist + part[u], 1u) : 0u;
        __syncthreads();
        const unsigned total = block_scan_hist(hist, off, P2, tmp);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (part[u] < 0) continue;
            const unsigned i = off[part[u]] + rank[u];
            stage[(size_t)i * RW] = h[u];
            if (RW > 1) stage[(size_t)i * RW + 1] = f1[u];
            if (RW > 2) stage[(size_t)i * RW + 2] = f2[u];
            // cells: the position is packed beside the partition (Cw < 2^17);
            // exact mode keeps the partition only (a hot partition can hold
            // more than 2^17 of a workgroup's records)
            dest[i] = MODE == JS_EXACT ? (uint32_t)part[u] : (uint32_t)part[u] << 17 | (cur[part[u]] + rank[u]);
        }
        __syncthreads();
        for (unsigned i = threadIdx.x; i < total; i += blockDim.x) {
            const uint32_t d = dest[i];
            int64_t *rec;
            if (MODE == JS_EXACT) {
                const int64_t pos = (int64_t)cur[d] + (i - off[d]);
                rec = a.cells + (a.xoff[(int64_t)d * G + blockIdx.x] + pos) * RW;
            } else if ((d & 0x1ffffu) < (unsigned)Cw) {
                rec = a.cells + (((int64_t)(d >> 17) * G + blockIdx.x) * Cw + (d & 0x1ffffu)) * RW;
            } else {
                overflow = 1;  // cell full: the host redoes the partitioning exactly
                continue;
            }
            rec[0] = stage[(size_t)i * RW];
            if (RW > 1) rec[1] = stage[(size_t)i * RW + 1];
            if (RW > 2) rec[2] = stage[(size_t)i * RW + 2];
        }
        __syncthreads();
        for (int i = threadIdx.x; i < P2; i += blockDim.x) {
            // cells: positions stay below Cw < 2^17 (they are packed in 17 bits)
            cur[i] = MODE == JS_EXACT ? cur[i] + hist[i] : min(cur[i] + hist[i], (unsigned)Cw);
            hist[i] = 0;
        }
        __syncthreads();
        raw = nxt;
    }
    __syncthreads();
    if (MODE == JS_COUNT) {
        for (int i = threadIdx.x; i < P2; i += blockDim.x) a.cell_cnt[(int64_t)i * G + blockIdx.x] = cur[i];
    } else if (MODE == JS_CELLS) {
        for (int i = threadIdx.x; i < P2; i += blockDim.x) a.cell_cnt[(int64_t)blockIdx.x * P2 + i] = cur[i];
        if (threadIdx.x == 0 && overflow) atomicExch(a.flags, 1ull);
    }
}

// One output row of the partitioned join, general form: probe payloads f1, f2
// from the record, the build value from the sub-table slot (`present` = a
// match; the slot holds the build column itself or the build row).
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
            if (ok) reg_add(o.kind, o.W, o.rscale, v, j.acc[i], j.hi[i], j.lo[i]);
        }
    }
}

// The aggregate ops of a join, resolved into (wave-uniform) registers once
// per workgroup: per op its kind, and where its value comes from (0 = none,
// 1 / 2 = probe payload field, 3 = the build value in the slot).
struct EmitOps {
    int n;
    int kind[JOIN_MAX_OPS];
    int src[JOIN_MAX_OPS];
    int special;  // bit i: op i's column holds Inf / NaN (has_flags), so values are classified
    int wide;     // bit i: op i is a two-word integer sum
    int scaled;   // bit i: op i's float sum scales its values by rscale[i] (sc > 0)
    double rscale[JOIN_MAX_OPS];
};
FGPU_DEV EmitOps emit_ops(const JoinPartArgs &a) {
    EmitOps e;
    e.n = __builtin_amdgcn_readfirstlane(a.pa.n_ops);
    int special = 0;
#pragma unroll
    for (int i = 0; i < JOIN_MAX_OPS; ++i) special |= (i < a.pa.n_ops && a.pa.ops[i].has_flags ? 1 : 0) << i;
    e.special = __builtin_amdgcn_readfirstlane(special);
    int wide = 0, scaled = 0;
#pragma unroll
    for (int i = 0; i < JOIN_MAX_OPS; ++i) {
        wide |= (i < a.pa.n_ops && a.pa.ops[i].kind == OP_ISUM && a.pa.ops[i].W == 2 ? 1 : 0) << i;
        scaled |= (i < a.pa.n_ops && a.pa.ops[i].kind == OP_FSUM && a.pa.ops[i].sc != 0 ? 1 : 0) << i;
        e.rscale[i] = a.pa.ops[i].rscale;
    }
    e.wide = __builtin_amdgcn_readfirstlane(wide);
    e.scaled = __builtin_amdgcn_readfirstlane(scaled);
#pragma unroll
    for (int i = 0; i < JOIN_MAX_OPS; ++i) {
        const JoinOp o = a.pa.ops[i];
        e.kind[i] = __builtin_amdgcn_readfirstlane(o.kind);
        const int src = i >= a.pa.n_ops || o.col < 0 ? 0 : o.side == 0 ? (a.field_of_op[i] == 2 ? 2 : 1) : 3;
        e.src[i] = __builtin_amdgcn_readfirstlane(src);
    }
    return e;
}
// Folds a group of R records per lane, fast form.  The op loop is outside, so
// each (wave-uniform) op kind is dispatched once per group rather than once per
// record, and records enter through selects: one that is not an output row
// adds the identity (0, or +0.0 to the float sum: the high word is never -0.0,
// so hi + 0.0 == hi and the error term is +0.0).  Every accumulator sees its
// records in the order q = 0 .. R-1.
template <int TYPE, int R>
FGPU_DEV void fold_group(const EmitOps &e, JoinAcc &j, const bool (&live)[R], const bool (&hit)[R],
                         const int64_t (&f1)[R], const int64_t (&f2)[R], const int64_t (&pay)[R]) {
    bool row[R], has_pay[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
        row[q] = live[q] && (TYPE == FGPU_INNER ? hit[q] : TYPE == FGPU_LEFT ? true : (TYPE == FGPU_SEMI) == hit[q]);
        has_pay[q] = row[q] && (TYPE == FGPU_INNER || (TYPE == FGPU_LEFT && hit[q]));
        j.rows += row[q] ? 1 : 0;
    }
#pragma unroll
    for (int i = 0; i < JOIN_MAX_OPS; ++i) {
        const int src = e.src[i];
        if (i >= e.n || src == 0) continue;
        bool on[R];
        int64_t v[R];
#pragma unroll
        for (int q = 0; q < R; ++q) {
            on[q] = src == 3 ? has_pay[q] : row[q];
            v[q] = src == 1 ? f1[q] : src == 2 ? f2[q] : pay[q];
        }
        int64_t &acc = j.acc[i];
        switch (e.kind[i]) {
            case OP_CNT:
#pragma unroll
                for (int q = 0; q < R; ++q) acc += on[q] ? 1 : 0;
                break;
            case OP_ISUM:
                if ((e.wide >> i) & 1) {  // two words: low halves in acc, high halves in hi's bits
                    int64_t h = __double_as_longlong(j.hi[i]);
#pragma unroll
                    for (int q = 0; q < R; ++q) {
                        uint64_t dl;
                        int64_t dh;
                        isum_digits(on[q] ? v[q] : 0, dl, dh);
                        acc = (int64_t)((uint64_t)acc + dl);
                        h += dh;
                    }
                    j.hi[i] = __longlong_as_double(h);
                } else {
#pragma unroll
                    for (int q = 0; q < R; ++q) acc = (int64_t)((uint64_t)acc + (on[q] ? (uint64_t)v[q] : 0ull));
                }
                break;
            case OP_IMIN:
#pragma unroll
                for (int q = 0; q < R; ++q) acc = on[q] && v[q] < acc ? v[q] : acc;
                break;
            case OP_IMAX:
#pragma unroll
                for (int q = 0; q < R; ++q) acc = on[q] && v[q] > acc ? v[q] : acc;
                break;
            case OP_FSUM:
                if ((e.scaled >> i) & 1) {  // values near the top of the double range
#pragma unroll
                    for (int q = 0; q < R; ++q) {
                        const uint64_t bits = (uint64_t)v[q];
                        const bool special = ((bits >> 52) & 0x7ff) == 0x7ff;
                        const uint64_t fl = (bits & 0x000fffffffffffffull) ? FL_NAN : (bits >> 63) ? FL_NINF : FL_PINF;
                        acc |= on[q] && special ? (int64_t)fl : 0;
                        const double x = on[q] && !special ? __longlong_as_double(v[q]) * e.rscale[i] : 0.0;
                        double sum, err;
                        two_sum(j.hi[i], x, sum, err);
                        j.hi[i] = sum;
                        j.lo[i] += err;
                    }
                } else if ((e.special >> i) & 1) {
#pragma unroll
                    for (int q = 0; q < R; ++q) {
                        const uint64_t bits = (uint64_t)v[q];
                        const bool special = ((bits >> 52) & 0x7ff) == 0x7ff;
                        const uint64_t fl = (bits & 0x000fffffffffffffull) ? FL_NAN : (bits >> 63) ? FL_NINF : FL_PINF;
                        acc |= on[q] && special ? (int64_t)fl : 0;
                        const double x = on[q] && !special ? __longlong_as_double(v[q]) : 0.0;
                        double sum, err;
                        two_sum(j.hi[i], x, sum, err);
                        j.hi[i] = sum;
                        j.lo[i] += err;
                    }
                } else {  // the column has no Inf / NaN: every value is finite
#pragma unroll
                    for (int q = 0; q < R; ++q) {
                        const double x = on[q] ? __longlong_as_double(v[q]) : 0.0;
                        double sum, err;
                        two_sum(j.hi[i], x, sum, err);
                        j.hi[i] = sum;
                        j.lo[i] += err;
                    }
                }
                break;
            case OP_FMIN:
            case OP_FMAX:
#pragma unroll
                for (int q = 0; q < R; ++q) {
                    const uint64_t bits = (uint64_t)v[q];
                    const bool nan = ((bits >> 52) & 0x7ff) == 0x7ff && (bits & 0x000fffffffffffffull);
                    const uint64_t o = f64_order(bits);
                    const bool take = on[q] && !nan &&
                                      (e.kind[i] == OP_FMIN ? o < (uint64_t)acc : o > (uint64_t)acc);
                    acc = take ? (int64_t)o : acc;
                }
                break;
        }
    }
}

// Records of the two side partitions: NULL keys (matching nothing) and the
// sentinel key, whose hash is EMPTY_KEY (matching the build side list).
template <int TYPE>
FGPU_DEV void probe_side_partitions(const JoinPartArgs &a, JoinAcc &j) {
    const int P = 1 << a.log_p, P2 = P + 2;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int w = wave; w < a.G; w += nw) {
        for (int sp = 0; sp < 2; ++sp) {
            unsigned n = a.cell_cnt[(int64_t)w * P2 + P + sp];
            n = n < (unsigned)a.Cw ? n : (unsigned)a.Cw;
            const int64_t *cell = a.cells + ((int64_t)(P + sp) * a.G + w) * a.Cw * a.RW;
            for (unsigned i = lane; i < n; i += 64) {
                const int64_t *rec = cell + (int64_t)i * a.RW;
                const int64_t f1 = a.RW > 1 ? rec[1] : 0, f2 = a.RW > 2 ? rec[2] : 0;
                int64_t m = 0;
                if (sp == 1) {  // sentinel key: every build row in the side list matches
                    for (int64_t q = 0; q < a.pa.n_min_key_rows; ++q) {
                        ++m;
                        if (TYPE == FGPU_INNER || TYPE == FGPU_LEFT) {
                            int64_t pay = a.pa.min_key_rows[q];
                            if (a.build_pay_col >= 0) pay = load_raw(a.pa.bcols, a.build_pay_col, pay);
                            part_emit(a, j, f1, f2, true, pay);
                        } else {
                            break;
                        }
                    }
                }
                if ((TYPE == FGPU_LEFT && m == 0) || (TYPE == FGPU_SEMI && m > 0) || (TYPE == FGPU_ANTI && m == 0))
                    part_emit(a, j, f1, f2, false, 0);
            }
        }
    }
}

// Side-partition records recs[0..n) (sp 0: NULL keys, 1: the sentinel key),
// shared by all threads of the workgroup.
template <int TYPE>
FGPU_DEV void side_records(const JoinPartArgs &a, JoinAcc &j, const int64_t *recs, int64_t n, int sp) {
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
        const int64_t *rec = recs + i * a.RW;
        const int64_t f1 = a.RW > 1 ? rec[1] : 0, f2 = a.RW > 2 ? rec[2] : 0;
        int64_t m = 0;
        if (sp == 1) {
            for (int64_t q = 0; q < a.pa.n_min_key_rows; ++q) {
                ++m;
                if (TYPE == FGPU_INNER || TYPE == FGPU_LEFT) {
                    int64_t pay = a.pa.min_key_rows[q];
                    if (a.build_pay_col >= 0) pay = load_raw(a.pa.bcols, a.build_pay_col, pay);
                    part_emit(a, j, f1, f2, true, pay);
                } else {
                    break;
                }
            }
        }
        if ((TYPE == FGPU_LEFT && m == 0) || (TYPE == FGPU_SEMI && m > 0) || (TYPE == FGPU_ANTI && m == 0))
            part_emit(a, j, f1, f2, false, 0);
    }
}

// Pass 2.  Workgroup b serves partition p = b % P: it copies the partition's
// sub-table into LDS (slot hashes and build values in two arrays) and the
// record counts of the partition's cells, then streams the records of its
// share of the cells and probes the LDS table.  The only HBM traffic is the
// records, read once and coalesced; the loads of the next group of records
// are in flight while the current one is probed.
//
// FAST (unique build keys, the build value inlined in the slot or not
// needed): each lane first resolves all R of its records -- the first probe
// of every record is issued at once, then only records whose chain goes on
// take further steps -- and folds the group afterwards, one dispatch per
// aggregate op for all R records (fold_group).
// A two-word record as one 16-byte vector (records are read once, non-temporal).
typedef long long RecPair __attribute__((ext_vector_type(2)));

template <int TYPE, bool FAST, int RW>
__global__ __launch_bounds__(1024) void k_jp_probe(JoinPartArgs a) {
    extern __shared__ uint64_t jp_lds[];
    const int P = 1 << a.log_p, P2 = P + 2;
    const int T = a.T, Cw = a.Cw, G = a.G;
    const bool vals = a.need_vals != 0;                     // an aggregate reads the build side
    int64_t *tk = (int64_t *)jp_lds;                        // T slot hashes
    int64_t *tv = tk + T;                                   // T build values / rows (if vals)
    uint64_t *lrow = jp_lds + (vals ? 2 : 1) * (size_t)T;   // 64 words
    unsigned *cnt = (unsigned *)(lrow + 64);                // G: records of each cell (or wave slice)
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    // Cell mode: workgroup b serves partition b % P, cells [w0, w1) of it.
    // Chunk mode (exact partitioning): workgroup b serves chunks[b] = {p, lo,
    // n}, records lo .. lo + n of the flat array; wave w takes slice w of it.
    const bool chunked = a.chunks != nullptr;
    int p, w0, w1;
    int64_t lo = 0, cn = 0, S = 0;
    if (chunked) {
        p = (int)a.chunks[3 * (int64_t)blockIdx.x];
        lo = a.chunks[3 * (int64_t)blockIdx.x + 1];
        cn = a.chunks[3 * (int64_t)blockIdx.x + 2];
        S = (cn + 64 * nw - 1) / (64 * nw) * 64;  // slice length, a multiple of 64
        w0 = 0;
        w1 = nw;
    } else {
        p = blockIdx.x & (P - 1);
        const int k = blockIdx.x >> a.log_p, K = gridDim.x >> a.log_p;
        w0 = (int)((int64_t)G * k / K);
        w1 = (int)((int64_t)G * (k + 1) / K);
    }
    if (p < P) {
        const longlong2 *src = (const longlong2 *)(a.sub + (int64_t)p * T * 2);
        for (int i = threadIdx.x; i < T; i += blockDim.x) {
            const longlong2 sl = src[i];
            tk[i] = sl.x;
            if (vals) tv[i] = sl.y;
        }
    }
    if (chunked) {
        for (int w = threadIdx.x; w < nw; w += blockDim.x) {
            const int64_t rest = cn - (int64_t)w * S;
            cnt[w] = (unsigned)(rest <= 0 ? 0 : rest < S ? rest : S);
        }
    } else {
        for (int w = threadIdx.x; w < G; w += blockDim.x) {
            const unsigned c = a.cell_cnt[(int64_t)w * P2 + p];
            cnt[w] = c < (unsigned)Cw ? c : (unsigned)Cw;
        }
    }
    JoinAcc j;
    acc_init(a.pa, j, lrow);  // synchronises the workgroup (table and counts are in place)
    if (p >= P) {  // chunk of the NULL-key or sentinel-key side partition
        side_records<TYPE>(a, j, a.cells + lo * RW, cn, p - P);
        acc_finish(a.pa, j, lrow);
        return;
    }
    const EmitOps e = emit_ops(a);
    const bool unique = a.pa.unique_keys != 0;
    const uint32_t tmask = (uint32_t)T - 1;
    constexpr int R = 4;  // records per lane per group
    // (w, i0): the wave's current cell and first record of the group
    int w = w0 + wave;
    unsigned i0 = 0, n = w < w1 ? cnt[w] : 0;
    while (w < w1 && n == 0) {
        w += nw;
        n = w < w1 ? cnt[w] : 0;
    }
    int64_t key[R], f1[R], f2[R];
    // Loads are issued unconditionally (a lane past the cell's end, or a wave
    // past its last cell, reads the first record of the partition's cells or
    // chunk, which always exists), so the number of loads in flight is fixed
    // and the wait before probing a group does not also wait for the next
    // group's loads.
    const int64_t *cell0 = chunked ? a.cells + lo * RW : a.cells + (int64_t)p * G * Cw * RW;
    auto load_group = [&](int cw, unsigned ci0, unsigned cn, int64_t (&kk)[R], int64_t (&g1)[R], int64_t (&g2)[R]) {
        const int64_t *cell = cw >= w1 ? cell0 : chunked ? cell0 + (int64_t)cw * S * RW : cell0 + (int64_t)cw * Cw * RW;
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const unsigned i = ci0 + 64 * q + lane;
            const int64_t *rec = i < cn ? cell + (int64_t)i * RW : cell0;
            if (RW == 1) {
                kk[q] = __builtin_nontemporal_load(rec);
                g1[q] = g2[q] = 0;
            } else if (RW == 2) {
                const RecPair w2 = __builtin_nontemporal_load((const RecPair *)rec);
                const longlong2 v = make_longlong2(w2.x, w2.y);
                kk[q] = v.x;
                g1[q] = v.y;
                g2[q] = 0;
            } else {
                kk[q] = rec[0];
                g1[q] = rec[1];
                g2[q] = rec[2];
            }
        }
    };
    load_group(w, i0, n, key, f1, f2);
    while (w < w1) {
        // next group: the rest of this cell, or the next non-empty cell
        int wn = w;
        unsigned in = i0 + 64 * R, nn = n;
        while (wn < w1 && in >= nn) {
            wn += nw;
            in = 0;
            nn = wn < w1 ? cnt[wn] : 0;
        }
        int64_t key2[R], g1[R], g2[R];
        load_group(wn, in, nn, key2, g1, g2);
        // keep every load of the next group ahead of the first use of this one
        __builtin_amdgcn_sched_barrier(0);
        bool live[R];
#pragma unroll
        for (int q = 0; q < R; ++q) live[q] = i0 + 64 * q + lane < n;
        if (FAST) {
            // one 32-byte bucket per record, read as two 16-byte halves in an
            // order that alternates with bucket bit 3, so the lanes of a wave
            // spread over all sixteen 16-byte LDS bank groups
            uint32_t si[R];
            bool hit[R], more[R];
#pragma unroll
            for (int q = 0; q < R; ++q) {
                const uint32_t b0 = ((uint32_t)key[q] << 2) & tmask;  // first slot of the bucket
                const uint32_t rot = (b0 >> 5) & 1;
                longlong2 kv[2];
#pragma unroll
                for (int c = 0; c < 2; ++c) kv[c] = *(const longlong2 *)(tk + b0 + 2 * ((c + rot) & 1));
                uint32_t match = 0, empty = 0;
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    const uint32_t sh = 2 * ((c + rot) & 1);
                    match |= (kv[c].x == key[q] ? 1u : 0u) << sh | (kv[c].y == key[q] ? 1u : 0u) << (sh + 1);
                    empty |= (kv[c].x == EMPTY_KEY ? 1u : 0u) << sh | (kv[c].y == EMPTY_KEY ? 1u : 0u) << (sh + 1);
                }
                hit[q] = match != 0;
                si[q] = b0 + (hit[q] ? (uint32_t)__builtin_ctz(match) : 3u);
                more[q] = live[q] && !hit[q] && empty == 0;  // full bucket without the key: read on
            }
#pragma unroll
            for (int q = 0; q < R; ++q) {
                while (__ballot(more[q])) {  // rare: walk on past a full bucket
                    if (more[q]) {
                        si[q] = (si[q] + 1) & tmask;
                        const int64_t x = tk[si[q]];
                        hit[q] = x == key[q];
                        more[q] = !hit[q] && x != EMPTY_KEY;
                    }
                }
            }
            // the build values of all R records are read before any is folded,
            // so their LDS latencies overlap (si is a slot of the table, hit or not)
            int64_t pay[R];
#pragma unroll
            for (int q = 0; q < R; ++q) pay[q] = (TYPE == FGPU_INNER || TYPE == FGPU_LEFT) && vals ? tv[si[q]] : 0;
            fold_group<TYPE, R>(e, j, live, hit, f1, f2, pay);
        } else {
#pragma unroll
            for (int q = 0; q < R; ++q) {
                if (!live[q]) continue;
                uint32_t si = ((uint32_t)key[q] << 2) & tmask;  // first slot of the home bucket
                int64_t m = 0;
                for (;;) {
                    const int64_t sk = tk[si];
                    if (sk == EMPTY_KEY) break;
                    if (sk == key[q]) {
                        ++m;
                        if (TYPE == FGPU_INNER || TYPE == FGPU_LEFT) {
                            part_emit(a, j, f1[q], f2[q], true, vals ? tv[si] : 0);
                            if (unique) break;
                        } else {
                            break;
                        }
                    }
                    si = (si + 1) & tmask;
                }
                if ((TYPE == FGPU_LEFT && m == 0) || (TYPE == FGPU_SEMI && m > 0) || (TYPE == FGPU_ANTI && m == 0))
                    part_emit(a, j, f1[q], f2[q], false, 0);
            }
        }
#pragma unroll
        for (int q = 0; q < R; ++q) {
            key[q] = key2[q];
            f1[q] = g1[q];
            f2[q] = g2[q];
        }
        w = wn;
        i0 = in;
        n = nn;
    }
    if (!chunked && blockIdx.x == 0) probe_side_partitions<TYPE>(a, j);
    acc_finish(a.pa, j, lrow);
}

static int scatter_npc(const JoinPartArgs &a) { return a.pa.pred.n == 0 ? 0 : a.pa.pred.n_pcols; }
// LDS of pass 1 with `rows` rows per thread: the staged tile (records and
// their destinations) and three per-partition counters.
static size_t scatter_lds_rows(const JoinPartArgs &a, int mode, int rows) {
    const size_t P2 = ((size_t)1 << a.log_p) + 2;
    const size_t tile = (size_t)JS_TPB * rows;
    const size_t staged = mode == JS_COUNT ? 0 : tile * a.RW * 8 + tile * 4;
    return staged + 3 * P2 * 4 + 16 * 4;
}
// Rows per thread per tile.  Counting: 8 when the predicate reads at most one
// column (the double-buffered loads then fit the registers), else 4.  Staging:
// the largest tile whose staging fits in LDS -- a larger tile gives each
// partition longer runs of consecutive records, so more stores leave in whole
// cache lines -- with 12 or 8 rows for one-word records when no predicate
// column is read and 6 or 4 for two-word records when at most one is (the
// registers hold no more), else 8 / 4 as counting.
int join_scatter_rows(const JoinPartArgs &a, int mode) {
    const int npc = scatter_npc(a);
    if (mode == JS_COUNT) return npc <= 1 ? 8 : 4;
    if ((npc == 0 && a.RW == 1) || (npc <= 1 && a.RW == 2)) {
        const int big = a.RW == 1 ? 12 : 6;
        if (scatter_lds_rows(a, mode, big) <= JOIN_LDS_MAX) return big;
    }
    return a.RW == 1 && npc <= 1 ? 8 : 4;
}
size_t join_scatter_lds(const JoinPartArgs &a, int mode) {
    return scatter_lds_rows(a, mode, join_scatter_rows(a, mode));
}
size_t join_probe_lds(const JoinPartArgs &a) {
    return (size_t)a.T * (a.need_vals ? 16 : 8) + 64 * 8 + (size_t)(a.G > 16 ? a.G : 16) * 4;
}

// Lets every pass-1 / pass-2 instantiation take up to JOIN_LDS_MAX bytes of
// dynamic LDS; done once per process (the attribute is per function).
template <bool FAST, int RW>
static void probe_fns(std::vector<const void *> &v) {
    v.push_back((const void *)k_jp_probe<FGPU_INNER, FAST, RW>);
    v.push_back((const void *)k_jp_probe<FGPU_LEFT, FAST, RW>);
    v.push_back((const void *)k_jp_probe<FGPU_SEMI, FAST, RW>);
    v.push_back((const void *)k_jp_probe<FGPU_ANTI, FAST, RW>);
}
template <int NPC, int MODE>
static void scatter_fns(std::vector<const void *> &v) {
    v.push_back((const void *)k_jp_scatter<NPC, 1, NPC <= 1 ? 8 : 4, MODE>);
    v.push_back((const void *)k_jp_scatter<NPC, 2, 4, MODE>);
    v.push_back((const void *)k_jp_scatter<NPC, 3, 4, MODE>);
    if constexpr (NPC == 0) v.push_back((const void *)k_jp_scatter<0, 1, 12, MODE>);
    if constexpr (NPC <= 1) v.push_back((const void *)k_jp_scatter<NPC, 2, 6, MODE>);
}
static hipError_t join_allow_lds_once() {
    static std::once_flag once;
    static hipError_t err = hipSuccess;
    std::call_once(once, [] {
        std::vector<const void *> fns;
        scatter_fns<0, JS_CELLS>(fns);
        scatter_fns<1, JS_CELLS>(fns);
        scatter_fns<2, JS_CELLS>(fns);
        scatter_fns<MAX_PCOLS, JS_CELLS>(fns);
        scatter_fns<0, JS_EXACT>(fns);
        scatter_fns<1, JS_EXACT>(fns);
        scatter_fns<2, JS_EXACT>(fns);
        scatter_fns<MAX_PCOLS, JS_EXACT>(fns);
        fns.push_back((const void *)k_jp_scatter<0, 1, 8, JS_COUNT>);
        fns.push_back((const void *)k_jp_scatter<1, 1, 8, JS_COUNT>);
        fns.push_back((const void *)k_jp_scatter<2, 1, 4, JS_COUNT>);
        fns.push_back((const void *)k_jp_scatter<MAX_PCOLS, 1, 4, JS_COUNT>);
        probe_fns<true, 1>(fns);
        probe_fns<true, 2>(fns);
        probe_fns<true, 3>(fns);
        probe_fns<false, 1>(fns);
        probe_fns<false, 2>(fns);
        probe_fns<false, 3>(fns);
        for (const void *f : fns) {
            const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)JOIN_LDS_MAX);
            if (e != hipSuccess && err == hipSuccess) err = e;
        }
    });
    return err;
}

template <int NPC, int MODE>
static void launch_scatter_npc(hipStream_t s, const JoinPartArgs &a, size_t lds) {
    const dim3 g((unsigned)a.G), b(JS_TPB);
    if constexpr (MODE == JS_COUNT) {
        hipLaunchKernelGGL((k_jp_scatter<NPC, 1, NPC <= 1 ? 8 : 4, JS_COUNT>), g, b, lds, s, a);
        return;
    }
    const int rows = join_scatter_rows(a, MODE);
    if constexpr (NPC == 0 && MODE != JS_COUNT) {
        if (a.RW == 1 && rows == 12) {
            hipLaunchKernelGGL((k_jp_scatter<0, 1, 12, MODE>), g, b, lds, s, a);
            return;
        }
    }
    if constexpr (NPC <= 1 && MODE != JS_COUNT) {
        if (a.RW == 2 && rows == 6) {
            hipLaunchKernelGGL((k_jp_scatter<NPC, 2, 6, MODE>), g, b, lds, s, a);
            return;
        }
    }
    if (a.RW == 1) hipLaunchKernelGGL((k_jp_scatter<NPC, 1, NPC <= 1 ? 8 : 4, MODE>), g, b, lds, s, a);
    else if (a.RW == 2) hipLaunchKernelGGL((k_jp_scatter<NPC, 2, 4, MODE>), g, b, lds, s, a);
    else hipLaunchKernelGGL((k_jp_scatter<NPC, 3, 4, MODE>), g, b, lds, s, a);
}
template <int MODE>
static void launch_scatter_mode(hipStream_t s, const JoinPartArgs &a, size_t lds) {
    const int n = scatter_npc(a);
    if (n == 0) launch_scatter_npc<0, MODE>(s, a, lds);
    else if (n == 1) launch_scatter_npc<1, MODE>(s, a, lds);
    else if (n == 2) launch_scatter_npc<2, MODE>(s, a, lds);
    else launch_scatter_npc<MAX_PCOLS, MODE>(s, a, lds);
}

hipError_t launch_join_part_scatter(hipStream_t s, const JoinPartArgs &a, int mode) {
    const size_t lds = join_scatter_lds(a, mode);
    if (lds > JOIN_LDS_MAX || a.log_p > 14 || a.RW < 1 || a.RW > 3 || a.Cw >= (1 << 17) ||
        (mode == JS_EXACT && !a.xoff))
        return hipErrorInvalidValue;
    if (hipError_t e = join_allow_lds_once()) return e;
    if (mode == JS_COUNT) launch_scatter_mode<JS_COUNT>(s, a, lds);
    else if (mode == JS_EXACT) launch_scatter_mode<JS_EXACT>(s, a, lds);
    else launch_scatter_mode<JS_CELLS>(s, a, lds);
    return hipGetLastError();
}

template <bool FAST, int RW>
static void launch_probe_rw(hipStream_t s, const JoinPartArgs &a, dim3 grid, size_t lds) {
    const dim3 b(1024);
    switch (a.pa.type) {
        case FGPU_INNER: hipLaunchKernelGGL((k_jp_probe<FGPU_INNER, FAST, RW>), grid, b, lds, s, a); break;
        case FGPU_LEFT: hipLaunchKernelGGL((k_jp_probe<FGPU_LEFT, FAST, RW>), grid, b, lds, s, a); break;
        case FGPU_SEMI: hipLaunchKernelGGL((k_jp_probe<FGPU_SEMI, FAST, RW>), grid, b, lds, s, a); break;
        default: hipLaunchKernelGGL((k_jp_probe<FGPU_ANTI, FAST, RW>), grid, b, lds, s, a); break;
    }
}
template <bool FAST>
static void launch_probe_fast(hipStream_t s, const JoinPartArgs &a, dim3 grid, size_t lds) {
    if (a.RW == 1) launch_probe_rw<FAST, 1>(s, a, grid, lds);
    else if (a.RW == 2) launch_probe_rw<FAST, 2>(s, a, grid, lds);
    else launch_probe_rw<FAST, 3>(s, a, grid, lds);
}

hipError_t launch_join_part_probe(hipStream_t s, const JoinPartArgs &a, int n_chunks) {
    const size_t lds = join_probe_lds(a);
    if (lds > JOIN_LDS_MAX || (n_chunks > 0 && !a.chunks) || (n_chunks == 0 && a.chunks)) return hipErrorInvalidValue;
    if (hipError_t e = join_allow_lds_once()) return e;
    // cell mode: a.split workgroups per partition; chunk mode: one per chunk
    const dim3 grid(n_chunks > 0 ? (unsigned)n_chunks : (unsigned)(a.split << a.log_p));
    // fast form: unique build keys, and the build value (if any) inlined in the slots
    bool build_rows = false;
    for (int i = 0; i < a.pa.n_ops; ++i) build_rows |= a.pa.ops[i].side == 1 && a.pa.ops[i].col >= 0;
    const bool fast = a.pa.unique_keys && (!build_rows || a.build_pay_col >= 0);
    if (fast) launch_probe_fast<true>(s, a, grid, lds);
    else launch_probe_fast<false>(s, a, grid, lds);
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
