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
        if (tile_dead(a.pred, base, std::min(base + tile, a.n_probe))) continue;  // zone maps
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

// The partition and record key of build key `key` (not NULL): false for the
// sentinel key of the hash64 form (its rows go to the side list).  In the
// compact form every build key is inside [kmin, kmax] by construction.
FGPU_DEV bool build_part(const JoinPartArgs &a, int64_t key, int &p, int64_t &h) {
    if (a.kw == 4) {
        const uint64_t m = cmix((uint64_t)key - (uint64_t)a.kmin, a.kbits, (1ull << a.kbits) - 1);
        p = (int)(m >> a.rbits);
        h = (int64_t)(m & ((1ull << a.rbits) - 1));
        return true;
    }
    h = (int64_t)hash64((uint64_t)key);
    p = part_of((uint64_t)h, a.log_p);
    return h != EMPTY_KEY;
}

__global__ void k_jp_build(JoinPartArgs a, JoinBuildArgs b) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const uint64_t tmask = (uint64_t)a.T - 1;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < b.n; r += stride) {
        if (b.key_valid && !((b.key_valid[r >> 6] >> (r & 63)) & 1)) continue;  // NULL joins nothing
        int p;
        int64_t h;
        if (!build_part(a, build_key_at(b, r), p, h)) {
            const unsigned long long k = atomicAdd(a.flags + 2, 1ull);
            if ((int64_t)k < a.min_key_cap) a.min_key_rows[k] = (uint32_t)r;
            continue;
        }
        // k_jp_count checked every sub-table has room
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
        int p;
        int64_t h;
        if (!build_part(a, build_key_at(b, r), p, h)) continue;
        atomicAdd(hist + p, 1u);
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

template <int NPC, int RW, int U, int MODE, int KW>
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
    // compact records: the mix's width and the record key's
    const uint64_t kmask = KW == 4 ? (1ull << a.kbits) - 1 : 0, rmask = KW == 4 ? (1ull << a.rbits) - 1 : 0;
    // a row whose key matches nothing (NULL, or outside the compact range) is
    // an output row only for LEFT and ANTI; INNER and SEMI drop it here
    const int nomatch = a.pa.type == FGPU_LEFT || a.pa.type == FGPU_ANTI ? P2 - 2 : -1;
    // issue returns true for a tile the zone maps rule out (nothing loaded)
    auto issue = [&](int64_t base, ScatterRaw<NPC, U> &t) -> bool {
        if (tile_dead(pa.pred, base, base + TILE < r1 ? base + TILE : r1)) {
            t = ScatterRaw<NPC, U>{};
            return true;
        }
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
        return false;
    };
    ScatterRaw<NPC, U> raw;
    bool raw_dead = false;
    if (r0 < r1) raw_dead = issue(r0, raw);
    for (int64_t base = r0; base < r1; base += TILE) {
        ScatterRaw<NPC, U> nxt;
        const bool nxt_dead = issue(base + TILE, nxt);
        int part[U];
        int64_t h[U], f1[U], f2[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t r = base + u * JS_TPB + threadIdx.x;
            f1[u] = f2[u] = 0;
            bool sel = r < r1 && !raw_dead;
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
            if (KW == 4) {  // keys outside the build range match nothing: the NULL-key partition
                const bool inr = key >= a.kmin && key <= a.kmax;
                const uint64_t m = cmix((uint64_t)key - (uint64_t)a.kmin, a.kbits, kmask);
                h[u] = (int64_t)(m & rmask);
                part[u] = !sel ? -1 : !ok || !inr ? nomatch : (int)(m >> a.rbits);
            } else {
                h[u] = (int64_t)hash64((uint64_t)key);
                part[u] = !sel ? -1 : !ok ? nomatch : h[u] == EMPTY_KEY ? P2 - 1 : part_of((uint64_t)h[u], log_p);
            }
        }
        if (!STAGE) {
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (part[u] >= 0) atomicAdd(cur + part[u], 1u);
            raw = nxt;
            raw_dead = nxt_dead;
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
            int64_t ri;  // record index
            if (MODE == JS_EXACT) {
                ri = a.xoff[(int64_t)d * G + blockIdx.x] + (int64_t)cur[d] + (i - off[d]);
            } else if ((d & 0x1ffffu) < (unsigned)Cw) {
                ri = ((int64_t)(d >> 17) * G + blockIdx.x) * Cw + (d & 0x1ffffu);
            } else {
                overflow = 1;  // cell full: the host redoes the partitioning exactly
                continue;
            }
            if (KW == 4) {
                uint32_t *rec = (uint32_t *)a.cells + ri * crec_words(RW, a.pw);
                rec[0] = (uint32_t)stage[(size_t)i * RW];
                if (RW == 2 && a.pw == 4) {
                    rec[1] = (uint32_t)((uint64_t)stage[(size_t)i * RW + 1] - (uint64_t)a.pmin);
                } else {
#pragma unroll
                    for (int f = 1; f < RW; ++f) {
                        const uint64_t x = (uint64_t)stage[(size_t)i * RW + f];
                        rec[2 * f - 1] = (uint32_t)x;
                        rec[2 * f] = (uint32_t)(x >> 32);
                    }
                }
            } else {
                int64_t *rec = a.cells + ri * RW;
                rec[0] = stage[(size_t)i * RW];
                if (RW > 1) rec[1] = stage[(size_t)i * RW + 1];
                if (RW > 2) rec[2] = stage[(size_t)i * RW + 2];
            }
        }
        __syncthreads();
        for (int i = threadIdx.x; i < P2; i += blockDim.x) {
            // cells: positions stay below Cw < 2^17 (they are packed in 17 bits)
            cur[i] = MODE == JS_EXACT ? cur[i] + hist[i] : min(cur[i] + hist[i], (unsigned)Cw);
            hist[i] = 0;
        }
        __syncthreads();
        raw = nxt;
        raw_dead = nxt_dead;
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

// The probe payloads of record ri (either record format).
// A compact record's payload f (0-based): two uint32 words, or (pw = 4) one
// uint32 offset from pmin.
FGPU_DEV int64_t crec_pay(const uint32_t *rec, int f, int pw, int64_t pmin) {
    if (pw == 4) return (int64_t)((uint64_t)pmin + rec[1 + f]);
    return (int64_t)((uint64_t)rec[2 * f + 1] | (uint64_t)rec[2 * f + 2] << 32);
}
FGPU_DEV void rec_pay(const JoinPartArgs &a, int64_t ri, int64_t &f1, int64_t &f2) {
    if (a.kw == 4) {
        const uint32_t *rec = (const uint32_t *)a.cells + ri * crec_words(a.RW, a.pw);
        f1 = a.RW > 1 ? crec_pay(rec, 0, a.pw, a.pmin) : 0;
        f2 = a.RW > 2 ? crec_pay(rec, 1, a.pw, a.pmin) : 0;
        return;
    }
    const int64_t *pp = a.cells + ri * a.RW + 1;
    f1 = a.RW > 1 ? pp[0] : 0;
    f2 = a.RW > 2 ? pp[1] : 0;
}

// One record of a side partition: sp 0 holds NULL keys (and, for compact
// records, keys outside the build range), which match nothing; sp 1 holds the
// sentinel key, whose hash is EMPTY_KEY, which matches the build side list.
template <int TYPE>
FGPU_DEV void side_record(const JoinPartArgs &a, JoinAcc &j, int64_t ri, int sp) {
    int64_t f1, f2;
    rec_pay(a, ri, f1, f2);
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

// Records of the two side partitions, cell mode (one workgroup takes them all).
template <int TYPE>
FGPU_DEV void probe_side_partitions(const JoinPartArgs &a, JoinAcc &j) {
    const int P = 1 << a.log_p, P2 = P + 2;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int w = wave; w < a.G; w += nw) {
        for (int sp = 0; sp < 2; ++sp) {
            unsigned n = a.cell_cnt[(int64_t)w * P2 + P + sp];
            n = n < (unsigned)a.Cw ? n : (unsigned)a.Cw;
            const int64_t cell = ((int64_t)(P + sp) * a.G + w) * a.Cw;  // first record index
            for (unsigned i = lane; i < n; i += 64) side_record<TYPE>(a, j, cell + i, sp);
        }
    }
}

// Side-partition records lo .. lo + n (sp 0: NULL keys, 1: the sentinel key),
// chunk mode, shared by all threads of the workgroup.
template <int TYPE>
FGPU_DEV void side_records(const JoinPartArgs &a, JoinAcc &j, int64_t lo, int64_t n, int sp) {
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) side_record<TYPE>(a, j, lo + i, sp);
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

template <int TYPE, bool FAST, int RW, int KW>
__global__ __launch_bounds__(1024) void k_jp_probe(JoinPartArgs a) {
    extern __shared__ uint64_t jp_lds[];
    const int P = 1 << a.log_p, P2 = P + 2;
    const int T = a.T, Cw = a.Cw, G = a.G;
    const bool vals = a.need_vals != 0;  // an aggregate reads the build side
    // LDS: the T slot keys (int64 hashes, or uint32 compact record keys), then
    // the T build values / rows (if vals), a 64-word state row and the counts
    int64_t *tk = (int64_t *)jp_lds;
    uint32_t *tk32 = (uint32_t *)jp_lds;
    int64_t *tv = KW == 4 ? (int64_t *)(tk32 + T) : tk + T;
    uint64_t *lrow = (uint64_t *)(tv + (vals ? T : 0));
    unsigned *cnt = (unsigned *)(lrow + 64);  // G: records of each cell (or wave slice)
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
            if (KW == 4) tk32[i] = sl.x == EMPTY_KEY ? EMPTY32 : (uint32_t)sl.x;
            else tk[i] = sl.x;
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
        side_records<TYPE>(a, j, lo, cn, p - P);
        acc_finish(a.pa, j, lrow);
        return;
    }
    const EmitOps e = emit_ops(a);
    const bool unique = a.pa.unique_keys != 0;
    const uint32_t tmask = (uint32_t)T - 1;
    // records per lane per group: compact records are 4 (+ 8 per payload)
    // bytes, so twice as many are kept in flight
    constexpr int R = KW == 4 ? 8 : 4;
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
    const int64_t cell0 = chunked ? lo : (int64_t)p * G * Cw;  // record index of the first record
    const uint32_t *k32 = (const uint32_t *)a.cells;
    auto load_group = [&](int cw, unsigned ci0, unsigned cn, int64_t (&kk)[R], int64_t (&g1)[R], int64_t (&g2)[R]) {
        const int64_t cell = cw >= w1 ? cell0 : chunked ? cell0 + (int64_t)cw * S : cell0 + (int64_t)cw * Cw;
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const unsigned i = ci0 + 64 * q + lane;
            const int64_t ri = i < cn ? cell + i : cell0;
            if (KW == 4) {
                const uint32_t *rec = k32 + ri * crec_words(RW, a.pw);
                kk[q] = (int64_t)__builtin_nontemporal_load(rec);
                g1[q] = RW > 1 ? crec_pay(rec, 0, a.pw, a.pmin) : 0;
                g2[q] = RW > 2 ? crec_pay(rec, 1, a.pw, a.pmin) : 0;
            } else if (RW == 1) {
                kk[q] = __builtin_nontemporal_load(a.cells + ri);
                g1[q] = g2[q] = 0;
            } else if (RW == 2) {
                const RecPair w2 = __builtin_nontemporal_load((const RecPair *)(a.cells + ri * 2));
                kk[q] = w2.x;
                g1[q] = w2.y;
                g2[q] = 0;
            } else {
                const int64_t *rec = a.cells + ri * 3;
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
            uint32_t si[R];
            bool hit[R], more[R];
            if (KW == 4) {
                // one 16-byte bucket of four uint32 keys per record
#pragma unroll
                for (int q = 0; q < R; ++q) {
                    const uint32_t rk = (uint32_t)key[q];
                    const uint32_t b0 = (rk << 2) & tmask;  // first slot of the bucket
                    const uint4 kv = *(const uint4 *)(tk32 + b0);
                    const uint32_t match = (kv.x == rk ? 1u : 0u) | (kv.y == rk ? 2u : 0u) |
                                           (kv.z == rk ? 4u : 0u) | (kv.w == rk ? 8u : 0u);
                    const bool full = kv.x != EMPTY32 && kv.y != EMPTY32 && kv.z != EMPTY32 && kv.w != EMPTY32;
                    hit[q] = match != 0;
                    si[q] = b0 + (hit[q] ? (uint32_t)__builtin_ctz(match) : 3u);
                    more[q] = live[q] && !hit[q] && full;  // full bucket without the key: read on
                }
#pragma unroll
                for (int q = 0; q < R; ++q) {
                    while (__ballot(more[q])) {  // rare: walk on past a full bucket
                        if (more[q]) {
                            si[q] = (si[q] + 1) & tmask;
                            const uint32_t x = tk32[si[q]];
                            hit[q] = x == (uint32_t)key[q];
                            more[q] = !hit[q] && x != EMPTY32;
                        }
                    }
                }
            } else {
                // one 32-byte bucket per record, read as two 16-byte halves in an
                // order that alternates with bucket bit 3, so the lanes of a wave
                // spread over all sixteen 16-byte LDS bank groups
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
                    const int64_t sk = KW == 4 ? (tk32[si] == EMPTY32 ? EMPTY_KEY : (int64_t)tk32[si]) : tk[si];
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

// ─── compact records in cells: lean pass 1 and pass 2 ──────────────────────
// The common shape -- compact record keys, a probe key column without NULLs,
// at most one probe payload and one predicate column, cells rather than exact
// partitioning, record indices below 2^32 -- runs these kernels.  Pass 1
// stages 4-byte record keys (so a tile holds twice the rows, and each
// partition's runs are twice as long) and keeps, per partition, the 32-bit
// index of its next record in this workgroup's cell, so a staged record is
// stored with no address arithmetic.  Pass 2 reads four records per lane with
// one 16-byte load, in groups of 256 records (cells hold several hundred) with
// two groups in flight, and looks each record up with one 16-byte LDS read of
// its bucket.
typedef unsigned int JcU4 __attribute__((ext_vector_type(4)));

template <int RW, int NPC, int U, bool K32, int PW>
__global__ __launch_bounds__(JS_TPB) void k_jc_scatter(JoinPartArgs a) {
    constexpr int RU = 1 + (RW - 1) * (PW == 4 ? 1 : 2);  // uint32 words per record
    extern __shared__ uint64_t jc_lds[];
    __shared__ int overflow;
    constexpr int TILE = JS_TPB * U;
    const JoinProbeArgs &pa = a.pa;
    const int P2 = (1 << a.log_p) + 2;
    const unsigned Cw = (unsigned)a.Cw;
    int64_t *spay = (int64_t *)jc_lds;                              // TILE * (RW - 1) payloads
    uint32_t *skey = (uint32_t *)(spay + (size_t)TILE * (RW - 1));  // TILE record keys
    uint32_t *dest = skey + TILE;                                   // TILE record indices (~0: cell full)
    unsigned *hist = dest + TILE;                                   // P2
    unsigned *off = hist + P2;                                      // P2
    unsigned *cur = off + P2;                                       // P2: records of this workgroup so far
    unsigned *cbase = cur + P2;                                     // P2: first record of this workgroup's cell
    unsigned *tmp = cbase + P2;                                     // 16
    // a row whose key matches nothing is an output row only for LEFT and ANTI
    const int nomatch = pa.type == FGPU_LEFT || pa.type == FGPU_ANTI ? P2 - 2 : -1;
    for (int i = threadIdx.x; i < P2; i += blockDim.x) {
        hist[i] = cur[i] = 0;
        cbase[i] = (unsigned)(((int64_t)i * a.G + blockIdx.x) * a.Cw);
    }
    if (threadIdx.x == 0) overflow = 0;
    __syncthreads();
    const int64_t r0 = (int64_t)blockIdx.x * a.rows_per_wg;
    const int64_t r1 = r0 + a.rows_per_wg < pa.n_probe ? r0 + a.rows_per_wg : pa.n_probe;
    const void *kcol = pa.pcols.v[pa.pkey];
    const int64_t *pcol = (const int64_t *)pa.pcols.v[a.pay_col[0]];
    const ColRef pc = col_ref(pa.pcols, pa.pred.pcol[0]);
    const uint64_t kmask = (1ull << a.kbits) - 1, rmask = (1ull << a.rbits) - 1;
    const int64_t kmin = a.kmin, kmax = a.kmax;
    const int kbits = a.kbits, rbits = a.rbits;
    char *recs = (char *)a.cells;  // packed records of RU uint32 words
    const uint64_t pmin = (uint64_t)a.pmin;
    struct Raw {
        int64_t key[U];
        int64_t pay[U];
        RawCell3 pc[U];
    };
    // rows base + u * JS_TPB + tid through uniform bases and 32-bit offsets;
    // a row past the range re-reads the range's last row (every load issued)
    // issue returns true for a tile the zone maps rule out (nothing loaded)
    auto issue = [&](int64_t base, Raw &t) -> bool {
        if (tile_dead(pa.pred, base, base + TILE < r1 ? base + TILE : r1)) {
            t = Raw{};
            return true;
        }
        const int64_t last = r1 - 1 - base;  // may be negative past the end: clamped below
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t want = (int64_t)(u * JS_TPB + threadIdx.x);
            const int64_t r = base + (want < last ? want : last);
            if (K32) t.key[u] = (int64_t)((const int32_t *)kcol)[r];
            else t.key[u] = ((const int64_t *)kcol)[r];
            if (RW > 1) t.pay[u] = pcol[r];
            if (NPC > 0) cell_issue3(pc, r, t.pc[u]);
        }
        return false;
    };
    Raw raw;
    bool raw_dead = false;
    if (r0 < r1) raw_dead = issue(r0, raw);
    for (int64_t base = r0; base < r1; base += TILE) {
        Raw nxt;
        const bool nxt_dead = issue(base + TILE, nxt);
        int part[U];
        uint32_t rec[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t r = base + u * JS_TPB + threadIdx.x;
            bool sel = r < r1 && !raw_dead;
            if (NPC > 0 && pa.pred.n > 0) {
                int64_t v[1];
                bool okj;
                cell_finish3(pc, r, raw.pc[u], v[0], okj);
                sel = sel && eval_reg<1>(pa.pred, v, okj ? 1u : 0u);
            }
            const int64_t key = raw.key[u];
            const bool inr = key >= kmin && key <= kmax;
            const uint64_t m = cmix((uint64_t)key - (uint64_t)kmin, kbits, kmask);
            rec[u] = (uint32_t)(m & rmask);
            part[u] = !sel ? -1 : !inr ? nomatch : (int)(m >> rbits);
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
            const unsigned pos = cur[part[u]] + rank[u];
            skey[i] = rec[u];
            if (RW > 1) spay[i] = raw.pay[u];
            dest[i] = pos < Cw ? cbase[part[u]] + pos : 0xFFFFFFFFu;
        }
        __syncthreads();
        for (unsigned i = threadIdx.x; i < total; i += JS_TPB) {
            const uint32_t d = dest[i];
            if (d == 0xFFFFFFFFu) {
                overflow = 1;  // cell full: the host redoes the partitioning exactly
                continue;
            }
            // byte offsets stay below 2^32 (checked by the host)
            uint32_t *rec = (uint32_t *)(recs + d * (uint32_t)(4 * RU));
            rec[0] = skey[i];
            if (RW > 1 && PW == 4) {
                rec[1] = (uint32_t)((uint64_t)spay[i] - pmin);
            } else if (RW > 1) {
                const uint64_t x = (uint64_t)spay[i];
                rec[1] = (uint32_t)x;
                rec[2] = (uint32_t)(x >> 32);
            }
        }
        __syncthreads();
        for (int i = threadIdx.x; i < P2; i += blockDim.x) {
            cur[i] = min(cur[i] + hist[i], Cw);
            hist[i] = 0;
        }
        __syncthreads();
        raw = nxt;
        raw_dead = nxt_dead;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < P2; i += blockDim.x) a.cell_cnt[(int64_t)blockIdx.x * P2 + i] = cur[i];
    if (threadIdx.x == 0 && overflow) atomicExch(a.flags, 1ull);
}

template <int TYPE, int RW, int PW>
__global__ __launch_bounds__(1024) void k_jc_probe(JoinPartArgs a) {
    constexpr int RU = 1 + (RW - 1) * (PW == 4 ? 1 : 2);  // uint32 words per record
    extern __shared__ uint64_t jp_lds[];
    const int P = 1 << a.log_p, P2 = P + 2;
    const int T = a.T, G = a.G;
    const unsigned Cw = (unsigned)a.Cw;
    const bool vals = a.need_vals != 0;
    uint32_t *tk32 = (uint32_t *)jp_lds;                  // T record keys
    int64_t *tv = (int64_t *)(tk32 + T);                  // T build values / rows (if vals)
    uint64_t *lrow = (uint64_t *)(tv + (vals ? T : 0));   // 64-word state row
    unsigned *cnt = (unsigned *)(lrow + 64);              // G record counts
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int p = blockIdx.x & (P - 1);
    const int k = blockIdx.x >> a.log_p, K = gridDim.x >> a.log_p;
    const int w0 = (int)((int64_t)G * k / K), w1 = (int)((int64_t)G * (k + 1) / K);
    {
        const longlong2 *src = (const longlong2 *)(a.sub + (int64_t)p * T * 2);
        for (int i = threadIdx.x; i < T; i += blockDim.x) {
            const longlong2 sl = src[i];
            tk32[i] = sl.x == EMPTY_KEY ? EMPTY32 : (uint32_t)sl.x;
            if (vals) tv[i] = sl.y;
        }
        for (int w = threadIdx.x; w < G; w += blockDim.x) {
            const unsigned c = a.cell_cnt[(int64_t)w * P2 + p];
            cnt[w] = c < Cw ? c : Cw;
        }
    }
    JoinAcc j;
    acc_init(a.pa, j, lrow);  // synchronises the workgroup (table and counts are in place)
    const EmitOps e = emit_ops(a);
    const uint32_t tmask = (uint32_t)T - 1;
    const unsigned pbase = (unsigned)((int64_t)p * G) * Cw;  // first record of the partition's cells
    const char *recs = (const char *)a.cells;  // packed records of RU uint32 words
    const uint64_t pmin = (uint64_t)a.pmin;
    // A wave's groups: cell w (w0 + wave, then every nw-th), records i0 .. i0 + 256
    // of it; lane l takes records i0 + 4l .. i0 + 4l + 3.
    struct Cur {
        int w;
        unsigned i0, n;
    };
    auto advance = [&](Cur c) {
        c.i0 += 256;
        while (c.w < w1 && c.i0 >= c.n) {
            c.w += nw;
            c.i0 = 0;
            c.n = c.w < w1 ? cnt[c.w] : 0;
        }
        return c;
    };
    struct Grp {
        JcU4 k;
        int64_t f[4];
    };
    // Loads are issued for every group slot (a lane past its cell's end, or
    // past the wave's last cell, reads the partition's first four records,
    // which exist), so the count of loads in flight is fixed.
    auto load = [&](const Cur &c, Grp &g) {
        const unsigned start = c.i0 + 4 * lane;
        const unsigned idx = c.w < w1 && start < c.n ? pbase + (unsigned)c.w * Cw + start : pbase;
        // four records: 16 bytes (keys only) or 48 bytes (keys and one payload),
        // 16-byte aligned as idx is a multiple of 4; byte offsets stay below
        // 2^32 (checked by the host)
        const JcU4 *v = (const JcU4 *)(recs + idx * (uint32_t)(4 * RU));
        if (RW == 1) {
            g.k = __builtin_nontemporal_load(v);
            g.f[0] = g.f[1] = g.f[2] = g.f[3] = 0;
        } else if (PW == 4) {
            // records (key, payload offset) x 4: x = k0 p0 k1 p1, y = k2 p2 k3 p3
            const JcU4 x = __builtin_nontemporal_load(v), y = __builtin_nontemporal_load(v + 1);
            g.k = JcU4{x.x, x.z, y.x, y.z};
            g.f[0] = (int64_t)(pmin + x.y);
            g.f[1] = (int64_t)(pmin + x.w);
            g.f[2] = (int64_t)(pmin + y.y);
            g.f[3] = (int64_t)(pmin + y.w);
        } else {
            const JcU4 x = __builtin_nontemporal_load(v), y = __builtin_nontemporal_load(v + 1),
                       z = __builtin_nontemporal_load(v + 2);
            // records (key, pay_lo, pay_hi) x 4: x = k0 l0 h0 k1, y = l1 h1 k2 l2, z = h2 k3 l3 h3
            g.k = JcU4{x.x, x.w, y.z, z.y};
            g.f[0] = (int64_t)((uint64_t)x.y | (uint64_t)x.z << 32);
            g.f[1] = (int64_t)((uint64_t)y.x | (uint64_t)y.y << 32);
            g.f[2] = (int64_t)((uint64_t)y.w | (uint64_t)z.x << 32);
            g.f[3] = (int64_t)((uint64_t)z.z | (uint64_t)z.w << 32);
        }
    };
    // advancing from "before the wave's first cell" finds its first non-empty cell
    Cur c0 = advance(Cur{w0 + wave - nw, 0u, 0u});
    Cur c1 = advance(c0), c2 = advance(c1);
    Grp g0, g1, g2;
    load(c0, g0);
    load(c1, g1);
    load(c2, g2);
    while (c0.w < w1) {
        const Cur c3 = advance(c2);
        Grp g3;
        load(c3, g3);
        // keep the loads of the three groups ahead in flight while this one is probed
        __builtin_amdgcn_sched_barrier(0);
        const uint32_t rk[4] = {g0.k.x, g0.k.y, g0.k.z, g0.k.w};
        bool live[4], hit[4], more[4];
        uint32_t si[4];
        bool any_more = false;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            live[q] = c0.i0 + 4 * lane + q < c0.n;
            const uint32_t b0 = (rk[q] << 2) & tmask;  // first slot of the bucket
            const JcU4 kv = *(const JcU4 *)(tk32 + b0);
            const bool h0 = kv.x == rk[q], h1 = kv.y == rk[q], h2 = kv.z == rk[q], h3 = kv.w == rk[q];
            hit[q] = h0 || h1 || h2 || h3;
            si[q] = b0 + (h0 ? 0u : h1 ? 1u : h2 ? 2u : 3u);
            const bool full = kv.x != EMPTY32 && kv.y != EMPTY32 && kv.z != EMPTY32 && kv.w != EMPTY32;
            more[q] = live[q] && !hit[q] && full;  // full bucket without the key: read on
            any_more = any_more || more[q];
        }
        if (__ballot(any_more)) {  // rare: walk on past a full bucket
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                while (__ballot(more[q])) {
                    if (more[q]) {
                        si[q] = (si[q] + 1) & tmask;
                        const uint32_t x = tk32[si[q]];
                        hit[q] = x == rk[q];
                        more[q] = !hit[q] && x != EMPTY32;
                    }
                }
            }
        }
        // plain local arrays: fold_group picks among them per op
        int64_t pay[4], f1[4], f2[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            pay[q] = (TYPE == FGPU_INNER || TYPE == FGPU_LEFT) && vals ? tv[si[q]] : 0;
            f1[q] = g0.f[q];
            f2[q] = 0;
        }
        fold_group<TYPE, 4>(e, j, live, hit, f1, f2, pay);
        c0 = c1;
        c1 = c2;
        c2 = c3;
        g0 = g1;
        g1 = g2;
        g2 = g3;
    }
    if (blockIdx.x == 0) probe_side_partitions<TYPE>(a, j);
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
    const size_t slot = (a.kw == 4 ? 4 : 8) + (a.need_vals ? 8 : 0);
    return (size_t)a.T * slot + 64 * 8 + (size_t)(a.G > 16 ? a.G : 16) * 4;
}

// Lets every pass-1 / pass-2 instantiation take up to JOIN_LDS_MAX bytes of
// dynamic LDS; done once per process (the attribute is per function).
template <bool FAST, int RW, int KW>
static void probe_fns(std::vector<const void *> &v) {
    v.push_back((const void *)k_jp_probe<FGPU_INNER, FAST, RW, KW>);
    v.push_back((const void *)k_jp_probe<FGPU_LEFT, FAST, RW, KW>);
    v.push_back((const void *)k_jp_probe<FGPU_SEMI, FAST, RW, KW>);
    v.push_back((const void *)k_jp_probe<FGPU_ANTI, FAST, RW, KW>);
}
template <int NPC, int MODE, int KW>
static void scatter_fns(std::vector<const void *> &v) {
    v.push_back((const void *)k_jp_scatter<NPC, 1, NPC <= 1 ? 8 : 4, MODE, KW>);
    v.push_back((const void *)k_jp_scatter<NPC, 2, 4, MODE, KW>);
    v.push_back((const void *)k_jp_scatter<NPC, 3, 4, MODE, KW>);
    if constexpr (NPC == 0) v.push_back((const void *)k_jp_scatter<0, 1, 12, MODE, KW>);
    if constexpr (NPC <= 1) v.push_back((const void *)k_jp_scatter<NPC, 2, 6, MODE, KW>);
}
template <int KW>
static void join_fns(std::vector<const void *> &fns) {
    scatter_fns<0, JS_CELLS, KW>(fns);
    scatter_fns<1, JS_CELLS, KW>(fns);
    scatter_fns<2, JS_CELLS, KW>(fns);
    scatter_fns<MAX_PCOLS, JS_CELLS, KW>(fns);
    scatter_fns<0, JS_EXACT, KW>(fns);
    scatter_fns<1, JS_EXACT, KW>(fns);
    scatter_fns<2, JS_EXACT, KW>(fns);
    scatter_fns<MAX_PCOLS, JS_EXACT, KW>(fns);
    fns.push_back((const void *)k_jp_scatter<0, 1, 8, JS_COUNT, KW>);
    fns.push_back((const void *)k_jp_scatter<1, 1, 8, JS_COUNT, KW>);
    fns.push_back((const void *)k_jp_scatter<2, 1, 4, JS_COUNT, KW>);
    fns.push_back((const void *)k_jp_scatter<MAX_PCOLS, 1, 4, JS_COUNT, KW>);
    probe_fns<true, 1, KW>(fns);
    probe_fns<true, 2, KW>(fns);
    probe_fns<true, 3, KW>(fns);
    probe_fns<false, 1, KW>(fns);
    probe_fns<false, 2, KW>(fns);
    probe_fns<false, 3, KW>(fns);
}
template <bool K32, int PW>
static void jc_fns2(std::vector<const void *> &v) {
    v.push_back((const void *)k_jc_scatter<2, 0, 8, K32, PW>);
    v.push_back((const void *)k_jc_scatter<2, 0, 4, K32, PW>);
    v.push_back((const void *)k_jc_scatter<2, 1, 6, K32, PW>);
    v.push_back((const void *)k_jc_scatter<2, 1, 4, K32, PW>);
}
template <bool K32>
static void jc_fns(std::vector<const void *> &v) {
    v.push_back((const void *)k_jc_scatter<1, 0, 8, K32, 8>);
    v.push_back((const void *)k_jc_scatter<1, 0, 4, K32, 8>);
    v.push_back((const void *)k_jc_scatter<1, 1, 8, K32, 8>);
    v.push_back((const void *)k_jc_scatter<1, 1, 4, K32, 8>);
    jc_fns2<K32, 4>(v);
    jc_fns2<K32, 8>(v);
}
template <int RW, int PW>
static void jc_probe_fns(std::vector<const void *> &v) {
    v.push_back((const void *)k_jc_probe<FGPU_INNER, RW, PW>);
    v.push_back((const void *)k_jc_probe<FGPU_LEFT, RW, PW>);
    v.push_back((const void *)k_jc_probe<FGPU_SEMI, RW, PW>);
    v.push_back((const void *)k_jc_probe<FGPU_ANTI, RW, PW>);
}
static hipError_t join_allow_lds_once() {
    static std::once_flag once;
    static hipError_t err = hipSuccess;
    std::call_once(once, [] {
        std::vector<const void *> fns;
        join_fns<8>(fns);
        join_fns<4>(fns);
        jc_fns<false>(fns);
        jc_fns<true>(fns);
        jc_probe_fns<1, 8>(fns);
        jc_probe_fns<2, 4>(fns);
        jc_probe_fns<2, 8>(fns);
        for (const void *f : fns) {
            const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)JOIN_LDS_MAX);
            if (e != hipSuccess && err == hipSuccess) err = e;
        }
    });
    return err;
}

template <int NPC, int MODE, int KW>
static void launch_scatter_npc(hipStream_t s, const JoinPartArgs &a, size_t lds) {
    const dim3 g((unsigned)a.G), b(JS_TPB);
    if constexpr (MODE == JS_COUNT) {
        hipLaunchKernelGGL((k_jp_scatter<NPC, 1, NPC <= 1 ? 8 : 4, JS_COUNT, KW>), g, b, lds, s, a);
        return;
    }
    const int rows = join_scatter_rows(a, MODE);
    if constexpr (NPC == 0 && MODE != JS_COUNT) {
        if (a.RW == 1 && rows == 12) {
            hipLaunchKernelGGL((k_jp_scatter<0, 1, 12, MODE, KW>), g, b, lds, s, a);
            return;
        }
    }
    if constexpr (NPC <= 1 && MODE != JS_COUNT) {
        if (a.RW == 2 && rows == 6) {
            hipLaunchKernelGGL((k_jp_scatter<NPC, 2, 6, MODE, KW>), g, b, lds, s, a);
            return;
        }
    }
    if (a.RW == 1) hipLaunchKernelGGL((k_jp_scatter<NPC, 1, NPC <= 1 ? 8 : 4, MODE, KW>), g, b, lds, s, a);
    else if (a.RW == 2) hipLaunchKernelGGL((k_jp_scatter<NPC, 2, 4, MODE, KW>), g, b, lds, s, a);
    else hipLaunchKernelGGL((k_jp_scatter<NPC, 3, 4, MODE, KW>), g, b, lds, s, a);
}
template <int MODE, int KW>
static void launch_scatter_mode(hipStream_t s, const JoinPartArgs &a, size_t lds) {
    const int n = scatter_npc(a);
    if (n == 0) launch_scatter_npc<0, MODE, KW>(s, a, lds);
    else if (n == 1) launch_scatter_npc<1, MODE, KW>(s, a, lds);
    else if (n == 2) launch_scatter_npc<2, MODE, KW>(s, a, lds);
    else launch_scatter_npc<MAX_PCOLS, MODE, KW>(s, a, lds);
}
template <int KW>
static void launch_scatter_kw(hipStream_t s, const JoinPartArgs &a, int mode, size_t lds) {
    if (mode == JS_COUNT) launch_scatter_mode<JS_COUNT, KW>(s, a, lds);
    else if (mode == JS_EXACT) launch_scatter_mode<JS_EXACT, KW>(s, a, lds);
    else launch_scatter_mode<JS_CELLS, KW>(s, a, lds);
}

// A compact record's key must leave the empty marker EMPTY32 free and its
// partition must take at least one bit of the mix.
static bool join_format_ok(const JoinPartArgs &a) {
    if (a.kw != 4) return true;  // hash64 records
    return a.kbits >= 1 && a.kbits <= 62 && a.rbits >= 1 && a.rbits <= 31 && a.rbits + a.log_p == a.kbits;
}

// The lean compact kernels' shape (see k_jc_scatter): compact records, a key
// column without NULLs, at most one 8-byte payload and one predicate column,
// and record byte offsets below 2^32.
static bool jc_records_ok(const JoinPartArgs &a) {
    const int64_t recs = (int64_t)a.G * ((1 << a.log_p) + 2) * a.Cw;
    return a.kw == 4 && a.RW <= 2 && recs * 4 * crec_words(a.RW, a.pw) < (int64_t(1) << 32);
}
static size_t jc_scatter_lds(const JoinPartArgs &a, int u) {
    const size_t tile = (size_t)JS_TPB * u, P2 = ((size_t)1 << a.log_p) + 2;
    return tile * (a.RW - 1) * 8 + tile * 8 + 4 * P2 * 4 + 16 * 4;
}
// Rows per thread of the lean pass 1 (0: the general kernel runs): the largest
// tile the registers hold (8 rows, 6 with both a payload and a predicate
// column) whose staging fits the LDS budget.
static int jc_scatter_rows(const JoinPartArgs &a) {
    if (!jc_records_ok(a) || scatter_npc(a) > 1) return 0;
    const JoinProbeArgs &pa = a.pa;
    const int kd = pa.pcols.dtype[pa.pkey];
    if (pa.pcols.valid[pa.pkey] || (kd != FGPU_INT64 && kd != FGPU_INT32)) return 0;
    if (a.RW > 1) {
        const int pd = pa.pcols.dtype[a.pay_col[0]];
        if (pd != FGPU_INT64 && pd != FGPU_FLOAT64) return 0;
    }
    const size_t budget = std::min<size_t>(JOIN_LDS_MAX, a.lds_budget > 0 ? (size_t)a.lds_budget : JOIN_LDS_MAX);
    const int big = a.RW > 1 && scatter_npc(a) > 0 ? 6 : 8;
    for (int u : {big, 4})
        if (jc_scatter_lds(a, u) <= budget) return u;
    return 0;
}
template <int RW, int NPC, bool K32, int PW>
static void launch_jc_scatter_u(hipStream_t s, const JoinPartArgs &a, int u) {
    const dim3 g((unsigned)a.G), b(JS_TPB);
    const size_t lds = jc_scatter_lds(a, u);
    constexpr int BIG = RW > 1 && NPC > 0 ? 6 : 8;
    if (u == BIG) hipLaunchKernelGGL((k_jc_scatter<RW, NPC, BIG, K32, PW>), g, b, lds, s, a);
    else hipLaunchKernelGGL((k_jc_scatter<RW, NPC, 4, K32, PW>), g, b, lds, s, a);
}
template <bool K32>
static void launch_jc_scatter(hipStream_t s, const JoinPartArgs &a, int u) {
    const int npc = scatter_npc(a);
    if (a.RW == 1) {
        if (npc == 0) launch_jc_scatter_u<1, 0, K32, 8>(s, a, u);
        else launch_jc_scatter_u<1, 1, K32, 8>(s, a, u);
    } else if (a.pw == 4) {
        if (npc == 0) launch_jc_scatter_u<2, 0, K32, 4>(s, a, u);
        else launch_jc_scatter_u<2, 1, K32, 4>(s, a, u);
    } else {
        if (npc == 0) launch_jc_scatter_u<2, 0, K32, 8>(s, a, u);
        else launch_jc_scatter_u<2, 1, K32, 8>(s, a, u);
    }
}

hipError_t launch_join_part_scatter(hipStream_t s, const JoinPartArgs &a, int mode) {
    if (mode == JS_CELLS && a.kw == 4 && a.Cw < (1 << 17) && join_format_ok(a)) {
        if (const int u = jc_scatter_rows(a)) {
            if (hipError_t e = join_allow_lds_once()) return e;
            if (a.pa.pcols.dtype[a.pa.pkey] == FGPU_INT32) launch_jc_scatter<true>(s, a, u);
            else launch_jc_scatter<false>(s, a, u);
            return hipGetLastError();
        }
    }
    const size_t lds = join_scatter_lds(a, mode);
    if (lds > JOIN_LDS_MAX || a.log_p > 14 || a.RW < 1 || a.RW > 3 || a.Cw >= (1 << 17) ||
        (mode == JS_EXACT && !a.xoff) || !join_format_ok(a))
        return hipErrorInvalidValue;
    if (hipError_t e = join_allow_lds_once()) return e;
    if (a.kw == 4) launch_scatter_kw<4>(s, a, mode, lds);
    else launch_scatter_kw<8>(s, a, mode, lds);
    return hipGetLastError();
}

template <bool FAST, int RW, int KW>
static void launch_probe_rw(hipStream_t s, const JoinPartArgs &a, dim3 grid, size_t lds) {
    const dim3 b(1024);
    switch (a.pa.type) {
        case FGPU_INNER: hipLaunchKernelGGL((k_jp_probe<FGPU_INNER, FAST, RW, KW>), grid, b, lds, s, a); break;
        case FGPU_LEFT: hipLaunchKernelGGL((k_jp_probe<FGPU_LEFT, FAST, RW, KW>), grid, b, lds, s, a); break;
        case FGPU_SEMI: hipLaunchKernelGGL((k_jp_probe<FGPU_SEMI, FAST, RW, KW>), grid, b, lds, s, a); break;
        default: hipLaunchKernelGGL((k_jp_probe<FGPU_ANTI, FAST, RW, KW>), grid, b, lds, s, a); break;
    }
}
template <bool FAST, int KW>
static void launch_probe_fast(hipStream_t s, const JoinPartArgs &a, dim3 grid, size_t lds) {
    if (a.RW == 1) launch_probe_rw<FAST, 1, KW>(s, a, grid, lds);
    else if (a.RW == 2) launch_probe_rw<FAST, 2, KW>(s, a, grid, lds);
    else launch_probe_rw<FAST, 3, KW>(s, a, grid, lds);
}

hipError_t launch_join_part_probe(hipStream_t s, const JoinPartArgs &a, int n_chunks) {
    const size_t lds = join_probe_lds(a);
    if (lds > JOIN_LDS_MAX || (n_chunks > 0 && !a.chunks) || (n_chunks == 0 && a.chunks) || !join_format_ok(a))
        return hipErrorInvalidValue;
    if (hipError_t e = join_allow_lds_once()) return e;
    // cell mode: a.split workgroups per partition; chunk mode: one per chunk
    const dim3 grid(n_chunks > 0 ? (unsigned)n_chunks : (unsigned)(a.split << a.log_p));
    // fast form: unique build keys, and the build value (if any) inlined in the slots
    bool build_rows = false;
    for (int i = 0; i < a.pa.n_ops; ++i) build_rows |= a.pa.ops[i].side == 1 && a.pa.ops[i].col >= 0;
    const bool fast = a.pa.unique_keys && (!build_rows || a.build_pay_col >= 0);
    if (a.kw == 4 && fast && n_chunks == 0 && jc_records_ok(a)) {  // the lean compact pass 2
        const dim3 b(1024);
        const int shape = a.RW == 1 ? 0 : a.pw == 4 ? 1 : 2;  // key only, 4-byte payload, 8-byte payload
        switch (a.pa.type * 4 + shape) {
            case FGPU_INNER * 4 + 0: hipLaunchKernelGGL((k_jc_probe<FGPU_INNER, 1, 8>), grid, b, lds, s, a); break;
            case FGPU_LEFT * 4 + 0: hipLaunchKernelGGL((k_jc_probe<FGPU_LEFT, 1, 8>), grid, b, lds, s, a); break;
            case FGPU_SEMI * 4 + 0: hipLaunchKernelGGL((k_jc_probe<FGPU_SEMI, 1, 8>), grid, b, lds, s, a); break;
            case FGPU_ANTI * 4 + 0: hipLaunchKernelGGL((k_jc_probe<FGPU_ANTI, 1, 8>), grid, b, lds, s, a); break;
            case FGPU_INNER * 4 + 1: hipLaunchKernelGGL((k_jc_probe<FGPU_INNER, 2, 4>), grid, b, lds, s, a); break;
            case FGPU_LEFT * 4 + 1: hipLaunchKernelGGL((k_jc_probe<FGPU_LEFT, 2, 4>), grid, b, lds, s, a); break;
            case FGPU_SEMI * 4 + 1: hipLaunchKernelGGL((k_jc_probe<FGPU_SEMI, 2, 4>), grid, b, lds, s, a); break;
            case FGPU_ANTI * 4 + 1: hipLaunchKernelGGL((k_jc_probe<FGPU_ANTI, 2, 4>), grid, b, lds, s, a); break;
            case FGPU_INNER * 4 + 2: hipLaunchKernelGGL((k_jc_probe<FGPU_INNER, 2, 8>), grid, b, lds, s, a); break;
            case FGPU_LEFT * 4 + 2: hipLaunchKernelGGL((k_jc_probe<FGPU_LEFT, 2, 8>), grid, b, lds, s, a); break;
            case FGPU_SEMI * 4 + 2: hipLaunchKernelGGL((k_jc_probe<FGPU_SEMI, 2, 8>), grid, b, lds, s, a); break;
            default: hipLaunchKernelGGL((k_jc_probe<FGPU_ANTI, 2, 8>), grid, b, lds, s, a); break;
        }
    } else if (a.kw == 4) {
        if (fast) launch_probe_fast<true, 4>(s, a, grid, lds);
        else launch_probe_fast<false, 4>(s, a, grid, lds);
    } else {
        if (fast) launch_probe_fast<true, 8>(s, a, grid, lds);
        else launch_probe_fast<false, 8>(s, a, grid, lds);
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
