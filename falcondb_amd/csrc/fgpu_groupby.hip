// fgpu_groupby.hip — fused WHERE + GROUP BY aggregation for gfx950.
//
// Replaces the reference's execute_select -> GroupByExec::execute pipeline
// (src/query/select.rs:105-160; src/query/aggregate.rs:535-1420): instead of
// materialising the filtered columns and building per-batch hash maps that a
// single thread merges, one kernel streams every row once, evaluates the WHERE
// predicate in registers, and folds the surviving rows into group states.
//
// Strategies (chosen on the host from upload-time column statistics):
//   GB_DENSE_LDS    key range (max-min+1, or dictionary size) small enough that
//                   every group's state fits in LDS: each workgroup aggregates
//                   privately in LDS and merges once into HBM.
//   GB_DENSE_GLOBAL key range moderate: state indexed by key - min in HBM,
//                   updated with device-scope atomics (no hashing, no CAS).
//   GB_HASH         open-addressing table in HBM (64-bit CAS on the key slot,
//                   linear probing) behind a per-workgroup LDS cache that absorbs
//                   the frequent keys (Zipf heavy hitters) before they reach HBM.
//   GB_SCALAR       no GROUP BY: per-thread register accumulators (exact integer,
//                   double-double float) reduced once per workgroup.
//
// Group state = S 64-bit words; word 0 counts rows; the DOp list says how each
// aggregate input updates its words.  Float SUM is exact (fixed-point words),
// so results do not depend on the order in which rows arrive.
#include "fgpu_state.h"

namespace fgpu {

// Loads the aggregate inputs of row r.
FGPU_DEV Inputs load_inputs(const GbArgs &a, int64_t r) {
    Inputs in{0, 0, 0, 0, 0, 0, 0};
    auto one = [&](int j, int64_t &dst) {
        if (j < a.n_in) {
            const int k = a.in_col[j];
            dst = load_raw(a.cols, k, r);
            if (load_valid(a.cols, k, r)) in.valid |= 1u << j;
        }
    };
    one(0, in.v0);
    one(1, in.v1);
    one(2, in.v2);
    one(3, in.v3);
    one(4, in.v4);
    one(5, in.v5);
    return in;
}

// Key of row r; returns false for the NULL group (int key NULL or INT64_MIN,
// which the reference folds together, src/query/aggregate.rs:867, :1207).
FGPU_DEV bool load_key(const GbArgs &a, int64_t r, int64_t &key) {
    if (a.key_is_utf8) {
        const uint32_t code = ((const uint32_t *)a.cols.v[a.key_col])[r];
        key = code == NULL_CODE ? (int64_t)a.utf8_null_code : (int64_t)code;
        return true;
    }
    if (!load_valid(a.cols, a.key_col, r)) return false;
    key = load_raw(a.cols, a.key_col, r);
    return key != EMPTY_KEY;
}

// Find-or-insert in the HBM table; -1 if the table is full.
FGPU_DEV int64_t hbm_slot(const GbArgs &a, int64_t key, uint64_t h) {
    uint64_t i = h & a.cap_mask;
    for (uint64_t p = 0; p <= a.cap_mask; ++p) {
        int64_t k = __hip_atomic_load(a.keys + i, __ATOMIC_RELAXED, DEV_SCOPE);
        if (k == key) return (int64_t)i;
        if (k == EMPTY_KEY) {
            int64_t expect = EMPTY_KEY;
            if (__hip_atomic_compare_exchange_strong(a.keys + i, &expect, key, __ATOMIC_RELAXED,
                                                     __ATOMIC_RELAXED, DEV_SCOPE))
                return (int64_t)i;
            if (expect == key) return (int64_t)i;
        }
        i = (i + 1) & a.cap_mask;
    }
    return -1;
}

// Find-or-insert in the LDS cache with a bounded number of probes; -1 = miss.
FGPU_DEV int lds_slot(int64_t *lkeys, int lmask, int limit, int64_t key, uint64_t h) {
    int i = (int)(h >> 40) & lmask;  // different hash bits than the HBM table
    for (int p = 0; p < limit; ++p) {
        int64_t k = __hip_atomic_load(lkeys + i, __ATOMIC_RELAXED, LDS_SCOPE);
        if (k == key) return i;
        if (k == EMPTY_KEY) {
            int64_t expect = EMPTY_KEY;
            if (__hip_atomic_compare_exchange_strong(lkeys + i, &expect, key, __ATOMIC_RELAXED,
                                                     __ATOMIC_RELAXED, LDS_SCOPE))
                return i;
            if (expect == key) return i;
        }
        i = (i + 1) & lmask;
    }
    return -1;
}

constexpr int GB_UNROLL = 4;  // rows per thread per tile (independent loads in flight)

// ─── dense keys, LDS-private partial states ─────────────────────────────────
template <int NPC>
__global__ void k_gb_dense_lds(GbArgs a) {
    extern __shared__ uint64_t lds[];
    const int S = a.ops.S;
    const int64_t groups = a.range + 1;  // + NULL group
    fill_ident(lds, groups, S, a.ops);
    __syncthreads();
    const GbCols gc = gb_cols(a);
    const int64_t tile = (int64_t)blockDim.x * GB_UNROLL;
    for (int64_t base = (int64_t)blockIdx.x * tile; base < a.n_rows; base += (int64_t)gridDim.x * tile) {
        if (tile_dead(a.pred, base, std::min(base + tile, a.n_rows))) continue;  // zone maps
        int64_t rows[GB_UNROLL];
        bool pass[GB_UNROLL];
#pragma unroll
        for (int u = 0; u < GB_UNROLL; ++u) rows[u] = base + u * blockDim.x + threadIdx.x;
        pred_tile<GB_UNROLL, NPC>(a.pred, a.cols, rows, a.n_rows, pass);
        // stage B: key and aggregate inputs of every passing row, all issued
        // before any is examined
        GbRow x[GB_UNROLL];
#pragma unroll
        for (int u = 0; u < GB_UNROLL; ++u) {
            x[u] = GbRow{};
            if (pass[u]) gb_issue(a, gc, rows[u], x[u]);
        }
        // the tile's rows folded op-major: each op's descriptor read once
        uint64_t *st[GB_UNROLL];
        Inputs in[GB_UNROLL];
#pragma unroll
        for (int u = 0; u < GB_UNROLL; ++u) {
            int64_t key = 0;
            in[u] = Inputs{0, 0, 0, 0, 0, 0, 0};
            const bool has_key = pass[u] && gb_finish(a, gc, rows[u], x[u], key, in[u]);
            st[u] = lds + (has_key ? (key - a.key_min) : a.range) * S;
        }
        update_rows<LDS_SCOPE, GB_UNROLL>(st, pass, a.ops, in);
    }
    __syncthreads();
    for (int64_t g = threadIdx.x; g < groups; g += blockDim.x)
        if (lds[g * S]) merge_state(a.state + (int64_t)(blockIdx.x % a.replicas) * a.rep_stride + g * S, lds + g * S, a.ops);
}

// ─── dense keys, HBM states ─────────────────────────────────────────────────
template <int NPC>
__global__ void k_gb_dense_global(GbArgs a) {
    const int S = a.ops.S;
    const int64_t tile = (int64_t)blockDim.x * GB_UNROLL;
    for (int64_t base = (int64_t)blockIdx.x * tile; base < a.n_rows; base += (int64_t)gridDim.x * tile) {
        if (tile_dead(a.pred, base, std::min(base + tile, a.n_rows))) continue;  // zone maps
        int64_t rows[GB_UNROLL];
        bool pass[GB_UNROLL];
#pragma unroll
        for (int u = 0; u < GB_UNROLL; ++u) rows[u] = base + u * blockDim.x + threadIdx.x;
        pred_tile<GB_UNROLL, NPC>(a.pred, a.cols, rows, a.n_rows, pass);
        // stage B: key and aggregate inputs of every passing row, loads first
        int64_t key[GB_UNROLL];
        bool has_key[GB_UNROLL];
        Inputs in[GB_UNROLL];
#pragma unroll
        for (int u = 0; u < GB_UNROLL; ++u) {
            key[u] = 0;
            has_key[u] = false;
            in[u] = Inputs{0, 0, 0, 0, 0, 0, 0};
            if (pass[u]) {
                if (a.key_col >= 0) has_key[u] = load_key(a, rows[u], key[u]);
                in[u] = load_inputs(a, rows[u]);
            }
        }
        uint64_t *st[GB_UNROLL];
#pragma unroll
        for (int u = 0; u < GB_UNROLL; ++u) st[u] = a.state + (has_key[u] ? (key[u] - a.key_min) : a.range) * S;
        update_rows<DEV_SCOPE, GB_UNROLL>(st, pass, a.ops, in);
    }
}

// ─── hash table in HBM behind an LDS cache ──────────────────────────────────
template <int NPC>
__global__ void k_gb_hash(GbArgs a) {
    extern __shared__ uint64_t lds[];
    const int S = a.ops.S;
    const int ls = a.lds_slots;
    int64_t *lkeys = (int64_t *)lds;
    uint64_t *lstate = lds + ls;
    for (int i = threadIdx.x; i < ls; i += blockDim.x) lkeys[i] = EMPTY_KEY;
    fill_ident(lstate, ls, S, a.ops);
    __syncthreads();
    const int64_t tile = (int64_t)blockDim.x * GB_UNROLL;
    for (int64_t base = (int64_t)blockIdx.x * tile; base < a.n_rows; base += (int64_t)gridDim.x * tile) {
        if (tile_dead(a.pred, base, std::min(base + tile, a.n_rows))) continue;  // zone maps
        int64_t rows[GB_UNROLL];
        bool pass[GB_UNROLL];
#pragma unroll
        for (int u = 0; u < GB_UNROLL; ++u) rows[u] = base + u * blockDim.x + threadIdx.x;
        pred_tile<GB_UNROLL, NPC>(a.pred, a.cols, rows, a.n_rows, pass);
        // stage B: key and aggregate inputs of every passing row, loads first
        int64_t key[GB_UNROLL];
        bool has_key[GB_UNROLL];
        Inputs in[GB_UNROLL];
#pragma unroll
        for (int u = 0; u < GB_UNROLL; ++u) {
            key[u] = 0;
            has_key[u] = false;
            in[u] = Inputs{0, 0, 0, 0, 0, 0, 0};
            if (pass[u]) {
                if (a.key_col >= 0) has_key[u] = load_key(a, rows[u], key[u]);
                in[u] = load_inputs(a, rows[u]);
            }
        }
#pragma unroll
        for (int u = 0; u < GB_UNROLL; ++u) {
            if (!pass[u]) continue;
            if (!has_key[u]) {  // NULL group: state row cap
                update_row<DEV_SCOPE>(a.state + (int64_t)(a.cap_mask + 1) * S, a.ops, in[u]);
                continue;
            }
            const uint64_t h = hash64((uint64_t)key[u]);
            const int li = ls ? lds_slot(lkeys, ls - 1, a.probe_limit, key[u], h) : -1;
            if (li >= 0) {
                update_row<LDS_SCOPE>(lstate + (int64_t)li * S, a.ops, in[u]);
            } else {
                const int64_t gi = hbm_slot(a, key[u], h);
                if (gi < 0) {  // table full: the host retries with a larger one
                    atomicExch(a.counters + 1, 1ull);
                    continue;
                }
                update_row<DEV_SCOPE>(a.state + gi * S, a.ops, in[u]);
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < ls; i += blockDim.x) {
        const int64_t key = lkeys[i];
        if (key == EMPTY_KEY || lstate[(int64_t)i * S] == 0) continue;
        const int64_t gi = hbm_slot(a, key, hash64((uint64_t)key));
        if (gi < 0) {
            atomicExch(a.counters + 1, 1ull);
            continue;
        }
        merge_state(a.state + gi * S, lstate + (int64_t)i * S, a.ops);
    }
}

// ─── no GROUP BY: register accumulators ─────────────────────────────────────
// Each thread folds its rows into registers (exact integer sums; double-double
// for float sums), then the workgroup merges them exactly into an LDS state
// row, which one pass of atomics adds to the HBM state row.
constexpr int SCALAR_OPS = 8;  // register accumulators per thread (the host checks)

template <int NPC>
__global__ void k_gb_scalar(GbArgs a) {
    __shared__ uint64_t lrow[64];
    const int S = a.ops.S;
    for (int i = threadIdx.x; i < S; i += blockDim.x) lrow[i] = a.ops.ident[i];
    __syncthreads();
    int64_t n_out = 0;
    int64_t acc[SCALAR_OPS];
    double hi[SCALAR_OPS], lo[SCALAR_OPS];
#pragma unroll
    for (int i = 0; i < SCALAR_OPS; ++i) {
        acc[i] = i < a.ops.n ? reg_ident(a.ops.op[i].kind) : 0;
        hi[i] = 0.0;
        lo[i] = 0.0;
    }
    const int64_t tile = (int64_t)blockDim.x * GB_UNROLL;
    for (int64_t base = (int64_t)blockIdx.x * tile; base < a.n_rows; base += (int64_t)gridDim.x * tile) {
        if (tile_dead(a.pred, base, std::min(base + tile, a.n_rows))) continue;  // zone maps
        int64_t rows[GB_UNROLL];
        bool pass[GB_UNROLL];
#pragma unroll
        for (int u = 0; u < GB_UNROLL; ++u) rows[u] = base + u * blockDim.x + threadIdx.x;
        pred_tile<GB_UNROLL, NPC>(a.pred, a.cols, rows, a.n_rows, pass);
        // stage B: key and aggregate inputs of every passing row, loads first
        int64_t key[GB_UNROLL];
        bool has_key[GB_UNROLL];
        Inputs in[GB_UNROLL];
#pragma unroll
        for (int u = 0; u < GB_UNROLL; ++u) {
            key[u] = 0;
            has_key[u] = false;
            in[u] = Inputs{0, 0, 0, 0, 0, 0, 0};
            if (pass[u]) {
                if (a.key_col >= 0) has_key[u] = load_key(a, rows[u], key[u]);
                in[u] = load_inputs(a, rows[u]);
            }
        }
#pragma unroll
        for (int u = 0; u < GB_UNROLL; ++u) {
            if (!pass[u]) continue;
            ++n_out;
#pragma unroll
            for (int i = 0; i < SCALAR_OPS; ++i) {
                if (i < a.ops.n) {
                    const DOp o = a.ops.op[i];
                    if ((in[u].valid >> o.col) & 1) reg_add(o, pick(in[u], o.col), acc[i], hi[i], lo[i]);
                }
            }
        }
    }
    // fold registers into the LDS row
    if (n_out) a_add<LDS_SCOPE>(lrow, (uint64_t)n_out);
#pragma unroll
    for (int i = 0; i < SCALAR_OPS; ++i)
        if (i < a.ops.n) fold_reg(lrow, a.ops.op[i], acc[i], hi[i], lo[i]);
    __syncthreads();
    if (threadIdx.x == 0) merge_state(a.state + (int64_t)(blockIdx.x % a.replicas) * a.rep_stride, lrow, a.ops);
}

// Folds replicas 1.. of the dense-LDS / scalar state rows into replica 0, one
// thread per group; a replica row with a zero row count holds identities only.
__global__ void k_gb_fold(GbArgs a, int64_t n_groups) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n_groups) return;
    const int S = a.ops.S;
    for (int r = 1; r < a.replicas; ++r) {
        const uint64_t *src = a.state + (int64_t)r * a.rep_stride + g * S;
        if (src[0]) merge_state(a.state + g * S, src, a.ops);
    }
}

// ─── emit: occupied groups -> output columns ────────────────────────────────
// Each workgroup covers EMIT_PER_THREAD * blockDim.x slots and reserves its
// output rows with one atomic.
constexpr int EMIT_PER_THREAD = 16;

__global__ void k_gb_emit(EmitArgs a) {
    const int64_t first = (int64_t)blockIdx.x * blockDim.x * EMIT_PER_THREAD + threadIdx.x;
    uint32_t have = 0;
#pragma unroll
    for (int k = 0; k < EMIT_PER_THREAD; ++k) {
        const int64_t i = first + (int64_t)k * blockDim.x;
        if (i >= a.n_slots) break;
        const uint64_t *st = a.state + i * a.S;
        bool h;
        if (a.mode == GB_SCALAR) h = true;
        else if (a.mode == GB_HASH) h = (i == a.n_slots - 1 || a.keys[i] != EMPTY_KEY) && st[0] > 0;
        else h = st[0] > 0;
        have |= h ? 1u << k : 0u;
    }
    unsigned total;
    const unsigned mine = block_excl_scan((unsigned)__popc(have), total);
    int64_t o = (int64_t)block_reserve(a.out_count, total) + mine;
    for (int k = 0; k < EMIT_PER_THREAD; ++k) {
        if (!((have >> k) & 1u)) continue;
        const int64_t i = first + (int64_t)k * blockDim.x;
        const uint64_t *st = a.state + i * a.S;
        if (a.out_key) {
            if (a.mode == GB_HASH) {
                const bool null_group = i == a.n_slots - 1;
                a.out_key[o] = null_group ? 0 : a.keys[i];
                a.out_key_valid[o] = null_group ? 0 : 1;
            } else {
                const bool null_group = i == a.n_slots - 1 && !a.key_is_utf8;
                a.out_key[o] = null_group ? 0 : i + a.key_min;
                a.out_key_valid[o] = null_group ? 0 : 1;
            }
        }
        emit_row(a.outs, st, o, a.out_val, a.out_valid);
        ++o;
    }
}

__global__ void k_state_init(int64_t *keys, int64_t n_keys, uint64_t *state, int64_t n_words,
                             DOps ops) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_keys; i += stride)
        keys[i] = EMPTY_KEY;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_words; i += stride)
        state[i] = ops.ident[i % ops.S];
}

__global__ void k_fill_u64(uint64_t *p, int64_t n, uint64_t v) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

hipError_t launch_fill_u64(hipStream_t s, uint64_t *p, int64_t n, uint64_t v) {
    if (n <= 0) return hipSuccess;
    const int grid = (int)std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill_u64, dim3(grid), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t launch_state_init(hipStream_t s, int64_t *keys, int64_t n_keys, uint64_t *state,
                             int64_t n_groups, const DOps &ops) {
    const int64_t n = std::max<int64_t>(n_keys, n_groups * ops.S);
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192));
    hipLaunchKernelGGL(k_state_init, dim3(grid), dim3(256), 0, s, keys, n_keys, state,
                       n_groups * ops.S, ops);
    return hipGetLastError();
}

hipError_t launch_gb_fold(hipStream_t s, const GbArgs &a, int64_t n_groups) {
    if (a.replicas <= 1 || n_groups <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gb_fold, dim3((unsigned)((n_groups + 255) / 256)), dim3(256), 0, s, a, n_groups);
    return hipGetLastError();
}

template <int NPC>
static void launch_gb_npc(hipStream_t s, int mode, const GbArgs &a, int grid, int block, size_t lds) {
    switch (mode) {
        case GB_DENSE_LDS: hipLaunchKernelGGL(k_gb_dense_lds<NPC>, dim3(grid), dim3(block), lds, s, a); break;
        case GB_DENSE_GLOBAL: hipLaunchKernelGGL(k_gb_dense_global<NPC>, dim3(grid), dim3(block), 0, s, a); break;
        case GB_HASH: hipLaunchKernelGGL(k_gb_hash<NPC>, dim3(grid), dim3(block), lds, s, a); break;
        default: hipLaunchKernelGGL(k_gb_scalar<NPC>, dim3(grid), dim3(block), 0, s, a); break;
    }
}

hipError_t launch_groupby(hipStream_t s, int mode, const GbArgs &a, int grid, int block,
                          size_t lds_bytes) {
    GbArgs b = a;
    if (b.replicas < 1) b.replicas = 1;  // one state table
    if (b.pred.n_pcols <= 1) launch_gb_npc<1>(s, mode, b, grid, block, lds_bytes);
    else if (b.pred.n_pcols == 2) launch_gb_npc<2>(s, mode, b, grid, block, lds_bytes);
    else launch_gb_npc<MAX_PCOLS>(s, mode, b, grid, block, lds_bytes);
    return hipGetLastError();
}

// GROUP BY over several key columns, or one Float64 key (the reference's slow
// path, src/query/aggregate.rs:1232-1346, which keys a group on the key
// values' display strings): one int64 key per row.  Float64 key: the value's
// bits + 1, with every NaN made one NaN (they all print "NaN") and -0.0 kept
// apart from 0.0 (they print "-0" and "0"); the + 1 keeps every value off
// INT64_MIN, which stands for a NULL key.  Several columns: the mixed-radix
// number of the columns' codes, code 0 for NULL, value - min + 1 for an
// integer, dictionary code + 1 for text (or, for a NULL text key, the code of
// the dictionary's "NULL" entry when it has one, as both print "NULL").
__global__ void k_composite_key(CompositeKeyArgs a) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < a.n_rows; r += stride) {
        if (a.float_key) {
            uint64_t bits = ((const uint64_t *)a.v[0])[r];
            const bool ok = !a.valid[0] || ((a.valid[0][r >> 6] >> (r & 63)) & 1);
            if (((bits >> 52) & 0x7ff) == 0x7ff && (bits & 0x000fffffffffffffull)) bits = CANON_NAN;
            a.out[r] = ok ? (int64_t)(bits + 1) : EMPTY_KEY;
            continue;
        }
        int64_t comp = 0;
        for (int i = 0; i < a.n; ++i) {
            const bool ok = !a.valid[i] || ((a.valid[i][r >> 6] >> (r & 63)) & 1);
            int64_t code = 0;
            if (a.dtype[i] == FGPU_UTF8) {
                const uint32_t c = ((const uint32_t *)a.v[i])[r];
                code = ok && c != NULL_CODE ? (int64_t)c + 1 : a.null_code[i];
            } else {
                const int64_t v = a.dtype[i] == FGPU_INT32 ? (int64_t)((const int32_t *)a.v[i])[r]
                                                           : ((const int64_t *)a.v[i])[r];
                code = ok ? v - a.kmin[i] + 1 : 0;
            }
            comp += code * a.mul[i];
        }
        a.out[r] = comp;
    }
}

hipError_t launch_composite_key(hipStream_t s, const CompositeKeyArgs &a, int grid) {
    hipLaunchKernelGGL(k_composite_key, dim3(grid), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_emit(hipStream_t s, const EmitArgs &a) {
    const int64_t per_wg = 256 * EMIT_PER_THREAD;
    const int grid = (int)std::max<int64_t>(1, (a.n_slots + per_wg - 1) / per_wg);
    hipLaunchKernelGGL(k_gb_emit, dim3(grid), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace fgpu
