// fgpu_batch.hip — many GROUP BY queries in two persistent kernels (see
// BatchQuery).  Each query is the same operator as fgpu_groupby
// (GroupByExec::execute, src/query/aggregate.rs:535-1420) over its own table;
// the batch only changes how the work is scheduled:
//
//   pass 1 (k_batch_scan): workgroups pull (query, row range) units from one
//     counter.  A query whose whole key range fits an LDS table aggregates
//     there and writes one state block per unit (BQ_DENSE); a wider query
//     writes (key, inputs) records into key-range partitions through LDS
//     cursors, so no global atomic is taken per row (BQ_PART).
//   pass 2 (k_batch_merge): workgroups pull (query, slot range) units, fold
//     the blocks or the partition's records into a dense LDS table (slot =
//     key - key_min: no hashing) and emit the groups with one output
//     reservation per unit.
//
// Both kernels run a grid of resident workgroups that exit when the unit
// counter passes the last unit, and the host captures the resets plus both
// launches in one hipGraph that it replays per run.
#include "fgpu_state.h"

namespace fgpu {

constexpr int BATCH_UNROLL = 4;  // 64-row steps per wave tile

FGPU_DEV int64_t rfl64(unsigned long long v) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Next unit index for the whole workgroup (thread 0 takes it from the counter).
FGPU_DEV int64_t batch_next(unsigned long long *counter, unsigned long long *slot) {
    if (threadIdx.x == 0) *slot = atomicAdd(counter, 1ull);
    __syncthreads();
    const int64_t u = rfl64(*slot);
    __syncthreads();
    return u;
}

FGPU_DEV int64_t batch_input(const GbArgs &a, int j, int64_t r, uint32_t &valid) {
    int64_t v = 0;
    if (j < a.n_in) {
        bool ok;
        load_cell(a.cols, a.in_col[j], r, v, ok);
        valid |= ok ? 1u << j : 0u;
    }
    return v;
}

FGPU_DEV Inputs batch_inputs(const GbArgs &a, int64_t r) {
    Inputs in;
    in.valid = 0;
    in.v0 = batch_input(a, 0, r, in.valid);
    in.v1 = batch_input(a, 1, r, in.valid);
    in.v2 = batch_input(a, 2, r, in.valid);
    in.v3 = batch_input(a, 3, r, in.valid);
    in.v4 = batch_input(a, 4, r, in.valid);
    in.v5 = batch_input(a, 5, r, in.valid);
    return in;
}

template <int NPC>
__global__ __launch_bounds__(1024) void k_batch_scan(const BatchQuery *__restrict__ qs,
                                                     const BatchUnit *__restrict__ units, BatchArgs A) {
    extern __shared__ uint64_t lds[];
    __shared__ unsigned long long unit_slot;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int nw = __builtin_amdgcn_readfirstlane((int)(blockDim.x >> 6));
    for (;;) {
        const int64_t ui = batch_next(A.work, &unit_slot);
        if (ui >= A.n_units) return;
        const BatchUnit U = units[ui];
        const BatchQuery &Q = qs[U.q];
        const GbArgs &a = Q.g;
        const int S = Q.S;
        const bool dense = Q.mode == BQ_DENSE;
        const OpList ops = op_list(a.ops, Q.has_valid_word != 0);
        uint64_t *null_st = lds;
        uint64_t *table = lds + S;                          // BQ_DENSE: range * S words
        unsigned int *cursor = (unsigned int *)(lds + S);   // BQ_PART: P cursors
        for (int i = threadIdx.x; i < S; i += blockDim.x) null_st[i] = a.ops.ident[i];
        if (dense) {
            fill_ident(table, Q.range, S, a.ops);
        } else {
            for (int i = threadIdx.x; i < Q.P; i += blockDim.x) cursor[i] = 0;
        }
        __syncthreads();
        int64_t *my_cells = dense ? nullptr : Q.cells + (int64_t)U.local * Q.P * Q.Cw * Q.RW;
        for (int64_t base = U.row0 + (int64_t)wave * 64 * BATCH_UNROLL; base < U.row1;
             base += (int64_t)nw * 64 * BATCH_UNROLL) {
            int64_t rows[BATCH_UNROLL];
            bool pass[BATCH_UNROLL];
#pragma unroll
            for (int u = 0; u < BATCH_UNROLL; ++u) rows[u] = base + u * 64 + lane;
            pred_tile<BATCH_UNROLL, NPC>(a.pred, a.cols, rows, U.row1, pass);
            int64_t key[BATCH_UNROLL];
            bool has_key[BATCH_UNROLL];
            Inputs in[BATCH_UNROLL];
#pragma unroll
            for (int u = 0; u < BATCH_UNROLL; ++u) {
                key[u] = 0;
                has_key[u] = false;
                in[u] = Inputs{0, 0, 0, 0, 0, 0, 0};
                if (pass[u]) {
                    bool ok;
                    load_cell(a.cols, a.key_col, rows[u], key[u], ok);
                    has_key[u] = ok && key[u] != EMPTY_KEY;  // NULL and the sentinel form the NULL group
                    in[u] = batch_inputs(a, rows[u]);
                }
            }
            // one row per iteration; the u-th values are picked with selects so
            // that update_row has a single call site (the loop stays rolled)
#pragma unroll 1
            for (int u = 0; u < BATCH_UNROLL; ++u) {
                bool ps = pass[0], hk = has_key[0];
                int64_t k = key[0];
                Inputs iu = in[0];
#pragma unroll
                for (int q = 1; q < BATCH_UNROLL; ++q)
                    if (u == q) {
                        ps = pass[q];
                        hk = has_key[q];
                        k = key[q];
                        iu = in[q];
                    }
                if (!ps) continue;
                const int64_t slot = k - Q.key_min;  // in [0, range): the range comes from the column's min / max
                uint64_t *st = !hk ? null_st : dense ? table + slot * S : nullptr;
                if (st) {
                    update_row<LDS_SCOPE>(st, ops, iu);
                    continue;
                }
                const int p = (int)(slot >> Q.shift);
                const unsigned pos = atomicAdd(cursor + p, 1u);
                if (pos >= (unsigned)Q.Cw) {
                    atomicExch(A.flags, 1ull);
                    continue;
                }
                int64_t *rec = my_cells + ((int64_t)p * Q.Cw + pos) * Q.RW;
                rec[0] = k;
                const int ni = a.n_in;
                if (ni > 0) rec[1] = iu.v0;
                if (ni > 1) rec[2] = iu.v1;
                if (ni > 2) rec[3] = iu.v2;
                if (ni > 3) rec[4] = iu.v3;
                if (ni > 4) rec[5] = iu.v4;
                if (ni > 5) rec[6] = iu.v5;
                if (Q.has_valid_word) rec[1 + ni] = iu.valid;
            }
        }
        __syncthreads();
        for (int i = threadIdx.x; i < S; i += blockDim.x) Q.null_blocks[(int64_t)U.local * S + i] = null_st[i];
        if (dense) {
            uint64_t *blk = Q.blocks + (int64_t)U.local * Q.range * S;
            for (int64_t i = threadIdx.x; i < Q.range * S; i += blockDim.x) blk[i] = table[i];
        } else {
            for (int i = threadIdx.x; i < Q.P; i += blockDim.x) Q.cell_cnt[(int64_t)U.local * Q.P + i] = cursor[i];
        }
        __syncthreads();  // the next unit re-initialises the LDS
    }
}

__global__ __launch_bounds__(1024) void k_batch_merge(const BatchQuery *__restrict__ qs,
                                                      const MergeUnit *__restrict__ units2, BatchArgs A) {
    extern __shared__ uint64_t lds[];
    __shared__ unsigned long long unit_slot;
    __shared__ uint64_t null_row[64];
    const int lane = threadIdx.x & 63;
    for (;;) {
        const int64_t mi = batch_next(A.work + 1, &unit_slot);
        if (mi >= A.n_units2) return;
        const MergeUnit M = units2[mi];
        const BatchQuery &Q = qs[M.q];
        const GbArgs &a = Q.g;
        const int S = Q.S;
        const OpList ops = op_list(a.ops, Q.has_valid_word != 0);
        const int n = (int)(M.k1 - M.k0);
        const int G = (int)Q.n_units;
        uint64_t *table = lds;                                   // n * S words
        unsigned int *pre = (unsigned int *)(lds + (int64_t)n * S);  // BQ_PART: G + 1 prefix
        fill_ident(table, n, S, a.ops);
        if (M.part >= 0)
            for (int g = threadIdx.x; g < G; g += blockDim.x) {
                const unsigned c = Q.cell_cnt[(int64_t)g * Q.P + M.part];
                pre[g] = c < (unsigned)Q.Cw ? c : (unsigned)Q.Cw;
            }
        if (M.null_group && threadIdx.x < S) null_row[threadIdx.x] = a.ops.ident[threadIdx.x];
        __syncthreads();
        if (M.part >= 0 && threadIdx.x < 64) {  // exclusive scan of the partition's cell counts
            const int per = (G + 63) / 64;
            unsigned s1 = 0;
            for (int k = 0; k < per; ++k) {
                const int g = lane * per + k;
                if (g < G) s1 += pre[g];
            }
            unsigned i1 = s1;
            for (int d = 1; d < 64; d <<= 1) {
                const unsigned t = __shfl_up(i1, d, 64);
                if (lane >= d) i1 += t;
            }
            unsigned e1 = i1 - s1;
            for (int k = 0; k < per; ++k) {
                const int g = lane * per + k;
                if (g < G) {
                    const unsigned c1 = pre[g];
                    pre[g] = e1;
                    e1 += c1;
                }
            }
            if (lane == 63) pre[G] = i1;
        }
        if (M.null_group && threadIdx.x == 0)
            for (int g = 0; g < G; ++g) merge_state<LDS_SCOPE>(null_row, Q.null_blocks + (int64_t)g * S, a.ops);
        __syncthreads();
        if (M.part < 0) {
            // BQ_DENSE: every unit's block over slots [k0, k1); a thread owns its slots
            for (int g = 0; g < G; ++g) {
                const uint64_t *blk = Q.blocks + ((int64_t)g * Q.range + M.k0) * S;
                for (int i = threadIdx.x; i < n; i += blockDim.x)
                    merge_state<LDS_SCOPE>(table + (int64_t)i * S, blk + (int64_t)i * S, a.ops);
            }
        } else {
            const unsigned n_rec = pre[G];
            const int ni = a.n_in;
            const int64_t slot0 = (int64_t)M.part << Q.shift;
            const int64_t cell_stride = (int64_t)Q.P * Q.Cw * Q.RW;
            const int64_t *part_cells = Q.cells + (int64_t)M.part * Q.Cw * Q.RW;
            for (unsigned i = threadIdx.x; i < n_rec; i += blockDim.x) {
                int lo = 0, hi = G;  // cell of record i
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (pre[mid] <= i) lo = mid;
                    else hi = mid;
                }
                const int64_t *rec = part_cells + lo * cell_stride + (int64_t)(i - pre[lo]) * Q.RW;
                Inputs in{0, 0, 0, 0, 0, 0, 0};
                if (ni > 0) in.v0 = rec[1];
                if (ni > 1) in.v1 = rec[2];
                if (ni > 2) in.v2 = rec[3];
                if (ni > 3) in.v3 = rec[4];
                if (ni > 4) in.v4 = rec[5];
                if (ni > 5) in.v5 = rec[6];
                in.valid = Q.has_valid_word ? (uint32_t)rec[1 + ni] : 0xffffffffu;
                const int64_t slot = rec[0] - Q.key_min - slot0;
                update_row<LDS_SCOPE>(table + slot * S, ops, in);
            }
        }
        __syncthreads();
        // emit the occupied slots (and the NULL group) with one reservation
        uint32_t have = 0;
        unsigned n_have = 0;
        for (int i = threadIdx.x, k = 0; i < n; i += blockDim.x, ++k) {
            const bool h = table[(int64_t)i * S] > 0;
            have |= h ? 1u << k : 0u;
            n_have += h ? 1u : 0u;
        }
        const bool emit_null = M.null_group && threadIdx.x == 0 && null_row[0] > 0;
        unsigned total;
        const unsigned mine = block_excl_scan(n_have + (emit_null ? 1u : 0u), total);
        int64_t o = (int64_t)block_reserve(Q.out_count, total) + mine;
        const int64_t key0 = Q.key_min + (M.part >= 0 ? ((int64_t)M.part << Q.shift) : M.k0);
        for (int i = threadIdx.x, k = 0; i < n; i += blockDim.x, ++k) {
            if (!((have >> k) & 1u)) continue;
            if (o >= Q.out_cap) {
                atomicExch(A.flags + 1, 1ull);
                break;
            }
            Q.out_key[o] = key0 + i;
            Q.out_key_valid[o] = 1;
            emit_row(Q.outs, table + (int64_t)i * S, o, Q.out_val, Q.out_valid);
            ++o;
        }
        if (emit_null) {
            if (o < Q.out_cap) {
                Q.out_key[o] = 0;
                Q.out_key_valid[o] = 0;
                emit_row(Q.outs, null_row, o, Q.out_val, Q.out_valid);
            } else {
                atomicExch(A.flags + 1, 1ull);
            }
        }
        __syncthreads();  // the next unit re-initialises the LDS
    }
}

hipError_t launch_batch_scan(hipStream_t s, const BatchQuery *qs, const BatchArgs &a, int npc, int grid,
                             int block, size_t lds) {
    if (npc <= 1) hipLaunchKernelGGL(k_batch_scan<1>, dim3(grid), dim3(block), lds, s, qs, a.units, a);
    else if (npc <= 2) hipLaunchKernelGGL(k_batch_scan<2>, dim3(grid), dim3(block), lds, s, qs, a.units, a);
    else hipLaunchKernelGGL(k_batch_scan<MAX_PCOLS>, dim3(grid), dim3(block), lds, s, qs, a.units, a);
    return hipGetLastError();
}

hipError_t launch_batch_merge(hipStream_t s, const BatchQuery *qs, const BatchArgs &a, int grid, int block,
                              size_t lds) {
    hipLaunchKernelGGL(k_batch_merge, dim3(grid), dim3(block), lds, s, qs, a.units2, a);
    return hipGetLastError();
}

// Raises the dynamic LDS limit of the kernels (outside any graph capture).
hipError_t batch_set_lds(size_t lds1, size_t lds2) {
    const void *scans[3] = {(const void *)k_batch_scan<1>, (const void *)k_batch_scan<2>,
                            (const void *)k_batch_scan<MAX_PCOLS>};
    for (const void *f : scans) {
        const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds1);
        if (e != hipSuccess) return e;
    }
    return hipFuncSetAttribute((const void *)k_batch_merge, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds2);
}

}  // namespace fgpu
