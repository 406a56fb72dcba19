// fgpu_gbpart.hip — radix-partitioned GROUP BY for high-cardinality keys on
// gfx950 (see GbPartArgs).  Every update on the hot path is an LDS operation:
// device-scope atomics are limited to one output reservation per workgroup.
//
// Traffic per input row: the scanned columns once; per cold row (a key the
// LDS cache did not hold) one record written in pass 1 and read in pass 2.
#include "fgpu_state.h"

namespace fgpu {

constexpr int GBP_UNROLL = 4;

FGPU_DEV int gbp_part(uint64_t h, int log_p) { return log_p ? (int)(h >> (64 - log_p)) : 0; }

// Find-or-insert in an LDS table with linear probing (at most `limit` probes);
// slot index from hash bits 40.. so it is independent of the partition bits.
FGPU_DEV int gbp_slot(int64_t *keys, int mask, int limit, int64_t key, uint64_t h) {
    int i = (int)(h >> 40) & mask;
    for (int p = 0; p < limit; ++p) {
        const int64_t k = __hip_atomic_load(keys + i, __ATOMIC_RELAXED, LDS_SCOPE);
        if (k == key) return i;
        if (k == EMPTY_KEY) {
            int64_t expect = EMPTY_KEY;
            if (__hip_atomic_compare_exchange_strong(keys + i, &expect, key, __ATOMIC_RELAXED,
                                                     __ATOMIC_RELAXED, LDS_SCOPE))
                return i;
            if (expect == key) return i;
        }
        i = (i + 1) & mask;
    }
    return -1;
}

// Lookup only (no insertion): slot of `key` within `limit` probes, or -1.
FGPU_DEV int gbp_find(const int64_t *keys, int mask, int limit, int64_t key, uint64_t h) {
    int i = (int)(h >> 40) & mask;
    for (int p = 0; p < limit; ++p) {
        const int64_t k = __hip_atomic_load(keys + i, __ATOMIC_RELAXED, LDS_SCOPE);
        if (k == key) return i;
        if (k == EMPTY_KEY) return -1;
        i = (i + 1) & mask;
    }
    return -1;
}

constexpr int GBP_DOOR_WORDS = 2048;  // doorkeeper bitmap: 65,536 bits of LDS

FGPU_DEV Inputs gbp_inputs(const GbArgs &a, int64_t r) {
    Inputs in{0, 0, 0, 0, 0, 0, 0};
    auto one = [&](int j, int64_t &dst) {
        if (j < a.n_in) {
            bool ok;
            load_cell(a.cols, a.in_col[j], r, dst, ok);
            if (ok) in.valid |= 1u << j;
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

FGPU_DEV bool gbp_key(const GbArgs &a, int64_t r, int64_t &key) {
    bool ok;
    load_cell(a.cols, a.key_col, r, key, ok);
    return ok && key != EMPTY_KEY;  // NULL and the sentinel value form the NULL group
}

template <int NPC>
__global__ void k_gbp_scan(GbPartArgs A) {
    extern __shared__ uint64_t lds[];
    const GbArgs &a = A.g;
    const int S = a.ops.S, H = A.H, P = 1 << A.log_p;
    int64_t *lkeys = (int64_t *)lds;
    uint64_t *lstate = lds + H;
    uint64_t *lnull = lstate + (int64_t)H * S;
    unsigned int *front = (unsigned int *)(lnull + S);
    unsigned int *back = front + P;
    unsigned int *door = back + P;  // a key enters the cache on its second sighting
    for (int i = threadIdx.x; i < H; i += blockDim.x) lkeys[i] = EMPTY_KEY;
    for (int i = threadIdx.x; i < H * S + S; i += blockDim.x) lstate[i] = a.ops.ident[i % S];
    for (int i = threadIdx.x; i < 2 * P + GBP_DOOR_WORDS; i += blockDim.x) front[i] = 0;
    __syncthreads();
    const int64_t r0 = (int64_t)blockIdx.x * A.rows_per_wg;
    const int64_t r1 = r0 + A.rows_per_wg < a.n_rows ? r0 + A.rows_per_wg : a.n_rows;
    int64_t *my = A.cells + (int64_t)blockIdx.x * P * A.Cw * A.RW;
    for (int64_t base = r0; base < r1; base += (int64_t)blockDim.x * GBP_UNROLL) {
        int64_t rows[GBP_UNROLL], key[GBP_UNROLL];
        bool pass[GBP_UNROLL], has_key[GBP_UNROLL];
        Inputs in[GBP_UNROLL];
#pragma unroll
        for (int u = 0; u < GBP_UNROLL; ++u) rows[u] = base + u * blockDim.x + threadIdx.x;
        pred_tile<GBP_UNROLL, NPC>(a.pred, a.cols, rows, r1, pass);
#pragma unroll
        for (int u = 0; u < GBP_UNROLL; ++u) {
            key[u] = 0;
            has_key[u] = false;
            in[u] = Inputs{0, 0, 0, 0, 0, 0, 0};
            if (pass[u]) {
                has_key[u] = gbp_key(a, rows[u], key[u]);
                in[u] = gbp_inputs(a, rows[u]);
            }
        }
#pragma unroll
        for (int u = 0; u < GBP_UNROLL; ++u) {
            if (!pass[u]) continue;
            if (!has_key[u]) {
                update_row<LDS_SCOPE>(lnull, a.ops, in[u]);
                continue;
            }
            const uint64_t h = hash64((uint64_t)key[u]);
            int li = H ? gbp_find(lkeys, H - 1, A.probe_limit, key[u], h) : -1;
            if (li < 0 && H) {
                const unsigned bit = 1u << (h & 31);
                const unsigned seen = atomicOr(door + ((h >> 5) & (GBP_DOOR_WORDS - 1)), bit) & bit;
                if (seen) li = gbp_slot(lkeys, H - 1, A.probe_limit, key[u], h);
            }
            if (li >= 0) {
                update_row<LDS_SCOPE>(lstate + (int64_t)li * S, a.ops, in[u]);
                continue;
            }
            const int p = gbp_part(h, A.log_p);
            const unsigned pos = atomicAdd(front + p, 1u);
            if (pos >= (unsigned)A.Cw) {
                atomicExch(A.flags, 1ull);
                continue;
            }
            int64_t *rec = my + ((int64_t)p * A.Cw + pos) * A.RW;
            rec[0] = key[u];
            const int ni = a.n_in;
            if (ni > 0) rec[1] = in[u].v0;
            if (ni > 1) rec[2] = in[u].v1;
            if (ni > 2) rec[3] = in[u].v2;
            if (ni > 3) rec[4] = in[u].v3;
            if (ni > 4) rec[5] = in[u].v4;
            if (ni > 5) rec[6] = in[u].v5;
            if (A.has_valid_word) rec[1 + ni] = in[u].valid;
        }
    }
    __syncthreads();
    // flush the cache: state records fill each cell from its back end
    for (int i = threadIdx.x; i < H; i += blockDim.x) {
        const int64_t k = lkeys[i];
        if (k == EMPTY_KEY || lstate[(int64_t)i * S] == 0) continue;
        const int p = gbp_part(hash64((uint64_t)k), A.log_p);
        const unsigned b = atomicAdd(back + p, 1u);
        const int64_t slot = (int64_t)A.Cw - (int64_t)(b + 1) * A.SR;
        const int64_t used = front[p] < (unsigned)A.Cw ? front[p] : A.Cw;
        if (slot < used) {
            atomicExch(A.flags, 1ull);
            continue;
        }
        int64_t *dst = my + ((int64_t)p * A.Cw + slot) * A.RW;
        dst[0] = k;
        for (int w = 0; w < S; ++w) dst[1 + w] = (int64_t)lstate[(int64_t)i * S + w];
    }
    __syncthreads();
    for (int p = threadIdx.x; p < P; p += blockDim.x) {
        A.cell_cnt[((int64_t)blockIdx.x * P + p) * 2] = front[p];
        A.cell_cnt[((int64_t)blockIdx.x * P + p) * 2 + 1] = back[p];
    }
    for (int w = threadIdx.x; w < S; w += blockDim.x) A.nullbuf[(int64_t)blockIdx.x * S + w] = lnull[w];
}

// One workgroup per partition: LDS table, every cell of the partition, emit.
__global__ void k_gbp_agg(GbPartArgs A) {
    extern __shared__ uint64_t lds[];
    const GbArgs &a = A.g;
    const int S = a.ops.S, T2 = A.T2, P = 1 << A.log_p;
    int64_t *tkeys = (int64_t *)lds;
    uint64_t *tstate = lds + T2;
    for (int i = threadIdx.x; i < T2; i += blockDim.x) tkeys[i] = EMPTY_KEY;
    for (int i = threadIdx.x; i < T2 * S; i += blockDim.x) tstate[i] = a.ops.ident[i % S];
    __syncthreads();
    const int p = blockIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int ni = a.n_in;
    for (int w = wave; w < A.G; w += nw) {
        const int64_t c = (int64_t)w * P + p;
        unsigned nf = A.cell_cnt[2 * c];
        nf = nf < (unsigned)A.Cw ? nf : (unsigned)A.Cw;
        const unsigned nb = A.cell_cnt[2 * c + 1];
        const int64_t *cell = A.cells + c * A.Cw * A.RW;
        for (unsigned i = lane; i < nf; i += 64) {
            const int64_t *rec = cell + (int64_t)i * A.RW;
            const int64_t key = rec[0];
            Inputs in{0, 0, 0, 0, 0, 0, 0};
            if (ni > 0) in.v0 = rec[1];
            if (ni > 1) in.v1 = rec[2];
            if (ni > 2) in.v2 = rec[3];
            if (ni > 3) in.v3 = rec[4];
            if (ni > 4) in.v4 = rec[5];
            if (ni > 5) in.v5 = rec[6];
            in.valid = A.has_valid_word ? (uint32_t)rec[1 + ni] : 0xffffffffu;
            const int li = gbp_slot(tkeys, T2 - 1, T2, key, hash64((uint64_t)key));
            if (li < 0) {
                atomicExch(A.flags + 1, 1ull);
                continue;
            }
            update_row<LDS_SCOPE>(tstate + (int64_t)li * S, a.ops, in);
        }
        for (unsigned k = lane; k < nb; k += 64) {
            const int64_t slot = (int64_t)A.Cw - (int64_t)(k + 1) * A.SR;
            if (slot < (int64_t)nf) continue;  // not written (cell full; flagged in pass 1)
            const int64_t *rec = cell + slot * A.RW;
            const int64_t key = rec[0];
            const int li = gbp_slot(tkeys, T2 - 1, T2, key, hash64((uint64_t)key));
            if (li < 0) {
                atomicExch(A.flags + 1, 1ull);
                continue;
            }
            merge_state<LDS_SCOPE>(tstate + (int64_t)li * S, (const uint64_t *)(rec + 1), a.ops);
        }
    }
    __syncthreads();
    for (int b = 0; b < T2; b += blockDim.x) {
        const int i = b + threadIdx.x;
        const bool have = i < T2 && tkeys[i] != EMPTY_KEY && tstate[(int64_t)i * S] > 0;
        const uint64_t ballot = __ballot(have);
        if (ballot == 0) continue;
        const int leader = __ffsll((unsigned long long)ballot) - 1;
        unsigned long long base = 0;
        if (lane == leader) base = atomicAdd(A.out_count, (unsigned long long)__popcll(ballot));
        base = __shfl(base, leader);
        if (!have) continue;
        const int64_t o = (int64_t)base + __popcll(ballot & ((1ull << lane) - 1));
        if (o >= A.out_cap) {
            atomicExch(A.flags + 2, 1ull);
            continue;
        }
        A.out_key[o] = tkeys[i];
        A.out_key_valid[o] = 1;
        emit_row(A.outs, tstate + (int64_t)i * S, o, A.out_val, A.out_valid);
    }
}

// The NULL group: merge every workgroup's partial state, emit one row.
__global__ void k_gbp_null(GbPartArgs A) {
    __shared__ uint64_t row[64];
    const DOps &ops = A.g.ops;
    for (int w = threadIdx.x; w < ops.S; w += blockDim.x) row[w] = ops.ident[w];
    __syncthreads();
    for (int g = threadIdx.x; g < A.G; g += blockDim.x)
        merge_state<LDS_SCOPE>(row, A.nullbuf + (int64_t)g * ops.S, ops);
    __syncthreads();
    if (threadIdx.x == 0 && row[0] > 0) {
        const int64_t o = (int64_t)atomicAdd(A.out_count, 1ull);
        if (o >= A.out_cap) {
            atomicExch(A.flags + 2, 1ull);
            return;
        }
        A.out_key[o] = 0;
        A.out_key_valid[o] = 0;
        emit_row(A.outs, row, o, A.out_val, A.out_valid);
    }
}

hipError_t launch_gbp_scan(hipStream_t s, const GbPartArgs &a, int block, size_t lds) {
    hipError_t e;
    if (a.g.pred.n_pcols <= 2) {
        e = hipFuncSetAttribute((const void *)k_gbp_scan<2>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_gbp_scan<2>, dim3(a.G), dim3(block), lds, s, a);
    } else {
        e = hipFuncSetAttribute((const void *)k_gbp_scan<MAX_PCOLS>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_gbp_scan<MAX_PCOLS>, dim3(a.G), dim3(block), lds, s, a);
    }
    return hipGetLastError();
}

hipError_t launch_gbp_agg(hipStream_t s, const GbPartArgs &a, int block, size_t lds) {
    hipError_t e = hipFuncSetAttribute((const void *)k_gbp_agg, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_gbp_agg, dim3(1 << a.log_p), dim3(block), lds, s, a);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_gbp_null, dim3(1), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace fgpu
