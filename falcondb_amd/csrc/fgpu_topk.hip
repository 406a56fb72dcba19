// fgpu_topk.hip — ORDER BY col [ASC | DESC] LIMIT k over the rows that pass a
// predicate (execute_sort_limit, src/query/sort.rs:59-115; try_typed_sort
// :209-330).  The order is the reference's typed one: the column's total
// order (integers with the sign bit flipped, floats through f64_ord_key
// :452-454), inverted for DESC, ties by row.  NULL rows are selected by a
// second scan (by row) and placed by the host.
//
// Each wave keeps its K best (key, row) pairs sorted across its lanes behind a
// threshold test, so almost every row costs one comparison; a workgroup's
// waves then merge into one list, and one final workgroup merges every
// workgroup's K candidates the same way.
#include "fgpu_state.h"

namespace fgpu {

struct TopPair {
    uint64_t key;
    int64_t row;
};

FGPU_DEV bool top_less(const TopPair &a, const TopPair &b) {
    return a.key < b.key || (a.key == b.key && a.row < b.row);
}

// ─── one wave's best K ───────────────────────────────────────────────────────
// The list is spread over the lanes, sorted: lane i holds the i-th best pair
// (lanes past the filled part hold the sentinel (~0, INT64_MAX)).  Lane K - 1
// holds the threshold a row must beat.  Rows are tested against it, and the
// few that beat it are inserted one at a time: the position is the number of
// lanes holding better pairs, and the lanes from there on shift up by one.
FGPU_DEV TopPair top_sentinel() { return TopPair{~0ull, INT64_MAX}; }

// Lane l's pair, in every lane (l is wave-uniform).
FGPU_DEV TopPair lane_pair(const TopPair &p, int l) {
    const uint32_t klo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)p.key, l);
    const uint32_t khi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(p.key >> 32), l);
    const uint32_t rlo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(uint64_t)p.row, l);
    const uint32_t rhi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)((uint64_t)p.row >> 32), l);
    return TopPair{(uint64_t)khi << 32 | klo, (int64_t)((uint64_t)rhi << 32 | rlo)};
}

// Inserts the wave-uniform pair c into the lane list `my`.
FGPU_DEV void wave_insert(TopPair &my, const TopPair &c) {
    const int pos = __popcll(__ballot(top_less(my, c)));
    const int lane = threadIdx.x & 63;
    TopPair prev;
    prev.key = __shfl_up(my.key, 1, 64);
    prev.row = __shfl_up(my.row, 1, 64);
    if (lane == pos) my = c;
    else if (lane > pos) my = prev;
}

// Offers each active lane's pair c to the wave's list; thr tracks lane K - 1.
// Every lane of the wave must call it (the ballots span the wave).
template <int K>
FGPU_DEV void wave_offer(TopPair &my, TopPair &thr, const TopPair &c, bool active) {
    uint64_t m = __ballot(active && top_less(c, thr));
    while (m) {
        const int l = __builtin_ctzll(m);
        wave_insert(my, lane_pair(c, l));
        thr = lane_pair(my, K - 1);
        m &= ~(1ull << l);
        m &= __ballot(active && top_less(c, thr));
    }
}

// The block's best K into out[0..K): every wave's list meets in LDS and
// wave 0 takes the others' pairs into its own.  Every thread must call it.
template <int K>
FGPU_DEV void block_best(TopPair &my, TopPair &thr, TopPair *out) {
    __shared__ TopPair lists[16][K];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if (lane < K) lists[wave][lane] = my;
    __syncthreads();
    if (wave == 0) {
        for (int w = 1; w < nw; ++w) {
            const TopPair c = lane < K ? lists[w][lane] : top_sentinel();
            wave_offer<K>(my, thr, c, lane < K);
        }
        if (lane < K) out[lane] = my;
    }
}

// Sort key of a value v of a column of type dtype (ascending order for the
// wanted direction).
FGPU_DEV uint64_t top_key_of(int dtype, int64_t v, int desc) {
    uint64_t k;
    if (dtype == FGPU_FLOAT64) {
        const uint64_t bits = (uint64_t)v;
        k = (bits >> 63) == 0 ? bits ^ (1ull << 63) : ~bits;
    } else {
        k = (uint64_t)v ^ (1ull << 63);
    }
    return desc ? ~k : k;
}

// Rows per thread in flight: 8, or 4 when the predicate reads several columns
// (their values for every row of the tile are held at once).
template <int NPC>
constexpr int topk_unroll() { return NPC > 1 ? 4 : 8; }

// null_mode 0: passing non-NULL rows by key; 1: passing NULL rows by row.
// A wave walks tiles of U * blockDim.x rows (its lanes' rows are
// u * blockDim.x + wave * 64 + lane), so the loop is uniform per wave.
template <int K, int NPC>
__global__ __launch_bounds__(256) void k_topk_scan(DCols cols, DPred pred, int col, int desc, int null_mode,
                                                   int64_t n, TopPair *cands) {
    constexpr int TOPK_UNROLL = topk_unroll<NPC>();
    TopPair my = top_sentinel(), thr = top_sentinel();
    const int lane = threadIdx.x & 63;
    const ColRef kc = col_ref(cols, col);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x * TOPK_UNROLL;
    for (int64_t w0 = (int64_t)blockIdx.x * blockDim.x * TOPK_UNROLL + (threadIdx.x & ~63); w0 < n; w0 += stride) {
        // zone maps: the wave's rows lie in [w0, w0 + (TOPK_UNROLL - 1) * blockDim.x + 64)
        if (tile_dead(pred, w0, std::min<int64_t>(w0 + (int64_t)(TOPK_UNROLL - 1) * blockDim.x + 64, n))) continue;
        int64_t rows[TOPK_UNROLL];
#pragma unroll
        for (int u = 0; u < TOPK_UNROLL; ++u) rows[u] = w0 + (int64_t)u * blockDim.x + lane;
        bool pass[TOPK_UNROLL];
        pred_tile<TOPK_UNROLL, NPC>(pred, cols, rows, n, pass);
        uint64_t key[TOPK_UNROLL];
        bool take[TOPK_UNROLL];
        // the key cells of every row of the tile are issued before any is examined
        RawCell3 raw[TOPK_UNROLL];
#pragma unroll
        for (int u = 0; u < TOPK_UNROLL; ++u) cell_issue3(kc, rows[u] < n ? rows[u] : n - 1, raw[u]);
#pragma unroll
        for (int u = 0; u < TOPK_UNROLL; ++u) {
            const int64_t r = rows[u] < n ? rows[u] : n - 1;
            int64_t v;
            bool ok;
            cell_finish3(kc, r, raw[u], v, ok);
            key[u] = null_mode ? 0ull : top_key_of(kc.dtype, v, desc);
            take[u] = pass[u] && (ok != (null_mode != 0));
        }
#pragma unroll
        for (int u = 0; u < TOPK_UNROLL; ++u) wave_offer<K>(my, thr, TopPair{key[u], rows[u]}, take[u]);
    }
    block_best<K>(my, thr, cands + (int64_t)blockIdx.x * K);
}

template <int K>
__global__ __launch_bounds__(256) void k_topk_final(const TopPair *cands, int64_t n_cands, TopPair *out) {
    TopPair my = top_sentinel(), thr = top_sentinel();
    const int lane = threadIdx.x & 63;
    for (int64_t i0 = threadIdx.x & ~63; i0 < n_cands; i0 += blockDim.x) {
        const int64_t i = i0 + lane;
        const bool a = i < n_cands;
        wave_offer<K>(my, thr, a ? cands[i] : top_sentinel(), a);
    }
    block_best<K>(my, thr, out);
}

template <int K>
static hipError_t topk_go(hipStream_t s, const DCols &c, const DPred &p, int col, int desc, int null_mode, int64_t n,
                          int grid, TopPair *cands, TopPair *out) {
    if (p.n_pcols <= 1)
        hipLaunchKernelGGL((k_topk_scan<K, 1>), dim3(grid), dim3(256), 0, s, c, p, col, desc, null_mode, n, cands);
    else
        hipLaunchKernelGGL((k_topk_scan<K, MAX_PCOLS>), dim3(grid), dim3(256), 0, s, c, p, col, desc, null_mode, n,
                           cands);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_topk_final<K>, dim3(1), dim3(256), 0, s, cands, (int64_t)grid * K, out);
    return hipGetLastError();
}

// out receives K pairs (the best first; unused ones are (~0, INT64_MAX)).
hipError_t launch_topk(hipStream_t s, const DCols &c, const DPred &p, int col, int desc, int null_mode, int64_t n,
                       int K, int grid, void *cands, void *out) {
    TopPair *cp = (TopPair *)cands, *op = (TopPair *)out;
    if (K <= 8) return topk_go<8>(s, c, p, col, desc, null_mode, n, grid, cp, op);
    if (K <= 16) return topk_go<16>(s, c, p, col, desc, null_mode, n, grid, cp, op);
    return topk_go<32>(s, c, p, col, desc, null_mode, n, grid, cp, op);
}

}  // namespace fgpu
