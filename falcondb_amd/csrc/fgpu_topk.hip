// fgpu_topk.hip — ORDER BY col [ASC | DESC] LIMIT k over the rows that pass a
// predicate (execute_sort_limit, src/query/sort.rs:59-115; try_typed_sort
// :209-330).  The order is the reference's typed one: the column's total
// order (integers with the sign bit flipped, floats through f64_ord_key
// :452-454), inverted for DESC, ties by row.  NULL rows are selected by a
// second scan (by row) and placed by the host.
//
// Each thread keeps its K best (key, row) pairs sorted in registers behind a
// threshold test, so almost every row costs one comparison; a workgroup then
// takes K rounds of a block-wide argmin over the threads' heads, and one final
// workgroup repeats that over every workgroup's K candidates.
#include "fgpu_state.h"

namespace fgpu {

struct TopPair {
    uint64_t key;
    int64_t row;
};

FGPU_DEV bool top_less(const TopPair &a, const TopPair &b) {
    return a.key < b.key || (a.key == b.key && a.row < b.row);
}

template <int K>
struct TopList {
    TopPair a[K];  // ascending; unused entries are (~0, INT64_MAX)
    FGPU_DEV void init() {
#pragma unroll
        for (int j = 0; j < K; ++j) a[j] = TopPair{~0ull, INT64_MAX};
    }
    FGPU_DEV void offer(const TopPair &c) {
        if (!top_less(c, a[K - 1])) return;
        a[K - 1] = c;
#pragma unroll
        for (int j = K - 1; j > 0; --j) {
            if (top_less(a[j], a[j - 1])) {
                const TopPair t = a[j];
                a[j] = a[j - 1];
                a[j - 1] = t;
            }
        }
    }
    FGPU_DEV void pop() {
#pragma unroll
        for (int j = 0; j < K - 1; ++j) a[j] = a[j + 1];
        a[K - 1] = TopPair{~0ull, INT64_MAX};
    }
};

// K rounds of a block-wide argmin over the threads' list heads; out[r] = r-th.
template <int K>
FGPU_DEV void top_block_merge(TopList<K> &l, TopPair *out) {
    __shared__ TopPair wbest[16];
    __shared__ int wtid[16];
    __shared__ int winner;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int r = 0; r < K; ++r) {
        TopPair b = l.a[0];
        int tid = threadIdx.x;
        for (int d = 32; d > 0; d >>= 1) {
            TopPair o;
            o.key = __shfl_xor(b.key, d, 64);
            o.row = __shfl_xor(b.row, d, 64);
            const int ot = __shfl_xor(tid, d, 64);
            if (top_less(o, b) || (!top_less(b, o) && ot < tid)) {
                b = o;
                tid = ot;
            }
        }
        if (lane == 0) {
            wbest[wave] = b;
            wtid[wave] = tid;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            TopPair m = wbest[0];
            int mt = wtid[0];
            for (int w = 1; w < nw; ++w)
                if (top_less(wbest[w], m)) {
                    m = wbest[w];
                    mt = wtid[w];
                }
            out[r] = m;
            winner = mt;
        }
        __syncthreads();
        if (threadIdx.x == winner) l.pop();
        __syncthreads();
    }
}

// Sort key of row r of the column (ascending order for the wanted direction).
FGPU_DEV uint64_t top_key(const DCols &c, int col, int64_t r, int desc) {
    const int64_t v = load_raw(c, col, r);
    uint64_t k;
    if (c.dtype[col] == FGPU_FLOAT64) {
        const uint64_t bits = (uint64_t)v;
        k = (bits >> 63) == 0 ? bits ^ (1ull << 63) : ~bits;
    } else {
        k = (uint64_t)v ^ (1ull << 63);
    }
    return desc ? ~k : k;
}

constexpr int TOPK_UNROLL = 4;

// null_mode 0: passing non-NULL rows by key; 1: passing NULL rows by row.
template <int K, int NPC>
__global__ __launch_bounds__(256) void k_topk_scan(DCols cols, DPred pred, int col, int desc, int null_mode,
                                                   int64_t n, TopPair *cands) {
    TopList<K> l;
    l.init();
    const int64_t stride = (int64_t)gridDim.x * blockDim.x * TOPK_UNROLL;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x * TOPK_UNROLL + threadIdx.x; base < n; base += stride) {
        int64_t rows[TOPK_UNROLL];
        bool pass[TOPK_UNROLL];
#pragma unroll
        for (int u = 0; u < TOPK_UNROLL; ++u) rows[u] = base + (int64_t)u * blockDim.x;
        pred_tile<TOPK_UNROLL, NPC>(pred, cols, rows, n, pass);
#pragma unroll
        for (int u = 0; u < TOPK_UNROLL; ++u) {
            if (!pass[u]) continue;
            const bool ok = load_valid(cols, col, rows[u]);
            if (null_mode ? ok : !ok) continue;
            l.offer(TopPair{null_mode ? 0ull : top_key(cols, col, rows[u], desc), rows[u]});
        }
    }
    top_block_merge<K>(l, cands + (int64_t)blockIdx.x * K);
}

template <int K>
__global__ __launch_bounds__(256) void k_topk_final(const TopPair *cands, int64_t n_cands, TopPair *out) {
    TopList<K> l;
    l.init();
    for (int64_t i = threadIdx.x; i < n_cands; i += blockDim.x) l.offer(cands[i]);
    __shared__ TopPair best[K];
    top_block_merge<K>(l, best);
    if (threadIdx.x < K) out[threadIdx.x] = best[threadIdx.x];
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
