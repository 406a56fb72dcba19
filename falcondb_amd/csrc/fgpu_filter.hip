// fgpu_filter.hip — WHERE predicate scans, stream compaction and upload-time
// column statistics for gfx950.
//
// Row layout per wavefront: lane l of a wave handles row 64*w + l of mask word
// w, so a 64-bit ballot of the predicate *is* the reference's LSB-first mask
// word (src/storage/column.rs:132-133; src/query/expr.rs:1076-1175).  Each
// wave walks FILTER_UNROLL words per step so several independent loads are in
// flight per lane.  COUNT(*) is a per-lane count reduced once per workgroup
// (db.rs:1439-1476 restated as one fused pass); compaction is an exclusive scan
// of per-word popcounts followed by an mbcnt-style in-word rank, which keeps
// the reference's row order (gather_by_bitmask, src/storage/column.rs:1077-1123).
#include "fgpu_launch.h"

namespace fgpu {

constexpr int FILTER_UNROLL = 8;

// Is any batch touched by mask words [w0, w0 + FILTER_UNROLL) live?  (A chunk
// lies inside one batch when the batch size is a multiple of its 512 rows.)
FGPU_DEV bool chunk_live(const DPred &p, int64_t w0, int64_t n) {
    const int64_t r0 = w0 << 6;
    const int64_t r1 = ((w0 + FILTER_UNROLL) << 6) < n ? ((w0 + FILTER_UNROLL) << 6) : n;
    for (int64_t b = r0 / p.batch_rows; b <= (r1 - 1) / p.batch_rows; ++b)
        if (p.live[b]) return true;
    return false;
}

// The predicate's columns, resolved once per kernel (uniform bases).
template <int NPC>
FGPU_DEV void pred_cols(const DCols &c, const DPred &p, ColRef (&pc)[NPC]) {
#pragma unroll
    for (int j = 0; j < NPC; ++j) pc[j] = col_ref(c, j < p.n_pcols ? p.pcol[j] : 0);
}

// Issues the loads of one cell of row tile0 + idx: one 64-bit or 32-bit load
// of the value from a uniform base (non-temporal: every value is read once),
// and the validity word when the column has NULLs.
FGPU_DEV void step_issue(const ColRef &c, int64_t tile0, uint32_t idx, RawCell3 &o) {
    if (c.dtype == FGPU_INT64 || c.dtype == FGPU_FLOAT64) {
        const uint64_t x = __builtin_nontemporal_load((const uint64_t *)c.base + tile0 + idx);
        o.lo = (uint32_t)x;
        o.hi = (uint32_t)(x >> 32);
    } else {
        o.lo = o.hi = __builtin_nontemporal_load((const uint32_t *)c.base + tile0 + idx);
    }
    o.vw = c.valid ? c.valid[(tile0 >> 5) + (idx >> 5)] : 0xffffffffu;
}

// The predicate on rows tile0 + 64 * u + lane (u < FILTER_UNROLL), tile0 a
// multiple of 64 below n.  Every cell of the step is issued (rows past n
// re-read row n - 1) before any is examined, so the loads of a step are all
// in flight together.
template <int NPC>
FGPU_DEV void pred_step(const DPred &p, const ColRef (&pc)[NPC], int64_t tile0, int64_t n, int lane,
                        bool (&pass)[FILTER_UNROLL]) {
    if (p.n == 0) {
#pragma unroll
        for (int u = 0; u < FILTER_UNROLL; ++u) pass[u] = tile0 + 64 * u + lane < n;
        return;
    }
    const int64_t last = n - 1 - tile0;
    RawCell3 raw[FILTER_UNROLL][NPC];
#pragma unroll
    for (int u = 0; u < FILTER_UNROLL; ++u) {
        const int64_t i = 64 * u + lane;
        const uint32_t idx = (uint32_t)(i < last ? i : last);
#pragma unroll
        for (int j = 0; j < NPC; ++j)
            if (j < p.n_pcols) step_issue(pc[j], tile0, idx, raw[u][j]);
    }
#pragma unroll
    for (int u = 0; u < FILTER_UNROLL; ++u) {
        const int64_t r = tile0 + 64 * u + lane;
        int64_t v[NPC];
        uint32_t ok = 0;
#pragma unroll
        for (int j = 0; j < NPC; ++j) {
            v[j] = 0;
            if (j < p.n_pcols) {
                bool okj;
                cell_finish3(pc[j], r, raw[u][j], v[j], okj);
                ok |= (okj ? 1u : 0u) << j;
            }
        }
        pass[u] = r < n && eval_reg<NPC>(p, v, ok);
    }
}

template <int NPC>
__global__ void k_filter_count(DCols c, DPred p, int64_t n, unsigned long long *out) {
    __shared__ unsigned long long part[BLOCK / 64];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    const int64_t gw = (int64_t)blockIdx.x * (blockDim.x >> 6) + wave;
    const int64_t n_words = (n + 63) >> 6;
    unsigned int cnt = 0;
    ColRef pc[NPC];
    pred_cols<NPC>(c, p, pc);
    for (int64_t w0 = gw * FILTER_UNROLL; w0 < n_words; w0 += waves * FILTER_UNROLL) {
        if (p.live && !chunk_live(p, w0, n)) continue;
        bool pass[FILTER_UNROLL];
        pred_step<NPC>(p, pc, w0 << 6, n, lane, pass);
#pragma unroll
        for (int u = 0; u < FILTER_UNROLL; ++u) cnt += pass[u] ? 1u : 0u;
    }
    // wave reduction
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off);
    if (lane == 0) part[wave] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += part[i];
        if (t) atomicAdd(out, t);
    }
}

template <int NPC>
__global__ void k_filter_mask(DCols c, DPred p, int64_t n, uint64_t *words) {
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    const int64_t gw = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int64_t n_words = (n + 63) >> 6;
    ColRef pc[NPC];
    pred_cols<NPC>(c, p, pc);
    for (int64_t w0 = gw * FILTER_UNROLL; w0 < n_words; w0 += waves * FILTER_UNROLL) {
        if (p.live && !chunk_live(p, w0, n)) {  // no row of these words can pass
            if (lane < FILTER_UNROLL && w0 + lane < n_words) words[w0 + lane] = 0;
            continue;
        }
        bool pass[FILTER_UNROLL];
        pred_step<NPC>(p, pc, w0 << 6, n, lane, pass);
#pragma unroll
        for (int u = 0; u < FILTER_UNROLL; ++u) {
            const uint64_t b = __ballot(pass[u]);
            if (lane == u && w0 + u < n_words) words[w0 + u] = b;
        }
    }
}

// ─── exclusive scans (block of 256 threads x 16 elements = 4096 per block) ──
constexpr int SCAN_PER_THREAD = 16;
constexpr int SCAN_TILE = BLOCK * SCAN_PER_THREAD;

struct PopSrc {
    const uint64_t *w;
    FGPU_DEV int64_t operator()(int64_t i) const { return __popcll(w[i]); }
};
struct U32Src {
    const uint32_t *c;
    FGPU_DEV int64_t operator()(int64_t i) const { return c[i]; }
};

// Block-wide exclusive scan of one int64 per thread; returns the block total.
FGPU_DEV int64_t block_excl_scan(int64_t v, int64_t &excl) {
    __shared__ int64_t wsum[BLOCK / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int64_t incl = v;
    for (int off = 1; off < 64; off <<= 1) {
        const int64_t y = __shfl_up(incl, off);
        if (lane >= off) incl += y;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    int64_t before = 0, total = 0;
    for (int i = 0; i < BLOCK / 64; ++i) {
        if (i < wave) before += wsum[i];
        total += wsum[i];
    }
    __syncthreads();
    excl = before + incl - v;
    return total;
}

template <class Src>
__global__ void k_scan_tiles(Src src, int64_t n, int64_t *tile_sums) {
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_PER_THREAD;
    int64_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_PER_THREAD; ++k)
        if (base + k < n) s += src(base + k);
    int64_t excl;
    const int64_t total = block_excl_scan(s, excl);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

// Single workgroup: exclusive scan of the tile sums in place; t[n_tiles] = total.
__global__ void k_scan_tile_sums(int64_t *t, int64_t n_tiles) {
    int64_t carry = 0;
    for (int64_t b = 0; b < n_tiles; b += BLOCK) {
        const int64_t i = b + threadIdx.x;
        const int64_t v = i < n_tiles ? t[i] : 0;
        int64_t excl;
        const int64_t total = block_excl_scan(v, excl);
        if (i < n_tiles) t[i] = carry + excl;
        carry += total;
    }
    if (threadIdx.x == 0) t[n_tiles] = carry;
}

template <class Src>
__global__ void k_scan_write(Src src, int64_t n, const int64_t *tile_offs, int64_t *offs) {
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_PER_THREAD;
    int64_t v[SCAN_PER_THREAD];
    int64_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_PER_THREAD; ++k) {
        v[k] = base + k < n ? src(base + k) : 0;
        s += v[k];
    }
    int64_t excl;
    block_excl_scan(s, excl);
    int64_t run = tile_offs[blockIdx.x] + excl;
#pragma unroll
    for (int k = 0; k < SCAN_PER_THREAD; ++k) {
        if (base + k < n) offs[base + k] = run;
        run += v[k];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
        const int64_t nt = gridDim.x;
        offs[n] = tile_offs[nt];
    }
}

template <class Src>
static hipError_t scan_impl(hipStream_t s, Src src, int64_t n, int64_t *offs, int64_t *tmp) {
    const int64_t tiles = std::max<int64_t>(1, (n + SCAN_TILE - 1) / SCAN_TILE);
    hipLaunchKernelGGL(k_scan_tiles<Src>, dim3((unsigned)tiles), dim3(BLOCK), 0, s, src, n, tmp);
    hipLaunchKernelGGL(k_scan_tile_sums, dim3(1), dim3(BLOCK), 0, s, tmp, tiles);
    hipLaunchKernelGGL(k_scan_write<Src>, dim3((unsigned)tiles), dim3(BLOCK), 0, s, src, n, tmp, offs);
    return hipGetLastError();
}

hipError_t launch_word_offsets(hipStream_t s, const uint64_t *words, int64_t n_words,
                               int64_t *offs, int64_t *block_tmp) {
    return scan_impl(s, PopSrc{words}, n_words, offs, block_tmp);
}
hipError_t launch_scan_u32(hipStream_t s, const uint32_t *cnt, int64_t n, int64_t *offs,
                           int64_t *block_tmp) {
    return scan_impl(s, U32Src{cnt}, n, offs, block_tmp);
}

// One wave per mask word: the lanes whose bit is set write their row at
// offs[word] + (number of set bits below the lane).
__global__ void k_gather(const uint64_t *words, const int64_t *offs, int64_t n_words,
                         const void *src, int elem_bytes, const uint64_t *src_valid, void *dst,
                         uint8_t *dst_valid) {
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t w = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); w < n_words;
         w += waves) {
        const uint64_t bits = words[w];
        if (!((bits >> lane) & 1)) continue;
        const int64_t r = (w << 6) + lane;
        const int64_t o = offs[w] + __popcll(bits & ((1ull << lane) - 1));
        if (elem_bytes == 8) ((int64_t *)dst)[o] = ((const int64_t *)src)[r];
        else ((int32_t *)dst)[o] = ((const int32_t *)src)[r];
        if (dst_valid) dst_valid[o] = src_valid ? (uint8_t)((src_valid[r >> 6] >> (r & 63)) & 1) : 1;
    }
}

// Shared scan: q COUNT(*) ... WHERE <col op literal> requests over one column
// answered by one pass.  Every loaded value is tested against every request's
// comparison; per-lane counters are reduced once per workgroup, one atomic per
// request.  The column is read once for all q requests instead of q times.
template <int QB>
__global__ void k_filter_count_shared(SharedCount a) {
    __shared__ unsigned int part[QB][BLOCK / 64];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    const int64_t gw = (int64_t)blockIdx.x * (blockDim.x >> 6) + wave;
    const int64_t n_words = (a.n + 63) >> 6;
    const ColRef cr{(const uint8_t *)a.values, a.valid, a.dtype};
    unsigned int cnt[QB];
#pragma unroll
    for (int q = 0; q < QB; ++q) cnt[q] = 0;
    for (int64_t w0 = gw * FILTER_UNROLL; w0 < n_words; w0 += waves * FILTER_UNROLL) {
        RawCell raw[FILTER_UNROLL];
        int64_t v[FILTER_UNROLL];
        bool ok[FILTER_UNROLL];
#pragma unroll
        for (int u = 0; u < FILTER_UNROLL; ++u) {
            const int64_t r = ((w0 + u) << 6) + lane;
            raw[u].bits = 0;
            raw[u].vword = 0;
            if (r < a.n) cell_issue(cr, r, raw[u]);
        }
#pragma unroll
        for (int u = 0; u < FILTER_UNROLL; ++u) {
            const int64_t r = ((w0 + u) << 6) + lane;
            cell_finish(cr, r, raw[u], v[u], ok[u]);
            ok[u] = ok[u] && r < a.n;
        }
#pragma unroll
        for (int q = 0; q < QB; ++q) {
            if (q < a.q) {
                const DPredNode nd = a.leaf[q];
#pragma unroll
                for (int u = 0; u < FILTER_UNROLL; ++u) {
                    const int64_t r = ((w0 + u) << 6) + lane;
                    // IS NULL passes rows whose validity bit is clear; rows past n never pass
                    cnt[q] += (r < a.n && leaf_reg(nd, v[u], ok[u])) ? 1u : 0u;
                }
            }
        }
    }
#pragma unroll
    for (int q = 0; q < QB; ++q) {
        unsigned int c = cnt[q];
        for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
        if (lane == 0) part[q][wave] = c;
    }
    __syncthreads();
    if ((int)threadIdx.x < a.q) {
        unsigned long long t = 0;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += part[threadIdx.x][i];
        if (t) atomicAdd(a.out + threadIdx.x, t);
    }
}

hipError_t launch_filter_count_shared(hipStream_t s, const SharedCount &a, int grid) {
    if (a.q < 1 || a.q > SHARED_MAX_Q) return hipErrorInvalidValue;
    if (a.q <= 8) hipLaunchKernelGGL(k_filter_count_shared<8>, dim3(grid), dim3(BLOCK), 0, s, a);
    else if (a.q <= 16) hipLaunchKernelGGL(k_filter_count_shared<16>, dim3(grid), dim3(BLOCK), 0, s, a);
    else hipLaunchKernelGGL(k_filter_count_shared<SHARED_MAX_Q>, dim3(grid), dim3(BLOCK), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_filter_count(hipStream_t s, const DCols &c, const DPred &p, int64_t n,
                               unsigned long long *d_count, int grid) {
    if (p.n_pcols <= 1) hipLaunchKernelGGL(k_filter_count<1>, dim3(grid), dim3(BLOCK), 0, s, c, p, n, d_count);
    else if (p.n_pcols == 2) hipLaunchKernelGGL(k_filter_count<2>, dim3(grid), dim3(BLOCK), 0, s, c, p, n, d_count);
    else hipLaunchKernelGGL(k_filter_count<MAX_PCOLS>, dim3(grid), dim3(BLOCK), 0, s, c, p, n, d_count);
    return hipGetLastError();
}
hipError_t launch_filter_mask(hipStream_t s, const DCols &c, const DPred &p, int64_t n,
                              uint64_t *d_words, int grid) {
    if (p.n_pcols <= 1) hipLaunchKernelGGL(k_filter_mask<1>, dim3(grid), dim3(BLOCK), 0, s, c, p, n, d_words);
    else if (p.n_pcols == 2) hipLaunchKernelGGL(k_filter_mask<2>, dim3(grid), dim3(BLOCK), 0, s, c, p, n, d_words);
    else hipLaunchKernelGGL(k_filter_mask<MAX_PCOLS>, dim3(grid), dim3(BLOCK), 0, s, c, p, n, d_words);
    return hipGetLastError();
}
// Gathers a column by row index (a materialised join): dst[i] = src[idx[i]]
// for 4- or 8-byte values; row i is NULL where idx[i] is NO_ROW (the
// NULL-extended side of a LEFT join) or the source row is NULL.  Each wave
// takes 64 consecutive rows, so the output validity word of those rows is one
// __ballot (dst_valid may be null when no row can be NULL).  A Utf8 column
// (utf8 != 0) writes the NULL code for its NULL rows.
__global__ void k_gather_idx(const void *src, int elem_bytes, int utf8, const uint64_t *src_valid,
                             const uint32_t *idx, int64_t n, void *dst, uint64_t *dst_valid) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t c = wave; c * 64 < n; c += nwaves) {
        const int64_t i = c * 64 + lane;
        bool ok = false;
        if (i < n) {
            const uint32_t r = idx[i];
            uint64_t v = utf8 ? NULL_CODE : 0;
            if (r != NO_ROW) {
                ok = src_valid ? ((src_valid[r >> 6] >> (r & 63)) & 1) != 0 : true;
                const uint64_t x = elem_bytes == 8 ? ((const uint64_t *)src)[r] : ((const uint32_t *)src)[r];
                if (utf8) ok = ok && (uint32_t)x != NULL_CODE;
                if (ok || !utf8) v = x;
            }
            if (elem_bytes == 8) ((uint64_t *)dst)[i] = v;
            else ((uint32_t *)dst)[i] = (uint32_t)v;
        }
        const uint64_t word = __ballot(ok);
        if (dst_valid && lane == 0) dst_valid[c] = word;
    }
}

hipError_t launch_gather_idx(hipStream_t s, const void *src, int elem_bytes, int utf8, const uint64_t *src_valid,
                             const uint32_t *idx, int64_t n, void *dst, uint64_t *dst_valid, int grid) {
    if (elem_bytes != 4 && elem_bytes != 8) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_gather_idx, dim3(grid), dim3(BLOCK), 0, s, src, elem_bytes, utf8, src_valid, idx, n, dst,
                       dst_valid);
    return hipGetLastError();
}

hipError_t launch_gather(hipStream_t s, const uint64_t *words, const int64_t *offs,
                         int64_t n_words, const void *src, int elem_bytes,
                         const uint64_t *src_valid, void *dst, uint8_t *dst_valid, int grid) {
    hipLaunchKernelGGL(k_gather, dim3(grid), dim3(BLOCK), 0, s, words, offs, n_words, src,
                       elem_bytes, src_valid, dst, dst_valid);
    return hipGetLastError();
}

// ─── column statistics ─────────────────────────────────────────────────────
__global__ void k_col_stats(const void *values, int dtype, const uint64_t *valid, int64_t n,
                            ColStatsDev *st) {
    __shared__ unsigned int hll[4096];
    for (int i = threadIdx.x; i < 4096; i += blockDim.x) hll[i] = 0;
    __syncthreads();
    long long imin = INT64_MAX, imax = INT64_MIN;
    unsigned long long fmin = ~0ull, fmax = 0ull, nonfinite = 0, nulls = 0;
    int emin = 1 << 20, emax = -(1 << 20);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride) {
        bool ok;
        int64_t v;
        if (dtype == FGPU_UTF8) {
            const uint32_t code = ((const uint32_t *)values)[r];
            ok = code != NULL_CODE;
            v = code;
        } else {
            ok = valid ? ((valid[r >> 6] >> (r & 63)) & 1) : true;
            v = dtype == FGPU_INT32 ? (int64_t)((const int32_t *)values)[r]
                                    : ((const int64_t *)values)[r];
        }
        if (!ok) {
            ++nulls;
            continue;
        }
        if (dtype == FGPU_FLOAT64) {
            const uint64_t bits = (uint64_t)v;
            const int be = (int)((bits >> 52) & 0x7ff);
            if (be == 0x7ff) {
                nonfinite |= (bits & 0x000fffffffffffffull) ? FL_NAN : (bits >> 63) ? FL_NINF : FL_PINF;
                if (bits & 0x000fffffffffffffull) continue;  // NaN: no min/max
            } else if ((bits & 0x7fffffffffffffffull) != 0) {
                const int e = be == 0 ? -1074 : be - 1075;
                emin = e < emin ? e : emin;
                emax = e > emax ? e : emax;
            }
            const uint64_t enc = f64_order(bits);
            fmin = enc < fmin ? enc : fmin;
            fmax = enc > fmax ? enc : fmax;
            v = (int64_t)bits;
        } else {
            imin = v < imin ? v : imin;
            imax = v > imax ? v : imax;
        }
        // HyperLogLog, p = 12
        const uint64_t h = hash64((uint64_t)v);
        const unsigned idx = (unsigned)(h >> 52);
        const uint64_t rest = (h << 12) | (1ull << 11);
        const unsigned rank = (unsigned)__clzll((long long)rest) + 1;
        atomicMax(&hll[idx], rank);
    }
    atomicMin(&st->imin, imin);
    atomicMax(&st->imax, imax);
    atomicMin(&st->fmin_enc, fmin);
    atomicMax(&st->fmax_enc, fmax);
    atomicMin(&st->emin, emin);
    atomicMax(&st->emax, emax);
    if (nonfinite) atomicOr(&st->nonfinite, nonfinite);
    if (nulls) atomicAdd(&st->nulls, nulls);
    __syncthreads();
    for (int i = threadIdx.x; i < 4096; i += blockDim.x)
        if (hll[i]) atomicMax(&st->hll[i], hll[i]);
}

// Batch-local dictionary codes to a table-wide dictionary: row r of batch b
// (row_off[b] <= r < row_off[b + 1]) maps code c to lut[lut_off[b] + c]
// (Utf8Column::concat_all's per-part remap, src/storage/column.rs:1576-1589).
// A code past its batch's dictionary becomes NULL and sets *bad.
__global__ void k_dict_remap(uint32_t *codes, int64_t n, const int64_t *row_off, int n_batches,
                             const uint32_t *lut, const int64_t *lut_off, unsigned long long *bad) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride) {
        const uint32_t c = codes[r];
        if (c == NULL_CODE) continue;
        int lo = 0, hi = n_batches;  // the last batch whose first row is <= r
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (row_off[mid] <= r) lo = mid;
            else hi = mid;
        }
        const int64_t base = lut_off[lo], len = lut_off[lo + 1] - base;
        if ((int64_t)c >= len) {
            atomicExch(bad, 1ull);
            codes[r] = NULL_CODE;
            continue;
        }
        codes[r] = lut[base + c];
    }
}

hipError_t launch_dict_remap(hipStream_t s, uint32_t *codes, int64_t n, const int64_t *row_off, int n_batches,
                             const uint32_t *lut, const int64_t *lut_off, unsigned long long *bad, int grid) {
    hipLaunchKernelGGL(k_dict_remap, dim3(grid), dim3(BLOCK), 0, s, codes, n, row_off, n_batches, lut, lut_off, bad);
    return hipGetLastError();
}

// Zone map, one batch per workgroup (BatchStats::compute / column_stats,
// src/storage/stats.rs:44-51, :143-178): bounds of the batch's valid values and
// its valid / NaN counts.  Float bounds skip NaN, as f64::min / f64::max do.
__global__ void k_batch_zones(const void *values, int dtype, const uint64_t *valid, int64_t n,
                              int64_t batch_rows, BatchZone *out) {
    __shared__ int64_t s_lo[BLOCK / 64], s_hi[BLOCK / 64];
    __shared__ unsigned long long s_nv[BLOCK / 64], s_nn[BLOCK / 64];
    const int64_t r0 = (int64_t)blockIdx.x * batch_rows;
    const int64_t r1 = r0 + batch_rows < n ? r0 + batch_rows : n;
    const bool is_f = dtype == FGPU_FLOAT64;
    // integers: signed bounds; floats: f64_order keys, compared as unsigned
    uint64_t lo = is_f ? ~0ull : (uint64_t)INT64_MAX, hi = is_f ? 0ull : (uint64_t)INT64_MIN;
    unsigned long long nv = 0, nn = 0;
    for (int64_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
        if (valid && !((valid[r >> 6] >> (r & 63)) & 1)) continue;
        ++nv;
        if (is_f) {
            const uint64_t bits = ((const uint64_t *)values)[r];
            if (((bits >> 52) & 0x7ff) == 0x7ff && (bits & 0x000fffffffffffffull)) {
                ++nn;
                continue;
            }
            const uint64_t k = f64_order(bits);
            lo = k < lo ? k : lo;
            hi = k > hi ? k : hi;
        } else {
            const int64_t v = dtype == FGPU_INT32 ? (int64_t)((const int32_t *)values)[r] : ((const int64_t *)values)[r];
            lo = v < (int64_t)lo ? (uint64_t)v : lo;
            hi = v > (int64_t)hi ? (uint64_t)v : hi;
        }
    }
    for (int d = 32; d > 0; d >>= 1) {
        const uint64_t l2 = __shfl_xor(lo, d, 64), h2 = __shfl_xor(hi, d, 64);
        if (is_f) {
            lo = l2 < lo ? l2 : lo;
            hi = h2 > hi ? h2 : hi;
        } else {
            lo = (int64_t)l2 < (int64_t)lo ? l2 : lo;
            hi = (int64_t)h2 > (int64_t)hi ? h2 : hi;
        }
        nv += __shfl_xor(nv, d, 64);
        nn += __shfl_xor(nn, d, 64);
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        s_lo[wave] = (int64_t)lo;
        s_hi[wave] = (int64_t)hi;
        s_nv[wave] = nv;
        s_nn[wave] = nn;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        BatchZone z{s_lo[0], s_hi[0], (int64_t)s_nv[0], (int64_t)s_nn[0]};
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
            if (is_f) {
                z.lo = (uint64_t)s_lo[w] < (uint64_t)z.lo ? s_lo[w] : z.lo;
                z.hi = (uint64_t)s_hi[w] > (uint64_t)z.hi ? s_hi[w] : z.hi;
            } else {
                z.lo = s_lo[w] < z.lo ? s_lo[w] : z.lo;
                z.hi = s_hi[w] > z.hi ? s_hi[w] : z.hi;
            }
            z.n_valid += (int64_t)s_nv[w];
            z.n_nan += (int64_t)s_nn[w];
        }
        out[blockIdx.x] = z;
    }
}

hipError_t launch_batch_zones(hipStream_t s, const void *values, int dtype, const uint64_t *valid, int64_t n,
                              int64_t batch_rows, BatchZone *out) {
    const int64_t nb = (n + batch_rows - 1) / batch_rows;
    hipLaunchKernelGGL(k_batch_zones, dim3((unsigned)nb), dim3(BLOCK), 0, s, values, dtype, valid, n, batch_rows, out);
    return hipGetLastError();
}

hipError_t launch_col_stats(hipStream_t s, const void *values, int dtype, const uint64_t *valid,
                            int64_t n, ColStatsDev *d_stats, int grid) {
    hipLaunchKernelGGL(k_col_stats, dim3(grid), dim3(BLOCK), 0, s, values, dtype, valid, n, d_stats);
    return hipGetLastError();
}

}  // namespace fgpu
