// fgpu_sort.hip — ORDER BY [... , ...] [LIMIT k] for any k and any number of
// sort keys (execute_sort_limit, src/query/sort.rs:59-115; try_typed_sort
// :209-330; compare_scalar :807-840).
//
// The rows that pass the WHERE are listed in row order, then ordered by stable
// LSD passes: for each sort key from the last to the first, a radix sort of
// the current order by the key's value, then one by its NULL flag (which puts
// NULL rows first or last).  Stable passes keep equal rows in the order of the
// previous pass, so the result is ordered by the keys in turn and then by row
// number.  The radix sorts are rocPRIM's device_radix_sort (stable LSD).
#include <rocprim/device/device_radix_sort.hpp>

#include "fgpu_launch.h"

namespace fgpu {

// Row numbers of the set bits of an LSB-first mask, in row order: word w's
// rows go to out[offs[w] ..].  One thread per word.
__global__ void k_mask_rows(const uint64_t *words, const int64_t *offs, int64_t n_words, uint32_t *out) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; w < n_words; w += stride) {
        uint64_t bits = words[w];
        int64_t o = offs[w];
        while (bits) {
            const int b = __builtin_ctzll(bits);
            bits &= bits - 1;
            out[o++] = (uint32_t)(w * 64 + b);
        }
    }
}

// The ordered key of one sort column at rows[i]: the column's total order as
// an unsigned integer (integers with the sign bit flipped, doubles through
// f64_ord_key, Utf8 codes through the dictionary's sort rank), inverted for
// DESC.  Int32 and Utf8 keys stay within 32 bits (DESC inverts those 32), so
// their sorts take 32 bits.  flag[i] is 1 for the rows that go after the
// others (NULL rows when NULLs go last, the non-NULL rows when NULLs go first).
__global__ void k_sort_key(const void *values, int dtype, const uint64_t *valid, const uint32_t *rank,
                           const uint32_t *rows, int64_t n, int desc, int nulls_last, uint64_t *key,
                           uint8_t *flag) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint32_t r = rows[i];
        bool ok = valid ? ((valid[r >> 6] >> (r & 63)) & 1) != 0 : true;
        uint64_t k = 0;
        switch (dtype) {
            case FGPU_INT32: k = (uint64_t)(((const uint32_t *)values)[r] ^ 0x80000000u); break;
            case FGPU_INT64: k = (uint64_t)((const int64_t *)values)[r] ^ (1ull << 63); break;
            case FGPU_FLOAT64: {
                const uint64_t b = ((const uint64_t *)values)[r];
                k = (b >> 63) == 0 ? b ^ (1ull << 63) : ~b;
                break;
            }
            default: {
                const uint32_t code = ((const uint32_t *)values)[r];
                ok = ok && code != NULL_CODE;
                k = ok ? rank[code] : 0;
                break;
            }
        }
        if (desc) k = (dtype == FGPU_INT32 || dtype == FGPU_UTF8) ? (~k & 0xffffffffull) : ~k;
        key[i] = ok ? k : 0;
        flag[i] = (uint8_t)(ok ? (nulls_last ? 0 : 1) : (nulls_last ? 1 : 0));
    }
}

// dst[i] = src[idx[i]] for 8-byte or 1-byte elements.
__global__ void k_take_u64(const uint64_t *src, const uint32_t *idx, int64_t n, uint64_t *dst) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[idx[i]];
}
__global__ void k_take_u8(const uint8_t *src, const uint32_t *idx, int64_t n, uint8_t *dst) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[idx[i]];
}
__global__ void k_take_u32(const uint32_t *src, const uint32_t *idx, int64_t n, uint32_t *dst) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[idx[i]];
}
__global__ void k_iota_u32(uint32_t *p, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = (uint32_t)i;
}

static int grid_of(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192)); }

hipError_t launch_mask_rows(hipStream_t s, const uint64_t *words, const int64_t *offs, int64_t n_words,
                            uint32_t *out) {
    hipLaunchKernelGGL(k_mask_rows, dim3(grid_of(n_words)), dim3(256), 0, s, words, offs, n_words, out);
    return hipGetLastError();
}
hipError_t launch_sort_key(hipStream_t s, const void *values, int dtype, const uint64_t *valid, const uint32_t *rank,
                           const uint32_t *rows, int64_t n, int desc, int nulls_last, uint64_t *key, uint8_t *flag) {
    hipLaunchKernelGGL(k_sort_key, dim3(grid_of(n)), dim3(256), 0, s, values, dtype, valid, rank, rows, n, desc,
                       nulls_last, key, flag);
    return hipGetLastError();
}
hipError_t launch_take_u64(hipStream_t s, const uint64_t *src, const uint32_t *idx, int64_t n, uint64_t *dst) {
    hipLaunchKernelGGL(k_take_u64, dim3(grid_of(n)), dim3(256), 0, s, src, idx, n, dst);
    return hipGetLastError();
}
hipError_t launch_take_u8(hipStream_t s, const uint8_t *src, const uint32_t *idx, int64_t n, uint8_t *dst) {
    hipLaunchKernelGGL(k_take_u8, dim3(grid_of(n)), dim3(256), 0, s, src, idx, n, dst);
    return hipGetLastError();
}
hipError_t launch_take_u32(hipStream_t s, const uint32_t *src, const uint32_t *idx, int64_t n, uint32_t *dst) {
    hipLaunchKernelGGL(k_take_u32, dim3(grid_of(n)), dim3(256), 0, s, src, idx, n, dst);
    return hipGetLastError();
}
hipError_t launch_iota_u32(hipStream_t s, uint32_t *p, int64_t n) {
    hipLaunchKernelGGL(k_iota_u32, dim3(grid_of(n)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

// Stable radix sorts of (key, value) pairs on bits [0, end_bit) of the key.
// With temp == nullptr only *temp_bytes is set.
hipError_t sort_pairs_u64(void *temp, size_t *temp_bytes, const uint64_t *k_in, uint64_t *k_out, const uint32_t *v_in,
                          uint32_t *v_out, int64_t n, int end_bit, hipStream_t s) {
    return rocprim::radix_sort_pairs(temp, *temp_bytes, k_in, k_out, v_in, v_out, (size_t)n, 0, end_bit, s);
}
hipError_t sort_pairs_u8(void *temp, size_t *temp_bytes, const uint8_t *k_in, uint8_t *k_out, const uint32_t *v_in,
                         uint32_t *v_out, int64_t n, hipStream_t s) {
    return rocprim::radix_sort_pairs(temp, *temp_bytes, k_in, k_out, v_in, v_out, (size_t)n, 0, 1, s);
}

}  // namespace fgpu
