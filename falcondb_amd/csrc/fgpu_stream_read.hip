// fgpu_stream_read.hip -- the achievable HBM read rate of this box for the
// GROUP BY scan's access pattern: the four C2 columns (key, v1, v2: 8 bytes;
// f: 4 bytes; 100M rows) read once with the scan's tiling (a contiguous row
// range per workgroup, waves on 64 * U-row tiles striding by the wave count).
// A STREAM-like reference point beside the 8 TB/s peak; it is a standalone
// program (`python bench.py --ceiling` runs it) and not part of the library.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#define CK(x)                                                                           \
    do {                                                                                \
        hipError_t e_ = (x);                                                            \
        if (e_ != hipSuccess) {                                                         \
            fprintf(stderr, "HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
            return 1;                                                                   \
        }                                                                               \
    } while (0)

template <int U>
__global__ __launch_bounds__(1024) void k_stream_read(const int64_t *key, const int64_t *v1, const int64_t *v2,
                                                      const int32_t *f, int64_t n, int64_t rows_per_wg,
                                                      uint64_t *out) {
    const int64_t r0 = (int64_t)blockIdx.x * rows_per_wg;
    const int64_t r1 = r0 + rows_per_wg < n ? r0 + rows_per_wg : n;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int nw = __builtin_amdgcn_readfirstlane((int)(blockDim.x >> 6));
    constexpr int TR = 64 * U;
    const int n_tiles = r1 > r0 ? (int)((r1 - r0 + TR - 1) / TR) : 0;
    uint64_t acc = 0;
    for (int t = wave; t < n_tiles; t += nw) {
        int64_t k[U], a[U], b[U];
        int32_t c[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            int64_t r = r0 + (int64_t)t * TR + u * 64 + lane;
            r = r < r1 ? r : r1 - 1;  // every index stays in [r0, r1)
            c[u] = f[r];
            k[u] = key[r];
            a[u] = v1[r];
            b[u] = v2[r];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) acc += c[u] < 40 ? (uint64_t)(k[u] ^ a[u] ^ b[u]) : 1ull;
    }
    out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = acc;  // keeps the loads
}

template <int U>
static int run(const int64_t *key, const int64_t *v1, const int64_t *v2, const int32_t *f, int64_t n,
               uint64_t *out, int cus, int block, int wpc, double &best) {
    const int G = cus * wpc;
    const int64_t rows_per_wg = ((n + G - 1) / G + 255) / 256 * 256;
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    for (int i = 0; i < 3; ++i)
        hipLaunchKernelGGL(k_stream_read<U>, dim3(G), dim3(block), 0, 0, key, v1, v2, f, n, rows_per_wg, out);
    CK(hipGetLastError());
    const int reps = 20;
    CK(hipEventRecord(e0));
    for (int i = 0; i < reps; ++i)
        hipLaunchKernelGGL(k_stream_read<U>, dim3(G), dim3(block), 0, 0, key, v1, v2, f, n, rows_per_wg, out);
    CK(hipEventRecord(e1));
    CK(hipEventSynchronize(e1));
    float ms = 0;
    CK(hipEventElapsedTime(&ms, e0, e1));
    ms /= reps;
    const double gbs = n * 28.0 / (ms * 1e-3) / 1e9;
    best = gbs > best ? gbs : best;
    printf("{\"rows_per_thread_step\": %d, \"block\": %d, \"workgroups_per_cu\": %d, \"ms\": %.4f, \"GBs\": %.1f}\n",
           U, block, wpc, ms, gbs);
    CK(hipEventDestroy(e0));
    CK(hipEventDestroy(e1));
    return 0;
}

int main() {
    const int64_t n = 100000000;
    int64_t *key, *v1, *v2;
    int32_t *f;
    uint64_t *out;
    CK(hipMalloc(&key, n * 8));
    CK(hipMalloc(&v1, n * 8));
    CK(hipMalloc(&v2, n * 8));
    CK(hipMalloc(&f, n * 4));
    CK(hipMemset(key, 1, n * 8));
    CK(hipMemset(v1, 2, n * 8));
    CK(hipMemset(v2, 3, n * 8));
    CK(hipMemset(f, 0, n * 4));
    hipDeviceProp_t pr;
    CK(hipGetDeviceProperties(&pr, 0));
    const int cus = pr.multiProcessorCount;
    CK(hipMalloc(&out, (size_t)cus * 2 * 1024 * 8));  // grids below: <= 2 * cus workgroups of <= 1024 threads
    double best = 0;
    for (int wpc : {1, 2}) {
        if (run<1>(key, v1, v2, f, n, out, cus, 1024, wpc, best)) return 1;
        if (run<2>(key, v1, v2, f, n, out, cus, 1024, wpc, best)) return 1;
        if (run<4>(key, v1, v2, f, n, out, cus, 1024, wpc, best)) return 1;
    }
    printf("{\"stream_read_ceiling_GBs\": %.1f, \"bytes\": %lld}\n", best, (long long)(n * 28));
    CK(hipFree(key));
    CK(hipFree(v1));
    CK(hipFree(v2));
    CK(hipFree(f));
    CK(hipFree(out));
    return 0;
}
