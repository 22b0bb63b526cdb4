// Streaming-bandwidth probe of the device: the rate a plain HBM-bound kernel of
// this project reaches, measured beside the stage kernels so their roofline
// fractions can be read against both the published peak and what a streaming
// kernel attains.  No counterpart in the reference (a measurement aid).
//
// Kernels: copy (b -> a), triad (a = b + s c) and read (sum of b), 16-byte
// vector accesses, four in flight per thread per iteration, grid-stride over
// a grid of a few workgroups per CU.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../../include/edus.h"

namespace {

typedef double v2d __attribute__((ext_vector_type(2)));
constexpr int kUnroll = 4;

template <bool NT>
__device__ __forceinline__ v2d ld(const v2d *p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}
template <bool NT>
__device__ __forceinline__ void st(v2d *p, v2d v) {
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// n: elements of 16 bytes, a multiple of kUnroll * blockDim.x * gridDim.x is not
// required (the tail is handled element by element).
template <bool NT>
__global__ __launch_bounds__(256) void k_copy(const v2d *__restrict__ b, v2d *__restrict__ a, long long n) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + (kUnroll - 1) * stride < n; i += kUnroll * stride) {
        v2d v[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; u++) v[u] = ld<NT>(b + i + u * stride);
#pragma unroll
        for (int u = 0; u < kUnroll; u++) st<NT>(a + i + u * stride, v[u]);
    }
    for (; i < n; i += stride) st<NT>(a + i, ld<NT>(b + i));
}

template <bool NT>
__global__ __launch_bounds__(256) void k_triad(const v2d *__restrict__ b, const v2d *__restrict__ c, v2d *__restrict__ a, double s,
                                               long long n) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + (kUnroll - 1) * stride < n; i += kUnroll * stride) {
        v2d x[kUnroll], y[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
            x[u] = ld<NT>(b + i + u * stride);
            y[u] = ld<NT>(c + i + u * stride);
        }
#pragma unroll
        for (int u = 0; u < kUnroll; u++) st<NT>(a + i + u * stride, x[u] + s * y[u]);
    }
    for (; i < n; i += stride) st<NT>(a + i, ld<NT>(b + i) + s * ld<NT>(c + i));
}

// Sum of b per workgroup into out[blockIdx.x] (keeps the loads live).
__global__ __launch_bounds__(256) void k_read(const v2d *__restrict__ b, double *__restrict__ out, long long n) {
    __shared__ double part[256];
    const long long stride = (long long)gridDim.x * blockDim.x;
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    v2d acc = {0.0, 0.0};
    for (; i + (kUnroll - 1) * stride < n; i += kUnroll * stride) {
        v2d v[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; u++) v[u] = b[i + u * stride];
#pragma unroll
        for (int u = 0; u < kUnroll; u++) acc += v[u];
    }
    for (; i < n; i += stride) acc += b[i];
    part[threadIdx.x] = acc.x + acc.y;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = part[0];
}

__global__ void k_fill(v2d *a, long long n, double v) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) a[i] = v2d{v, v + 1.0};
}

}  // namespace

extern "C" {

// Copy, triad and read rates (GB/s, 1e9 bytes) over arrays of `bytes` bytes each:
// the best of several workgroups-per-CU choices, each the mean of `reps` timed
// launches after one untimed launch.  Moves 2, 3 and 1 x bytes per launch.
int edus_stream_probe(int64_t bytes, int32_t reps, double *copy_gbs, double *triad_gbs, double *read_gbs) {
    if (bytes < (1 << 20) || reps < 1) return EDUS_ERR_INVALID;
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return EDUS_ERR_NODEVICE;
    const long long n = bytes / 16;
    v2d *a = nullptr, *b = nullptr, *c = nullptr;
    double *out = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    auto done = [&](int r) {
        if (s) (void)hipStreamSynchronize(s);
        (void)hipFree(a); (void)hipFree(b); (void)hipFree(c); (void)hipFree(out);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (s) (void)hipStreamDestroy(s);
        return r;
    };
    const int maxblk = 64 * ncu;
    if (hipMalloc(&a, n * 16) != hipSuccess || hipMalloc(&b, n * 16) != hipSuccess || hipMalloc(&c, n * 16) != hipSuccess ||
        hipMalloc(&out, (size_t)maxblk * sizeof(double)) != hipSuccess)
        return done(EDUS_ERR_NOMEM);
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess)
        return done(EDUS_ERR_HIP);
    hipLaunchKernelGGL(k_fill, dim3(4 * ncu), dim3(256), 0, s, a, n, 0.0);
    hipLaunchKernelGGL(k_fill, dim3(4 * ncu), dim3(256), 0, s, b, n, 1.0);
    hipLaunchKernelGGL(k_fill, dim3(4 * ncu), dim3(256), 0, s, c, n, 2.0);
    // kinds: copy (nontemporal), copy, triad (nontemporal), triad, read
    double best[3] = {0, 0, 0};
    const int per_cu[4] = {4, 8, 16, 64};
    for (int q = 0; q < 4 && rc == 0; q++) {
        const dim3 grid(per_cu[q] * ncu), block(256);
        for (int kind = 0; kind < 5; kind++) {
            for (int r = -1; r < reps; r++) {
                if (r == 0 && hipEventRecord(e0, s) != hipSuccess) rc = EDUS_ERR_HIP;
                switch (kind) {
                    case 0: hipLaunchKernelGGL(k_copy<true>, grid, block, 0, s, (const v2d *)b, a, n); break;
                    case 1: hipLaunchKernelGGL(k_copy<false>, grid, block, 0, s, (const v2d *)b, a, n); break;
                    case 2: hipLaunchKernelGGL(k_triad<true>, grid, block, 0, s, (const v2d *)b, (const v2d *)c, a, 0.5, n); break;
                    case 3: hipLaunchKernelGGL(k_triad<false>, grid, block, 0, s, (const v2d *)b, (const v2d *)c, a, 0.5, n); break;
                    default: hipLaunchKernelGGL(k_read, grid, block, 0, s, (const v2d *)b, out, n); break;
                }
            }
            float ms = 0;
            if (hipGetLastError() != hipSuccess || hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                rc = EDUS_ERR_HIP;
                break;
            }
            const int slot = kind < 2 ? 0 : kind < 4 ? 1 : 2;
            const double moved = (double)n * 16 * (slot == 0 ? 2 : slot == 1 ? 3 : 1);
            const double gbs = moved * reps / (ms * 1e-3) / 1e9;
            if (gbs > best[slot]) best[slot] = gbs;
        }
    }
    if (rc) return done(rc);
    if (copy_gbs) *copy_gbs = best[0];
    if (triad_gbs) *triad_gbs = best[1];
    if (read_gbs) *read_gbs = best[2];
    return done(0);
}

}  // extern "C"
