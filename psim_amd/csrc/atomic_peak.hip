// atomic_peak: attainable atomic-add throughput of the card, the peak that the transport kernel's
// tally atomics are compared against.  Three patterns, each with every lane of every wave adding
// to its own address in every iteration (no two lanes of a wave share an address):
//   lds_f64   ds_add_f64 into a 1024-entry LDS array per block (the window tallies' flux)
//   lds_i32   ds_add_u32 into a 1024-entry LDS array per block (the window tallies' count)
//   hbm_f64   global FP64 atomic add into an array of 16 Mi doubles (the tallies of a batch that
//             does not fit LDS)
// Usage: atomic_peak [iterations]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

constexpr int BLOCK = 256;
constexpr int LDS_ENTRIES = 1024;  // a power of two: index & (LDS_ENTRIES - 1)
constexpr size_t HBM_ENTRIES = size_t(1) << 24;  // a power of two

__global__ void __launch_bounds__(BLOCK) lds_f64_kernel(double* __restrict__ out, int iters) {
    __shared__ double buf[LDS_ENTRIES];
    for (int i = threadIdx.x; i < LDS_ENTRIES; i += BLOCK) { buf[i] = 0.; }
    __syncthreads();
    const double v = 1.0 + 1e-3 * threadIdx.x;
    for (int k = 0; k < iters; ++k) { atomicAdd(&buf[(threadIdx.x + k * BLOCK) & (LDS_ENTRIES - 1)], v); }
    __syncthreads();
    if (threadIdx.x == 0) { out[blockIdx.x] = buf[0]; }
}

__global__ void __launch_bounds__(BLOCK) lds_i32_kernel(double* __restrict__ out, int iters) {
    __shared__ int buf[LDS_ENTRIES];
    for (int i = threadIdx.x; i < LDS_ENTRIES; i += BLOCK) { buf[i] = 0; }
    __syncthreads();
    const int v = 1 + (int)(threadIdx.x & 3);
    for (int k = 0; k < iters; ++k) { atomicAdd(&buf[(threadIdx.x + k * BLOCK) & (LDS_ENTRIES - 1)], v); }
    __syncthreads();
    if (threadIdx.x == 0) { out[blockIdx.x] = (double)buf[0]; }
}

__global__ void __launch_bounds__(BLOCK) hbm_f64_kernel(double* __restrict__ buf, int iters) {
    const size_t gid = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    const size_t threads = (size_t)gridDim.x * BLOCK;
    const double v = 1.0;
    for (int k = 0; k < iters; ++k) { atomicAdd(&buf[(gid + (size_t)k * threads) & (HBM_ENTRIES - 1)], v); }
}

#define CHECK(x)                                                                     \
    do {                                                                             \
        hipError_t e = (x);                                                          \
        if (e != hipSuccess) {                                                       \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e));              \
            return 1;                                                                \
        }                                                                            \
    } while (0)

int main(int argc, char** argv) {
    const int iters = argc > 1 ? std::atoi(argv[1]) : 4096;
    if (iters <= 0) {
        std::fprintf(stderr, "iterations must be positive\n");
        return 1;
    }
    hipDeviceProp_t prop;
    CHECK(hipGetDeviceProperties(&prop, 0));
    const int grid = prop.multiProcessorCount * 8;  // as many blocks as the transport kernel
    double *out = nullptr, *buf = nullptr;
    CHECK(hipMalloc(&out, sizeof(double) * (size_t)grid));
    CHECK(hipMalloc(&buf, sizeof(double) * HBM_ENTRIES));
    CHECK(hipMemset(buf, 0, sizeof(double) * HBM_ENTRIES));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    const double adds = (double)grid * BLOCK * iters;
    double rate[3] = { 0., 0., 0. };
    for (int which = 0; which < 3; ++which) {
        for (int rep = 0; rep < 2; ++rep) {  // the first launch warms up; the second is timed
            CHECK(hipEventRecord(e0, 0));
            if (which == 0) {
                hipLaunchKernelGGL(lds_f64_kernel, dim3(grid), dim3(BLOCK), 0, 0, out, iters);
            } else if (which == 1) {
                hipLaunchKernelGGL(lds_i32_kernel, dim3(grid), dim3(BLOCK), 0, 0, out, iters);
            } else {
                hipLaunchKernelGGL(hbm_f64_kernel, dim3(grid), dim3(BLOCK), 0, 0, buf, iters / 16 > 0 ? iters / 16 : 1);
            }
            CHECK(hipGetLastError());
            CHECK(hipEventRecord(e1, 0));
            CHECK(hipEventSynchronize(e1));
        }
        float ms = 0.f;
        CHECK(hipEventElapsedTime(&ms, e0, e1));
        const double n = which == 2 ? (double)grid * BLOCK * (iters / 16 > 0 ? iters / 16 : 1) : adds;
        rate[which] = n / (ms / 1e3);
    }
    std::printf("{\"lds_f64_adds_per_s\": %.4e, \"lds_i32_adds_per_s\": %.4e, \"hbm_f64_adds_per_s\": %.4e, "
                "\"blocks\": %d, \"threads_per_block\": %d, \"iterations\": %d}\n",
        rate[0], rate[1], rate[2], grid, BLOCK, iters);
    CHECK(hipFree(out));
    CHECK(hipFree(buf));
    return 0;
}
