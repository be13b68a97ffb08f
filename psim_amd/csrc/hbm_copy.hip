// hbm_copy: a plain device-to-device copy of a known size, used to calibrate the rocprofv3
// FETCH_SIZE / WRITE_SIZE counters (bytes seen per byte moved) and to measure the attainable
// HBM bandwidth of the card.  Usage: hbm_copy [MiB] [repeats]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void __launch_bounds__(256) copy_kernel(const double4* __restrict__ src, double4* __restrict__ dst, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) { dst[i] = src[i]; }
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
    const size_t mib = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 1024;
    const int reps = argc > 2 ? std::atoi(argv[2]) : 5;
    const size_t bytes = mib << 20;
    const size_t n = bytes / sizeof(double4);
    double4 *a = nullptr, *b = nullptr;
    CHECK(hipMalloc(&a, bytes));
    CHECK(hipMalloc(&b, bytes));
    CHECK(hipMemset(a, 1, bytes));
    hipDeviceProp_t prop;
    CHECK(hipGetDeviceProperties(&prop, 0));
    const int grid = prop.multiProcessorCount * 16;
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    hipLaunchKernelGGL(copy_kernel, dim3(grid), dim3(256), 0, 0, a, b, n);
    CHECK(hipDeviceSynchronize());
    CHECK(hipEventRecord(e0, 0));
    for (int r = 0; r < reps; ++r) { hipLaunchKernelGGL(copy_kernel, dim3(grid), dim3(256), 0, 0, a, b, n); }
    CHECK(hipEventRecord(e1, 0));
    CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    const double gbps = 2.0 * (double)bytes * reps / (ms / 1e3) / 1e9;
    std::printf("{\"bytes_read_per_launch\": %zu, \"bytes_written_per_launch\": %zu, \"launches\": %d, "
                "\"ms_per_launch\": %.4f, \"copy_GBps\": %.1f}\n",
        bytes, bytes, reps + 1, ms / reps, gbps);
    CHECK(hipFree(a));
    CHECK(hipFree(b));
    return 0;
}
