// HBM copy probe: calibrates the FETCH_SIZE / WRITE_SIZE counters read by
// rocprofv3 against a copy of known size, and reports the copy bandwidth.
//
//   emerald_amd/bandwidth_probe [MiB] [reps]
//
// One launch of probe_copy reads MiB MiB and writes MiB MiB (float4 grid-stride
// copy); the counter scale factor of a profile is (MiB * 1024) / counter_KiB.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ __launch_bounds__(256) void probe_copy(float4 const* __restrict__ src,
                                                  float4* __restrict__ dst, size_t n) {
    size_t const stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        dst[i] = src[i];
    }
}

#define CHECK(x)                                                                   \
    do {                                                                           \
        hipError_t e_ = (x);                                                       \
        if (e_ != hipSuccess) {                                                    \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));          \
            return 1;                                                              \
        }                                                                          \
    } while (0)

int main(int argc, char** argv) {
    size_t const mib = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 1024;
    int const reps = argc > 2 ? std::atoi(argv[2]) : 10;
    size_t const bytes = mib << 20;
    size_t const n = bytes / sizeof(float4);
    float4* a = nullptr;
    float4* b = nullptr;
    CHECK(hipMalloc(&a, bytes));
    CHECK(hipMalloc(&b, bytes));
    CHECK(hipMemset(a, 1, bytes));
    CHECK(hipMemset(b, 0, bytes));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    int const blocks = 256 * 8 * 4;
    hipLaunchKernelGGL(probe_copy, dim3(blocks), dim3(256), 0, 0, a, b, n);  // warm-up
    CHECK(hipDeviceSynchronize());
    CHECK(hipEventRecord(e0, 0));
    for (int r = 0; r < reps; ++r) {
        hipLaunchKernelGGL(probe_copy, dim3(blocks), dim3(256), 0, 0, a, b, n);
    }
    CHECK(hipEventRecord(e1, 0));
    CHECK(hipEventSynchronize(e1));
    float ms = 0.0f;
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    double const s = ms / 1e3 / reps;
    std::printf("probe_copy %zu MiB: %.3f ms per launch, %.1f GB/s (read+write), "
                "expected FETCH_SIZE %zu KiB, WRITE_SIZE %zu KiB per launch\n",
                mib, s * 1e3, 2.0 * bytes / s / 1e9, bytes >> 10, bytes >> 10);
    CHECK(hipFree(a));
    CHECK(hipFree(b));
    return 0;
}
