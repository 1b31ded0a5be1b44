// Plain device-to-device copy of a known number of bytes, used to calibrate the HBM byte
// counters (FETCH_SIZE, WRITE_SIZE) of the profiler: the copy reads and writes exactly
// n bytes per launch, so counter/n is the scale factor for every other kernel's reading.
//
//   tf_copy_calib [MiB]   (default 256; three launches of k_copy16 after one fill)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ __launch_bounds__(256) void k_fill16(uint4* p, size_t n16)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
        p[i] = make_uint4((unsigned)i, 1u, 2u, 3u);
}

__global__ __launch_bounds__(256) void k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n16)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

#define CHECK(x)                                                                         \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess) {                                                          \
            std::fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(e_));          \
            return 1;                                                                    \
        }                                                                                \
    } while (0)

int main(int argc, char** argv)
{
    const long mib = argc > 1 ? std::atol(argv[1]) : 256;
    if (mib <= 0 || mib > 16384) {
        std::fprintf(stderr, "size must be 1..16384 MiB\n");
        return 2;
    }
    const size_t n = (size_t)mib << 20, n16 = n / 16;
    uint4 *a = nullptr, *b = nullptr;
    CHECK(hipMalloc(&a, n));
    CHECK(hipMalloc(&b, n));
    const int grid = 256 * 8 * 4;  // 4 workgroups per CU, grid-stride over the buffer
    hipLaunchKernelGGL(k_fill16, dim3(grid), dim3(256), 0, 0, a, n16);
    CHECK(hipGetLastError());
    hipEvent_t t0, t1;
    CHECK(hipEventCreate(&t0));
    CHECK(hipEventCreate(&t1));
    for (int rep = 0; rep < 3; ++rep) {
        CHECK(hipEventRecord(t0, 0));
        hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, 0, a, b, n16);
        CHECK(hipGetLastError());
        CHECK(hipEventRecord(t1, 0));
        CHECK(hipEventSynchronize(t1));
        float ms = 0.f;
        CHECK(hipEventElapsedTime(&ms, t0, t1));
        std::printf("{\"kernel\": \"k_copy16\", \"bytes_read\": %zu, \"bytes_written\": %zu, \"ms\": %.4f, "
                    "\"GBps_read_plus_write\": %.1f}\n", n, n, ms, 2.0 * n / (ms * 1e-3) / 1e9);
    }
    CHECK(hipFree(a));
    CHECK(hipFree(b));
    return 0;
}
