// valu_peak: measured 32-bit integer VALU ceiling of the device, the denominator the
// sweep kernels are judged against.  Every lane runs chains of dependent integer add/sub/max
// triples (the operations of the cell update) over 16 independent accumulators; the
// kernel is launched with 1..4 waves per SIMD.  Prints one JSON line.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>

namespace {

constexpr int kAcc = 16, kIters = 4096;

__global__ __launch_bounds__(256) void k_valu(int* out, int seed, int inc) {
    int acc[kAcc];
#pragma unroll
    for (int k = 0; k < kAcc; ++k) acc[k] = seed + k * (int)threadIdx.x;
    for (int it = 0; it < kIters; ++it) {
#pragma unroll
        for (int k = 0; k < kAcc; ++k) acc[k] = max(acc[k] + inc, acc[(k + 5) % kAcc] - inc);
    }
    int r = 0;
#pragma unroll
    for (int k = 0; k < kAcc; ++k) r ^= acc[k];
    if (r == 0x7fffffff) out[threadIdx.x] = r;  // keeps the chains alive
}

}  // namespace

int main() {
    hipDeviceProp_t pr;
    if (hipGetDeviceProperties(&pr, 0) != hipSuccess) {
        fprintf(stderr, "valu_peak: no HIP device\n");
        return 2;
    }
    int* out = nullptr;
    if (hipMalloc(&out, 1024 * sizeof(int)) != hipSuccess) return 2;
    hipEvent_t a, b;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    printf("{\"device_cus\": %d, \"clock_khz\": %d, \"per_waves_per_simd\": {", pr.multiProcessorCount, pr.clockRate);
    for (int wps = 1; wps <= 4; wps *= 2) {
        const int blocks = pr.multiProcessorCount * wps;  // 256 threads = one wave per SIMD
        hipLaunchKernelGGL(k_valu, dim3(blocks), dim3(256), 0, 0, out, 1, 3);
        (void)hipDeviceSynchronize();
        (void)hipEventRecord(a, 0);
        for (int r = 0; r < 5; ++r) hipLaunchKernelGGL(k_valu, dim3(blocks), dim3(256), 0, 0, out, r, 3);
        (void)hipEventRecord(b, 0);
        (void)hipEventSynchronize(b);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, a, b);
        // three VALU operations (add, sub, max) per accumulator per iteration and lane
        const double ops = 5.0 * blocks * 256.0 * kIters * kAcc * 3.0;
        printf("%s\"%d\": %.3f", wps > 1 ? ", " : "", wps, ops / (ms * 1e-3) / 1e12);
    }
    printf("}, \"unit\": \"int32 TOP/s\"}\n");
    (void)hipFree(out);
    return 0;
}
