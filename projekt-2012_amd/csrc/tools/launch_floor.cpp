// launch_floor: the time per launch of back-to-back kernels on one stream, each waiting for
// the one before (the order the sweeps' one-launch-per-anti-diagonal design relies on).  An
// empty kernel gives the dispatch floor every sweep launch pays; a kernel whose waves chase
// two dependent loads (a span, then a region record, as make_tile does) and store one value
// per lane gives the floor with a sweep launch's set-up reads.  One, 256 and 1024
// workgroups of 256 threads.  Prints one JSON line.
#include <hip/hip_runtime.h>
#include <cstdio>

namespace {

struct Args {
    const int* spans;
    const int4* regions;
    int* out;
    int n;
};

__global__ __launch_bounds__(256) void k_empty(Args) {}

__global__ __launch_bounds__(256) void k_chain(Args a) {
    const int sp = a.spans[0];
    const int4 r = a.regions[sp];
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t < a.n) a.out[t] = r.x + t;
}

}  // namespace

int main() {
    int *spans = nullptr, *out = nullptr;
    int4* regions = nullptr;
    constexpr int kMaxBlocks = 1024, kReps = 20000;
    if (hipMalloc(&spans, sizeof(int)) != hipSuccess || hipMalloc(&regions, sizeof(int4)) != hipSuccess ||
        hipMalloc(&out, kMaxBlocks * 256 * sizeof(int)) != hipSuccess) {
        fprintf(stderr, "launch_floor: no HIP device or out of memory\n");
        return 2;
    }
    if (hipMemset(spans, 0, sizeof(int)) != hipSuccess || hipMemset(regions, 0, sizeof(int4)) != hipSuccess) return 2;
    hipStream_t s;
    hipEvent_t e0, e1;
    if (hipStreamCreate(&s) != hipSuccess || hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        return 2;
    printf("{\"unit\": \"us per launch\", \"launches\": %d", kReps);
    for (int kind = 0; kind < 2; ++kind) {
        printf(", \"%s\": {", kind ? "chain" : "empty");
        const int blocks_list[3] = {1, 256, kMaxBlocks};
        for (int bi = 0; bi < 3; ++bi) {
            const int blocks = blocks_list[bi];
            const Args a{spans, regions, out, blocks * 256};
            float ms = 0.f;
            for (int pass = 0; pass < 2; ++pass) {  // the first pass warms up
                if (hipEventRecord(e0, s) != hipSuccess) return 2;
                for (int r = 0; r < kReps; ++r) {
                    if (kind) hipLaunchKernelGGL(k_chain, dim3(blocks), dim3(256), 0, s, a);
                    else hipLaunchKernelGGL(k_empty, dim3(blocks), dim3(256), 0, s, a);
                }
                if (hipGetLastError() != hipSuccess || hipEventRecord(e1, s) != hipSuccess ||
                    hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
                    return 2;
            }
            printf("%s\"%d_workgroups\": %.3f", bi ? ", " : "", blocks, ms * 1000.f / kReps);
        }
        printf("}");
    }
    printf("}\n");
    return 0;
}
