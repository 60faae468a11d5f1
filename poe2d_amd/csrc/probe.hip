// Device bandwidth probe: plain grid-stride copy, read and write kernels over float4, timed with
// hipEvents.  A measurement helper for the roofline figures (the achievable HBM rate of simple
// streams on this device, and the calibration of the FETCH_SIZE / WRITE_SIZE counters against
// a copy of known size); not part of the time loop.
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/poe2d.h"
#include "common.h"

using poe2d::fail;

namespace {

__global__ void __launch_bounds__(256) copy_f4(const float4 *__restrict__ a, float4 *__restrict__ b, size_t n)
{
    for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) b[i] = a[i];
}

__global__ void __launch_bounds__(256) read_f4(const float4 *__restrict__ a, float *sink, size_t n)
{
    float s = 0.f;
    for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const float4 v = a[i];
        s += v.x + v.y + v.z + v.w;
    }
    if (s == -1.2345f) sink[0] = s;  // keeps the loads; the buffers hold zeros, so never taken
}

__global__ void __launch_bounds__(256) write_f4(float4 *__restrict__ b, size_t n)
{
    for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        b[i] = make_float4(1.f, 2.f, 3.f, 4.f);
}

}  // namespace

extern "C" poe2d_status poe2d_stream_probe(size_t bytes, int reps, double *copy_gbps, double *read_gbps,
                                           double *write_gbps)
{
    if (bytes < (1u << 20) || bytes % 16 || reps < 1 || !copy_gbps || !read_gbps || !write_gbps)
        return fail(POE2D_E_INVALID, "stream_probe: bytes must be a multiple of 16 and >= 1 MiB, reps >= 1");
    float4 *a = nullptr, *b = nullptr;
    float *sink = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    poe2d_status st = POE2D_OK;
    auto chk = [&](hipError_t e, const char *what) {
        if (e != hipSuccess && st == POE2D_OK)
            st = fail(e == hipErrorOutOfMemory ? POE2D_E_NOMEM : POE2D_E_HIP,
                      std::string("stream_probe: ") + what + ": " + hipGetErrorString(e));
        return e == hipSuccess;
    };
    const size_t n = bytes / 16;
    if (chk(hipMalloc(&a, bytes), "hipMalloc") && chk(hipMalloc(&b, bytes), "hipMalloc") &&
        chk(hipMalloc(&sink, sizeof(float)), "hipMalloc") && chk(hipMemset(a, 0, bytes), "hipMemset") &&
        chk(hipMemset(b, 0, bytes), "hipMemset") && chk(hipEventCreate(&e0), "hipEventCreate") &&
        chk(hipEventCreate(&e1), "hipEventCreate")) {
        double best[3] = {0, 0, 0};
        for (int grid : {1024, 2048, 4096, 8192, 16384}) {
            for (int kind = 0; kind < 3 && st == POE2D_OK; ++kind) {
                auto launch = [&]() {
                    if (kind == 0) copy_f4<<<grid, 256>>>(a, b, n);
                    else if (kind == 1) read_f4<<<grid, 256>>>(a, sink, n);
                    else write_f4<<<grid, 256>>>(b, n);
                };
                launch();
                if (!chk(hipDeviceSynchronize(), "warm-up")) break;
                chk(hipEventRecord(e0), "hipEventRecord");
                for (int r = 0; r < reps; ++r) launch();
                chk(hipEventRecord(e1), "hipEventRecord");
                if (!chk(hipEventSynchronize(e1), "hipEventSynchronize")) break;
                float ms = 0.f;
                chk(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
                const double moved = kind == 0 ? 2.0 * bytes : 1.0 * bytes;
                const double gbps = moved * reps / (ms * 1e-3) / 1e9;
                if (gbps > best[kind]) best[kind] = gbps;
            }
        }
        *copy_gbps = best[0];
        *read_gbps = best[1];
        *write_gbps = best[2];
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (sink) (void)hipFree(sink);
    return st;
}
