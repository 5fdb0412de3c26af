// calib.hip — HBM streaming kernels and their timing entry points
// (libicempm_calib.so, declared in include/icempm_calib.h).  Each thread moves
// four 16-byte vectors per step with all loads issued before the stores; the
// known byte counts of these kernels give the device's streaming ceiling and
// the factor that turns the profiler's FETCH_SIZE / WRITE_SIZE into HBM bytes.
#include <hip/hip_runtime.h>

#include <string>

#include "icempm.h"
#include "icempm_calib.h"

namespace {

std::string g_error;

__global__ __launch_bounds__(256) void k_copy4(const float4 *__restrict__ a, float4 *__restrict__ b, size_t n4) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n4; i += 4 * stride) {
        const float4 v0 = a[i], v1 = a[i + stride], v2 = a[i + 2 * stride], v3 = a[i + 3 * stride];
        b[i] = v0;
        b[i + stride] = v1;
        b[i + 2 * stride] = v2;
        b[i + 3 * stride] = v3;
    }
    for (; i < n4; i += stride) b[i] = a[i];
}

// Read-only: sums the buffer; one word per thread is written only when the
// sum equals a value the data never produces (never, for finite input).
__global__ __launch_bounds__(256) void k_read4(const float4 *__restrict__ a, size_t n4, float *__restrict__ sink) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t t = i;
    float acc = 0.f;
    for (; i + 3 * stride < n4; i += 4 * stride) {
        const float4 v0 = a[i], v1 = a[i + stride], v2 = a[i + 2 * stride], v3 = a[i + 3 * stride];
        acc += (v0.x + v0.y + v0.z + v0.w) + (v1.x + v1.y + v1.z + v1.w) + (v2.x + v2.y + v2.z + v2.w) +
               (v3.x + v3.y + v3.z + v3.w);
    }
    for (; i < n4; i += stride) acc += a[i].x + a[i].y + a[i].z + a[i].w;
    if (acc == -1.2345e30f) sink[t] = acc;
}

// Write-only.
__global__ __launch_bounds__(256) void k_write4(float4 *__restrict__ b, size_t n4) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) b[i] = z;
}

}  // namespace

extern "C" {

int ice_stream_bench(int device, size_t bytes, int reps, int mode, double *ms_per_pass) {
    g_error.clear();
    if (!ms_per_pass || reps <= 0 || bytes < 1024 || mode < 0 || mode > 2) {
        g_error = "ice_stream_bench: ms_per_pass required, reps > 0, bytes >= 1024, mode 0..2";
        return ICE_E_ARG;
    }
    *ms_per_pass = 0.0;
    if (hipSetDevice(device) != hipSuccess) { g_error = "hipSetDevice failed"; return ICE_E_DEVICE; }
    const size_t n4 = bytes / 16;
    const unsigned blocks = 8192, threads = 256;
    float4 *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = ICE_OK;
    auto ok = [&](hipError_t e) {
        if (e != hipSuccess && rc == ICE_OK) { g_error = hipGetErrorString(e); rc = ICE_E_HIP; }
        return rc == ICE_OK;
    };
    // b: the copy destination, or the read kernel's sink (one float per thread)
    const size_t b_bytes = mode == 1 ? (size_t)blocks * threads * sizeof(float) : n4 * 16;
    auto launch = [&]() {
        if (mode == 0) hipLaunchKernelGGL(k_copy4, dim3(blocks), dim3(threads), 0, st, a, b, n4);
        else if (mode == 1) hipLaunchKernelGGL(k_read4, dim3(blocks), dim3(threads), 0, st, a, n4, (float *)b);
        else hipLaunchKernelGGL(k_write4, dim3(blocks), dim3(threads), 0, st, a, n4);
    };
    if (ok(hipMalloc(&a, n4 * 16)) && ok(hipMalloc(&b, b_bytes)) && ok(hipStreamCreate(&st)) &&
        ok(hipEventCreate(&e0)) && ok(hipEventCreate(&e1)) && ok(hipMemsetAsync(a, 0, n4 * 16, st))) {
        launch();   // warm-up
        ok(hipEventRecord(e0, st));
        for (int r = 0; r < reps; r++) launch();
        ok(hipGetLastError());
        ok(hipEventRecord(e1, st));
        float ms = 0.f;
        if (ok(hipEventSynchronize(e1)) && ok(hipEventElapsedTime(&ms, e0, e1))) *ms_per_pass = ms / reps;
    }
    if (st) (void)hipStreamSynchronize(st);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}

int ice_copy_bench(int device, size_t bytes, int reps, double *ms_per_copy) {
    return ice_stream_bench(device, bytes, reps, 0, ms_per_copy);
}

const char *ice_calib_last_error(void) { return g_error.c_str(); }

}  // extern "C"
