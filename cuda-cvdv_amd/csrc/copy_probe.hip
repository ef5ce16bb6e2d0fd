// copy_probe.hip - achievable-bandwidth references for the FFT passes.
//
// Copies of one 2^28-amplitude complex64 state (2 GiB read + 2 GiB written)
// with exactly the access pattern of the FFT tiles of a [14,14] state and no
// arithmetic: the ceilings the FFT kernels' HBM rates are compared with.  A
// plain vectorised copy is the calibration kernel for the FETCH_SIZE /
// WRITE_SIZE counters (known bytes per launch).  Measurement only; nothing in
// the gate path calls this.

#include <cstdint>

#include "cvdv_internal.h"
#include "../../include/cvdv_hip.h"

namespace {

constexpr uint64_t kAmps = 1ull << 28;

__global__ void k_probe_copy4(const float4* __restrict__ in, float4* __restrict__ out, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = in[i];
}

// Strided four-step tile: 128 rows at stride 2^21 x 32 consecutive columns;
// block t covers columns [32t, 32t + 32) (65536 blocks).
__global__ void __launch_bounds__(256) k_probe_strided(const float2* in, float2* out) {
    const uint64_t lo = (uint64_t)blockIdx.x * 32;
    float2 v[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int e = threadIdx.x + k * 256;
        v[k] = in[lo + ((uint64_t)(e >> 5) << 21) + (e & 31)];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int e = threadIdx.x + k * 256;
        out[lo + ((uint64_t)(e >> 5) << 21) + (e & 31)] = v[k];
    }
}

// Transposing four-step tile: rows at stride 2^14 of block h2 (base h2 * 2^21)
// written at stride 2^21 from base h2 * 2^14.
__global__ void __launch_bounds__(256) k_probe_transpose(const float2* in, float2* out) {
    const uint64_t lo = (uint64_t)(blockIdx.x & 511) * 32, h2 = blockIdx.x >> 9;
    float2 v[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int e = threadIdx.x + k * 256;
        v[k] = in[(h2 << 21) + lo + ((uint64_t)(e >> 5) << 14) + (e & 31)];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int e = threadIdx.x + k * 256;
        out[(h2 << 14) + lo + ((uint64_t)(e >> 5) << 21) + (e & 31)] = v[k];
    }
}

// Mid-digit tile of the same register: 128 rows at stride 2^14 x 32 columns,
// block t = (h2, column block), in place.
__global__ void __launch_bounds__(256) k_probe_mid(const float2* in, float2* out) {
    const uint64_t lo = (uint64_t)(blockIdx.x & 511) * 32, h2 = blockIdx.x >> 9;
    float2 v[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int e = threadIdx.x + k * 256;
        v[k] = in[(h2 << 21) + lo + ((uint64_t)(e >> 5) << 14) + (e & 31)];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int e = threadIdx.x + k * 256;
        out[(h2 << 21) + lo + ((uint64_t)(e >> 5) << 14) + (e & 31)] = v[k];
    }
}

// One contiguous row of 2^14 per workgroup (1024 threads x 16 elements),
// grid-stride over the 2^14 rows.
__global__ void __launch_bounds__(1024) k_probe_row(const float2* in, float2* out, int nRows) {
    for (int t = blockIdx.x; t < nRows; t += gridDim.x) {
        const uint64_t base = (uint64_t)t << 14;
        float2 v[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) v[k] = in[base + threadIdx.x + k * 1024];
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 16; ++k) out[base + threadIdx.x + k * 1024] = v[k];
        __syncthreads();
    }
}

}  // namespace

extern "C" {

int cvdvCopyProbe(int pattern, int reps, double* msOut) {
    if (!msOut || reps < 1 || pattern < 0 || pattern >= CVDV_PROBE_COUNT) {
        cvdv_set_error(CVDV_ERR_ARG, "cvdvCopyProbe: pattern out of range, reps < 1 or no output");
        return CVDV_ERR_ARG;
    }
    float2 *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = CVDV_OK;
    try {
        HIPCHK(hipMalloc(&a, kAmps * sizeof(float2)));
        HIPCHK(hipMalloc(&b, kAmps * sizeof(float2)));
        HIPCHK(hipMemset(a, 0, kAmps * sizeof(float2)));
        HIPCHK(hipMemset(b, 0, kAmps * sizeof(float2)));
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        int dev = 0, cus = 0;
        HIPCHK(hipGetDevice(&dev));
        HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        auto launch = [&] {
            switch (pattern) {
                case CVDV_PROBE_COPY:
                    hipLaunchKernelGGL(k_probe_copy4, dim3(cus * 4), dim3(256), 0, 0, (const float4*)a, (float4*)b,
                                       kAmps / 2);
                    break;
                case CVDV_PROBE_MEMCPY:
                    HIPCHK(hipMemcpyAsync(b, a, kAmps * sizeof(float2), hipMemcpyDeviceToDevice, 0));
                    break;
                case CVDV_PROBE_STRIDED:
                    hipLaunchKernelGGL(k_probe_strided, dim3(65536), dim3(256), 0, 0, a, a);
                    break;
                case CVDV_PROBE_TRANSPOSE:
                    hipLaunchKernelGGL(k_probe_transpose, dim3(65536), dim3(256), 0, 0, a, b);
                    break;
                case CVDV_PROBE_MID:
                    hipLaunchKernelGGL(k_probe_mid, dim3(65536), dim3(256), 0, 0, a, a);
                    break;
                default:
                    hipLaunchKernelGGL(k_probe_row, dim3(cus), dim3(1024), 0, 0, a, a, 16384);
                    break;
            }
            HIPCHK(hipGetLastError());
        };
        launch();
        HIPCHK(hipDeviceSynchronize());
        HIPCHK(hipEventRecord(e0, 0));
        for (int r = 0; r < reps; ++r) launch();
        HIPCHK(hipEventRecord(e1, 0));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        *msOut = (double)ms / reps;
    } catch (const CvdvError& e) {
        cvdv_set_error(e.code, e.msg);
        rc = e.code;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}

}  // extern "C"
