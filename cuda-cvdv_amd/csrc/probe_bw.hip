// Copy-bandwidth probe for the FFT passes on one 2^28-amplitude complex64 state
// (2 GiB): a plain float4 copy, and copies with exactly the access pattern of the
// FFT tiles of a [14,14] state (no arithmetic).  These are the achievable-rate
// ceilings the FFT kernels are compared with.
//
//   make -C cuda-cvdv_amd/csrc probe && cuda-cvdv_amd/hipcvdv/probe_bw
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#define CK(x)                                                                         \
    do {                                                                              \
        hipError_t e_ = (x);                                                          \
        if (e_ != hipSuccess) {                                                       \
            std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
            return 1;                                                                 \
        }                                                                             \
    } while (0)

static constexpr uint64_t kN = 1ull << 28;  // complex64 amplitudes

__global__ void k_copy4(const float4* __restrict__ in, float4* __restrict__ out, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = in[i];
}

// Four-step pass 1 of the strided register: 128 rows at stride 2^21 x 32 columns;
// tile t covers columns [32t, 32t + 32), 65536 tiles.
__global__ void __launch_bounds__(256) k_pass1(const float2* in, float2* out) {
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

// Four-step pass 2: rows at stride 2^14 of block h2 (base h2 * 2^21), written
// transposed (rows at stride 2^21, base h2 * 2^14).
__global__ void __launch_bounds__(256) k_pass2(const float2* in, float2* out) {
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

// One contiguous row of 2^14 per tile, 1024 threads x 16 elements, grid-stride over rows.
__global__ void __launch_bounds__(1024) k_row(const float2* in, float2* out, int nRows) {
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

int main() {
    float2 *a = nullptr, *b = nullptr;
    CK(hipMalloc(&a, kN * 8));
    CK(hipMalloc(&b, kN * 8));
    CK(hipMemset(a, 0, kN * 8));
    CK(hipMemset(b, 0, kN * 8));
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    int cus = 0;
    CK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0));
    std::printf("CUs %d; every line moves 2 GiB in + 2 GiB out\n", cus);
    auto timeit = [&](const char* name, auto launch) -> int {
        launch();
        CK(hipGetLastError());
        CK(hipDeviceSynchronize());
        CK(hipEventRecord(e0));
        for (int r = 0; r < 10; ++r) launch();
        CK(hipEventRecord(e1));
        CK(hipEventSynchronize(e1));
        float ms = 0;
        CK(hipEventElapsedTime(&ms, e0, e1));
        ms /= 10;
        std::printf("%-56s %8.3f ms %7.0f GB/s\n", name, ms, 2.0 * kN * 8 / (ms * 1e-3) / 1e9);
        std::fflush(stdout);
        return 0;
    };
    int rc = 0;
    rc |= timeit("float4 copy, grid 4 x CUs x 256", [&] {
        hipLaunchKernelGGL(k_copy4, dim3(cus * 4), dim3(256), 0, 0, (const float4*)a, (float4*)b, kN / 2);
    });
    rc |= timeit("hipMemcpy device to device", [&] { (void)hipMemcpyAsync(b, a, kN * 8, hipMemcpyDeviceToDevice, 0); });
    rc |= timeit("four-step pass 1 shape (128 x 256 B rows, stride 16 MiB)",
                 [&] { hipLaunchKernelGGL(k_pass1, dim3(65536), dim3(256), 0, 0, a, a); });
    rc |= timeit("four-step pass 2 shape (transposing)",
                 [&] { hipLaunchKernelGGL(k_pass2, dim3(65536), dim3(256), 0, 0, a, b); });
    rc |= timeit("one-row shape (128 KiB row per workgroup, 1 per CU)",
                 [&] { hipLaunchKernelGGL(k_row, dim3(cus), dim3(1024), 0, 0, a, a, 16384); });
    CK(hipFree(a));
    CK(hipFree(b));
    return rc;
}
