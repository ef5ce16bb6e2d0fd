// fft_bfly.h - radix-2/4/8 DFT butterflies shared by the tile FFT (fft.hip)
// and the batched-circuit interpreter (batch.hip).  Forward sign e^{-2 pi i/N};
// INV selects e^{+2 pi i/N}.
#pragma once

#include "cvdv_internal.h"

// Workgroup barrier that orders LDS traffic only.  __syncthreads() also waits
// for every outstanding global access (one counter covers loads and stores on
// this architecture), which would stall prefetched loads or stores still in
// flight at every Stockham exchange.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <bool INV, typename C>
__device__ __forceinline__ C mul_w(C a, C w) {
    return INV ? cmulc(a, w) : cmul(a, w);
}

template <bool INV, typename C>
__device__ __forceinline__ C rot_mi(C a) {  // a * (-i) forward, a * (+i) inverse
    C r;
    if (INV) { r.x = -a.y; r.y = a.x; } else { r.x = a.y; r.y = -a.x; }
    return r;
}

template <typename C> __device__ __forceinline__ C cadd(C a, C b) { C r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { C r; r.x = a.x - b.x; r.y = a.y - b.y; return r; }

template <bool INV, typename C>
__device__ __forceinline__ void dft2(C* x) {
    const C a = x[0], b = x[1];
    x[0] = cadd(a, b);
    x[1] = csub(a, b);
}

template <bool INV, typename C>
__device__ __forceinline__ void dft4(C& x0, C& x1, C& x2, C& x3) {
    const C t0 = cadd(x0, x2), t1 = csub(x0, x2);
    const C t2 = cadd(x1, x3), t3 = rot_mi<INV>(csub(x1, x3));
    x0 = cadd(t0, t2);
    x2 = csub(t0, t2);
    x1 = cadd(t1, t3);
    x3 = csub(t1, t3);
}

// v / sqrt(2).  In FP32 the constant is carried as hi + lo: a single rounded
// constant is 1.2e-8 low, which shrinks the norm on every radix-8 stage.  The
// FMA takes the exact product v*hi and adds the rounded small term v*lo, so the
// result is v/sqrt(2) rounded once (the other order, fma(v, lo, v*hi), rounds
// v*hi first and the small term can no longer change the result).
__device__ __forceinline__ double mul_rsqrt2(double v) { return v * 0.70710678118654752440; }
__device__ __forceinline__ float mul_rsqrt2(float v) {
    const float hi = 0.707106769084930419921875f;  // nearest float below 1/sqrt(2)
    const float lo = 1.2101617485882343e-08f;      // 1/sqrt(2) - hi
    return fmaf(v, hi, v * lo);
}

template <bool INV, typename C>
__device__ __forceinline__ void dft8(C* x) {
    C a0 = x[0], a1 = x[2], a2 = x[4], a3 = x[6];
    C b0 = x[1], b1 = x[3], b2 = x[5], b3 = x[7];
    dft4<INV>(a0, a1, a2, a3);
    dft4<INV>(b0, b1, b2, b3);
    // W8^1 = (1 - i)/sqrt2, W8^2 = -i, W8^3 = (-1 - i)/sqrt2 (forward)
    C w1b, w3b;
    if (!INV) {
        w1b.x = mul_rsqrt2(b1.x + b1.y); w1b.y = mul_rsqrt2(b1.y - b1.x);
        w3b.x = mul_rsqrt2(b3.y - b3.x); w3b.y = -mul_rsqrt2(b3.x + b3.y);
    } else {
        w1b.x = mul_rsqrt2(b1.x - b1.y); w1b.y = mul_rsqrt2(b1.y + b1.x);
        w3b.x = -mul_rsqrt2(b3.x + b3.y); w3b.y = mul_rsqrt2(b3.x - b3.y);
    }
    const C w2b = rot_mi<INV>(b2);
    x[0] = cadd(a0, b0); x[4] = csub(a0, b0);
    x[1] = cadd(a1, w1b); x[5] = csub(a1, w1b);
    x[2] = cadd(a2, w2b); x[6] = csub(a2, w2b);
    x[3] = cadd(a3, w3b); x[7] = csub(a3, w3b);
}

template <int RP, bool INV, typename C>
__device__ __forceinline__ void dft_r(C* x) {
    if (RP == 2) dft2<INV>(x);
    else if (RP == 4) dft4<INV>(x[0], x[1], x[2], x[3]);
    else dft8<INV>(x);
}

