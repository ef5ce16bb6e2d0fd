// diag_fast.h - pending diagonal phases compiled for fast per-element use.
//
// Before a kernel that loads the state runs, its pending DiagProg is split:
//   * every single-register polynomial term (and every Rz bit term) becomes a
//     table of e^{2 pi i phase(j)} over that register's local index, evaluated
//     in FP64 (sincospi) by k_make_tables and stored in the state's precision;
//     a control bit conjugates the looked-up entry;
//   * bilinear q1*q2 terms (and polynomial terms of registers too large to
//     tabulate) stay in a residual DiagProg evaluated per element (FP64 angle
//     with range reduction, then sin/cos);
//   * the global phase is one FP64 complex factor.
// Applying the program to an element is then a few cached table reads and
// complex multiplies instead of an FP64 angle evaluation and sin/cos per term.
#pragma once

#include "cvdv_internal.h"

struct DiagTab {
    int sh, nb, ctrl;
    uint32_t off;  // first entry in the table arena
};

struct FastDiag {
    int ntab;
    int usegen;
    double2 gphase;
    const void* arena;
    DiagTab tab[CVDV_MAX_TERMS];
    DiagProg gen;
};

// Mixed precision: the angle is already reduced to [-1/2, 1/2] revolutions in
// FP64, so the hardware sin/cos (which take revolutions) apply directly.  Their
// error is ~2 ulp but |e|^2 - 1 averages -6.5e-8, which would shrink the norm
// on every phase pass.  One Newton step k = (3 - |e|^2)/2 removes the bias; it
// runs in FP64 because in FP32 a factor within 6e-8 of 1 rounds to exactly 1.
// The final rounding of c k and s k to FP32 is unbiased.
__device__ __forceinline__ float2 unit_rev_fast(double f, float) {
    const float ff = (float)f;
    const double s = (double)__builtin_amdgcn_sinf(ff), c = (double)__builtin_amdgcn_cosf(ff);
    const double k = fma(-0.5, fma(c, c, s * s), 1.5);
    return make_float2((float)(c * k), (float)(s * k));
}
__device__ __forceinline__ double2 unit_rev_fast(double f, double) { return unit_rev(f, 0.0); }

template <typename C>
__device__ __forceinline__ C apply_fast(const FastDiag& f, uint64_t g, C v) {
    using R = decltype(v.x);
    C fac = cmk<C>(f.gphase.x, f.gphase.y);
    const C* ar = (const C*)f.arena;
#pragma unroll
    for (int i = 0; i < CVDV_MAX_TERMS; ++i) {  // constant indices: see diag_rev
        if (i < f.ntab) {
            const DiagTab& t = f.tab[i];
            C e = ar[t.off + ((g >> t.sh) & ((1ull << t.nb) - 1ull))];
            if (t.ctrl >= 0 && ((g >> t.ctrl) & 1ull)) e.y = -e.y;
            fac = cmul(fac, e);
        }
    }
    if (f.usegen) fac = cmul(fac, unit_rev_fast(diag_rev(f.gen, g), R(0)));
    return cmul(v, fac);
}

// Host side (gates.hip): build the tables on ctx->stream and return the
// compiled program; `active` is false when the program is the identity.
FastDiag compile_diag(CVDVContext* ctx, const DiagProg& p, bool& active);
