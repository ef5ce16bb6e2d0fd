// diag_fast.h - pending diagonal phases compiled for fast per-element use.
//
// Before a kernel that loads the state runs, the part of the pending DiagProg
// that does not commute with it is split into:
//   * tables: every single-register polynomial term (and every Rz bit term)
//     becomes a table of e^{2 pi i phase(j)} over that register's local index,
//     evaluated in FP64 (sincospi) by k_make_tables and rounded once to the
//     state's precision; a control bit conjugates the looked-up entry.  The
//     global phase and, for an FFT, the 1/sqrt(N) normalisation are folded
//     into an unconditional table in FP64, so no rounded constant factor is
//     applied per element;
//   * generic terms: bilinear q1*q2 terms (and polynomial terms of registers
//     too large to tabulate), evaluated per element as one FP64 angle with
//     range reduction, then sin/cos;
//   * a one-entry table carries the global phase and scale when no
//     unconditional table exists.
// Capacities are fixed (CVDV_FAST_TABS tables, CVDV_FAST_GEN generic terms); a
// program that does not fit is applied by a separate FP64 pass first.
#pragma once

#include "cvdv_internal.h"

#define CVDV_FAST_TABS 6
#define CVDV_FAST_GEN 4

struct DiagTab {
    int sh, nb, ctrl;
    uint32_t off;  // first entry in the table arena
};

struct FastDiag {
    int ntab, ngen;
    double genrev;   // constant revolutions added to the generic terms' angle
    const void* arena;
    DiagTab tab[CVDV_FAST_TABS];
    DiagTerm gen[CVDV_FAST_GEN];
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
__device__ __forceinline__ C tab_entry(const void* arena, const DiagTab& t, uint64_t g) {
    C e = ((const C*)arena)[t.off + ((g >> t.sh) & ((1ull << t.nb) - 1ull))];
    if (t.ctrl >= 0 && ((g >> t.ctrl) & 1ull)) e.y = -e.y;
    return e;
}

// Phase factor of element g.  The loops are unrolled to constant indices so a
// FastDiag passed as a kernel argument is read in place (a runtime index would
// make the compiler copy the whole argument into per-lane scratch memory).
template <typename C>
__device__ __forceinline__ C fast_factor(const FastDiag& f, uint64_t g) {
    using R = decltype(C{}.x);
    C fac = cmk<C>(1.0, 0.0);
    if (f.ntab > 0) fac = tab_entry<C>(f.arena, f.tab[0], g);
#pragma unroll
    for (int i = 1; i < CVDV_FAST_TABS; ++i)
        if (i < f.ntab) fac = cmul(fac, tab_entry<C>(f.arena, f.tab[i], g));
    if (f.ngen > 0) {
        double r = f.genrev;
#pragma unroll
        for (int i = 0; i < CVDV_FAST_GEN; ++i)
            if (i < f.ngen) r += cvdv_frac(diag_term_rev(f.gen[i], g));
        const C u = unit_rev_fast(cvdv_frac(r), R(0));
        fac = f.ntab > 0 ? cmul(fac, u) : u;
    }
    return fac;
}

template <typename C>
__device__ __forceinline__ C apply_fast(const FastDiag& f, uint64_t g, C v) {
    return cmul(v, fast_factor<C>(f, g));
}

// Host side (gates.hip): build the tables on ctx->stream and return the
// compiled program; `active` is false when the program is the identity.
// scale != 1 is folded into an unconditional table when one exists
// (*scaleFolded tells the caller whether it still has to apply it).
FastDiag compile_diag(CVDVContext* ctx, const DiagProg& p, bool& active, double scale = 1.0,
                      bool* scaleFolded = nullptr);
