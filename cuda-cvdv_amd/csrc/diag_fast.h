// diag_fast.h - pending diagonal phases compiled for fast per-element use.
//
// Before a kernel that loads the state runs, the part of the pending DiagProg
// that does not commute with it is split into:
//   * tables: single-register polynomial terms (and Rz bit terms) become
//     tables of e^{2 pi i phase} over that register's local index, evaluated in
//     FP64 (sincospi) by k_make_tables and rounded once to the state's
//     precision.  Terms over the same register share one table, with a
//     sub-table per combination of their control bits (choice bits); one-bit
//     terms on different bits share a table over the combinations of those
//     bits; a lone conditional term keeps one table whose entry its control
//     bit conjugates.  An element reads one entry per table.  The global phase
//     and, for an FFT, the 1/sqrt(N) normalisation are folded in FP64 into a
//     table that every element reads unconjugated, so no rounded constant
//     factor is applied per element;
//   * generic terms: bilinear q1*q2 terms (and polynomial terms of registers
//     too large to tabulate), evaluated per element as one FP64 angle with
//     range reduction, then sin/cos;
//   * a one-entry table carries the global phase and scale when no table
//     is read by every element.
// Capacities are fixed (CVDV_FAST_TABS tables, CVDV_FAST_GEN generic terms).
// Tables beyond them are applied by table passes of their own first; too many
// generic terms, by one FP64 pass of the whole program.
#pragma once

#include "cvdv_internal.h"

#define CVDV_FAST_TABS 6
#define CVDV_FAST_GEN 4
#define CVDV_TAB_SEL 4        // choice bits of one table
#define CVDV_TAB_MERGE_BITS 16  // a table shared by several terms holds at most 2^16 entries

// Entry of element g: off + local(g over [sh, sh+nb)) + (choice(g) << nb), where
// bit j < nsel of the choice is bit ((sel >> 8j) & 63) of g; ctrl >= 0
// conjugates the entry when that bit of g is set.
struct DiagTab {
    int sh, nb, ctrl;
    uint32_t off;  // first entry in the table arena
    int nsel;
    uint32_t sel;  // choice bit positions, one per byte
};

__host__ __device__ inline int tab_sel(const DiagTab& t, int j) { return (int)((t.sel >> (8 * j)) & 63u); }

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

__device__ __forceinline__ uint32_t tab_choice(const DiagTab& t, uint64_t g) {
    uint32_t c = 0;
#pragma unroll
    for (int j = 0; j < CVDV_TAB_SEL; ++j)
        if (j < t.nsel) c |= (uint32_t)((g >> tab_sel(t, j)) & 1ull) << j;
    return c << t.nb;
}

template <typename C>
__device__ __forceinline__ C tab_entry(const void* arena, const DiagTab& t, uint64_t g) {
    uint32_t i = t.off + (uint32_t)((g >> t.sh) & ((1ull << t.nb) - 1ull));
    if (t.nsel > 0) i += tab_choice(t, g);
    C e = ((const C*)arena)[i];
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
// tileBits: index bits that vary within one tile of a consumer that reads its
// tables through a per-tile base; no choice bit is placed on them, so such a
// consumer computes the choice once per tile.
FastDiag compile_diag(CVDVContext* ctx, const DiagProg& p, bool& active, double scale = 1.0,
                      bool* scaleFolded = nullptr, uint64_t tileBits = 0);

// p with its linear polynomial terms over exactly the field [sh, sh + nb)
// (control bit outside it) marked DT_LINF, as many as the generic-term slots
// allow: a consumer that folds such terms per butterfly (the one-row FFT tile)
// evaluates them in place instead of reading a table per element.
DiagProg diag_fold_linear(const DiagProg& p, int sh, int nb);

// compile_diag(p, ..., tileBits) launches nothing but the table build: no
// FP64 pass and no table pass of its own (with a table slot to spare).
bool diag_fits(const DiagProg& p, uint64_t tileBits);

// Largest table arena compile_diag can need for registers of at most maxNb qubits.
size_t diag_arena_bound(int maxNb, size_t elemBytes);
