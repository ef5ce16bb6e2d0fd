// batch.hip - thousands of small independent circuits in one launch.
//
// A batch owns one recording context per circuit (ctx_create_recording).  The
// gate entry points of gates.hip run unchanged on those contexts - same
// decompositions, same deferred phase programs, same lazy layout - but append
// ops (FFT along a register field, 2x2 qubit gate, conditional parity, phase
// pass) to the context's recording instead of launching.
//
// cvdvBatchRun executes every recorded circuit in ONE kernel launch (captured
// once in a hipGraph and replayed): one workgroup per circuit, circuits ordered
// by estimated cost (largest first, so the hardware dispatcher's in-order
// assignment acts as a longest-job-first schedule).  A workgroup
//   1. builds its initial product state from the per-register arrays,
//   2. interprets its op list with a workgroup barrier between ops, the state
//      held in LDS when it fits the launch's LDS budget, otherwise in a global
//      work area (a few hundred KiB per circuit, cache resident),
//   3. stores the state in canonical layout (pending phases applied, lazy
//      reflections and register swaps undone) to the batch output.
// FFTs run in place: each Stockham stage loads its butterflies' inputs into
// registers, crosses a barrier, and writes the outputs, so no ping-pong buffer
// is needed and an LDS-resident state may use the whole LDS budget.
//
// Readouts (marginal of one register per circuit, norms, single states) read
// the canonical output with a second kernel, one workgroup per circuit, in
// fixed order with compensated FP64 sums (bitwise reproducible).

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "cvdv_internal.h"
#include "diag_fast.h"
#include "fft_bfly.h"
#include "../../include/cvdv_hip.h"

#define BATCH_MAX_REG 8
#define BATCH_MAX_REG_QUBITS 12  // largest register (FFT length 2^12 = BATCH_EMAX * 512)
#define BATCH_MAX_QUBITS 20      // largest circuit (2^20 amplitudes)
#define BATCH_EMAX 8             // FFT elements held per thread per stage
#define BATCH_MAX_THREADS 512

// ---------------------------------------------------------------- device data
// Polynomial phase terms over one register field, merged on the host: for
// every sign pattern s of up to three control bits, one cubic
// c[s][0] m + c[s][1] m^2 + c[s][2] m^3 (revolutions) that sums the group's
// terms with their signs (a term whose control bit is set enters negated).
// An element then evaluates one cubic per field instead of one per term.
#define BATCH_GROUP_CTL 3
#define BATCH_MAX_GROUPS 2
struct BGroup {
    int sh, nb, nctl, pad;
    int ctl[4];
    double c[1 << BATCH_GROUP_CTL][3];
};

struct BOp {
    int kind, sh, nb, bit;
    int ctrlFlip, tgtRefl, inv, tfirst;
    int tn, gfirst, gn, pad;   // load program: generic terms [tfirst, +tn), groups [gfirst, +gn)
    double grev;      // global phase of the op's load program, revolutions
    double outScale;
    double2 m[4];
};

struct BCirc {
    uint64_t outOff;   // canonical output state (elements)
    uint64_t workOff;  // global work state (elements), used when lds == 0
    uint32_t opBegin, opEnd;
    uint32_t initOff;  // c128 entry of register 0's initial array; registers follow in order
    int T, nreg, lds;
    int ftFirst, ftN;  // trailing phase program, applied at the final store
    int fgFirst, fgN;
    double fgrev;
    uint8_t nq[BATCH_MAX_REG], flw[BATCH_MAX_REG], psh[BATCH_MAX_REG], refl[BATCH_MAX_REG];
};

struct BatchArgs {
    const BCirc* circ;
    const int* order;
    const BOp* ops;
    const DiagTerm* terms;
    const BGroup* groups;
    const double2* init;
    const void* tw;
    void* out;
    void* work;
    int twOff[BATCH_MAX_REG_QUBITS + 1];  // stage-table offset for FFT length 2^nb
    uint32_t ldsCap;  // elements of dynamic LDS per workgroup (staging for states in the work area)
};

// ---------------------------------------------------------------- device helpers
// An op's load program as staged in LDS: merged polynomial groups, the other
// (generic) terms, the global phase.  tab (optional): the whole program
// tabulated in LDS, entry (s << tnb) | j for the control pattern s (bit k:
// control bit (tctl >> 8k) & 255) and the field index j (bits [tsh,
// tsh + tnb)); each entry is the factor the evaluation below computes for
// that (s, j).
struct SProg {
    const DiagTerm* t;
    const BGroup* gr;
    int tn, gn;
    double grev;
    const void* tab;
    int tsh, tnb, tnctl, tctl;
    __device__ __forceinline__ bool active() const { return tn > 0 || gn > 0 || grev != 0.0; }
};

// Revolutions of group G at control pattern s and coordinate m, reduced mod 1.
// Contraction off: the product is rounded on its own before the reduction, so
// the per-element evaluation and a table entry (tabulate) agree bit for bit
// wherever the compiler inlines them.
__device__ __forceinline__ double group_rev(const BGroup& G, int s, double m) {
#pragma clang fp contract(off)
    const double y = fma(fma(G.c[s][2], m, G.c[s][1]), m, G.c[s][0]) * m;
    return y - rint(y);
}

// a * b with a fixed choice of fused products (contraction off), the same
// wherever it is inlined: a factor from a table and one computed in place give
// the same product bit for bit, and so do the twiddle products of the two
// Stockham stage forms (bstage, fstage), whichever a launch shape runs.
template <typename C>
__device__ __forceinline__ C cmul_fixed(C a, C b) {
#pragma clang fp contract(off)
    C r;
    r.x = fma(a.x, b.x, -(a.y * b.y));
    r.y = fma(a.x, b.y, a.y * b.x);
    return r;
}

template <typename C>
__device__ __forceinline__ C prog_apply(SProg P, uint32_t g, C v) {
    using R = decltype(v.x);
    if (P.tab) {
        uint32_t s = 0;
        for (int k = 0; k < P.tnctl; ++k) s |= ((g >> ((P.tctl >> (8 * k)) & 255)) & 1u) << k;
        return cmul_fixed(v, ((const C*)P.tab)[(s << P.tnb) | ((g >> P.tsh) & ((1u << P.tnb) - 1u))]);
    }
    double r = P.grev;
    for (int k = 0; k < P.gn; ++k) {
        const BGroup& G = P.gr[k];
        int s = 0;
        for (int j = 0; j < G.nctl; ++j) s |= (int)((g >> G.ctl[j]) & 1u) << j;
        const double m = (double)((int)(2u * ((g >> G.sh) & ((1u << G.nb) - 1u))) + 1 - (1 << G.nb));
        r += group_rev(G, s, m);
    }
    for (int i = 0; i < P.tn; ++i) r += cvdv_frac(diag_term_rev(P.t[i], g));
    return cmul_fixed(v, unit_rev_fast(cvdv_frac(r), R(0)));
}

__device__ __forceinline__ double2 scale_by(double2 v, double s, float, float) {
    return make_double2(v.x * s, v.y * s);
}
// FP32: the scale is applied as hi + lo, exact product v*hi plus the rounded
// v*lo in one FMA, so a rounded 1/sqrt(N) does not bias the norm.
__device__ __forceinline__ float2 scale_by(float2 v, double, float hi, float lo) {
    return make_float2(fmaf(v.x, hi, v.x * lo), fmaf(v.y, hi, v.y * lo));
}

// A state or staging buffer in LDS is stored with one pad slot after every 8
// elements (index i at i + i/8).  The Stockham stages' first two radix-8
// exchanges write at strides of 8 and 64 elements; padded, those become 9 and
// 72, which fall on distinct 8-byte LDS slots across 32 lanes instead of 16-
// and 8-way bank conflicts.  States in global memory are not padded.
#define BATCH_LDS_PAD 3
__host__ __device__ constexpr uint32_t lds_padded(uint32_t n) { return n + (n >> BATCH_LDS_PAD); }

template <typename C, bool PAD>
struct BState {
    C* p;
    __device__ __forceinline__ C& operator[](uint32_t i) const { return p[PAD ? i + (i >> BATCH_LDS_PAD) : i]; }
};

// Barrier between the phases of an op.  A state in LDS needs only LDS ordering;
// a state in the global work area also needs this thread's stores complete.
template <bool GST>
__device__ __forceinline__ void bsync() {
    if constexpr (GST) __syncthreads();
    else lds_barrier();
}

// One in-place Stockham stage of radix R over a chunk of 2^lgC rows of the
// register field [sh, sh+nb).  Row rho (0 <= rho < 2^(T-nb)) is the flat index
// with the field removed; the chunk is rows rho0 .. rho0 + 2^lgC - 1.
// Butterfly b of a row (sub-length L = 2^lgL done) reads rows b + t N/R and
// writes (b / L) R L + (b mod L) + t L, t < R.  GST: the state is in global
// memory (else in LDS).  This general form serves fields above the low bits
// and short fields; fstage below covers the rest.
template <int R, bool GST, typename C, typename S>
__device__ __forceinline__ void bstage(S st, int sh, int nb, int lgC, uint32_t rho0, int lgL, const C* tw,
                                       bool first, bool last, bool inv, double sc, float schi, float sclo,
                                       SProg P, bool usePre, int tid, int nt) {
    constexpr int LGR = R == 8 ? 3 : (R == 4 ? 2 : 1);
    constexpr int NBF = BATCH_EMAX / R;
    const int lgQ = nb - LGR;
    const uint32_t nbf = 1u << (lgC + lgQ);
    const uint32_t lomask = (1u << sh) - 1u;
    const uint32_t cmask = (1u << lgC) - 1u;
    C x[NBF][R];
#pragma unroll
    for (int k = 0; k < NBF; ++k) {
        const uint32_t q = (uint32_t)tid + (uint32_t)k * nt;
        if (q < nbf) {
            const uint32_t rho = rho0 + (q & cmask);
            const uint32_t b = q >> lgC;
            const uint32_t base = ((rho >> sh) << (sh + nb)) | (rho & lomask);
#pragma unroll
            for (int t = 0; t < R; ++t) {
                const uint32_t g = base | ((b + ((uint32_t)t << lgQ)) << sh);
                C v = st[g];
                if (first && usePre) v = prog_apply(P, g, v);  // the load program, on the first load
                if (first && inv) v.y = -v.y;  // inverse = conj . forward . conj
                x[k][t] = v;
            }
        }
    }
    bsync<GST>();
#pragma unroll
    for (int k = 0; k < NBF; ++k) {
        const uint32_t q = (uint32_t)tid + (uint32_t)k * nt;
        if (q < nbf) {
            const uint32_t rho = rho0 + (q & cmask);
            const uint32_t b = q >> lgC;
            const uint32_t base = ((rho >> sh) << (sh + nb)) | (rho & lomask);
            const uint32_t p = b & ((1u << lgL) - 1u);
            if (lgL > 0) {
                const C* w = tw + p * (R - 1);
#pragma unroll
                for (int t = 1; t < R; ++t) x[k][t] = cmul_fixed(x[k][t], w[t - 1]);
            }
            dft_r<R, false>(x[k]);
            const uint32_t d = ((b >> lgL) << (lgL + LGR)) + p;
#pragma unroll
            for (int t = 0; t < R; ++t) {
                C v = x[k][t];
                if (last) {
                    if (inv) v.y = -v.y;
                    v = scale_by(v, sc, schi, sclo);
                }
                st[base | ((d + ((uint32_t)t << lgL)) << sh)] = v;
            }
        }
    }
    bsync<GST>();
}

// Centred-DFT core along the field: load program, plain FFT (sign by inv),
// output scale.  Rows are processed in chunks of at most BATCH_EMAX * nt
// elements; rows are independent, so chunks need no synchronisation between
// them beyond the per-stage barriers.
// The Stockham stages of one chunk of rows (radix 8, then one radix 4 or 2).
template <bool GST, typename C, typename S>
__device__ __forceinline__ void bstages(S st, int sh, int nb, int lgC, uint32_t rho0, const C* tw, const BOp& op,
                                        float schi, float sclo, SProg P, bool usePre, int tid, int nt) {
    int rem = nb, lgL = 0, off = 0;
    while (rem > 0) {
        const int r = rem >= 3 ? 3 : rem;
        const bool first = lgL == 0, last = rem == r;
        if (r == 3)
            bstage<8, GST>(st, sh, nb, lgC, rho0, lgL, tw + off, first, last, op.inv, op.outScale, schi, sclo, P,
                           usePre, tid, nt);
        else if (r == 2)
            bstage<4, GST>(st, sh, nb, lgC, rho0, lgL, tw + off, first, last, op.inv, op.outScale, schi, sclo, P,
                           usePre, tid, nt);
        else
            bstage<2, GST>(st, sh, nb, lgC, rho0, lgL, tw + off, first, last, op.inv, op.outScale, schi, sclo, P,
                           usePre, tid, nt);
        off += (1 << lgL) * ((1 << r) - 1);
        lgL += r;
        rem -= r;
    }
}

// The Stockham stages with the FFT length 2^NB fixed at compile time, for a
// field in the low bits of an LDS-held state or staging buffer (padded layout:
// row rho of the chunk starts at rho (N + N/8), element j of a row at j + j/8).
// Lanes take consecutive butterflies of a row, so loads are contiguous; each
// butterfly computes one padded base and reaches its R elements at
// compile-time offsets (input stride N/R and output stride L are multiples of
// 8, or L = 1 with R = 8 and an 8-aligned base).  Stage LGL of the plan: radix
// 8 while at least 3 bits remain, then one radix 4 or 2; the twiddles of the
// stages before it (all radix 8) take 2^LGL - 1 table entries.
template <int NB, int LGL, typename C>
__device__ __forceinline__ void fstage(C* s, int lgC, uint32_t rho0, const C* tw, bool inv, double sc, float schi,
                                       float sclo, SProg P, bool usePre, int tid, int nt) {
    constexpr int REM = NB - LGL;
    constexpr int LGR = REM >= 3 ? 3 : REM;
    constexpr int R = 1 << LGR;
    constexpr bool FIRST = LGL == 0, LAST = REM == LGR;
    constexpr int LGQ = NB - LGR, Q = 1 << LGQ, L = 1 << LGL;
    constexpr uint32_t ROWP = (1u << NB) + (1u << NB) / 8;
    constexpr int INS = Q + Q / 8;                  // padded input stride
    constexpr int OUTS = L == 1 ? 1 : L + L / 8;   // padded output stride
    static_assert(Q % 8 == 0 && (L == 1 ? R == 8 : L % 8 == 0), "padding must stay linear");
    constexpr int NBF = BATCH_EMAX / R;
    const uint32_t nbf = 1u << (lgC + LGQ);
    // an opaque copy of the thread id per stage: the stages' addresses would
    // otherwise all be computed up front and held across the whole chain
    int ti = tid;
    asm volatile("" : "+v"(ti));
    // complex64: the stage's twiddles (global, cache resident) are loaded with
    // its inputs, before the barrier, so their latency is not exposed after it
    // (complex128 has no registers left for them)
    constexpr bool TWPRE = LGL > 0 && sizeof(C) == 8;
    C x[NBF][R];
    C w[TWPRE ? NBF : 1][R];
#pragma unroll
    for (int k = 0; k < NBF; ++k) {
        const uint32_t q = (uint32_t)ti + (uint32_t)k * nt;
        if (q < nbf) {
            const uint32_t rho = q >> LGQ, b = q & (Q - 1);
            if constexpr (TWPRE) {
                const C* wp = tw + ((1 << LGL) - 1) + (b & (L - 1)) * (R - 1);
#pragma unroll
                for (int t = 1; t < R; ++t) w[k][t] = wp[t - 1];
            }
            const C* src = s + rho * ROWP + b + (b >> 3);
#pragma unroll
            for (int t = 0; t < R; ++t) {
                C v = src[t * INS];
                if constexpr (FIRST) {
                    if (usePre) v = prog_apply(P, ((rho0 + rho) << NB) | (b + (uint32_t)t * Q), v);
                    if (inv) v.y = -v.y;  // inverse = conj . forward . conj
                }
                x[k][t] = v;
            }
        }
    }
    lds_barrier();
#pragma unroll
    for (int k = 0; k < NBF; ++k) {
        const uint32_t q = (uint32_t)ti + (uint32_t)k * nt;
        if (q < nbf) {
            const uint32_t rho = q >> LGQ, b = q & (Q - 1);
            const uint32_t p = b & (L - 1);
            if constexpr (TWPRE) {
#pragma unroll
                for (int t = 1; t < R; ++t) x[k][t] = cmul_fixed(x[k][t], w[k][t]);
            } else if constexpr (LGL > 0) {
                const C* wp = tw + ((1 << LGL) - 1) + p * (R - 1);
#pragma unroll
                for (int t = 1; t < R; ++t) x[k][t] = cmul_fixed(x[k][t], wp[t - 1]);
            }
            dft_r<R, false>(x[k]);
            const uint32_t d = ((b >> LGL) << (LGL + LGR)) + p;
            C* dst = s + rho * ROWP + d + (d >> 3);
#pragma unroll
            for (int t = 0; t < R; ++t) {
                C v = x[k][t];
                if constexpr (LAST) {
                    if (inv) v.y = -v.y;
                    v = scale_by(v, sc, schi, sclo);
                }
                dst[t * OUTS] = v;
            }
        }
    }
    lds_barrier();
    if constexpr (!LAST) fstage<NB, LGL + LGR, C>(s, lgC, rho0, tw, inv, sc, schi, sclo, P, usePre, tid, nt);
}

// Runs the fixed-length stages for 2^6 .. 2^12 points; false for other lengths.
template <typename C>
__device__ __forceinline__ bool fstages(int nb, C* s, int lgC, uint32_t rho0, const C* tw, const BOp& op,
                                        float schi, float sclo, SProg P, bool usePre, int tid, int nt) {
    switch (nb) {
#define BATCH_FSTAGES_CASE(NB) \
    case NB: fstage<NB, 0, C>(s, lgC, rho0, tw, op.inv, op.outScale, schi, sclo, P, usePre, tid, nt); return true;
        BATCH_FSTAGES_CASE(6)
        BATCH_FSTAGES_CASE(7)
        BATCH_FSTAGES_CASE(8)
        BATCH_FSTAGES_CASE(9)
        BATCH_FSTAGES_CASE(10)
        BATCH_FSTAGES_CASE(11)
        BATCH_FSTAGES_CASE(12)
#undef BATCH_FSTAGES_CASE
        default: return false;
    }
}

// A state in the global work area (L2-resident) takes each chunk of rows
// through the workgroup's LDS when it has room (stage): copied in with the
// load program applied at each element's flat index, transformed there, copied
// back; one L2 round trip per chunk instead of one per Stockham stage.  A chunk
// holds at most BATCH_EMAX elements per thread (the host sizes the launch so),
// and each thread issues all its loads (stores) of a chunk before using them.
// Chunks cover disjoint rows, so a chunk's stores may still be in flight while
// the next chunk loads; the op's closing barrier completes them.
template <bool GST, typename C, typename S>
__device__ __forceinline__ void bfft(S st, int T, const BOp& op, const C* tw, SProg P, int tid,
                                     int nt, int lgEN, BState<C, true> stage, uint32_t stageCap) {
    const int nb = op.nb, sh = op.sh;
    const int lgRows = T - nb;
    const int lgC = min(lgRows, lgEN - nb);
    const uint32_t nChunks = 1u << (lgRows - lgC);
    const bool usePre = P.active();
    const float schi = (float)op.outScale;
    const float sclo = (float)(op.outScale - (double)schi);
    const uint32_t chunk = 1u << (nb + lgC);
    if (stage.p && lds_padded(chunk) <= stageCap) {
        const uint32_t lomask = (1u << sh) - 1u, jmask = (1u << nb) - 1u;
        auto flat = [&](uint32_t rho0, uint32_t e) {
            const uint32_t rho = rho0 + (e >> nb);
            return ((rho >> sh) << (sh + nb)) | (rho & lomask) | ((e & jmask) << sh);
        };
        for (uint32_t ch = 0; ch < nChunks; ++ch) {
            const uint32_t rho0 = ch << lgC;
            C v[BATCH_EMAX];
#pragma unroll
            for (int k = 0; k < BATCH_EMAX; ++k) {
                const uint32_t e = (uint32_t)tid + (uint32_t)k * nt;
                if (e < chunk) v[k] = st[flat(rho0, e)];
            }
#pragma unroll
            for (int k = 0; k < BATCH_EMAX; ++k) {
                const uint32_t e = (uint32_t)tid + (uint32_t)k * nt;
                if (e < chunk) stage[e] = usePre ? prog_apply(P, flat(rho0, e), v[k]) : v[k];
            }
            lds_barrier();
            if (!fstages<C>(nb, stage.p, lgC, 0u, tw, op, schi, sclo, P, false, tid, nt))
                bstages<false, C>(stage, 0, nb, lgC, 0u, tw, op, schi, sclo, P, false, tid, nt);
#pragma unroll
            for (int k = 0; k < BATCH_EMAX; ++k) {
                const uint32_t e = (uint32_t)tid + (uint32_t)k * nt;
                if (e < chunk) v[k] = stage[e];
            }
#pragma unroll
            for (int k = 0; k < BATCH_EMAX; ++k) {
                const uint32_t e = (uint32_t)tid + (uint32_t)k * nt;
                if (e < chunk) st[flat(rho0, e)] = v[k];
            }
            lds_barrier();  // the next chunk overwrites the staging buffer
        }
        return;
    }
    // the load program is applied by the first stage as it loads each element
    for (uint32_t ch = 0; ch < nChunks; ++ch) {
        const uint32_t rho0 = ch << lgC;
        if constexpr (!GST) {
            const uint32_t rowp = lds_padded(1u << nb);
            if (sh == 0 && fstages<C>(nb, st.p + rho0 * rowp, lgC, rho0, tw, op, schi, sclo, P, usePre, tid, nt))
                continue;
        }
        bstages<GST, C>(st, sh, nb, lgC, rho0, tw, op, schi, sclo, P, usePre, tid, nt);
    }
}

// Qubit gate and conditional parity: the same pair maps and FP64 matrix
// arithmetic as k_qubit / k_cparity (gates.hip); each pair belongs to one thread.
template <typename C, typename S>
__device__ __forceinline__ void bqubit(S st, uint32_t size, const BOp& op, SProg P, int tid, int nt) {
    const uint32_t lowmask = (1u << op.bit) - 1u;
    const bool usePre = P.active();
    for (uint32_t p = tid; p < size / 2; p += nt) {
        const uint32_t i0 = ((p & ~lowmask) << 1) | (p & lowmask);
        const uint32_t i1 = i0 | (1u << op.bit);
        C a = st[i0], b = st[i1];
        if (usePre) {
            a = prog_apply(P, i0, a);
            b = prog_apply(P, i1, b);
        }
        const double2 ad = make_double2((double)a.x, (double)a.y);
        const double2 bd = make_double2((double)b.x, (double)b.y);
        const double2 t00 = cmul(op.m[0], ad), t01 = cmul(op.m[1], bd);
        const double2 t10 = cmul(op.m[2], ad), t11 = cmul(op.m[3], bd);
        st[i0] = cmk<C>(t00.x + t01.x, t00.y + t01.y);
        st[i1] = cmk<C>(t10.x + t11.x, t10.y + t11.y);
    }
}

template <typename C, typename S>
__device__ __forceinline__ void bcparity(S st, uint32_t size, const BOp& op, SProg P, int tid, int nt) {
    const int top = op.sh + op.nb - 1;
    const uint32_t lowmask = (1u << top) - 1u;
    const uint32_t fmask = ((1u << op.nb) - 1u) << op.sh;
    const bool usePre = P.active();
    for (uint32_t p = tid; p < size / 2; p += nt) {
        const uint32_t i0 = ((p & ~lowmask) << 1) | (p & lowmask);
        const uint32_t i1 = i0 ^ fmask;
        C a = st[i0], b = st[i1];
        if (usePre) {
            a = prog_apply(P, i0, a);
            b = prog_apply(P, i1, b);
        }
        const uint32_t lower = op.tgtRefl ? i1 : i0;
        const int cb = (int)((lower >> op.bit) & 1u) ^ op.ctrlFlip;
        if (cb) {
            st[i0] = b;
            st[i1] = a;
        } else if (usePre) {
            st[i0] = a;
            st[i1] = b;
        }
    }
}

template <typename C, typename S>
__device__ __forceinline__ void bdiag(S st, uint32_t size, SProg P, int tid, int nt) {
    for (uint32_t g = tid; g < size; g += nt) st[g] = prog_apply(P, g, st[g]);
}

// Stage a program's generic terms and groups in LDS (one entry per thread).
__device__ __forceinline__ SProg load_prog(DiagTerm* sT, BGroup* sG, const BatchArgs& a, int tfirst, int tn,
                                           int gfirst, int gn, double grev, int tid) {
    if (tid < tn) sT[tid] = a.terms[tfirst + tid];
    if (tid >= CVDV_MAX_TERMS && tid < CVDV_MAX_TERMS + gn) sG[tid - CVDV_MAX_TERMS] = a.groups[gfirst + tid - CVDV_MAX_TERMS];
    SProg P{};
    P.t = sT;
    P.gr = sG;
    P.tn = tn;
    P.gn = gn;
    P.grev = grev;
    return P;
}

// A program of one polynomial group and nothing else, over a state of 2^T
// amplitudes: its factor depends on the group's field index and control bits
// only, so its 2^(nb + nctl) distinct factors are tabulated in the free LDS
// (tab, cap entries) when that at least halves the evaluations.  The entries
// are computed exactly as prog_apply evaluates them.  Call after the program
// is staged (and its barrier); it ends with a barrier when it builds.
template <typename C>
__device__ __forceinline__ void tabulate(SProg& P, int T, C* tab, uint32_t cap, int tid, int nt) {
    using R = decltype(C{}.x);
    if (P.gn != 1 || P.tn != 0 || !tab) return;
    const BGroup& G = P.gr[0];
    const int nb = G.nb, nctl = G.nctl;
    const uint32_t n = 1u << (nb + nctl);
    if (nb + nctl + 1 > T || n > cap) return;
    for (uint32_t e = tid; e < n; e += nt) {
        const uint32_t sg = e >> nb, j = e & ((1u << nb) - 1u);
        const double m = (double)((int)(2u * j) + 1 - (1 << nb));
        double r = P.grev;
        r += group_rev(G, (int)sg, m);
        tab[e] = unit_rev_fast(cvdv_frac(r), R(0));
    }
    P.tab = tab;
    P.tsh = G.sh;
    P.tnb = nb;
    P.tnctl = nctl;
    P.tctl = 0;
    for (int k = 0; k < nctl; ++k) P.tctl |= G.ctl[k] << (8 * k);
    lds_barrier();
}

// tab, tabCap: LDS the circuit leaves free (after its state or staging buffer),
// where an op's program may be tabulated.
template <bool GST, typename C, typename S>
__device__ __forceinline__ void run_circuit(S st, const BCirc& cd, const BatchArgs& a, DiagTerm* sT, BGroup* sG,
                                            int tid, int nt, int lgEN, BState<C, true> stage, C* tab,
                                            uint32_t tabCap) {
    const uint32_t size = 1u << cd.T;
    // 1. product state in the canonical layout (k_init_product order: register 0 first)
    for (uint32_t g = tid; g < size; g += nt) {
        uint32_t o = cd.initOff;  // register r's array follows register r-1's
        double2 p = a.init[o + ((g >> cd.flw[0]) & ((1u << cd.nq[0]) - 1u))];
        for (int r = 1; r < cd.nreg; ++r) {
            o += 1u << cd.nq[r - 1];
            p = cmul(p, a.init[o + ((g >> cd.flw[r]) & ((1u << cd.nq[r]) - 1u))]);
        }
        st[g] = cmk<C>(p.x, p.y);
    }
    bsync<GST>();
    // 2. ops
    for (uint32_t i = cd.opBegin; i < cd.opEnd; ++i) {
        const BOp& op = a.ops[i];
        SProg P = load_prog(sT, sG, a, op.tfirst, op.tn, op.gfirst, op.gn, op.grev, tid);
        if (op.tn > 0 || op.gn > 0) lds_barrier();
        tabulate<C>(P, cd.T, tab, tabCap, tid, nt);
        switch (op.kind) {
            case BOP_FFT:
                bfft<GST, C>(st, cd.T, op, (const C*)a.tw + a.twOff[op.nb], P, tid, nt, lgEN, stage, a.ldsCap);
                break;
            case BOP_QUBIT: bqubit<C>(st, size, op, P, tid, nt); break;
            case BOP_CPARITY: bcparity<C>(st, size, op, P, tid, nt); break;
            default: bdiag<C>(st, size, P, tid, nt); break;
        }
        bsync<GST>();
    }
    // 3. canonical store with the trailing program
    const SProg F = load_prog(sT, sG, a, cd.ftFirst, cd.ftN, cd.fgFirst, cd.fgN, cd.fgrev, tid);
    if (cd.ftN > 0 || cd.fgN > 0) lds_barrier();
    const bool useFinal = F.active();
    C* out = (C*)a.out + cd.outOff;
    for (uint32_t g = tid; g < size; g += nt) {
        C v = st[g];
        if (useFinal) v = prog_apply(F, g, v);
        uint32_t o = 0;
        for (int r = 0; r < cd.nreg; ++r) {
            const uint32_t mask = (1u << cd.nq[r]) - 1u;
            uint32_t j = (g >> cd.psh[r]) & mask;
            if (cd.refl[r]) j = mask - j;
            o |= j << cd.flw[r];
        }
        out[o] = v;
    }
}

template <typename C>
__global__ __launch_bounds__(BATCH_MAX_THREADS) void k_batch(BatchArgs a, int lgEN) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ DiagTerm sT[CVDV_MAX_TERMS];
    __shared__ BGroup sG[BATCH_MAX_GROUPS];
    const BCirc& cd = a.circ[a.order[blockIdx.x]];
    const int tid = threadIdx.x, nt = blockDim.x;
    // the LDS after the circuit's state (or its largest staged chunk) holds
    // tabulated programs
    const uint32_t used = lds_padded(cd.lds ? (1u << cd.T) : (1u << lgEN));
    C* tab = used < a.ldsCap ? (C*)smem + used : nullptr;
    const uint32_t tabCap = used < a.ldsCap ? a.ldsCap - used : 0u;
    if (cd.lds)
        run_circuit<false, C>(BState<C, true>{(C*)smem}, cd, a, sT, sG, tid, nt, lgEN, BState<C, true>{nullptr},
                              tab, tabCap);
    else
        run_circuit<true, C>(BState<C, false>{(C*)a.work + cd.workOff}, cd, a, sT, sG, tid, nt, lgEN,
                             BState<C, true>{(C*)smem}, tab, tabCap);
}

// ---------------------------------------------------------------- readout kernel
struct KS {
    double s, c;
};
__device__ __forceinline__ void ks_add(KS& a, double v) {  // Neumaier
    const double t = a.s + v;
    if (fabs(a.s) >= fabs(v)) a.c += (a.s - t) + v;
    else a.c += (v - t) + a.s;
    a.s = t;
}

// res[binOff[c] + j] = sum over the elements of circuit c whose register `reg`
// (canonical layout) has value j of |psi|^2; reg < 0: the norm.  Fixed order.
template <typename C>
__global__ __launch_bounds__(256) void k_batch_marginal(const BCirc* circ, const void* outv, int reg,
                                                        const uint64_t* binOff, double* res) {
    __shared__ KS red[256];
    const BCirc& cd = circ[blockIdx.x];
    const C* st = (const C*)outv + cd.outOff;
    const int tid = threadIdx.x;
    const int nbr = reg < 0 ? 0 : cd.nq[reg];
    const int f = reg < 0 ? 0 : cd.flw[reg];
    const uint32_t B = 1u << nbr;
    const uint32_t E = 1u << (cd.T - nbr);
    const uint32_t fmask = (1u << f) - 1u;
    double* o = res + binOff[blockIdx.x];
    if (B >= 256) {
        for (uint32_t j = tid; j < B; j += 256) {
            KS k{0.0, 0.0};
            for (uint32_t e = 0; e < E; ++e) {
                const C v = st[((e >> f) << (f + nbr)) | (j << f) | (e & fmask)];
                ks_add(k, (double)v.x * v.x + (double)v.y * v.y);
            }
            o[j] = k.s + k.c;
        }
        return;
    }
    const uint32_t tpb = 256u / B;
    const uint32_t j = tid / tpb, l = tid % tpb;
    KS k{0.0, 0.0};
    for (uint32_t e = l; e < E; e += tpb) {
        const C v = st[((e >> f) << (f + nbr)) | (j << f) | (e & fmask)];
        ks_add(k, (double)v.x * v.x + (double)v.y * v.y);
    }
    red[tid] = k;
    __syncthreads();
    for (uint32_t w = tpb / 2; w >= 1; w /= 2) {
        if (l < w) {
            KS x = red[tid], y = red[tid + w];
            ks_add(x, y.s);
            x.c += y.c;
            red[tid] = x;
        }
        __syncthreads();
    }
    if (l == 0) o[j] = red[tid].s + red[tid].c;
}

// ---------------------------------------------------------------- host side
struct CVDVBatch {
    int prec = CVDV_C128;
    int n = 0;
    std::vector<CVDVContext*> circ;
    hipStream_t stream = nullptr;
    int threads = 256;             // launch shape (cvdvBatchSetLaunch / autotune)
    int ldsBytes = 32768;
    // device buffers
    void* dOut = nullptr;
    size_t outElems = 0;
    std::vector<uint64_t> outOff;  // per circuit, elements
    void* dWork = nullptr;
    size_t workElems = 0;
    double2* dInit = nullptr;
    size_t initElems = 0;
    bool initSet = false;
    BCirc* dCirc = nullptr;
    int* dOrder = nullptr;
    BOp* dOps = nullptr;
    size_t opsCap = 0;
    DiagTerm* dTerms = nullptr;
    size_t termsCap = 0;
    BGroup* dGroups = nullptr;
    size_t groupsCap = 0;
    void* dTw = nullptr;
    int twOff[BATCH_MAX_REG_QUBITS + 1] = {0};
    double* dRes = nullptr;
    size_t resCap = 0;
    uint64_t* dBinOff = nullptr;
    // compile state
    bool compiled = false;
    int compiledThreads = 0, compiledLds = 0;
    std::vector<size_t> seenOps;
    std::vector<DiagProg> seenPending;
    std::vector<std::vector<int>> seenLayout;  // pshift and refl of every register
    size_t totalOps = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    size_t elemBytes() const { return prec == CVDV_C128 ? 16 : 8; }
};

static void bfree(void*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

template <typename T>
static void bgrow(T*& p, size_t& cap, size_t need) {
    if (need <= cap && p) return;
    if (p) (void)hipFree(p);
    p = nullptr;
    HIPCHK(hipMalloc((void**)&p, std::max<size_t>(need, 1) * sizeof(T)));
    cap = need;
}

static void drop_graph(CVDVBatch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

static int ilog2(uint64_t v) {
    int r = 0;
    while ((1ull << r) < v) ++r;
    return r;
}

// The lazy layout of a circuit: parity and register swaps change only this,
// and the compiled circuit carries it (the canonical store undoes it).
static std::vector<int> layout_of(const CVDVContext* c) {
    std::vector<int> v(2 * c->numReg);
    for (int r = 0; r < c->numReg; ++r) {
        v[2 * r] = c->pshift[r];
        v[2 * r + 1] = c->refl[r];
    }
    return v;
}

static bool recording_changed(CVDVBatch* b) {
    if (!b->compiled || b->compiledThreads != b->threads || b->compiledLds != b->ldsBytes) return true;
    for (int i = 0; i < b->n; ++i) {
        const CVDVContext* c = b->circ[i];
        if (c->rec->ops.size() != b->seenOps[i]) return true;
        if (std::memcmp(&c->pending, &b->seenPending[i], sizeof(DiagProg)) != 0) return true;
        if (layout_of(c) != b->seenLayout[i]) return true;
    }
    return false;
}

// Flatten every recording into device arrays and capture the launch.
static void batch_compile(CVDVBatch* b) {
    drop_graph(b);
    const size_t eb = b->elemBytes();
    // the launch shape must hold one FFT row per chunk
    int maxNb = 0;
    for (auto* c : b->circ)
        for (const BatchOp& op : c->rec->ops)
            if (op.kind == BOP_FFT) maxNb = std::max(maxNb, op.nb);
    while ((BATCH_EMAX << ilog2(b->threads)) < (1 << maxNb) && b->threads < BATCH_MAX_THREADS) b->threads *= 2;
    const int lgEN = ilog2(BATCH_EMAX) + ilog2(b->threads);
    const size_t ldsElems = std::min<size_t>((size_t)b->ldsBytes, device_max_lds() - 1024) / eb;

    std::vector<BCirc> hc(b->n);
    std::vector<BOp> hops;
    std::vector<DiagTerm> hterms;
    std::vector<double> cost(b->n);
    size_t workElems = 0;
    uint32_t initOff = 0;
    std::vector<BGroup> hgroups;
    // A program's polynomial terms merge into at most BATCH_MAX_GROUPS groups (one
    // per register field, up to BATCH_GROUP_CTL control bits each); the rest stay
    // generic terms.
    auto add_terms = [&](const DiagProg& p, int& first, int& n, int& gfirst, int& gn) {
        first = (int)hterms.size();
        gfirst = (int)hgroups.size();
        n = 0;
        gn = 0;
        std::vector<BGroup> gs;
        std::vector<std::vector<const DiagTerm*>> mem;
        for (int t = 0; t < p.n; ++t) {
            const DiagTerm& T = p.t[t];
            if (T.c1 == 0.0 && T.c2 == 0.0 && T.c3 == 0.0) continue;
            int k = -1;
            if (T.kind == DT_POLY) {
                for (int j = 0; j < (int)gs.size(); ++j)
                    if (gs[j].sh == T.sh1 && gs[j].nb == T.nb1) k = j;
                if (k < 0 && (int)gs.size() < BATCH_MAX_GROUPS) {
                    BGroup G{};
                    G.sh = T.sh1;
                    G.nb = T.nb1;
                    gs.push_back(G);
                    mem.emplace_back();
                    k = (int)gs.size() - 1;
                }
                if (k >= 0 && T.ctrl >= 0) {
                    bool have = false;
                    for (int j = 0; j < gs[k].nctl; ++j) have |= gs[k].ctl[j] == T.ctrl;
                    if (!have) {
                        if (gs[k].nctl == BATCH_GROUP_CTL) k = -1;
                        else gs[k].ctl[gs[k].nctl++] = T.ctrl;
                    }
                }
            }
            if (k >= 0) {
                mem[k].push_back(&T);
            } else {
                hterms.push_back(T);
                ++n;
            }
        }
        for (size_t k = 0; k < gs.size(); ++k) {
            BGroup& G = gs[k];
            for (int sgn = 0; sgn < (1 << BATCH_GROUP_CTL); ++sgn) {
                double c[3] = {0.0, 0.0, 0.0};
                for (const DiagTerm* T : mem[k]) {
                    double f = 1.0;
                    for (int j = 0; j < G.nctl; ++j)
                        if (T->ctrl == G.ctl[j] && ((sgn >> j) & 1)) f = -1.0;
                    c[0] += f * T->c1;
                    c[1] += f * T->c2;
                    c[2] += f * T->c3;
                }
                for (int q = 0; q < 3; ++q) G.c[sgn][q] = c[q];
            }
            hgroups.push_back(G);
            ++gn;
        }
    };
    for (int i = 0; i < b->n; ++i) {
        const CVDVContext* c = b->circ[i];
        BCirc& d = hc[i];
        std::memset(&d, 0, sizeof(d));
        d.outOff = b->outOff[i];
        d.T = c->totalQbt;
        d.nreg = c->numReg;
        d.initOff = initOff;
        for (int r = 0; r < c->numReg; ++r) {
            d.nq[r] = (uint8_t)c->nq[r];
            d.flw[r] = (uint8_t)c->flw[r];
            d.psh[r] = (uint8_t)c->pshift[r];
            d.refl[r] = (uint8_t)c->refl[r];
            initOff += 1u << c->nq[r];
        }
        const size_t size = c->totalSize;
        d.lds = lds_padded((uint32_t)size) <= ldsElems ? 1 : 0;  // padded LDS layout
        if (!d.lds) {
            d.workOff = workElems;
            workElems += size;
        }
        d.opBegin = (uint32_t)hops.size();
        double cst = 2.0;
        for (const BatchOp& op : c->rec->ops) {
            BOp o{};
            o.kind = op.kind;
            o.sh = op.sh;
            o.nb = op.nb;
            o.bit = op.bit;
            o.ctrlFlip = op.ctrlFlip;
            o.tgtRefl = op.tgtRefl;
            o.inv = op.inv;
            o.outScale = op.outScale;
            for (int k = 0; k < 4; ++k) o.m[k] = op.m[k];
            if (op.prog >= 0) {
                const DiagProg& p = c->rec->progs[op.prog];
                add_terms(p, o.tfirst, o.tn, o.gfirst, o.gn);
                o.grev = p.grev;
            }
            hops.push_back(o);
            cst += op.kind == BOP_FFT ? 1.0 + (op.nb + 2) / 3 : 1.0;
        }
        d.opEnd = (uint32_t)hops.size();
        add_terms(c->pending, d.ftFirst, d.ftN, d.fgFirst, d.fgN);
        d.fgrev = c->pending.grev;
        // global-mode circuits pay cache traffic instead of LDS traffic
        cost[i] = cst * (double)size * (d.lds ? 1.0 : 2.0);
    }
    std::vector<int> order(b->n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost[x] > cost[y]; });

    // twiddles: stage tables for every FFT length present
    std::vector<uint8_t> tw;
    for (int nb = 1; nb <= BATCH_MAX_REG_QUBITS; ++nb) {
        b->twOff[nb] = (int)(tw.size() / eb);
        if (nb > maxNb) continue;
        const std::vector<uint8_t> t = fft_stage_host(nb, b->prec);
        tw.insert(tw.end(), t.begin(), t.end());
    }
    bfree(b->dTw);
    HIPCHK(hipMalloc(&b->dTw, std::max<size_t>(tw.size(), 16)));
    if (!tw.empty()) HIPCHK(hipMemcpy(b->dTw, tw.data(), tw.size(), hipMemcpyHostToDevice));

    bfree((void*&)b->dCirc);
    HIPCHK(hipMalloc((void**)&b->dCirc, hc.size() * sizeof(BCirc)));
    HIPCHK(hipMemcpy(b->dCirc, hc.data(), hc.size() * sizeof(BCirc), hipMemcpyHostToDevice));
    bfree((void*&)b->dOrder);
    HIPCHK(hipMalloc((void**)&b->dOrder, order.size() * sizeof(int)));
    HIPCHK(hipMemcpy(b->dOrder, order.data(), order.size() * sizeof(int), hipMemcpyHostToDevice));
    bgrow(b->dOps, b->opsCap, hops.size());
    if (!hops.empty()) HIPCHK(hipMemcpy(b->dOps, hops.data(), hops.size() * sizeof(BOp), hipMemcpyHostToDevice));
    bgrow(b->dTerms, b->termsCap, hterms.size());
    if (!hterms.empty())
        HIPCHK(hipMemcpy(b->dTerms, hterms.data(), hterms.size() * sizeof(DiagTerm), hipMemcpyHostToDevice));
    bgrow(b->dGroups, b->groupsCap, hgroups.size());
    if (!hgroups.empty())
        HIPCHK(hipMemcpy(b->dGroups, hgroups.data(), hgroups.size() * sizeof(BGroup), hipMemcpyHostToDevice));
    if (workElems > b->workElems || !b->dWork) {
        bfree(b->dWork);
        HIPCHK(hipMalloc(&b->dWork, std::max<size_t>(workElems, 1) * eb));
        b->workElems = workElems;
    }
    b->totalOps = hops.size();

    // capture the launch once; cvdvBatchRun replays it
    BatchArgs a{};
    a.circ = b->dCirc;
    a.order = b->dOrder;
    a.ops = b->dOps;
    a.terms = b->dTerms;
    a.groups = b->dGroups;
    a.init = b->dInit;
    a.tw = b->dTw;
    a.out = b->dOut;
    a.work = b->dWork;
    for (int i = 0; i <= BATCH_MAX_REG_QUBITS; ++i) a.twOff[i] = b->twOff[i];
    bool anyLds = false, anyWork = false;
    for (const BCirc& d : hc) {
        anyLds |= d.lds != 0;
        anyWork |= d.lds == 0;
    }
    // circuits in the work area stage their FFT chunks (at most EMAX * threads
    // elements, padded) in the same dynamic LDS
    const size_t stageElems = lds_padded(1u << lgEN);
    const size_t shm = std::max(anyLds ? ldsElems * eb : 0, anyWork ? stageElems * eb : 0);
    a.ldsCap = (uint32_t)(shm / eb);
    if (b->prec == CVDV_C128)
        HIPCHK(hipFuncSetAttribute((const void*)k_batch<double2>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)std::max<size_t>(shm, 1)));
    else
        HIPCHK(hipFuncSetAttribute((const void*)k_batch<float2>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)std::max<size_t>(shm, 1)));
    HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    if (b->prec == CVDV_C128)
        hipLaunchKernelGGL(k_batch<double2>, dim3(b->n), dim3(b->threads), shm, b->stream, a, lgEN);
    else
        hipLaunchKernelGGL(k_batch<float2>, dim3(b->n), dim3(b->threads), shm, b->stream, a, lgEN);
    const hipError_t le = hipGetLastError();
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(b->stream, &g);
    HIPCHK(le);
    HIPCHK(ce);
    b->graph = g;
    HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));

    b->seenOps.resize(b->n);
    b->seenPending.resize(b->n);
    b->seenLayout.resize(b->n);
    for (int i = 0; i < b->n; ++i) {
        b->seenOps[i] = b->circ[i]->rec->ops.size();
        b->seenPending[i] = b->circ[i]->pending;
        b->seenLayout[i] = layout_of(b->circ[i]);
    }
    b->compiled = true;
    b->compiledThreads = b->threads;
    b->compiledLds = b->ldsBytes;
}

static void batch_release(CVDVBatch* b) {
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    bfree(b->dOut);
    bfree(b->dWork);
    bfree((void*&)b->dInit);
    bfree((void*&)b->dCirc);
    bfree((void*&)b->dOrder);
    bfree((void*&)b->dOps);
    bfree((void*&)b->dTerms);
    bfree((void*&)b->dGroups);
    bfree(b->dTw);
    bfree((void*&)b->dRes);
    bfree((void*&)b->dBinOff);
    for (auto* c : b->circ) {
        delete c->rec;
        delete c;
    }
    b->circ.clear();
    if (b->stream) (void)hipStreamDestroy(b->stream);
    b->stream = nullptr;
}

#define BATCH_TRY try {
#define BATCH_CATCH(ret)                                                     \
    }                                                                        \
    catch (const CvdvError& e) {                                             \
        cvdv_set_error(e.code, e.msg);                                       \
        return ret;                                                          \
    }                                                                        \
    catch (const std::bad_alloc&) {                                          \
        cvdv_set_error(CVDV_ERR_MEMORY, "host allocation failed");           \
        return ret;                                                          \
    }

static bool batch_ok(CVDVBatch* b) {
    if (!b || b->n == 0) {
        cvdv_set_error(CVDV_ERR_STATE, "invalid batch handle");
        return false;
    }
    return true;
}

static void batch_run(CVDVBatch* b) {
    if (!b->initSet) throw CvdvError{CVDV_ERR_STATE, "cvdvBatchRun: call cvdvBatchInitFromSeparable first"};
    if (recording_changed(b)) batch_compile(b);
    HIPCHK(hipGraphLaunch(b->exec, b->stream));
}

// per-circuit marginal of `reg` (canonical), results concatenated on the host
static void batch_marginal(CVDVBatch* b, int reg, double* out) {
    std::vector<uint64_t> off(b->n + 1, 0);
    for (int i = 0; i < b->n; ++i) {
        const CVDVContext* c = b->circ[i];
        if (reg >= c->numReg)
            throw CvdvError{CVDV_ERR_ARG, "cvdvBatchMeasure: circuit " + std::to_string(i) + " has no register " +
                                              std::to_string(reg)};
        off[i + 1] = off[i] + (reg < 0 ? 1 : (1ull << c->nq[reg]));
    }
    if (!b->compiled) throw CvdvError{CVDV_ERR_STATE, "batch has not been run"};
    bgrow(b->dRes, b->resCap, off[b->n]);
    bfree((void*&)b->dBinOff);
    HIPCHK(hipMalloc((void**)&b->dBinOff, off.size() * sizeof(uint64_t)));
    HIPCHK(hipMemcpyAsync(b->dBinOff, off.data(), off.size() * sizeof(uint64_t), hipMemcpyHostToDevice, b->stream));
    if (b->prec == CVDV_C128)
        hipLaunchKernelGGL(k_batch_marginal<double2>, dim3(b->n), dim3(256), 0, b->stream, b->dCirc,
                           (const void*)b->dOut, reg, b->dBinOff, b->dRes);
    else
        hipLaunchKernelGGL(k_batch_marginal<float2>, dim3(b->n), dim3(256), 0, b->stream, b->dCirc,
                           (const void*)b->dOut, reg, b->dBinOff, b->dRes);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, b->dRes, off[b->n] * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
}

extern "C" {

CVDVBatch* cvdvBatchCreate(int numCircuits, const int* numRegs, const int* numQubitsFlat, int precision) {
    if (numCircuits <= 0 || !numRegs || !numQubitsFlat) {
        cvdv_set_error(CVDV_ERR_ARG, "cvdvBatchCreate: need at least one circuit");
        return nullptr;
    }
    if (precision != CVDV_C128 && precision != CVDV_MIXED) {
        cvdv_set_error(CVDV_ERR_ARG, "cvdvBatchCreate: precision must be 0 (complex128) or 1 (mixed)");
        return nullptr;
    }
    CVDVBatch* b = new CVDVBatch();
    BATCH_TRY
    b->prec = precision;
    size_t k = 0, out = 0;
    for (int i = 0; i < numCircuits; ++i) {
        const int nr = numRegs[i];
        if (nr < 1 || nr > BATCH_MAX_REG)
            throw CvdvError{CVDV_ERR_ARG, "cvdvBatchCreate: circuit " + std::to_string(i) +
                                              ": 1 to 8 registers per circuit"};
        int tot = 0;
        for (int r = 0; r < nr; ++r) {
            const int q = numQubitsFlat[k + r];
            if (q < 0 || q > BATCH_MAX_REG_QUBITS)
                throw CvdvError{CVDV_ERR_ARG, "cvdvBatchCreate: circuit " + std::to_string(i) +
                                                  ": register qubit count out of range [0, 12]"};
            tot += q;
        }
        if (tot > BATCH_MAX_QUBITS)
            throw CvdvError{CVDV_ERR_ARG, "cvdvBatchCreate: circuit " + std::to_string(i) +
                                              ": more than 20 qubits in total"};
        CVDVContext* c = ctx_create_recording(nr, numQubitsFlat + k, precision);
        if (!c) throw CvdvError{CVDV_ERR_ARG, "cvdvBatchCreate: invalid circuit " + std::to_string(i)};
        b->circ.push_back(c);
        b->outOff.push_back(out);
        out += c->totalSize;
        for (int r = 0; r < nr; ++r) b->initElems += (size_t)1 << numQubitsFlat[k + r];
        k += nr;
    }
    b->n = numCircuits;
    b->outElems = out;
    HIPCHK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    HIPCHK(hipMalloc(&b->dOut, out * b->elemBytes()));
    HIPCHK(hipMalloc((void**)&b->dInit, b->initElems * sizeof(double2)));
    return b;
    }
    catch (const CvdvError& e) {
        cvdv_set_error(e.code, e.msg);
    }
    catch (...) {
        cvdv_set_error(CVDV_ERR_MEMORY, "cvdvBatchCreate: allocation failed");
    }
    b->n = (int)b->circ.size();
    batch_release(b);
    delete b;
    return nullptr;
}

void cvdvBatchDestroy(CVDVBatch* b) {
    if (!b) return;
    batch_release(b);
    delete b;
}

int cvdvBatchNumCircuits(CVDVBatch* b) { return b ? b->n : 0; }

CVDVContext* cvdvBatchCircuit(CVDVBatch* b, int i) {
    if (!batch_ok(b)) return nullptr;
    if (i < 0 || i >= b->n) {
        cvdv_set_error(CVDV_ERR_ARG, "cvdvBatchCircuit: index out of range: " + std::to_string(i));
        return nullptr;
    }
    return b->circ[i];
}

size_t cvdvBatchStateSize(CVDVBatch* b, int i) {
    if (!batch_ok(b) || i < 0 || i >= b->n) return 0;
    return b->circ[i]->totalSize;
}

size_t cvdvBatchNumOps(CVDVBatch* b) {
    if (!b) return 0;
    size_t s = 0;
    for (auto* c : b->circ) s += c->rec->ops.size();
    return s;
}

void cvdvBatchInitFromSeparable(CVDVBatch* b, const double* hostC128) {
    if (!batch_ok(b) || !hostC128) return;
    BATCH_TRY
    HIPCHK(hipMemcpy(b->dInit, hostC128, b->initElems * sizeof(double2), hipMemcpyHostToDevice));
    b->initSet = true;
    BATCH_CATCH()
}

void cvdvBatchSetLaunch(CVDVBatch* b, int threads, int ldsBytes) {
    if (!batch_ok(b)) return;
    if (threads != 256 && threads != 512) {
        cvdv_set_error(CVDV_ERR_ARG, "cvdvBatchSetLaunch: threads must be 256 or 512");
        return;
    }
    if (ldsBytes < 0 || (size_t)ldsBytes > device_max_lds() - 1024) {
        cvdv_set_error(CVDV_ERR_ARG, "cvdvBatchSetLaunch: LDS budget out of range");
        return;
    }
    b->threads = threads;
    b->ldsBytes = ldsBytes;
}

void cvdvBatchGetLaunch(CVDVBatch* b, int* threads, int* ldsBytes) {
    if (!b) return;
    if (threads) *threads = b->threads;
    if (ldsBytes) *ldsBytes = b->ldsBytes;
}

void cvdvBatchRun(CVDVBatch* b) {
    if (!batch_ok(b)) return;
    BATCH_TRY batch_run(b); BATCH_CATCH()
}

void cvdvBatchSynchronize(CVDVBatch* b) {
    if (!batch_ok(b)) return;
    BATCH_TRY HIPCHK(hipStreamSynchronize(b->stream)); BATCH_CATCH()
}

// Time `reps` runs of every launch shape in a small grid and keep the fastest.
// Returns the chosen time per run in microseconds (negative on error).
double cvdvBatchAutotune(CVDVBatch* b, int reps) {
    if (!batch_ok(b)) return -1.0;
    BATCH_TRY
    if (!b->initSet) throw CvdvError{CVDV_ERR_STATE, "cvdvBatchAutotune: call cvdvBatchInitFromSeparable first"};
    if (reps < 1) reps = 1;
    const int threadsOpts[2] = {256, 512};
    // budgets that hold padded states of 2^11 .. 2^14 complex64 (2^10 .. 2^13 complex128)
    const int ldsOpts[5] = {0, (int)lds_padded(16384), (int)lds_padded(32768), (int)lds_padded(65536),
                            (int)lds_padded(131072)};
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    double best = 1e300;
    int bt = b->threads, bl = b->ldsBytes;
    for (int t : threadsOpts)
        for (int l : ldsOpts) {
            if ((size_t)l > device_max_lds() - 1024) continue;
            b->threads = t;
            b->ldsBytes = l;
            batch_run(b);  // compile + warm
            if (b->threads != t) continue;  // raised to fit the largest FFT: covered by another shape
            HIPCHK(hipEventRecord(e0, b->stream));
            for (int r = 0; r < reps; ++r) HIPCHK(hipGraphLaunch(b->exec, b->stream));
            HIPCHK(hipEventRecord(e1, b->stream));
            HIPCHK(hipEventSynchronize(e1));
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            const double us = 1000.0 * ms / reps;
            if (us < best) {
                best = us;
                bt = t;
                bl = l;
            }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->threads = bt;
    b->ldsBytes = bl;
    batch_run(b);
    HIPCHK(hipStreamSynchronize(b->stream));
    return best;
    BATCH_CATCH(-1.0)
}

void cvdvBatchMeasure(CVDVBatch* b, int regIdx, double* out) {
    if (!batch_ok(b) || !out) return;
    if (regIdx < 0) {
        cvdv_set_error(CVDV_ERR_ARG, "cvdvBatchMeasure: invalid register index: " + std::to_string(regIdx));
        return;
    }
    BATCH_TRY batch_marginal(b, regIdx, out); BATCH_CATCH()
}

void cvdvBatchNorms(CVDVBatch* b, double* out) {
    if (!batch_ok(b) || !out) return;
    BATCH_TRY batch_marginal(b, -1, out); BATCH_CATCH()
}

void cvdvBatchGetState(CVDVBatch* b, int i, double* realOut, double* imagOut) {
    if (!batch_ok(b) || !realOut || !imagOut) return;
    if (i < 0 || i >= b->n) {
        cvdv_set_error(CVDV_ERR_ARG, "cvdvBatchGetState: index out of range: " + std::to_string(i));
        return;
    }
    BATCH_TRY
    if (!b->compiled) throw CvdvError{CVDV_ERR_STATE, "batch has not been run"};
    const size_t n = b->circ[i]->totalSize;
    const size_t eb = b->elemBytes();
    std::vector<uint8_t> h(n * eb);
    HIPCHK(hipMemcpyAsync(h.data(), (const uint8_t*)b->dOut + b->outOff[i] * eb, n * eb, hipMemcpyDeviceToHost,
                          b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    for (size_t g = 0; g < n; ++g) {
        if (b->prec == CVDV_C128) {
            const double2 v = ((const double2*)h.data())[g];
            realOut[g] = v.x;
            imagOut[g] = v.y;
        } else {
            const float2 v = ((const float2*)h.data())[g];
            realOut[g] = v.x;
            imagOut[g] = v.y;
        }
    }
    BATCH_CATCH()
}

}  // extern "C"
