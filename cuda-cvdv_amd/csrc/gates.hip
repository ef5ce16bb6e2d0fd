// gates.hip - gate compiler and the element-wise / pair kernels.
//
// Every C entry point issues the same decomposition as the reference
// (src/cuda/gates.cu:19-666, same turn-count rounding, same 1e-12 / 1e-15
// guards), but diagonal factors are not launched one by one: they are added
// to the context's pending DiagProg and applied by the next kernel that loads
// the state (an FFT pass, a qubit-pair pass, a flush).  Parity and register
// swap are recorded in the lazy layout map instead of moving data.

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <string>

#include "cvdv_internal.h"
#include "diag_fast.h"
#include "../../include/cvdv_hip.h"

static const double PI = 3.14159265358979323846;
static const double SQRT2 = 1.41421356237309504880;

// ---------------------------------------------------------------- kernels
// The element-wise and pair kernels give each thread this many elements (pairs)
// per iteration, their loads issued together.
#define PAIRS_PER_THREAD 4

template <typename C>
__global__ void k_diag(C* __restrict__ s, uint64_t total, FastDiag f) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t g0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g0 < total; g0 += PAIRS_PER_THREAD * stride) {
        C v[PAIRS_PER_THREAD];
#pragma unroll
        for (int k = 0; k < PAIRS_PER_THREAD; ++k)
            if (g0 + k * stride < total) v[k] = s[g0 + k * stride];
#pragma unroll
        for (int k = 0; k < PAIRS_PER_THREAD; ++k)
            if (g0 + k * stride < total) s[g0 + k * stride] = apply_fast(f, g0 + k * stride, v[k]);
    }
}

// A whole program evaluated per element in FP64 (programs beyond the compiled
// capacities; rare): the angle in FP64, then unit_rev_fast's sin/cos.
template <typename C>
__global__ void k_diag_prog(C* __restrict__ s, uint64_t total, DiagProg p) {
    using R = decltype(C{}.x);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t g0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g0 < total; g0 += PAIRS_PER_THREAD * stride) {
        C v[PAIRS_PER_THREAD];
#pragma unroll
        for (int k = 0; k < PAIRS_PER_THREAD; ++k)
            if (g0 + k * stride < total) v[k] = s[g0 + k * stride];
#pragma unroll
        for (int k = 0; k < PAIRS_PER_THREAD; ++k)
            if (g0 + k * stride < total)
                s[g0 + k * stride] = cmul(v[k], unit_rev_fast(diag_rev(p, g0 + k * stride), R(0)));
    }
}

// One term of a table: a polynomial over the table's field (vsel < 0), or a
// one-bit term whose bit is choice bit vsel (Rz on a qubit: c1 * (+-1); a
// polynomial over a one-qubit register: m = +-1); a control on choice bit csel
// negates its angle.
struct TabMember {
    int kind, vsel, csel, pad;
    double c1, c2, c3;
};

struct TabMake {
    int nb;          // field width; entries = 2^(nb + nsel)
    int m0, nm;      // members [m0, m0 + nm)
    uint32_t off;
    double g;        // constant phase folded into this table (revolutions)
    double scale;    // constant magnitude folded into this table
};

struct TabMakeArgs {
    int n;
    uint32_t total;
    TabMake t[CVDV_FAST_TABS];
    TabMember m[CVDV_MAX_TERMS];
};

// e^{2 pi i phase(entry)} for every table entry: each member's angle in FP64,
// reduced and summed, then FP64 sin/cos, rounded once to the state's precision.
template <typename C>
__global__ void k_make_tables(C* __restrict__ arena, TabMakeArgs a) {
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < a.total; e += gridDim.x * blockDim.x) {
        int i = 0;
        while (i + 1 < a.n && e >= a.t[i + 1].off) ++i;
        const TabMake& t = a.t[i];
        const uint32_t r = e - t.off;
        const uint32_t loc = r & ((1u << t.nb) - 1u), choice = r >> t.nb;
        const double mf = (double)((int64_t)(2 * loc) + 1 - ((int64_t)1 << t.nb));
        double v = 0.0;
        for (int k = t.m0; k < t.m0 + t.nm; ++k) {
            const TabMember& M = a.m[k];
            double mv;
            if (M.vsel >= 0) {
                const double m = ((choice >> M.vsel) & 1u) ? 1.0 : -1.0;
                mv = M.kind == DT_ZBIT ? M.c1 * m : fma(fma(M.c3, m, M.c2), m, M.c1) * m;
            } else {
                mv = M.kind == DT_ZBIT ? M.c1 * mf : fma(fma(M.c3, mf, M.c2), mf, M.c1) * mf;
            }
            if (M.csel >= 0 && ((choice >> M.csel) & 1u)) mv = -mv;
            v += cvdv_frac(mv);
        }
        const double2 u = unit_rev(cvdv_frac(cvdv_frac(v) + t.g), 0.0);
        arena[e] = cmk<C>(u.x * t.scale, u.y * t.scale);
    }
}

// 2x2 unitary on the pair (bit = 0, bit = 1); pending phases applied on load.
// The matrix product runs in FP64 for both precisions, so a complex64 state is
// rounded once per gate and never sees a rounded (non-unitary) matrix.
template <typename C>
__global__ void k_qubit(C* __restrict__ s, uint64_t npairs, int bit, double2 m00, double2 m01,
                        double2 m10, double2 m11, FastDiag p, int usep) {
    const uint64_t lowmask = (1ull << bit) - 1ull;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    // PAIRS_PER_THREAD pairs per iteration, all loaded before any is used: a
    // single pair per thread leaves too few bytes in flight to cover HBM latency
    for (uint64_t p0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p0 < npairs;
         p0 += PAIRS_PER_THREAD * stride) {
        C a[PAIRS_PER_THREAD], b[PAIRS_PER_THREAD];
#pragma unroll
        for (int k = 0; k < PAIRS_PER_THREAD; ++k) {
            const uint64_t pid = p0 + k * stride;
            if (pid < npairs) {
                const uint64_t i0 = ((pid & ~lowmask) << 1) | (pid & lowmask);
                a[k] = s[i0];
                b[k] = s[i0 | (1ull << bit)];
            }
        }
#pragma unroll
        for (int k = 0; k < PAIRS_PER_THREAD; ++k) {
            const uint64_t pid = p0 + k * stride;
            if (pid >= npairs) continue;
            const uint64_t i0 = ((pid & ~lowmask) << 1) | (pid & lowmask);
            const uint64_t i1 = i0 | (1ull << bit);
            C va = a[k], vb = b[k];
            if (usep) {
                va = apply_fast(p, i0, va);
                vb = apply_fast(p, i1, vb);
            }
            const double2 ad = make_double2((double)va.x, (double)va.y);
            const double2 bd = make_double2((double)vb.x, (double)vb.y);
            const double2 t00 = cmul(m00, ad), t01 = cmul(m01, bd), t10 = cmul(m10, ad), t11 = cmul(m11, bd);
            s[i0] = cmk<C>(t00.x + t01.x, t00.y + t01.y);
            s[i1] = cmk<C>(t10.x + t11.x, t10.y + t11.y);
        }
    }
}

// Conditional reflection j <-> N-1-j of the field [sh, sh+nb) where the
// control bit of the logically lower partner is 1 (reference semantics,
// src/cuda/kernels/gates.cuh:134-157 and torchCvdv.py:122-137).
template <typename C>
__global__ void k_cparity(C* __restrict__ s, uint64_t npairs, int sh, int nb, int ctrlPos,
                          int ctrlFlip, int tgtRefl, FastDiag p, int usep) {
    const int top = sh + nb - 1;
    const uint64_t lowmask = (1ull << top) - 1ull;
    const uint64_t fmask = ((1ull << nb) - 1ull) << sh;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t p0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p0 < npairs;
         p0 += PAIRS_PER_THREAD * stride) {
        C a[PAIRS_PER_THREAD], b[PAIRS_PER_THREAD];
#pragma unroll
        for (int k = 0; k < PAIRS_PER_THREAD; ++k) {
            const uint64_t pid = p0 + k * stride;
            if (pid < npairs) {
                const uint64_t i0 = ((pid & ~lowmask) << 1) | (pid & lowmask);  // top field bit 0
                a[k] = s[i0];
                b[k] = s[i0 ^ fmask];
            }
        }
#pragma unroll
        for (int k = 0; k < PAIRS_PER_THREAD; ++k) {
            const uint64_t pid = p0 + k * stride;
            if (pid >= npairs) continue;
            const uint64_t i0 = ((pid & ~lowmask) << 1) | (pid & lowmask);
            const uint64_t i1 = i0 ^ fmask;
            C va = a[k], vb = b[k];
            if (usep) {
                va = apply_fast(p, i0, va);
                vb = apply_fast(p, i1, vb);
            }
            const uint64_t lower = tgtRefl ? i1 : i0;
            const int cb = (int)((lower >> ctrlPos) & 1ull) ^ ctrlFlip;
            if (cb) {
                s[i0] = vb;
                s[i1] = va;
            } else if (usep) {
                s[i0] = va;
                s[i1] = vb;
            }
        }
    }
}

static unsigned grid_for(uint64_t total, unsigned block = 256) {
    uint64_t b = (total + block - 1) / block;
    if (b > 65536ull * 8) b = 65536ull * 8;
    return (unsigned)(b ? b : 1);
}

static inline bool prog_active(const DiagProg& p) { return p.n > 0 || p.grev != 0.0; }

// ---------------------------------------------------------------- pending program
DiagProg ctx_take_pending(CVDVContext* ctx) {
    DiagProg p = ctx->pending;
    ctx->pending = DiagProg{};
    return p;
}

static bool term_is_identity(const DiagTerm& t) {
    return t.c1 == 0.0 && t.c2 == 0.0 && t.c3 == 0.0;
}

// The whole program as one FP64 pass over the state (programs that exceed the
// compiled capacities).
static void diag_prog_pass(CVDVContext* ctx, const DiagProg& p) {
    fft_settle_for(ctx, p);
    const uint64_t n = ctx->totalSize;
    const int pid = prof_begin(ctx, KK_DIAG);
    if (ctx->prec == CVDV_C128)
        hipLaunchKernelGGL(k_diag_prog<double2>, dim3(grid_for((n + PAIRS_PER_THREAD - 1) / PAIRS_PER_THREAD)), dim3(256), 0, ctx->stream,
                           (double2*)ctx->state(), n, p);
    else
        hipLaunchKernelGGL(k_diag_prog<float2>, dim3(grid_for((n + PAIRS_PER_THREAD - 1) / PAIRS_PER_THREAD)), dim3(256), 0, ctx->stream,
                           (float2*)ctx->state(), n, p);
    HIPCHK(hipGetLastError());
    prof_end(ctx, pid);
}

// Grow the phase-table arena (synchronises: earlier launches may still read it).
void arena_reserve(CVDVContext* ctx, size_t need) {
    if (need <= ctx->arenaBytes) return;
    if (ctx->capturing)
        throw CvdvError{CVDV_ERR_UNSUPPORTED, "phase tables outgrew the arena reserved for the graph capture"};
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (ctx->arena) {
        if (ctx->graphs > 0) ctx->retired.push_back(ctx->arena);  // a captured graph may replay into it
        else HIPCHK(hipFree(ctx->arena));
    }
    ctx->arena = nullptr;
    ctx->arenaBytes = 0;
    const size_t sz = need < (size_t)(1 << 20) ? (size_t)(1 << 20) : need;
    HIPCHK(hipMalloc(&ctx->arena, sz));
    ctx->arenaBytes = sz;
}

static bool term_tabulated(const DiagTerm& t) {
    return (t.kind == DT_POLY && t.nb1 <= 20) || t.kind == DT_ZBIT;
}

// A program with more tabulated terms than one kernel holds (a conditional
// term per control qubit, for instance): diagonal terms commute, so the
// tabulated terms beyond the first CVDV_FAST_TABS - 1 go through table passes
// of their own, CVDV_FAST_TABS at a time, and the rest (with the slot left for
// the constant phase and scale) stays fused in the caller's kernel.  A table
// pass reads one table entry per amplitude; the FP64 pass it replaces
// evaluates every term's angle and sin/cos per amplitude.
static FastDiag compile_spilled(CVDVContext* ctx, const DiagProg& p, bool& active, double scale,
                                bool* scaleFolded, uint64_t tileBits) {
    DiagProg head{}, pass{};
    head.grev = p.grev;
    int nh = 0;
    for (int i = 0; i < p.n; ++i) {
        const DiagTerm& t = p.t[i];
        if (term_is_identity(t)) continue;
        if (!term_tabulated(t) || nh < CVDV_FAST_TABS - 1) {
            nh += term_tabulated(t) ? 1 : 0;
            head.t[head.n++] = t;
            continue;
        }
        pass.t[pass.n++] = t;
        if (pass.n == CVDV_FAST_TABS) {
            launch_diag(ctx, pass);
            pass = DiagProg{};
        }
    }
    if (pass.n) launch_diag(ctx, pass);
    return compile_diag(ctx, head, active, scale, scaleFolded, tileBits);
}

// Tabulated terms grouped into tables (at most CVDV_FAST_TABS + 1 groups are
// formed; more means the program does not fit).
struct TabGroup {
    int sh, nb, ctrl, nsel;
    int solo;  // a term whose choice bit would vary within a consumer's tile
    int sel[CVDV_TAB_SEL];
    int nm;
    int mem[CVDV_MAX_TERMS];
};

static int sel_index(const TabGroup& g, int bit) {
    for (int j = 0; j < g.nsel; ++j)
        if (g.sel[j] == bit) return j;
    return -1;
}

// Terms over the same field of two or more bits share a table when their
// control bits lie outside the field; one-bit terms (and their controls) share
// a field-less table over the combinations of their bits.  A table keeps at
// most CVDV_TAB_SEL choice bits and 2^CVDV_TAB_MERGE_BITS entries, and no
// choice bit among tileBits (such a term keeps a table of its own).
static int group_tables(const DiagProg& p, TabGroup* g, uint64_t tileBits) {
    int ng = 0;
    auto add_sel = [](TabGroup& G, int bit) {
        if (bit >= 0 && sel_index(G, bit) < 0) G.sel[G.nsel++] = bit;
    };
    for (int i = 0; i < p.n; ++i) {
        const DiagTerm& t = p.t[i];
        if (term_is_identity(t) || !term_tabulated(t)) continue;
        const bool oneBit = t.kind == DT_ZBIT || t.nb1 == 1;
        const bool solo = (oneBit && ((tileBits >> t.sh1) & 1ull)) || (t.ctrl >= 0 && ((tileBits >> t.ctrl) & 1ull));
        int k = solo ? ng : 0;
        for (; k < ng; ++k) {
            TabGroup& G = g[k];
            if (G.solo) continue;
            int need = 0;
            if (oneBit) {
                if (G.nb != 0) continue;
                need = (sel_index(G, t.sh1) < 0) + (t.ctrl >= 0 && t.ctrl != t.sh1 && sel_index(G, t.ctrl) < 0);
            } else {
                if (G.sh != t.sh1 || G.nb != t.nb1) continue;
                if (t.ctrl >= t.sh1 && t.ctrl < t.sh1 + t.nb1) continue;
                need = t.ctrl >= 0 && sel_index(G, t.ctrl) < 0;
            }
            if (G.nsel + need <= CVDV_TAB_SEL && G.nb + G.nsel + need <= CVDV_TAB_MERGE_BITS) break;
        }
        if (k == ng) {
            if (ng == CVDV_FAST_TABS + 1) return ng;
            g[ng] = TabGroup{};
            g[ng].sh = oneBit ? 0 : t.sh1;
            g[ng].nb = oneBit ? 0 : t.nb1;
            g[ng].ctrl = -1;
            g[ng].solo = solo;
            ++ng;
        }
        TabGroup& G = g[k];
        if (oneBit) add_sel(G, t.sh1);
        if (t.ctrl >= 0) add_sel(G, t.ctrl);
        G.mem[G.nm++] = i;
    }
    // A lone term keeps the plain form: its field indexes the table and its
    // control conjugates the entry.
    for (int k = 0; k < ng; ++k) {
        TabGroup& G = g[k];
        if (G.nm != 1) continue;
        const DiagTerm& t = p.t[G.mem[0]];
        G.sh = t.sh1;
        G.nb = t.kind == DT_ZBIT ? 1 : t.nb1;
        G.ctrl = t.ctrl;
        G.nsel = 0;
    }
    return ng;
}

bool diag_fits(const DiagProg& p, uint64_t tileBits) {
    int ngen = 0;
    for (int i = 0; i < p.n; ++i) ngen += !term_is_identity(p.t[i]) && !term_tabulated(p.t[i]);
    if (ngen > CVDV_FAST_GEN) return false;
    TabGroup grp[CVDV_FAST_TABS + 1];
    // one table slot stays free for the constant phase and scale
    return group_tables(p, grp, tileBits) < CVDV_FAST_TABS;
}

size_t diag_arena_bound(int maxNb, size_t elemBytes) {
    const int plain = std::min(maxNb, 20);
    const int merged = std::min(maxNb + CVDV_TAB_SEL, CVDV_TAB_MERGE_BITS);
    return ((size_t)CVDV_FAST_TABS * ((size_t)1 << std::max(plain, merged)) + 1) * elemBytes;
}

FastDiag compile_diag(CVDVContext* ctx, const DiagProg& p, bool& active, double scale, bool* scaleFolded,
                      uint64_t tileBits) {
    FastDiag f{};
    if (scaleFolded) *scaleFolded = false;
    active = false;
    int ngen = 0;
    for (int i = 0; i < p.n; ++i) ngen += !term_is_identity(p.t[i]) && !term_tabulated(p.t[i]);
    if (ngen > CVDV_FAST_GEN) {
        // too many generic terms: one FP64 pass now, nothing left to apply
        diag_prog_pass(ctx, p);
        return FastDiag{};
    }
    TabGroup grp[CVDV_FAST_TABS + 1];
    const int ng = group_tables(p, grp, tileBits);
    if (ng > CVDV_FAST_TABS) return compile_spilled(ctx, p, active, scale, scaleFolded, tileBits);
    for (int i = 0; i < p.n; ++i)
        if (!term_is_identity(p.t[i]) && !term_tabulated(p.t[i])) f.gen[f.ngen++] = p.t[i];

    TabMakeArgs a{};
    uint32_t off = 0;
    int nm = 0;
    for (int k = 0; k < ng; ++k) {
        const TabGroup& G = grp[k];
        DiagTab T{};
        T.sh = G.sh;
        T.nb = G.nb;
        T.ctrl = G.ctrl;
        T.off = off;
        T.nsel = G.nsel;
        for (int j = 0; j < G.nsel; ++j) T.sel |= (uint32_t)G.sel[j] << (8 * j);
        f.tab[f.ntab++] = T;
        a.t[a.n++] = TabMake{G.nb, nm, G.nm, off, 0.0, 1.0};
        for (int j = 0; j < G.nm; ++j) {
            const DiagTerm& t = p.t[G.mem[j]];
            TabMember M{};
            M.kind = t.kind;
            M.vsel = G.nsel > 0 && (t.kind == DT_ZBIT || t.nb1 == 1) && G.nb == 0 ? sel_index(G, t.sh1) : -1;
            M.csel = G.nsel > 0 && t.ctrl >= 0 ? sel_index(G, t.ctrl) : -1;
            M.c1 = t.c1;
            M.c2 = t.c2;
            M.c3 = t.c3;
            a.m[nm++] = M;
        }
        off += 1u << (G.nb + G.nsel);
    }
    // The global phase and the scale are folded into a table that every
    // element reads unconjugated (exact in FP64 before the single rounding); a
    // rounded constant factor would scale the norm by the same |g|^2 - 1 on
    // every pass.  Without such a table the phase goes to the generic terms'
    // angle and the scale back to the caller (which applies it unbiased), or a
    // lone conditional term takes its control as a choice bit and carries both,
    // or, last, a one-entry table does.
    const double grev = cvdv_frac(p.grev);
    int host = -1;
    for (int i = 0; i < a.n; ++i)
        if (f.tab[i].ctrl < 0) {
            host = i;
            break;
        }
    if (host < 0 && (grev != 0.0 || scale != 1.0) && (f.ngen == 0 || (scale != 1.0 && !scaleFolded))) {
        for (int i = 0; i < a.n; ++i) {
            DiagTab& T = f.tab[i];
            if (a.t[i].nm != 1 || T.ctrl < 0 || ((tileBits >> T.ctrl) & 1ull) ||
                T.nb + 1 > CVDV_TAB_MERGE_BITS)
                continue;
            // entries double from here on: shift the later tables
            const uint32_t grow = 1u << T.nb;
            for (int j = i + 1; j < a.n; ++j) {
                f.tab[j].off += grow;
                a.t[j].off += grow;
            }
            off += grow;
            T.sel = (uint32_t)T.ctrl;
            T.nsel = 1;
            T.ctrl = -1;
            a.m[a.t[i].m0].csel = 0;
            host = i;
            break;
        }
    }
    if (host < 0 && (grev != 0.0 || scale != 1.0)) {
        if (f.ngen > 0 && (scale == 1.0 || scaleFolded)) {
            f.genrev = grev;
        } else if (f.ntab < CVDV_FAST_TABS) {
            host = f.ntab;
            f.tab[f.ntab++] = DiagTab{0, 0, -1, off, 0, 0u};
            a.t[a.n++] = TabMake{0, nm, 0, off, 0.0, 1.0};
            off += 1;
        } else {
            return compile_spilled(ctx, p, active, scale, scaleFolded, tileBits);
        }
    }
    if (host >= 0) {
        a.t[host].g = grev;
        a.t[host].scale = scale;
        if (scaleFolded) *scaleFolded = true;
    }
    active = f.ntab > 0 || f.ngen > 0;
    if (a.n) {
        a.total = off;
        arena_reserve(ctx, (size_t)off * ctx->elemBytes());
        const unsigned blocks = (unsigned)((off + 255) / 256);
        if (ctx->prec == CVDV_C128)
            hipLaunchKernelGGL(k_make_tables<double2>, dim3(blocks), dim3(256), 0, ctx->stream,
                               (double2*)ctx->arena, a);
        else
            hipLaunchKernelGGL(k_make_tables<float2>, dim3(blocks), dim3(256), 0, ctx->stream,
                               (float2*)ctx->arena, a);
        HIPCHK(hipGetLastError());
    }
    f.arena = ctx->arena;
    return f;
}

// Bits of the flat index a term depends on.
static uint64_t term_bits(const DiagTerm& t) {
    auto field = [](int sh, int nb) { return (((uint64_t)1 << nb) - 1ull) << sh; };
    uint64_t m;
    if (t.kind == DT_POLY || t.kind == DT_LINF) m = field(t.sh1, t.nb1);
    else if (t.kind == DT_BILIN) m = field(t.sh1, t.nb1) | field(t.sh2, t.nb2);
    else m = 1ull << t.sh1;
    if (t.ctrl >= 0) m |= 1ull << t.ctrl;
    return m;
}

DiagProg diag_fold_linear(const DiagProg& p, int sh, int nb) {
    DiagProg q = p;
    int ngen = 0;
    for (int i = 0; i < q.n; ++i) ngen += !term_is_identity(q.t[i]) && !term_tabulated(q.t[i]);
    for (int i = 0; i < q.n && ngen < CVDV_FAST_GEN; ++i) {
        DiagTerm& t = q.t[i];
        if (t.kind == DT_POLY && t.sh1 == sh && t.nb1 == nb && t.c2 == 0.0 && t.c3 == 0.0 && !term_is_identity(t) &&
            !(t.ctrl >= sh && t.ctrl < sh + nb)) {
            t.kind = DT_LINF;
            ++ngen;
        }
    }
    return q;
}

void fft_settle_for(CVDVContext* ctx, const DiagProg& p, uint64_t except) {
    if (!ctx->defer.any()) return;
    uint64_t m = 0;
    for (int i = 0; i < p.n; ++i)
        if (!term_is_identity(p.t[i])) m |= term_bits(p.t[i]);
    fft_settle_bits(ctx, m & ~except);
}

// The pending terms that do not commute with an operation mixing the index
// bits in `mix` (an FFT along a field, a 2x2 gate on a bit, a conditional
// reflection of a field).  A diagonal term that depends on none of those bits
// has the same value on every element the operation mixes, so it stays
// pending; the global phase goes with any non-empty taken program.
static DiagProg ctx_take_involving(CVDVContext* ctx, uint64_t mix) {
    DiagProg take{}, keep{};
    for (int i = 0; i < ctx->pending.n; ++i) {
        const DiagTerm& t = ctx->pending.t[i];
        if (term_is_identity(t)) continue;
        if (term_bits(t) & mix) take.t[take.n++] = t;
        else keep.t[keep.n++] = t;
    }
    if (take.n > 0) take.grev = ctx->pending.grev;
    else keep.grev = ctx->pending.grev;
    ctx->pending = keep;
    return take;
}

void rec_op(CVDVContext* ctx, BatchOp op, const DiagProg& p) {
    op.prog = -1;
    if (prog_active(p)) {
        op.prog = (int)ctx->rec->progs.size();
        ctx->rec->progs.push_back(p);
    }
    ctx->rec->ops.push_back(op);
}

bool refuse_recording(CVDVContext* ctx) {
    if (ctx && ctx->rec) {
        cvdv_set_error(CVDV_ERR_STATE, "batch circuit context: run the batch and read it through the batch API");
        return true;
    }
    if (ctx && ctx->capturing) {
        cvdv_set_error(CVDV_ERR_STATE, "the context is capturing a graph: end the capture first");
        return true;
    }
    return false;
}

void launch_diag(CVDVContext* ctx, const DiagProg& pp) {
    if (ctx->rec) {
        BatchOp op{};
        op.kind = BOP_DIAG;
        rec_op(ctx, op, pp);
        return;
    }
    fft_settle_for(ctx, pp);
    const uint64_t n = ctx->totalSize;
    bool act;
    const FastDiag p = compile_diag(ctx, pp, act);
    const int pid = prof_begin(ctx, KK_DIAG);
    if (ctx->prec == CVDV_C128)
        hipLaunchKernelGGL(k_diag<double2>, dim3(grid_for((n + PAIRS_PER_THREAD - 1) / PAIRS_PER_THREAD)), dim3(256), 0, ctx->stream,
                           (double2*)ctx->state(), n, p);
    else
        hipLaunchKernelGGL(k_diag<float2>, dim3(grid_for((n + PAIRS_PER_THREAD - 1) / PAIRS_PER_THREAD)), dim3(256), 0, ctx->stream,
                           (float2*)ctx->state(), n, p);
    HIPCHK(hipGetLastError());
    prof_end(ctx, pid);
}

void ctx_flush(CVDVContext* ctx) {
    if (!ctx->rec) fft_settle(ctx);
    if (!prog_active(ctx->pending)) return;
    DiagProg p = ctx_take_pending(ctx);
    launch_diag(ctx, p);
}

static bool same_slot(const DiagTerm& a, const DiagTerm& b) {
    return a.kind == b.kind && a.sh1 == b.sh1 && a.nb1 == b.nb1 && a.sh2 == b.sh2 &&
           a.nb2 == b.nb2 && a.ctrl == b.ctrl;
}

static void prog_add(DiagProg& p, const DiagTerm& t) {
    for (int i = 0; i < p.n; ++i) {
        if (same_slot(p.t[i], t)) {
            p.t[i].c1 += t.c1;
            p.t[i].c2 += t.c2;
            p.t[i].c3 += t.c3;
            return;
        }
    }
    p.t[p.n++] = t;
}

void ctx_push_term(CVDVContext* ctx, const DiagTerm& t0) {
    DiagTerm t = t0;
    if (t.kind == DT_BILIN && t.sh2 < t.sh1) {
        std::swap(t.sh1, t.sh2);
        std::swap(t.nb1, t.nb2);
    }
    bool found = false;
    for (int i = 0; i < ctx->pending.n; ++i) found |= same_slot(ctx->pending.t[i], t);
    if (!found && ctx->pending.n == CVDV_MAX_TERMS) {
        // apply what is pending (settling a half-done transform only if it reads its field)
        const DiagProg p = ctx_take_pending(ctx);
        launch_diag(ctx, p);
    }
    prog_add(ctx->pending, t);
}

// A logical diagonal term over register r (radians per x^k) -> physical term.
static DiagTerm poly_term(CVDVContext* ctx, int r, double c1, double c2, double c3, int ctrlReg,
                          int ctrlQubit) {
    DiagTerm t{};
    t.kind = DT_POLY;
    t.sh1 = ctx->pshift[r];
    t.nb1 = ctx->nq[r];
    const double N = (double)((uint64_t)1 << ctx->nq[r]);
    const double dx = ctx->dx[r];
    // x = m dx / 2; revolutions = angle / (2 pi); dx^2 = 2 pi / N exactly.
    t.c1 = c1 * dx / (4.0 * PI);
    t.c2 = c2 / (4.0 * N);
    t.c3 = c3 * dx * dx * dx / (16.0 * PI);
    if (ctx->refl[r]) {  // logical m = -physical m
        t.c1 = -t.c1;
        t.c3 = -t.c3;
    }
    t.ctrl = -1;
    if (ctrlReg >= 0) {
        t.ctrl = (int)ctx_ctrl_bit(ctx, ctrlReg, ctrlQubit);
        if (ctx->refl[ctrlReg]) {  // logical control bit = NOT physical bit
            t.c1 = -t.c1;
            t.c2 = -t.c2;
            t.c3 = -t.c3;
        }
    }
    return t;
}

static DiagTerm bilin_term(CVDVContext* ctx, int r1, int r2, double coeff, int ctrlReg,
                           int ctrlQubit) {
    DiagTerm t{};
    t.kind = DT_BILIN;
    t.sh1 = ctx->pshift[r1];
    t.nb1 = ctx->nq[r1];
    t.sh2 = ctx->pshift[r2];
    t.nb2 = ctx->nq[r2];
    if (ctx->nq[r1] == ctx->nq[r2])
        t.c1 = coeff / (4.0 * (double)((uint64_t)1 << ctx->nq[r1]));
    else
        t.c1 = coeff * ctx->dx[r1] * ctx->dx[r2] / (8.0 * PI);
    if (ctx->refl[r1] ^ ctx->refl[r2]) t.c1 = -t.c1;
    t.ctrl = -1;
    if (ctrlReg >= 0) {
        t.ctrl = (int)ctx_ctrl_bit(ctx, ctrlReg, ctrlQubit);
        if (ctx->refl[ctrlReg]) t.c1 = -t.c1;
    }
    return t;
}

// ---------------------------------------------------------------- primitive ops
void ctx_ft(CVDVContext* ctx, int reg, bool inverse) {
    const int nb = ctx->nq[reg];
    if (nb == 0) return;  // N = 1: the transform is the identity
    const uint64_t N = (uint64_t)1 << nb;
    const double sgn = inverse ? -1.0 : 1.0;
    // Centred DFT: F[k,j] = N^-1/2 exp(-2 pi i m_j m_k / (4N)) =
    //   e^{2 pi i (N-1)^2/(4N)} D FFT D / sqrt(N),  D = diag e^{2 pi i (N-1) m/(4N)}
    // (gates.cu:19-67: pre phase, FFT, post phase, 1/sqrt(N), global phase).
    DiagTerm tw{};
    tw.kind = DT_POLY;
    tw.sh1 = ctx->pshift[reg];
    tw.nb1 = nb;
    tw.ctrl = -1;
    tw.c1 = sgn * (double)(N - 1) / (4.0 * (double)N);
    DiagProg pre = ctx_take_involving(ctx, (((uint64_t)1 << nb) - 1ull) << ctx->pshift[reg]);
    // other registers' half-done transforms whose fields these phases read
    if (!ctx->rec) fft_settle_for(ctx, pre, (((uint64_t)1 << nb) - 1ull) << ctx->pshift[reg]);
    if (pre.n == CVDV_MAX_TERMS) {
        bool found = false;
        for (int i = 0; i < pre.n; ++i) found |= same_slot(pre.t[i], tw);
        if (!found) {
            launch_diag(ctx, pre);
            pre = DiagProg{};
        }
    }
    prog_add(pre, tw);
    if (ctx->rec) {
        BatchOp op{};
        op.kind = BOP_FFT;
        op.sh = ctx->pshift[reg];
        op.nb = nb;
        op.inv = inverse ? 1 : 0;
        op.outScale = 1.0 / std::sqrt((double)N);
        rec_op(ctx, op, pre);
    } else {
        fft_register(ctx, ctx->pshift[reg], nb, inverse, pre, 1.0 / std::sqrt((double)N));
    }
    const double gphase = (double)(N - 1) * (double)(N - 1) / (4.0 * (double)N);
    ctx->pending.grev = cvdv_frac(ctx->pending.grev + sgn * gphase);
    ctx_push_term(ctx, tw);
}

static void qubit_gate(CVDVContext* ctx, int reg, int qubit, const double m[8]) {
    // m = {m00.re, m00.im, m01.re, m01.im, m10..., m11...}
    double mm[8];
    for (int i = 0; i < 8; ++i) mm[i] = m[i];
    if (ctx->refl[reg]) {  // logical bit = NOT physical bit: X M X
        const double t[8] = {m[6], m[7], m[4], m[5], m[2], m[3], m[0], m[1]};
        for (int i = 0; i < 8; ++i) mm[i] = t[i];
    }
    const int bit = (int)ctx_ctrl_bit(ctx, reg, qubit);
    if (ctx->rec) {
        BatchOp op{};
        op.kind = BOP_QUBIT;
        op.bit = bit;
        for (int i = 0; i < 4; ++i) op.m[i] = make_double2(mm[2 * i], mm[2 * i + 1]);
        rec_op(ctx, op, ctx_take_involving(ctx, 1ull << bit));
        return;
    }
    fft_settle_bits(ctx, 1ull << bit);
    const DiagProg taken = ctx_take_involving(ctx, 1ull << bit);
    fft_settle_for(ctx, taken);
    bool act;
    const FastDiag p = compile_diag(ctx, taken, act);
    const int usep = act ? 1 : 0;
    const uint64_t np = ctx->totalSize / 2;
    const double2 m00 = make_double2(mm[0], mm[1]), m01 = make_double2(mm[2], mm[3]);
    const double2 m10 = make_double2(mm[4], mm[5]), m11 = make_double2(mm[6], mm[7]);
    const int pid = prof_begin(ctx, KK_QUBIT);
    if (ctx->prec == CVDV_C128)
        hipLaunchKernelGGL(k_qubit<double2>, dim3(grid_for((np + PAIRS_PER_THREAD - 1) / PAIRS_PER_THREAD)), dim3(256), 0, ctx->stream,
                           (double2*)ctx->state(), np, bit, m00, m01, m10, m11, p, usep);
    else
        hipLaunchKernelGGL(k_qubit<float2>, dim3(grid_for((np + PAIRS_PER_THREAD - 1) / PAIRS_PER_THREAD)), dim3(256), 0, ctx->stream,
                           (float2*)ctx->state(), np, bit, m00, m01, m10, m11, p, usep);
    HIPCHK(hipGetLastError());
    prof_end(ctx, pid);
}

static void zbit(CVDVContext* ctx, int reg, int qubit, double theta) {
    // Rz(theta) = diag(e^{-i theta/2}, e^{i theta/2}) (gates.cuh:88-91)
    DiagTerm t{};
    t.kind = DT_ZBIT;
    t.sh1 = (int)ctx_ctrl_bit(ctx, reg, qubit);
    t.nb1 = 1;
    t.ctrl = -1;
    t.c1 = theta / (4.0 * PI);
    if (ctx->refl[reg]) t.c1 = -t.c1;
    ctx_push_term(ctx, t);
}

static void pauli(CVDVContext* ctx, int reg, int qubit, int axis, double theta) {
    if (axis != 0 && axis != 1) {  // gates.cuh:88: any other axis value is Z
        zbit(ctx, reg, qubit, theta);
        return;
    }
    const double c = std::cos(theta / 2.0), s = std::sin(theta / 2.0);
    if (axis == 0) {  // Rx = [[c, -is], [-is, c]]
        const double m[8] = {c, 0, 0, -s, 0, -s, c, 0};
        qubit_gate(ctx, reg, qubit, m);
    } else {  // Ry = [[c, -s], [s, c]]
        const double m[8] = {c, 0, -s, 0, s, 0, c, 0};
        qubit_gate(ctx, reg, qubit, m);
    }
}

static void hadamard(CVDVContext* ctx, int reg, int qubit) {
    const double h = 1.0 / SQRT2;
    const double m[8] = {h, 0, h, 0, h, 0, -h, 0};
    qubit_gate(ctx, reg, qubit, m);
}

static void parity(CVDVContext* ctx, int reg) { ctx->refl[reg] ^= 1; }

static void cparity(CVDVContext* ctx, int tgt, int creg, int cq) {
    const int nb = ctx->nq[tgt];
    if (nb == 0) return;
    const int cpos = (int)ctx_ctrl_bit(ctx, creg, cq);
    if (ctx->rec) {
        BatchOp op{};
        op.kind = BOP_CPARITY;
        op.sh = ctx->pshift[tgt];
        op.nb = nb;
        op.bit = cpos;
        op.ctrlFlip = ctx->refl[creg];
        op.tgtRefl = ctx->refl[tgt];
        rec_op(ctx, op, ctx_take_involving(ctx, (((uint64_t)1 << nb) - 1ull) << ctx->pshift[tgt]));
        return;
    }
    fft_settle_bits(ctx, ((((uint64_t)1 << nb) - 1ull) << ctx->pshift[tgt]) | (1ull << cpos));
    const DiagProg taken = ctx_take_involving(ctx, (((uint64_t)1 << nb) - 1ull) << ctx->pshift[tgt]);
    fft_settle_for(ctx, taken);
    bool act;
    const FastDiag p = compile_diag(ctx, taken, act);
    const int usep = act ? 1 : 0;
    const uint64_t np = ctx->totalSize / 2;
    const int pid = prof_begin(ctx, KK_CPARITY);
    if (ctx->prec == CVDV_C128)
        hipLaunchKernelGGL(k_cparity<double2>, dim3(grid_for((np + PAIRS_PER_THREAD - 1) / PAIRS_PER_THREAD)), dim3(256), 0, ctx->stream,
                           (double2*)ctx->state(), np, ctx->pshift[tgt], nb, cpos, ctx->refl[creg],
                           ctx->refl[tgt], p, usep);
    else
        hipLaunchKernelGGL(k_cparity<float2>, dim3(grid_for((np + PAIRS_PER_THREAD - 1) / PAIRS_PER_THREAD)), dim3(256), 0, ctx->stream,
                           (float2*)ctx->state(), np, ctx->pshift[tgt], nb, cpos, ctx->refl[creg],
                           ctx->refl[tgt], p, usep);
    HIPCHK(hipGetLastError());
    prof_end(ctx, pid);
}

static void swap_regs(CVDVContext* ctx, int r1, int r2) {
    std::swap(ctx->pshift[r1], ctx->pshift[r2]);
    std::swap(ctx->refl[r1], ctx->refl[r2]);
}

static void phase_poly(CVDVContext* ctx, int r, double c1, double c2, double c3, int creg = -1,
                       int cq = 0) {
    ctx_push_term(ctx, poly_term(ctx, r, c1, c2, c3, creg, cq));
}

static void phase_bilin(CVDVContext* ctx, int r1, int r2, double c, int creg = -1, int cq = 0) {
    ctx_push_term(ctx, bilin_term(ctx, r1, r2, c, creg, cq));
}

// ---------------------------------------------------------------- validation
#define CVDV_TRY try {
#define CVDV_CATCH()                                                         \
    }                                                                        \
    catch (const CvdvError& e) {                                             \
        cvdv_set_error(e.code, e.msg);                                       \
    }                                                                        \
    catch (...) {                                                            \
        cvdv_set_error(CVDV_ERR_HIP, "unexpected internal error");           \
    }

static bool ready(CVDVContext* ctx) {
    if (!ctx) return false;
    if (!ctx->initialized) {
        cvdv_set_error(CVDV_ERR_STATE, "State not initialized");
        return false;
    }
    return true;
}

static bool reg_ok(CVDVContext* ctx, int r, const char* what = "Invalid register index") {
    if (r < 0 || r >= ctx->numReg) {
        cvdv_set_error(CVDV_ERR_ARG, std::string(what) + ": " + std::to_string(r));
        return false;
    }
    return true;
}

static bool ctrl_ok(CVDVContext* ctx, int creg, int cq) {
    if (!reg_ok(ctx, creg, "Invalid control register index")) return false;
    if (cq < 0 || cq >= ctx->nq[creg]) {
        cvdv_set_error(CVDV_ERR_ARG, "Invalid control qubit index: " + std::to_string(cq));
        return false;
    }
    return true;
}

static bool qubit_ok(CVDVContext* ctx, int r, int q) {
    if (!reg_ok(ctx, r)) return false;
    if (q < 0 || q >= ctx->nq[r]) {
        cvdv_set_error(CVDV_ERR_ARG, "Invalid qubit index: " + std::to_string(q));
        return false;
    }
    return true;
}

static bool two_regs_ok(CVDVContext* ctx, int r1, int r2) {
    if (r1 < 0 || r1 >= ctx->numReg || r2 < 0 || r2 >= ctx->numReg) {
        cvdv_set_error(CVDV_ERR_ARG, "Invalid register indices: " + std::to_string(r1) + ", " +
                                         std::to_string(r2));
        return false;
    }
    return true;
}

// ---------------------------------------------------------------- composite helpers
static void rotation_small(CVDVContext* ctx, int r, double th) {  // gates.cu:572-584
    if (std::fabs(th) < 1e-15) return;
    const double tq = std::tan(th / 2.0), st = std::sin(th);
    phase_poly(ctx, r, 0, -0.5 * tq, 0);
    ctx_ft(ctx, r, false);
    phase_poly(ctx, r, 0, -0.5 * st, 0);
    ctx_ft(ctx, r, true);
    phase_poly(ctx, r, 0, -0.5 * tq, 0);
}

static void crotation_small(CVDVContext* ctx, int t, int c, int q, double th) {  // gates.cu:598-611
    if (std::fabs(th) < 1e-15) return;
    const double tq = std::tan(th / 2.0), st = std::sin(th);
    phase_poly(ctx, t, 0, -0.5 * tq, 0, c, q);
    ctx_ft(ctx, t, false);
    phase_poly(ctx, t, 0, -0.5 * st, 0, c, q);
    ctx_ft(ctx, t, true);
    phase_poly(ctx, t, 0, -0.5 * tq, 0, c, q);
}

static void bs_small(CVDVContext* ctx, int r1, int r2, double th, int c = -1, int q = 0) {
    // gates.cu:614-636 (and 640-666 with a control)
    if (std::fabs(th) < 1e-15) return;
    const double cq = -std::tan(0.25 * th), cp = -std::sin(0.5 * th);
    // The two transforms act on different registers and commute; the first one
    // of each pair applies the pending q1*q2 (p1*p2) phase.  A register whose
    // transform keeps a half pending goes first: the other register's pass then
    // finds no pending term on its bits, and its two transforms join one chain of
    // fused passes (a pending half must be completed before any pass whose
    // phases read its bits).  Otherwise the register with the lower physical
    // field goes first, since its rows are contiguous, where the phase reduces to
    // one affine term per row instead of a per-element evaluation.
    const bool d1 = fft_defers(ctx, ctx->pshift[r1], ctx->nq[r1]);
    const bool d2 = fft_defers(ctx, ctx->pshift[r2], ctx->nq[r2]);
    int fa = ctx->pshift[r1] <= ctx->pshift[r2] ? r1 : r2;
    if (d1 != d2) fa = d1 ? r1 : r2;
    const int fb = fa == r1 ? r2 : r1;
    phase_bilin(ctx, r1, r2, cq, c, q);
    ctx_ft(ctx, fa, false);
    ctx_ft(ctx, fb, false);
    phase_bilin(ctx, r1, r2, cp, c, q);
    ctx_ft(ctx, fa, true);
    ctx_ft(ctx, fb, true);
    phase_bilin(ctx, r1, r2, cq, c, q);
}

// ---------------------------------------------------------------- C ABI
extern "C" {

void cvdvFtQ2P(CVDVContext* ctx, int regIdx) {
    if (!ready(ctx) || !reg_ok(ctx, regIdx)) return;
    CVDV_TRY ctx_ft(ctx, regIdx, false); CVDV_CATCH()
}

void cvdvFtP2Q(CVDVContext* ctx, int regIdx) {
    if (!ready(ctx) || !reg_ok(ctx, regIdx)) return;
    CVDV_TRY ctx_ft(ctx, regIdx, true); CVDV_CATCH()
}

void cvdvDisplacement(CVDVContext* ctx, int regIdx, double betaRe, double betaIm) {
    // gates.cu:124-161; the global phase e^{-i Re Im} is dropped there too
    if (!ready(ctx) || !reg_ok(ctx, regIdx)) return;
    CVDV_TRY
    if (std::fabs(betaIm) > 1e-12) phase_poly(ctx, regIdx, SQRT2 * betaIm, 0, 0);
    if (std::fabs(betaRe) > 1e-12) {
        ctx_ft(ctx, regIdx, false);
        phase_poly(ctx, regIdx, -SQRT2 * betaRe, 0, 0);
        ctx_ft(ctx, regIdx, true);
    }
    CVDV_CATCH()
}

void cvdvConditionalDisplacement(CVDVContext* ctx, int targetReg, int ctrlReg, int ctrlQubit,
                                 double alphaRe, double alphaIm) {
    // gates.cu:163-204
    if (!ready(ctx) || !reg_ok(ctx, targetReg, "Invalid target register index") ||
        !ctrl_ok(ctx, ctrlReg, ctrlQubit))
        return;
    CVDV_TRY
    if (std::fabs(alphaIm) > 1e-12) phase_poly(ctx, targetReg, SQRT2 * alphaIm, 0, 0, ctrlReg, ctrlQubit);
    if (std::fabs(alphaRe) > 1e-12) {
        ctx_ft(ctx, targetReg, false);
        phase_poly(ctx, targetReg, -SQRT2 * alphaRe, 0, 0, ctrlReg, ctrlQubit);
        ctx_ft(ctx, targetReg, true);
    }
    CVDV_CATCH()
}

void cvdvPauliRotation(CVDVContext* ctx, int regIdx, int qubitIdx, int axis, double theta) {
    // gates.cu:206-219
    if (!ready(ctx) || !qubit_ok(ctx, regIdx, qubitIdx)) return;
    CVDV_TRY pauli(ctx, regIdx, qubitIdx, axis, theta); CVDV_CATCH()
}

void cvdvHadamard(CVDVContext* ctx, int regIdx, int qubitIdx) {
    // gates.cu:221-233
    if (!ready(ctx) || !qubit_ok(ctx, regIdx, qubitIdx)) return;
    CVDV_TRY hadamard(ctx, regIdx, qubitIdx); CVDV_CATCH()
}

void cvdvParity(CVDVContext* ctx, int regIdx) {
    // gates.cu:235-246: j -> N-1-j, recorded in the layout map
    if (!ready(ctx) || !reg_ok(ctx, regIdx)) return;
    parity(ctx, regIdx);
}

void cvdvConditionalParity(CVDVContext* ctx, int targetReg, int ctrlReg, int ctrlQubit) {
    // gates.cu:248-268
    if (!ready(ctx) || !reg_ok(ctx, targetReg, "Invalid target register index") ||
        !ctrl_ok(ctx, ctrlReg, ctrlQubit))
        return;
    CVDV_TRY cparity(ctx, targetReg, ctrlReg, ctrlQubit); CVDV_CATCH()
}

void cvdvSwapRegisters(CVDVContext* ctx, int reg1, int reg2) {
    // gates.cu:270-286: recorded in the layout map
    if (!ready(ctx) || !two_regs_ok(ctx, reg1, reg2)) return;
    if (ctx->nq[reg1] != ctx->nq[reg2]) {
        cvdv_set_error(CVDV_ERR_ARG, "SWAP requires registers with same number of qubits");
        return;
    }
    if (reg1 == reg2) return;
    swap_regs(ctx, reg1, reg2);
}

void cvdvPhaseSquare(CVDVContext* ctx, int regIdx, double t) {
    // gates.cu:288-300: e^{i t x^2}
    if (!ready(ctx) || !reg_ok(ctx, regIdx)) return;
    CVDV_TRY phase_poly(ctx, regIdx, 0, t, 0); CVDV_CATCH()
}

void cvdvPhaseCubic(CVDVContext* ctx, int regIdx, double t) {
    // gates.cu:302-314: e^{i t x^3}
    if (!ready(ctx) || !reg_ok(ctx, regIdx)) return;
    CVDV_TRY phase_poly(ctx, regIdx, 0, 0, t); CVDV_CATCH()
}

void cvdvRotation(CVDVContext* ctx, int regIdx, double theta) {
    // gates.cu:316-348: quarter turns by rint (half-even), then the 3-shear remainder
    if (!ready(ctx) || !reg_ok(ctx, regIdx)) return;
    CVDV_TRY
    const double theta0 = std::rint(theta / (PI / 2.0)) * (PI / 2.0);
    const double rem = theta - theta0;
    int qt = (int)std::rint(theta0 / (PI / 2.0));
    qt = ((qt % 4) + 4) % 4;
    if (qt == 1) ctx_ft(ctx, regIdx, false);
    else if (qt == 2) parity(ctx, regIdx);
    else if (qt == 3) ctx_ft(ctx, regIdx, true);
    rotation_small(ctx, regIdx, rem);
    CVDV_CATCH()
}

void cvdvConditionalRotation(CVDVContext* ctx, int targetReg, int ctrlReg, int ctrlQubit,
                             double theta) {
    // gates.cu:350-392
    if (!ready(ctx) || !reg_ok(ctx, targetReg, "Invalid target register index") ||
        !ctrl_ok(ctx, ctrlReg, ctrlQubit))
        return;
    CVDV_TRY
    const double theta0 = std::rint(theta / (PI / 2.0)) * (PI / 2.0);
    const double rem = theta - theta0;
    int qt = (int)std::rint(theta0 / (PI / 2.0));
    qt = ((qt % 4) + 4) % 4;
    if (qt == 1) {
        ctx_ft(ctx, targetReg, false);
        cparity(ctx, targetReg, ctrlReg, ctrlQubit);
        zbit(ctx, ctrlReg, ctrlQubit, PI / 2);
    } else if (qt == 2) {
        parity(ctx, targetReg);
        zbit(ctx, ctrlReg, ctrlQubit, PI);
    } else if (qt == 3) {
        ctx_ft(ctx, targetReg, true);
        cparity(ctx, targetReg, ctrlReg, ctrlQubit);
        zbit(ctx, ctrlReg, ctrlQubit, -PI / 2);
    }
    crotation_small(ctx, targetReg, ctrlReg, ctrlQubit, rem);
    CVDV_CATCH()
}

void cvdvSqueeze(CVDVContext* ctx, int regIdx, double r) {
    // gates.cu:394-421
    if (!ready(ctx) || !reg_ok(ctx, regIdx)) return;
    CVDV_TRY
    const double eh = std::exp(0.5 * r), emh = std::exp(-0.5 * r);
    const double sq = std::sqrt(std::fabs(std::exp(r) - 1.0));
    const double sign = (r >= 0) ? 1.0 : -1.0;
    ctx_ft(ctx, regIdx, false);
    phase_poly(ctx, regIdx, 0, -0.5 * eh * sq, 0);
    ctx_ft(ctx, regIdx, true);
    phase_poly(ctx, regIdx, 0, 0.5 * emh * sq * sign, 0);
    ctx_ft(ctx, regIdx, false);
    phase_poly(ctx, regIdx, 0, 0.5 * emh * sq, 0);
    ctx_ft(ctx, regIdx, true);
    phase_poly(ctx, regIdx, 0, -0.5 * eh * sq * sign, 0);
    CVDV_CATCH()
}

void cvdvConditionalSqueeze(CVDVContext* ctx, int targetReg, int ctrlReg, int ctrlQubit, double r) {
    // gates.cu:423-459 (r = 0 gives NaN, as in the reference)
    if (!ready(ctx) || !reg_ok(ctx, targetReg, "Invalid target register index") ||
        !ctrl_ok(ctx, ctrlReg, ctrlQubit))
        return;
    CVDV_TRY
    const double ch = std::cosh(r), sh = std::sinh(r);
    const double s = std::sqrt(2.0 * std::fabs(std::sinh(0.5 * r)));
    phase_poly(ctx, targetReg, 0, 0.5 * s * ch, 0);
    phase_poly(ctx, targetReg, 0, -0.5 * s * sh, 0, ctrlReg, ctrlQubit);
    ctx_ft(ctx, targetReg, false);
    phase_poly(ctx, targetReg, 0, 0.5 * (ch - 1) / s, 0);
    phase_poly(ctx, targetReg, 0, 0.5 * sh / s, 0, ctrlReg, ctrlQubit);
    ctx_ft(ctx, targetReg, true);
    phase_poly(ctx, targetReg, 0, -0.5 * s, 0);
    ctx_ft(ctx, targetReg, false);
    phase_poly(ctx, targetReg, 0, 0.5 * (ch - 1) / s, 0);
    phase_poly(ctx, targetReg, 0, -0.5 * sh / s, 0, ctrlReg, ctrlQubit);
    ctx_ft(ctx, targetReg, true);
    CVDV_CATCH()
}

void cvdvBeamSplitter(CVDVContext* ctx, int reg1, int reg2, double theta) {
    // gates.cu:461-502: half turns by round (half away from zero)
    if (!ready(ctx) || !two_regs_ok(ctx, reg1, reg2)) return;
    if (reg1 == reg2) {
        cvdv_set_error(CVDV_ERR_ARG, "Beam splitter requires two different registers");
        return;
    }
    CVDV_TRY
    const double theta0 = std::round(theta / PI) * PI;
    const double rem = theta - theta0;
    int ht = (int)std::round(theta0 / PI);
    ht = ((ht % 4) + 4) % 4;
    if ((ht == 1 || ht == 3) && ctx->nq[reg1] != ctx->nq[reg2]) {
        cvdv_set_error(CVDV_ERR_ARG, "SWAP requires registers with same number of qubits");
        return;
    }
    if (ht == 1) {
        ctx_ft(ctx, reg1, false);
        ctx_ft(ctx, reg2, false);
        swap_regs(ctx, reg1, reg2);
    } else if (ht == 2) {
        parity(ctx, reg1);
        parity(ctx, reg2);
    } else if (ht == 3) {
        ctx_ft(ctx, reg1, true);
        ctx_ft(ctx, reg2, true);
        swap_regs(ctx, reg1, reg2);
    }
    bs_small(ctx, reg1, reg2, rem);
    CVDV_CATCH()
}

void cvdvConditionalBeamSplitter(CVDVContext* ctx, int reg1, int reg2, int ctrlReg, int ctrlQubit,
                                 double theta) {
    // gates.cu:504-547
    if (!ready(ctx) || !two_regs_ok(ctx, reg1, reg2) || !ctrl_ok(ctx, ctrlReg, ctrlQubit)) return;
    if (reg1 == reg2) {
        cvdv_set_error(CVDV_ERR_ARG, "Conditional beam splitter requires two different registers");
        return;
    }
    CVDV_TRY
    const double theta0 = std::round(theta / PI) * PI;
    const double rem = theta - theta0;
    int ht = (int)std::round(theta0 / PI);
    ht = ((ht % 4) + 4) % 4;
    if ((ht == 1 || ht == 3) && ctx->nq[reg1] != ctx->nq[reg2]) {
        cvdv_set_error(CVDV_ERR_ARG, "SWAP requires registers with same number of qubits");
        return;
    }
    if (ht == 1 || ht == 3) {
        ctx_ft(ctx, reg1, ht == 3);
        ctx_ft(ctx, reg2, ht == 3);
        cparity(ctx, reg1, ctrlReg, ctrlQubit);
        cparity(ctx, reg2, ctrlReg, ctrlQubit);
        swap_regs(ctx, reg1, reg2);
    } else if (ht == 2) {
        parity(ctx, reg1);
        parity(ctx, reg2);
    }
    bs_small(ctx, reg1, reg2, rem, ctrlReg, ctrlQubit);
    CVDV_CATCH()
}

void cvdvQ1Q2Gate(CVDVContext* ctx, int reg1, int reg2, double coeff) {
    // gates.cu:549-565: e^{i coeff q1 q2}
    if (!ready(ctx) || !two_regs_ok(ctx, reg1, reg2)) return;
    if (reg1 == reg2) {
        cvdv_set_error(CVDV_ERR_ARG, "Cannot apply Q1Q2 gate to the same register");
        return;
    }
    CVDV_TRY phase_bilin(ctx, reg1, reg2, coeff); CVDV_CATCH()
}

}  // extern "C"
