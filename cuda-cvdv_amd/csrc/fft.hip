// fft.hip - register-axis FFT for gfx950 (replaces the cuFFT plans and the
// outer-block exec loop of src/cuda/context.cu:48-65 and gates.cu:37-52).
//
// A register field [sh, sh+n) of the flat index splits every element index as
//     g = hi * (N*S) + j * S + lo,   S = 2^sh, N = 2^n,
// and the FFT runs over j for every (hi, lo).  A workgroup stages a tile of
// C consecutive lo columns x M rows x Hb consecutive hi blocks in LDS, runs a
// Stockham FFT there (radix 8/4/2 butterflies in registers, per-stage twiddle
// tables generated in FP64), and writes the tile back:
//   * one HBM pass when the tile fits in LDS: C*M elements with C = 32 (c64)
//     or 16 (c128) columns (256 bytes per row) for a strided register, or a
//     whole contiguous register of up to 2^14 (c64) / 2^13 (c128) points in one
//     workgroup (persistent and prefetching for c64 rows of 2^13 and 2^14);
//   * otherwise the four-step split N = N1 * N2 (j = j1 + N1 j2):
//       pass 1 (in place):  FFT over j2;
//       pass 2 (out of place, transposing): times W_N^{j1 k2} on load, FFT
//       over j1, written at k2 + N2 k1.
// Pending diagonal phases are applied (compiled, see diag_fast.h) as elements
// are loaded by the first pass and 1/sqrt(N) as they are stored by the last.
// Every tile dimension is a power of two: indices use shifts and masks only.

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "cvdv_internal.h"
#include "diag_fast.h"
#include "fft_bfly.h"

static const long double PI_L = 3.141592653589793238462643383279502884L;

// ---------------------------------------------------------------- twiddles
struct TwKey {
    int logM;
    int prec;
    bool operator<(const TwKey& o) const { return logM != o.logM ? logM < o.logM : prec < o.prec; }
};

struct TwiddleCache {
    std::map<TwKey, void*> stage;  // per-stage Stockham tables for length M
    std::map<TwKey, void*> full;   // W_N^e, e in [0, N)
};

void twiddle_cache_free(TwiddleCache* tc) {
    for (auto& kv : tc->stage) (void)hipFree(kv.second);
    for (auto& kv : tc->full) (void)hipFree(kv.second);
    delete tc;
}

// exp(-2 pi i e / N) from an 80-bit evaluation, rounded once to double.
static void host_twiddle(uint64_t e, uint64_t N, double& c, double& s) {
    const long double a = 2.0L * PI_L * (long double)e / (long double)N;
    c = (double)cosl(a);
    s = (double)-sinl(a);
}

struct StagePlan {
    int nst = 0;
    int rad[16];
    int L[16];
    int off[16];
    int total = 0;  // table entries
};

static StagePlan stage_plan(int logM) {
    StagePlan p;
    int rem = logM, L = 1, off = 0;
    while (rem > 0) {
        const int r = rem >= 3 ? 3 : rem;  // radix 8, then a final 4 or 2
        const int R = 1 << r;
        p.rad[p.nst] = R;
        p.L[p.nst] = L;
        p.off[p.nst] = off;
        off += L * (R - 1);
        L *= R;
        rem -= r;
        p.nst++;
    }
    p.total = off;
    return p;
}

// FP32 twiddles: every entry is one of the floats adjacent to (or equal to) the
// exact value, chosen by error diffusion so that sum(|w|^2 - 1) over each
// stage stays near zero.  Rounding each entry to nearest leaves a mean
// |w|^2 - 1 of order -1e-8, which shrinks the norm by ~2e-8 per FFT.
static void round_unit_diffused(const double* re, const double* im, size_t n, float2* out) {
    double acc = 0.0;
    for (size_t i = 0; i < n; ++i) {
        const float c0 = (float)re[i], s0 = (float)im[i];
        const float cs[3] = {nextafterf(c0, -2.f), c0, nextafterf(c0, 2.f)};
        const float ss[3] = {nextafterf(s0, -2.f), s0, nextafterf(s0, 2.f)};
        double bestErr = 0.0, bestScore = 1e300;
        float bc = c0, bs = s0;
        for (int a = 0; a < 3; ++a) {
            if (std::fabs((double)cs[a] - re[i]) > std::fabs((double)nextafterf(c0, 2.f) - (double)c0)) continue;
            for (int b = 0; b < 3; ++b) {
                if (std::fabs((double)ss[b] - im[i]) > std::fabs((double)nextafterf(s0, 2.f) - (double)s0)) continue;
                const double err = (double)cs[a] * cs[a] + (double)ss[b] * ss[b] - 1.0;
                const double score = std::fabs(acc + err);
                if (score < bestScore) {
                    bestScore = score;
                    bestErr = err;
                    bc = cs[a];
                    bs = ss[b];
                }
            }
        }
        acc += bestErr;
        out[i] = make_float2(bc, bs);
    }
}

// Host copy of a table in the state's precision (FP32: error-diffused per segment).
static std::vector<uint8_t> round_table(const std::vector<double>& re, const std::vector<double>& im, int prec,
                                        const std::vector<size_t>& segments = {}) {
    const size_t n = re.size();
    std::vector<uint8_t> bytes;
    if (prec == CVDV_C128) {
        bytes.resize(n * sizeof(double2));
        double2* h = (double2*)bytes.data();
        for (size_t i = 0; i < n; ++i) h[i] = make_double2(re[i], im[i]);
    } else {
        bytes.resize(n * sizeof(float2));
        float2* h = (float2*)bytes.data();
        size_t start = 0;
        std::vector<size_t> seg = segments;
        seg.push_back(n);
        for (size_t end : seg) {
            if (end > start) round_unit_diffused(re.data() + start, im.data() + start, end - start, h + start);
            start = end;
        }
    }
    return bytes;
}

static void* upload_table(const std::vector<double>& re, const std::vector<double>& im, int prec,
                          const std::vector<size_t>& segments = {}) {
    const std::vector<uint8_t> h = round_table(re, im, prec, segments);
    void* d = nullptr;
    HIPCHK(hipMalloc(&d, h.size() ? h.size() : 16));
    if (!h.empty()) HIPCHK(hipMemcpy(d, h.data(), h.size(), hipMemcpyHostToDevice));
    return d;
}

static TwiddleCache* cache(CVDVContext* ctx) {
    if (!ctx->tw) ctx->tw = new TwiddleCache();
    return ctx->tw;
}

// Stockham stage twiddles for length 2^logM: stage s (radix R, sub-length L)
// holds W_{RL}^{ij} at off[s] + j (R-1) + i - 1, j < L, 1 <= i < R.
static void stage_values(int logM, std::vector<double>& re, std::vector<double>& im, std::vector<size_t>& segs) {
    const StagePlan p = stage_plan(logM);
    const uint64_t M = 1ull << logM;
    re.assign(p.total, 0.0);
    im.assign(p.total, 0.0);
    for (int s = 0; s < p.nst; ++s) {
        const int R = p.rad[s], L = p.L[s];
        for (int j = 0; j < L; ++j)
            for (int i = 1; i < R; ++i) {
                // W_{R L}^{i j} = W_M^{i j M / (R L)}
                const uint64_t e = (uint64_t)i * j * (M / ((uint64_t)R * L));
                host_twiddle(e, M, re[p.off[s] + j * (R - 1) + i - 1], im[p.off[s] + j * (R - 1) + i - 1]);
            }
    }
    segs.clear();
    for (int st = 1; st < p.nst; ++st) segs.push_back((size_t)p.off[st]);
}

std::vector<uint8_t> fft_stage_host(int logM, int prec) {
    std::vector<double> re, im;
    std::vector<size_t> segs;
    stage_values(logM, re, im, segs);
    return round_table(re, im, prec, segs);
}

const void* fft_stage_table(TwiddleCache*& tcp, int logM, int prec) {
    if (!tcp) tcp = new TwiddleCache();
    TwiddleCache* tc = tcp;
    auto it = tc->stage.find({logM, prec});
    if (it != tc->stage.end()) return it->second;
    std::vector<double> re, im;
    std::vector<size_t> segs;
    stage_values(logM, re, im, segs);
    void* d = upload_table(re, im, prec, segs);
    tc->stage[{logM, prec}] = d;
    return d;
}

// Split twiddles for the one-row tiles, which keep all twiddles in LDS.  Stage s
// (radix R, sub-length L): L <= 64 stores W_{RL}^{p i} directly (p < L); larger L
// stores W_{RL/64}^{a i} (a < L/64) and W_{RL}^{b i} (b < 64), and
// W_{RL}^{p i} = W_{RL/64}^{(p>>6) i} W_{RL}^{(p&63) i}.  Layout per stage: the
// direct segment, or the a segment then the b segment, each i-major (p fastest,
// so the lanes of a wave read consecutive LDS words).
static void split_values(int logM, std::vector<double>& re, std::vector<double>& im, std::vector<size_t>& segs) {
    const StagePlan p = stage_plan(logM);
    re.clear();
    im.clear();
    segs.clear();
    auto put = [&](uint64_t e, uint64_t M) {
        double c, s;
        host_twiddle(e, M, c, s);
        re.push_back(c);
        im.push_back(s);
    };
    for (int st = 1; st < p.nst; ++st) {
        const uint64_t R = (uint64_t)p.rad[st], L = (uint64_t)p.L[st], RL = R * L;
        if (L <= 64) {
            segs.push_back(re.size());
            for (uint64_t i = 1; i < R; ++i)
                for (uint64_t j = 0; j < L; ++j) put(i * j, RL);
        } else {
            segs.push_back(re.size());
            for (uint64_t i = 1; i < R; ++i)
                for (uint64_t a = 0; a < L / 64; ++a) put(i * a, RL / 64);
            segs.push_back(re.size());
            for (uint64_t i = 1; i < R; ++i)
                for (uint64_t b = 0; b < 64; ++b) put(i * b, RL);
        }
    }
    if (!segs.empty() && segs[0] == 0) segs.erase(segs.begin());
}

static const void* split_table(CVDVContext* ctx, int logM, int prec) {
    TwiddleCache* tc = cache(ctx);
    const TwKey key{logM + 1000, prec};
    auto it = tc->stage.find(key);
    if (it != tc->stage.end()) return it->second;
    if (ctx->capturing) throw CvdvError{CVDV_ERR_UNSUPPORTED, "FFT table not prepared before the graph capture"};
    std::vector<double> re, im;
    std::vector<size_t> segs;
    split_values(logM, re, im, segs);
    void* d = upload_table(re, im, prec, segs);
    tc->stage[key] = d;
    return d;
}

static void* stage_table(CVDVContext* ctx, int logM, int prec) {
    if (ctx->capturing && (!ctx->tw || !ctx->tw->stage.count({logM, prec})))
        throw CvdvError{CVDV_ERR_UNSUPPORTED, "FFT table not prepared before the graph capture"};
    return (void*)fft_stage_table(ctx->tw, logM, prec);
}

// W_N^e for e in [0, N) in the given precision (FP32: error-diffused rounding,
// as for the stage tables).  The direct DFT always uses the FP64 table.
static const void* full_table(CVDVContext* ctx, int logN, int prec = CVDV_C128) {
    TwiddleCache* tc = cache(ctx);
    auto it = tc->full.find({logN, prec});
    if (it != tc->full.end()) return it->second;
    if (ctx->capturing) throw CvdvError{CVDV_ERR_UNSUPPORTED, "FFT table not prepared before the graph capture"};
    const uint64_t N = 1ull << logN;
    std::vector<double> re(N), im(N);
    for (uint64_t e = 0; e < N; ++e) host_twiddle(e, N, re[e], im[e]);
    void* d = upload_table(re, im, prec);
    tc->full[{logN, prec}] = d;
    return d;
}

// Device properties the planner asks for on every launch, cached per device.
// Contexts may run on several host threads: the caches are guarded.
static std::mutex g_devCacheMu;

size_t device_max_lds() {
    static std::map<int, size_t> cached;
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_devCacheMu);
    auto it = cached.find(dev);
    if (it != cached.end()) return it->second;
    int v = 0;
    HIPCHK(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    cached[dev] = (size_t)v;
    return (size_t)v;
}

// ---------------------------------------------------------------- pass descriptor
struct FftPass {
    int lgM, lgC, lgHb;            // tile = [Hb][M][C] (lgM also selects the kernel instance)
    int lgNLoBlk, lgNH2Blk;        // tiles = nH1 * nH2Blk * nLoBlk
    uint64_t nTiles;
    int rowFastOut;                // linear order of the staged store after the column index
    int vecOut;                    // 1: pairs of complex64 stored as one 16-byte access
    int directOut;                 // 1: last stage stores straight to global (coalesced rows)
    int tw4, tw4Shift, tw4HiBits, tw4LogN;
    int tw4Uni;                    // 1: the twiddle's second index is the same over the tile
    int usePre, useScale;
    int conjIn, conjOut;           // inverse = conj . forward . conj, split over the passes
    int conjLoad;                  // conjugate as loaded, before the four-step twiddle (the
                                   // second half of an inverse whose first half ran alone)
    int lgInH1, lgInH2, lgInR;     // every stride is a power of two
    int lgOutH1, lgOutH2, lgOutR;
    double outScale;  // applied as exact v*hi + rounded v*lo (one FMA): an FP32 store stays unbiased
    int genSteps;     // 1: every generic term is affine in the tile's row index (the host checked)
};

// The value lane `lane` of the wave holds (`lane` the same on every lane, and
// that lane active).
__device__ __forceinline__ float2 lane_value(float2 v, int lane) {
    return make_float2(__int_as_float(__builtin_amdgcn_readlane(__float_as_int(v.x), lane)),
                       __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v.y), lane)));
}
__device__ __forceinline__ double2 lane_value(double2 v, int lane) {
    auto rd = [lane](double d) {
        const long long b = __double_as_longlong(d);
        const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, lane);
        const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), lane);
        return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
    };
    return make_double2(rd(v.x), rd(v.y));
}

// LDS layouts of the Stockham exchanges (padding breaks power-of-two strides).
// LAY 0, every exchange of every tile unless below: one pad entry per 32.
// One-row complex64 tiles of 2^13-2^14 points write exchange 0 at a stride of 8
// entries and exchange 1 in runs of 8 lanes 64 entries apart, which puts two
// lanes of every 16 on one LDS bank under LAY 0: there exchange 0 takes one pad
// entry per 16 (LAY 1) and exchange 1 eight per 64 (LAY 2).
template <int LAY>
__device__ __forceinline__ int lay_ix(int i) {
    return LAY == 0 ? i + (i >> 5) : (LAY == 1 ? i + (i >> 4) : i + ((i >> 6) << 3));
}
template <int LAY> constexpr int lay_period() { return LAY == 0 ? 32 : (LAY == 1 ? 16 : 64); }
template <int LAY> constexpr int lay_padded(int stride) {
    return stride + (LAY == 0 ? stride / 32 : (LAY == 1 ? stride / 16 : stride / 8));
}
// The layout of exchange XS (stage XS's outputs, read by stage XS + 1).
template <bool ONE, int XS> constexpr int ex_lay() { return !ONE ? 0 : (XS == 0 ? 1 : (XS == 1 ? 2 : 0)); }

__device__ __forceinline__ int lds_ix(int i) { return lay_ix<0>(i); }

// Padded index of i0 + m * STRIDE.  When STRIDE is a multiple of the layout's
// period the padding is linear (adding it leaves the bits below the period
// alone), so every element is the padded base plus a compile-time offset (an
// immediate in the LDS access).
template <int STRIDE, int LAY = 0>
__device__ __forceinline__ int lds_at(int i0, int pbase, int m) {
    if constexpr (STRIDE % lay_period<LAY>() == 0) return pbase + m * lay_padded<LAY>(STRIDE);
    else return lay_ix<LAY>(i0 + m * STRIDE);
}

// Stage plan of a length-2^LGM transform: radix-8 stages, then one radix-4 or
// radix-2 stage.  Stage s has sub-length L = 8^s and its twiddles start at
// entry off(s) of the stage table (W_{RL}^{ij} at off + j (R-1) + i - 1), the
// layout stage_values() writes.
template <int LGM> struct Plan {
    static constexpr int NST = LGM / 3 + (LGM % 3 ? 1 : 0);
    static constexpr int lgr(int s) { return s < LGM / 3 ? 3 : LGM % 3; }
    static constexpr int off(int s) {
        int o = 0, L = 1;
        for (int i = 0; i < s; ++i) {
            o += L * ((1 << lgr(i)) - 1);
            L <<= lgr(i);
        }
        return o;
    }
};

// Offsets of the split twiddle table (split_values) for stage s of a length-2^LGM
// transform: the direct segment (L <= 64) or the a and b segments.
template <int LGM> struct SplitPlan {
    static constexpr int seg(int s) {  // entries of stage s
        const int R = 1 << Plan<LGM>::lgr(s), L = 1 << (3 * s);
        return s == 0 ? 0 : (L <= 64 ? L * (R - 1) : (L / 64) * (R - 1) + 64 * (R - 1));
    }
    static constexpr int off(int s) {
        int o = 0;
        for (int i = 0; i < s; ++i) o += seg(i);
        return o;
    }
    static constexpr int total() { return off(Plan<LGM>::NST); }
};

// ---------------------------------------------------------------- butterflies (fft_bfly.h)

// Tile element e (linear) -> (c, row, hb) in the given order.
template <int LGM>
__device__ __forceinline__ void tile_coords(int e, int lgC, int lgHb, int rowFast, int& c, int& row, int& hb) {
    c = e & ((1 << lgC) - 1);
    const int rest = e >> lgC;
    const int lgLo = rowFast ? LGM : lgHb;  // selects, not branches on references (no scratch)
    const int lo = rest & ((1 << lgLo) - 1), hi = rest >> lgLo;
    row = rowFast ? lo : hi;
    hb = rowFast ? hi : lo;
}

// Stockham stage geometry.  Butterfly b of a stage of radix RP (sub-length
// L = 2^lgL already transformed) takes rows q + i*M/RP of column c in block hb
// and writes rows ob + m*L, ob = hb*M + (q/L)*RP*L + (q mod L); every thread
// owns E/RP butterflies b = tid + k*nt.
struct BflyPos {
    int c, q, hb;
};

template <int RP, int LGM>
__device__ __forceinline__ BflyPos bfly_pos(int b, int lgC) {
    constexpr int LGR = RP == 8 ? 3 : (RP == 4 ? 2 : 1);
    constexpr int lgMq = LGM - LGR;
    BflyPos p;
    p.c = b & ((1 << lgC) - 1);
    const int rest = b >> lgC;
    p.q = rest & ((1 << lgMq) - 1);
    p.hb = rest >> lgMq;
    return p;
}

// Twiddle the inputs of the butterflies in x (stage S, radix RP) and transform.
// TWL: the twiddles are the split table staged in LDS (one-row tiles), else the
// stage table in global memory.
template <int RP, int E, int LGM, int LGC, int S, bool TWL, typename C>
__device__ __forceinline__ void stage_compute(C* x, const C* __restrict__ tw, int tid, int nt) {
    constexpr int lgC = LGC;
    constexpr int NB = E / RP;
    constexpr int lgL = 3 * S;
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        if (lgL > 0) {
            const BflyPos p = bfly_pos<RP, LGM>(tid + k * nt, lgC);
            const int pq = p.q & ((1 << lgL) - 1);
            if constexpr (!TWL) {
                const C* t = tw + Plan<LGM>::off(S) + pq * (RP - 1);
#pragma unroll
                for (int i = 1; i < RP; ++i) x[k * RP + i] = cmul(x[k * RP + i], t[i - 1]);
            } else if constexpr (lgL <= 6) {
                const C* t = tw + SplitPlan<LGM>::off(S) + pq;
#pragma unroll
                for (int i = 1; i < RP; ++i) x[k * RP + i] = cmul(x[k * RP + i], t[(i - 1) << lgL]);
            } else {
                constexpr int LA = (1 << lgL) / 64;
                constexpr int OA = SplitPlan<LGM>::off(S);
                constexpr int OB = OA + LA * (RP - 1);
                const C* ta = tw + OA + (pq >> 6);
                const C* tb = tw + OB + (pq & 63);
#pragma unroll
                for (int i = 1; i < RP; ++i)
                    x[k * RP + i] = cmul(x[k * RP + i], cmul(ta[(i - 1) * LA], tb[(i - 1) * 64]));
            }
        }
        dft_r<RP, false>(x + k * RP);
    }
}

// Outputs of stage S (radix RP) -> LDS in Stockham order (exchange S; ONE: the
// one-row layouts).
template <int RP, int E, int LGM, int LGC, int S, bool ONE = false, typename C>
__device__ __forceinline__ void stage_write(C* lds, const C* x, int tid, int nt) {
    constexpr int LAY = ex_lay<ONE, S>();
    constexpr int NB = E / RP;
    constexpr int LGR = RP == 8 ? 3 : (RP == 4 ? 2 : 1);
    constexpr int lgL = 3 * S;
    constexpr int STRIDE = 1 << (lgL + LGC);
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        const BflyPos p = bfly_pos<RP, LGM>(tid + k * nt, LGC);
        const int ob = (p.hb << LGM) + ((p.q >> lgL) << (LGR + lgL)) + (p.q & ((1 << lgL) - 1));
        const int i0 = (ob << LGC) + p.c, pb = lay_ix<LAY>(i0);
        if constexpr (LAY != 0) {
            // one-row exchanges 0 and 1: i0 = 8q with entries i0 + m (period 16),
            // or i0 mod 64 < 8 with entries i0 + 8m (period 64): all within i0's period
#pragma unroll
            for (int m = 0; m < RP; ++m) lds[pb + m * STRIDE] = x[k * RP + m];
        } else {
#pragma unroll
            for (int m = 0; m < RP; ++m) lds[lds_at<STRIDE, LAY>(i0, pb, m)] = x[k * RP + m];
        }
    }
}

// Inputs of a stage of radix RP <- LDS (exchange XS).
template <int RP, int E, int LGM, int LGC, int XS = 0, bool ONE = false, typename C>
__device__ __forceinline__ void stage_read(const C* lds, C* x, int tid, int nt) {
    constexpr int LAY = ex_lay<ONE, XS>();
    constexpr int NB = E / RP;
    constexpr int LGR = RP == 8 ? 3 : (RP == 4 ? 2 : 1);
    constexpr int lgMq = LGM - LGR;
    constexpr int STRIDE = 1 << (lgMq + LGC);
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        const BflyPos p = bfly_pos<RP, LGM>(tid + k * nt, LGC);
        const int i0 = (((p.hb << LGM) + p.q) << LGC) + p.c, pb = lay_ix<LAY>(i0);
#pragma unroll
        for (int i = 0; i < RP; ++i) x[k * RP + i] = lds[lds_at<STRIDE, LAY>(i0, pb, i)];
    }
}

// The compiled pending program, staged in LDS so its runtime-length loops read
// LDS instead of forcing a scratch copy of the kernel argument.
struct PreLds {
    DiagTab tab[CVDV_FAST_TABS];
    DiagTerm gen[CVDV_FAST_GEN];
};

// The generic terms over a one-row tile (only the transformed field
// [fs, fs + fn) varies inside it): when every term is affine in that field's
// coordinate m, their sum is a·m + b (mod 1) with per-tile a, b, and one FP64
// fma per element replaces the per-term evaluation.  a is reduced mod 1 (m is
// an integer), so a·m stays below 2^(fn-1) and keeps ~2^-40 absolute accuracy.
struct GenFold {
    double a, b;
    int ok;
};

__device__ __forceinline__ bool field_overlaps(int sh, int nb, int fs, int fn) {
    return sh < fs + fn && fs < sh + nb;
}

__device__ __forceinline__ GenFold fold_generic(const PreLds& P, const FastDiag& pre, uint64_t g0, int fs, int fn) {
    GenFold f{0.0, pre.genrev, 1};
    for (int i = 0; i < pre.ngen; ++i) {
        const DiagTerm& T = P.gen[i];
        if (T.ctrl >= 0 && T.ctrl >= fs && T.ctrl < fs + fn) f.ok = 0;
        const double sg = (T.ctrl >= 0 && ((g0 >> T.ctrl) & 1ull)) ? -1.0 : 1.0;
        const bool in1 = T.sh1 == fs && T.nb1 == fn;
        if (T.kind == DT_BILIN) {
            const bool in2 = T.sh2 == fs && T.nb2 == fn;
            if ((!in1 && field_overlaps(T.sh1, T.nb1, fs, fn)) || (!in2 && field_overlaps(T.sh2, T.nb2, fs, fn)) ||
                (in1 && in2))
                f.ok = 0;
            else if (in1)
                f.a += cvdv_frac(sg * T.c1 * diag_field_m(g0, T.sh2, T.nb2));
            else if (in2)
                f.a += cvdv_frac(sg * T.c1 * diag_field_m(g0, T.sh1, T.nb1));
            else
                f.b += cvdv_frac(sg * diag_term_rev(T, g0));
        } else if (T.kind == DT_POLY || T.kind == DT_LINF) {
            if (in1 && T.c2 == 0.0 && T.c3 == 0.0)
                f.a += cvdv_frac(sg * T.c1);
            else if (field_overlaps(T.sh1, T.nb1, fs, fn))
                f.ok = 0;
            else
                f.b += cvdv_frac(sg * diag_term_rev(T, g0));
        } else {
            if (T.sh1 >= fs && T.sh1 < fs + fn) f.ok = 0;
            else f.b += cvdv_frac(sg * diag_term_rev(T, g0));
        }
    }
    f.a = cvdv_frac(f.a);
    f.b = cvdv_frac(f.b);
    return f;
}

// Phase of the generic (untabulated) terms at element g.
template <typename C>
__device__ __forceinline__ C gen_factor(const PreLds& P, const FastDiag& pre, uint64_t g) {
    using R = decltype(C{}.x);
    double r = pre.genrev;
    for (int i = 0; i < pre.ngen; ++i) r += cvdv_frac(diag_term_rev(P.gen[i], g));
    return unit_rev_fast(cvdv_frac(r), R(0));
}

// Stage-0 inputs straight from global memory (issued early: a tile ahead).
template <int RP, int E, int LGM, int LGC, typename C>
__device__ __forceinline__ void stage0_load(const C* __restrict__ in, C* x, const FftPass& d, uint64_t inBase,
                                            int tid, int nt) {
    constexpr int NB = E / RP;
    constexpr int LGR = RP == 8 ? 3 : (RP == 4 ? 2 : 1);
    constexpr int lgMq = LGM - LGR;
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        const BflyPos p = bfly_pos<RP, LGM>(tid + k * nt, LGC);
        const uint64_t g0 = inBase + ((uint64_t)p.hb << d.lgInH2) + ((uint64_t)p.q << d.lgInR) + p.c;
#pragma unroll
        for (int i = 0; i < RP; ++i) x[k * RP + i] = in[g0 + ((uint64_t)(i << lgMq) << d.lgInR)];
    }
}

// Pending phases, the four-step twiddle W_N^{j1 k2} of a second pass (row j1 =
// the transformed index, k2 = (g >> tw4Shift) mod 2^tw4HiBits; TW4: the
// instance may carry one) and the input conjugation of an inverse on the loaded
// inputs.  TM, how the phase tables are read:
//   TM_ROWS   every table is read by row (the host checked, table_by_row()), one
//             load per element at a row offset (one-row tiles);
//   TM_ANY    tables of several columns stage their product by row in LDS when
//             every table allows it, else each element takes its flat index;
//   TM_STAGED every table is read by row (the host checked, tile_tables_by_row())
//             and is staged by row; generic terms go through the free exchange
//             buffer (complex64: their row steps per column and row block;
//             complex128: each element's factor from a rolled loop), which
//             keeps their FP64 temporaries off the other elements' registers.
// TM_NONE: the pass applies no phases (only the four-step twiddle, if TW4).
enum { TM_ROWS = 0, TM_ANY = 1, TM_STAGED = 2, TM_NONE = 3 };
template <int RP, int E, int LGM, int LGC, bool TW4, int TM, typename C>
__device__ __forceinline__ void stage0_pre(C* x, const FftPass& d, uint64_t inBase, const PreLds& P,
                                           const FastDiag& pre, const C* __restrict__ tw4, int tid, int nt,
                                           C* sRow) {
    constexpr int NB = E / RP;
    constexpr int LGR = RP == 8 ? 3 : (RP == 4 ? 2 : 1);
    constexpr int lgMq = LGM - LGR;
    if (TW4 && d.tw4) {
        // second four-step pass: straight-line, so these loads issue with the
        // tile's input loads and share their wait
        if (d.tw4Uni) {
            // k2 is the tile's (no tile index bit selects it), so W^{k2 (q + i M/RP)} =
            // W^{k2 q} W^{k2 i M/RP} takes one gathered and RP-1 tile-uniform
            // (scalar) loads per butterfly instead of RP gathers
            const uint32_t nmask = (1u << d.tw4LogN) - 1u;
            const uint32_t k2 = (uint32_t)(inBase >> d.tw4Shift) & ((1u << d.tw4HiBits) - 1u);
            C st[RP];
#pragma unroll
            for (int i = 1; i < RP; ++i) st[i] = tw4[(k2 * (uint32_t)(i << lgMq)) & nmask];
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                const BflyPos p = bfly_pos<RP, LGM>(tid + k * nt, LGC);
                const C b = tw4[(k2 * (uint32_t)p.q) & nmask];
                x[k * RP] = cmul(x[k * RP], b);
#pragma unroll
                for (int i = 1; i < RP; ++i) x[k * RP + i] = cmul(x[k * RP + i], cmul(b, st[i]));
            }
        } else {
            // k2 differs within the tile (between row blocks or columns): per
            // butterfly, gather W^{k2 q} and
            // w = W^{k2 M/RP}; the powers w^i come from a multiplication tree of
            // depth log2(RP)
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                const BflyPos p = bfly_pos<RP, LGM>(tid + k * nt, LGC);
                const uint64_t g0 = inBase + ((uint64_t)p.hb << d.lgInH2) + ((uint64_t)p.q << d.lgInR) + p.c;
                const uint32_t nmask = (1u << d.tw4LogN) - 1u;
                const uint32_t k2 = (uint32_t)(g0 >> d.tw4Shift) & ((1u << d.tw4HiBits) - 1u);
                const C b = tw4[(k2 * (uint32_t)p.q) & nmask];
                C st[RP];
                st[1] = tw4[(k2 << lgMq) & nmask];
#pragma unroll
                for (int i = 2; i < RP; ++i) st[i] = cmul(st[i / 2], st[i - i / 2]);
                x[k * RP] = cmul(x[k * RP], b);
#pragma unroll
                for (int i = 1; i < RP; ++i) x[k * RP + i] = cmul(x[k * RP + i], cmul(b, st[i]));
            }
        }
    }
    if (TM != TM_NONE && d.usePre) {
        int tt = tid;
        auto elem = [&](int k, int i) {
            const BflyPos p = bfly_pos<RP, LGM>(tt + k * nt, LGC);
            return inBase + ((uint64_t)p.hb << d.lgInH2) + ((uint64_t)p.q << d.lgInR) + p.c +
                   ((uint64_t)(i << lgMq) << d.lgInR);
        };
        // Table-major: the E entries of one table are loaded together, so a wave
        // has E loads in flight instead of one dependent load per element.  The
        // first table is applied in straight-line code, so its loads issue with
        // the tile's input loads and share their wait (a loop that touches x
        // would wait for x's loads before issuing its first table load).
        // A table whose field misses the tile's column and row-block bits, and
        // holds all or none of its row bits, is read at base + (row << s) with a
        // per-tile base and conjugation: 32-bit indices instead of each element's
        // 64-bit flat index.
        const int rowLo = d.lgInR, rowHi = d.lgInR + LGM;
        auto hits = [](int sh, int nb, int lo, int n) { return n > 0 && sh < lo + n && lo < sh + nb; };
        // Where a table is read through the tile's base: the base, the row shift
        // (31: the row does not index it) and the conjugation.
        struct TabAt {
            DiagTab T;
            bool fast, cj;
            int s;
            const C* tb;
        };
        auto tab_at = [&](const DiagTab& TL) {
            TabAt A;
            // the descriptor is read from LDS into vector registers; it is the
            // same for every lane, so keep it in scalar registers
            DiagTab& T = A.T;
            T.sh = __builtin_amdgcn_readfirstlane(TL.sh);
            T.nb = __builtin_amdgcn_readfirstlane(TL.nb);
            T.ctrl = __builtin_amdgcn_readfirstlane(TL.ctrl);
            T.off = (uint32_t)__builtin_amdgcn_readfirstlane((int)TL.off);
            T.nsel = __builtin_amdgcn_readfirstlane(TL.nsel);
            T.sel = (uint32_t)__builtin_amdgcn_readfirstlane((int)TL.sel);
            const bool rowIn = T.sh <= rowLo && rowHi <= T.sh + T.nb;
            auto varies = [&](int b) {
                return b >= 0 && (b < LGC || (b >= rowLo && b < rowHi) ||
                                  (d.lgHb > 0 && b >= d.lgInH2 && b < d.lgInH2 + d.lgHb));
            };
            bool ctrlVar = varies(T.ctrl);
#pragma unroll
            for (int j = 0; j < CVDV_TAB_SEL; ++j)
                if (j < T.nsel) ctrlVar = ctrlVar || varies(tab_sel(T, j));
            A.fast = !hits(T.sh, T.nb, 0, LGC) && !hits(T.sh, T.nb, d.lgInH2, d.lgHb) &&
                     (rowIn || !hits(T.sh, T.nb, rowLo, LGM)) && !ctrlVar;
            // choice bits are tile-uniform here: they only move the tile's base
            A.tb = (const C*)pre.arena + T.off + (uint32_t)((inBase >> T.sh) & ((1ull << T.nb) - 1ull)) +
                   (T.nsel > 0 ? tab_choice(T, inBase) : 0u);
            A.s = rowIn ? rowLo - T.sh : 31;
            A.cj = T.ctrl >= 0 && ((inBase >> T.ctrl) & 1ull);
            return A;
        };
        auto apply_table = [&](const DiagTab& TL) {
            const TabAt A = tab_at(TL);
            const DiagTab& T = A.T;
            const bool fast = A.fast;
            constexpr int CH = sizeof(C) == 16 ? (RP < 8 ? 2 : 4) : 8;  // entries in flight per batch
            if (TM == TM_ROWS || fast) {
                const C* tb = A.tb;
                const int s = A.s;
                const bool cj = A.cj;
#pragma unroll
                for (int j0 = 0; j0 < E; j0 += CH) {
                    C e[CH];
#pragma unroll
                    for (int j = 0; j < CH; ++j) {
                        const BflyPos p = bfly_pos<RP, LGM>(tt + ((j0 + j) / RP) * nt, LGC);
                        const uint32_t row = (uint32_t)(p.q + (((j0 + j) % RP) << lgMq));
                        e[j] = tb[s < 31 ? (row << s) : 0u];
                        if (cj) e[j].y = -e[j].y;
                    }
#pragma unroll
                    for (int j = 0; j < CH; ++j) x[j0 + j] = cmul(x[j0 + j], e[j]);
                }
            } else if constexpr (TM == TM_ANY) {
                // per element; the choice bits are tile-uniform (the tables were
                // compiled for this tiling), so the choice is one per tile
                const C* tb = (const C*)pre.arena + T.off + (T.nsel > 0 ? tab_choice(T, inBase) : 0u);
#pragma unroll
                for (int j0 = 0; j0 < E; j0 += CH) {
                    C e[CH];
#pragma unroll
                    for (int j = 0; j < CH; ++j) {
                        const uint64_t g = elem((j0 + j) / RP, (j0 + j) % RP);
                        e[j] = tb[(uint32_t)((g >> T.sh) & ((1ull << T.nb) - 1ull))];
                        if (T.ctrl >= 0 && ((g >> T.ctrl) & 1ull)) e[j].y = -e[j].y;
                    }
#pragma unroll
                    for (int j = 0; j < CH; ++j) x[j0 + j] = cmul(x[j0 + j], e[j]);
                }
            }
        };
        // Generic terms affine in the row index (complex64 staged instances always;
        // otherwise when d.genSteps): a butterfly's rows are q + i M/RP with its
        // column and row block (position j) fixed, so its phases are P(q) + i D_j.
        // The steps e^{2 pi i i D_j} go to the free exchange buffer, RP-1 per
        // position, one sin/cos each; a butterfly then takes one FP64 evaluation,
        // one sin/cos e0 and RP-1 products e0 * step.  The caller puts a barrier
        // before the buffer is written again.
        auto row_steps = [&]() {
            using RR = decltype(C{}.x);
            const int nj = (nt * E) >> LGM;
#pragma unroll 1
            for (int s = tt; s < nj * (RP - 1); s += nt) {
                const int j = s / (RP - 1), i = s - j * (RP - 1) + 1;
                const uint64_t g0 =
                    inBase + ((uint64_t)(j >> LGC) << d.lgInH2) + (uint64_t)(j & ((1 << LGC) - 1));
                const uint64_t g1 = g0 + ((uint64_t)(1 << lgMq) << d.lgInR);
                double dl = 0.0;
                for (int t = 0; t < pre.ngen; ++t)
                    dl += cvdv_frac(diag_term_rev(P.gen[t], g1) - diag_term_rev(P.gen[t], g0));
                sRow[s] = unit_rev_fast(cvdv_frac((double)i * cvdv_frac(dl)), RR(0));
            }
            lds_barrier();
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                const BflyPos p = bfly_pos<RP, LGM>(tt + k * nt, LGC);
                const uint64_t g0 = elem(k, 0);
                double p0 = pre.genrev;
                for (int t = 0; t < pre.ngen; ++t) p0 += cvdv_frac(diag_term_rev(P.gen[t], g0));
                const C e0 = unit_rev_fast(cvdv_frac(p0), RR(0));
                const C* st = sRow + (p.c | (p.hb << LGC)) * (RP - 1) - 1;
                x[k * RP] = cmul(x[k * RP], e0);
#pragma unroll
                for (int i = 1; i < RP; ++i) x[k * RP + i] = cmul(x[k * RP + i], cmul(e0, st[i]));
            }
        };
        bool staged = false;
        if constexpr (LGC > 0 && TM != TM_ROWS) {
            // Tiles of several columns: when every table is read by row, their
            // product over the tile's M rows is staged in LDS (one table load per
            // row instead of one per element), and each element takes one LDS
            // read and one product.  The exchange buffer is free before the
            // stages; a barrier hands it back.
            bool byRow = pre.ntab > 0;
            if constexpr (TM == TM_ANY)
                for (int t = 0; t < pre.ntab; ++t) byRow = byRow && tab_at(P.tab[t]).fast;
            if (byRow) {
                for (int r = tid; r < (1 << LGM); r += nt) {
                    C f = cmk<C>(1.0, 0.0);
                    for (int t = 0; t < pre.ntab; ++t) {
                        const TabAt A = tab_at(P.tab[t]);
                        C e = A.tb[A.s < 31 ? ((uint32_t)r << A.s) : 0u];
                        if (A.cj) e.y = -e.y;
                        f = t == 0 ? e : cmul(f, e);
                    }
                    sRow[r] = f;
                }
                lds_barrier();
#pragma unroll
                for (int k = 0; k < NB; ++k) {
                    const BflyPos p = bfly_pos<RP, LGM>(tt + k * nt, LGC);
#pragma unroll
                    for (int i = 0; i < RP; ++i) x[k * RP + i] = cmul(x[k * RP + i], sRow[p.q + (i << lgMq)]);
                }
                lds_barrier();
                staged = true;
            }
        }
        if (TM != TM_STAGED && !staged) {
            if (pre.ntab > 0) apply_table(P.tab[0]);
            for (int t = 1; t < pre.ntab; ++t) {
                // opaque: the element indices are recomputed per table, not held
                asm volatile("" : "+v"(tt));
                apply_table(P.tab[t]);
            }
        }
        if constexpr (TM == TM_STAGED && LGC > 0) {
            if (pre.ngen > 0) {
                using RR = decltype(C{}.x);
                if (sizeof(C) == 8 || d.genSteps) {
                    row_steps();
                } else {
                    // complex128 with terms that are not affine in the rows: one
                    // element per iteration, its terms evaluated exactly; its factor
                    // goes to its own slot of the exchange buffer (element-major, so
                    // a wave's writes are consecutive), then every element takes its
                    // product.
#pragma unroll 1
                    for (int k = 0; k < NB; ++k) {
#pragma unroll 1
                        for (int i = 0; i < RP; ++i) {
                            const uint64_t g = elem(k, i);
                            double r = pre.genrev;
                            for (int t = 0; t < pre.ngen; ++t) r += cvdv_frac(diag_term_rev(P.gen[t], g));
                            sRow[(k * RP + i) * nt + tt] = unit_rev_fast(cvdv_frac(r), RR(0));
                        }
                    }
#pragma unroll
                    for (int j = 0; j < E; ++j) x[j] = cmul(x[j], sRow[j * nt + tt]);
                }
                lds_barrier();  // the stages that follow write other threads' slots
            }
        } else if (pre.ngen > 0) {
            asm volatile("" : "+v"(tt));
            // one-row tiles: the terms are affine in the row coordinate
            GenFold gf{0.0, 0.0, 0};
            if (LGC == 0 && d.lgHb == 0) gf = fold_generic(P, pre, inBase, d.lgInR, LGM);
            if (gf.ok) {
                using R = decltype(C{}.x);
                // A butterfly's rows are q + i M/RP, so its phases are e^{2 pi i (a m0 + b)}
                // times the tile-wide steps e^{2 pi i a 2 i M/RP}: one sin/cos per
                // butterfly, and RP-1 per tile (uniform across the wave).
                C st[RP];
                if (nt >= 64) {
                    // full waves: lane i evaluates step i, the others read it
                    const int li = (int)__builtin_amdgcn_mbcnt_lo(~0u, 0u) & (RP - 1);
                    const C mine = unit_rev_fast(cvdv_frac(gf.a * (double)(2 * (li << lgMq))), R(0));
#pragma unroll
                    for (int i = 1; i < RP; ++i) st[i] = lane_value(mine, i);
                } else {
#pragma unroll
                    for (int i = 1; i < RP; ++i)
                        st[i] = unit_rev_fast(cvdv_frac(gf.a * (double)(2 * (i << lgMq))), R(0));
                }
#pragma unroll
                for (int k = 0; k < NB; ++k) {
                    const BflyPos p = bfly_pos<RP, LGM>(tt + k * nt, LGC);
                    const double m0 = (double)(2 * p.q + 1 - (1 << LGM));
                    const C e0 = unit_rev_fast(cvdv_frac(fma(gf.a, m0, gf.b)), R(0));
                    x[k * RP] = cmul(x[k * RP], e0);
#pragma unroll
                    for (int i = 1; i < RP; ++i) x[k * RP + i] = cmul(x[k * RP + i], cmul(e0, st[i]));
                }
            } else if (TM != TM_ROWS && d.genSteps) {
                // tiles of several columns or row blocks
                row_steps();
                lds_barrier();  // the stages that follow write the exchange buffer
            } else {
#pragma unroll
                for (int k = 0; k < NB; ++k)
#pragma unroll
                    for (int i = 0; i < RP; ++i)
                        x[k * RP + i] = cmul(x[k * RP + i], gen_factor<C>(P, pre, elem(k, i)));
            }
        }
        if (d.conjIn) {
#pragma unroll
            for (int i = 0; i < E; ++i) x[i].y = -x[i].y;
        }
    } else if (d.conjIn) {
#pragma unroll
        for (int i = 0; i < E; ++i) x[i].y = -x[i].y;
    }
}

// Output scale and conjugation (inverse), applied to a thread's transformed
// elements before either store path.  Behind uniform branches: the compiler
// would otherwise select per element and pay for the scale even when it is off.
template <int E, typename R, typename C>
__device__ __forceinline__ void out_finish(C* x, const FftPass& d, R scale, R scaleLo) {
    if (d.useScale) {
#pragma unroll
        for (int i = 0; i < E; ++i) {
            x[i].x = fma(x[i].x, scale, x[i].x * scaleLo);
            x[i].y = fma(x[i].y, scale, x[i].y * scaleLo);
        }
    }
    if (d.conjOut) {
#pragma unroll
        for (int i = 0; i < E; ++i) x[i].y = -x[i].y;
    }
}

// Last-stage outputs straight to global memory (rows q + m*M/RP of column c).
template <int RP, int E, int LGM, int LGC, typename R, typename C>
__device__ __forceinline__ void last_store(C* __restrict__ out, const C* v, const FftPass& d, uint64_t outBase,
                                           R scale, R scaleLo, int tid, int nt) {
    constexpr int NB = E / RP;
    constexpr int LGR = RP == 8 ? 3 : (RP == 4 ? 2 : 1);
    constexpr int lgMq = LGM - LGR;
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        const BflyPos p = bfly_pos<RP, LGM>(tid + k * nt, LGC);
        const uint64_t g0 = outBase + ((uint64_t)p.hb << d.lgOutH2) + p.c;
#pragma unroll
        for (int m = 0; m < RP; ++m) {
            const int row = p.q + (m << lgMq);
            out[g0 + ((uint64_t)row << d.lgOutR)] = v[k * RP + m];
        }
    }
}

// Stages 1 .. NST-1: exchange the previous stage's outputs through LDS, then
// twiddle and transform.
template <int LGM, int LGC, int S, int E, bool TWL, bool ONE = false, typename C>
__device__ __forceinline__ void run_stages(C* x, C* lds, const C* __restrict__ tw, int tid, int nt) {
    if constexpr (S < Plan<LGM>::NST) {
        constexpr int RPprev = 1 << Plan<LGM>::lgr(S - 1);
        constexpr int RP = 1 << Plan<LGM>::lgr(S);
        stage_write<RPprev, E, LGM, LGC, S - 1, ONE>(lds, x, tid, nt);
        lds_barrier();
        stage_read<RP, E, LGM, LGC, S - 1, ONE>(lds, x, tid, nt);
        lds_barrier();
        stage_compute<RP, E, LGM, LGC, S, TWL>(x, tw, tid, nt);
        run_stages<LGM, LGC, S + 1, E, TWL, ONE>(x, lds, tw, tid, nt);
    }
}

// One-row complex64 tiles of 2^13-2^14 points lay out their first two exchanges
// apart from the others (ex_lay); their LDS ends at tile + tile/8.
template <typename R, int LGM, int LGC>
constexpr bool one_row_layouts() {
    return sizeof(R) == 4 && LGC == 0 && LGM >= 13 && Plan<LGM>::lgr(0) == 3 && Plan<LGM>::lgr(1) == 3;
}

// PF: persistent workgroups that prefetch the next tile (one-row tiles).
// Workgroup size bound: one-row tiles of 2^12 points and more take up to 1024
// threads; every other tile fits 64 KiB of LDS, i.e. at most 512 (complex64) or
// 256 (complex128) threads at 8 elements each.
template <typename R, int LGM, int LGC>
constexpr int tile_threads_max() {
    return (LGC == 0 && LGM >= 12) ? 1024 : (sizeof(R) == 4 ? 512 : 256);
}

template <typename R, int LGM, int LGC, int E, bool PF>
__global__ void __launch_bounds__((tile_threads_max<R, LGM, LGC>()))
k_fft_tile(const typename CT<R>::type* in, typename CT<R>::type* out, FftPass d, FastDiag pre,
           const typename CT<R>::type* __restrict__ tw, const typename CT<R>::type* __restrict__ tw4) {
    using C = typename CT<R>::type;
    constexpr int NST = Plan<LGM>::NST;
    constexpr int R0 = 1 << Plan<LGM>::lgr(0);
    constexpr int RL = 1 << Plan<LGM>::lgr(NST - 1);
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ PreLds P;
    C* lds = reinterpret_cast<C*>(smem_raw);
    // persistent one-row tiles are launched with 2^LGM / E threads (the host
    // checks): a compile-time count lets the butterflies of one thread, which sit
    // nt apart, share the twiddles their rows have in common
    const int tid = threadIdx.x, nt = PF ? (1 << LGM) / E : (int)blockDim.x;
    constexpr int lgC = LGC;  // the tile's column count is part of the instance
    const int lgHb = d.lgHb;
    const int tile = 1 << (LGM + lgC + lgHb);
    // one-row complex64 tiles of 2^13-2^14 points: exchanges 0 and 1 in their own
    // layouts, the largest of which ends at tile + tile/8 (the host sizes LDS so)
    constexpr bool ONE = one_row_layouts<R, LGM, LGC>();
    const int ldsEnd = ONE ? tile + (tile >> 3) : lds_ix(tile);

    if (d.usePre) {
        // constant indices: the argument is read in place (one entry per thread,
        // or all by thread 0 in a tile with fewer threads than entries)
        if (nt >= CVDV_FAST_TABS + CVDV_FAST_GEN) {
#pragma unroll
            for (int i = 0; i < CVDV_FAST_TABS; ++i)
                if (tid == i && i < pre.ntab) P.tab[i] = pre.tab[i];
#pragma unroll
            for (int i = 0; i < CVDV_FAST_GEN; ++i)
                if (tid == CVDV_FAST_TABS + i && i < pre.ngen) P.gen[i] = pre.gen[i];
        } else if (tid == 0) {
#pragma unroll
            for (int i = 0; i < CVDV_FAST_TABS; ++i)
                if (i < pre.ntab) P.tab[i] = pre.tab[i];
#pragma unroll
            for (int i = 0; i < CVDV_FAST_GEN; ++i)
                if (i < pre.ngen) P.gen[i] = pre.gen[i];
        }
        __syncthreads();
    }

    const R scale = (R)d.outScale;
    const R scaleLo = (R)(d.outScale - (double)scale);
    // One-row tiles keep every twiddle in LDS after the tile (split table, ~11 KiB):
    // no global loads remain between a prefetch and its use.
    const C* twl = tw;
    if constexpr (PF) {
        C* sTw = lds + ldsEnd;
        for (int e = tid; e < SplitPlan<LGM>::total(); e += nt) sTw[e] = tw[e];
        __syncthreads();
        twl = sTw;
    }

    auto tile_base = [&](uint64_t t, uint64_t& lo0, uint64_t& inBase, uint64_t& outBase) {
        const uint64_t loBlk = t & ((1ull << d.lgNLoBlk) - 1);
        const uint64_t r = t >> d.lgNLoBlk;
        const uint64_t h2Blk = r & ((1ull << d.lgNH2Blk) - 1);
        const uint64_t h1 = r >> d.lgNH2Blk;
        lo0 = loBlk << lgC;
        inBase = (h1 << d.lgInH1) + ((h2Blk << lgHb) << d.lgInH2) + lo0;
        outBase = (h1 << d.lgOutH1) + ((h2Blk << lgHb) << d.lgOutH2) + lo0;
    };
    // Phases of the loaded tile t (their table loads are global loads).
    auto prephase = [&](C* x, uint64_t t, int ti) {
        uint64_t lo0, inBase, outBase;
        tile_base(t, lo0, inBase, outBase);
        stage0_pre<R0, E, LGM, LGC, !PF, PF ? TM_ROWS : TM_ANY>(x, d, inBase, P, pre, tw4, ti, nt, lds);
    };
    // Transform the phased tile in registers (the last stage's outputs stay in x).
    auto stages = [&](C* x, int ti, const C* twt) {
        stage_compute<R0, E, LGM, LGC, 0, PF>(x, twt, ti, nt);
        run_stages<LGM, LGC, 1, E, PF, ONE>(x, lds, twt, ti, nt);
    };
    // Store the transformed tile t.
    // Staged store, first half: the last stage's outputs into LDS in natural order.
    auto stage_out = [&](C* x, int ti) {
        stage_write<RL, E, LGM, LGC, NST - 1>(lds, x, ti, nt);
        lds_barrier();
    };
    // Staged store, second half: LDS to global memory in the output's contiguous order.
    auto store_lds = [&](uint64_t t, int ti) {
        uint64_t lo0, inBase, outBase;
        tile_base(t, lo0, inBase, outBase);
        auto value = [&](int c, int row, int hb) {
            return lds[lds_ix((((hb << LGM) + row) << lgC) + c)];
        };
        if (sizeof(C) == 8 && d.vecOut) {
#pragma unroll
            for (int j = 0; j < E / 2; ++j) {
                const int e = 2 * ti + j * 2 * nt;
                int c, row, hb, c1, row1, hb1;
                tile_coords<LGM>(e, lgC, lgHb, d.rowFastOut, c, row, hb);
                tile_coords<LGM>(e + 1, lgC, lgHb, d.rowFastOut, c1, row1, hb1);
                const C v0 = value(c, row, hb);
                const C v1 = value(c1, row1, hb1);
                const uint64_t g = outBase + ((uint64_t)hb << d.lgOutH2) + ((uint64_t)row << d.lgOutR) + c;
                *reinterpret_cast<float4*>(out + g) = make_float4((float)v0.x, (float)v0.y, (float)v1.x, (float)v1.y);
            }
        } else {
#pragma unroll
            for (int j = 0; j < E; ++j) {
                const int e = ti + j * nt;
                int c, row, hb;
                tile_coords<LGM>(e, lgC, lgHb, d.rowFastOut, c, row, hb);
                out[outBase + ((uint64_t)hb << d.lgOutH2) + ((uint64_t)row << d.lgOutR) + c] = value(c, row, hb);
            }
        }
    };
    auto store = [&](C* x, uint64_t t, int ti) {
        out_finish<E>(x, d, scale, scaleLo);
        if (d.directOut) {
            uint64_t lo0, inBase, outBase;
            tile_base(t, lo0, inBase, outBase);
            last_store<RL, E, LGM, LGC>(out, x, d, outBase, scale, scaleLo, ti, nt);
            return;
        }
        stage_out(x, ti);
        store_lds(t, ti);
    };

    if constexpr (!PF) {
        // one tile per workgroup
        const uint64_t t = blockIdx.x;
        uint64_t lo0, inBase, outBase;
        tile_base(t, lo0, inBase, outBase);
        C x[E];
        stage0_load<R0, E, LGM, LGC>(in, x, d, inBase, tid, nt);
        if (d.conjLoad) {
#pragma unroll
            for (int i = 0; i < E; ++i) x[i].y = -x[i].y;
        }
        prephase(x, t, tid);
        // an opaque copy of the thread id for the stages and the store: their
        // index math would otherwise be computed early and held across the phases
        // (8-element tiles; the 16-element one-row tile has no register left for it)
        int t2 = tid;
        if constexpr (E == 8) asm volatile("" : "+v"(t2));
        stages(x, t2, twl);
        store(x, t, t2);
    } else {
        // Persistent workgroups (tile t = blockIdx.x, + gridDim.x, ...): the next
        // tile's inputs are loaded into registers before the current tile is
        // transformed, so HBM reads overlap the LDS stages (the exchanges use
        // LDS-only barriers, which leave those loads in flight).  Used for the
        // one-row tiles that fill a CU's LDS alone.
        uint64_t t = blockIdx.x;
        C xn[E];
        if (t < d.nTiles) {
            uint64_t lo0, inBase, outBase;
            tile_base(t, lo0, inBase, outBase);
            stage0_load<R0, E, LGM, LGC>(in, xn, d, inBase, tid, nt);
        }
        while (t < d.nTiles) {
            // Opaque per-iteration copies: index math derived from the thread id
            // is the same for every tile; hoisted out of the loop it would pin
            // dozens of registers across iterations.  The LDS twiddles are
            // addressed from the shared array so that their loads stay LDS loads
            // (a generic pointer would also count on the memory counter).
            int ti = tid;
            int twOff = ldsEnd;
            asm volatile("" : "+v"(ti), "+v"(twOff));
            const C* twt = lds + twOff;
            C x[E];
#pragma unroll
            for (int i = 0; i < E; ++i) x[i] = xn[i];
            // (no conjLoad here: it belongs to four-step halves, and the
            // one-row tiles sit at 128 VGPRs already)
            // the phase-table loads wait on the memory counter, which completes in
            // order: issue the next tile's loads after them
            prephase(x, t, ti);
            const uint64_t tn = t + gridDim.x;
            if (tn < d.nTiles) {
                uint64_t lo0, inBase, outBase;
                tile_base(tn, lo0, inBase, outBase);
                stage0_load<R0, E, LGM, LGC>(in, xn, d, inBase, ti, nt);
            }
            stages(x, ti, twt);
            store(x, t, ti);
            lds_barrier();  // the next tile's first exchange overwrites LDS
            t = tn;
        }
    }
}

// Two half transforms of one register field in one in-place pass (FtDefer):
// the pending half of the previous transform (the four-step twiddle on load,
// a length-M transform over the tile's rows), the pending phases, then the
// first half of the next transform (a transform over the same rows), stored
// back in place.  Each half of an inverse is conj . forward half . conj (so a
// half-done inverse is a linear map of the state and commutes with work on
// other bits): conjMid conjugates before and after the pending half, dD.conjIn
// and d.conjOut before and after the next one.
// The phases are evaluated at the natural index of the completed transform:
// dD is the tile geometry seen in that index (row and row-block shifts), and
// swapN > 0 exchanges the two swapN-bit digits at swapSh of the tile base (the
// in-place L half leaves the register's digits exchanged).  Between the halves
// the elements go through LDS from the last stage's order to stage 0's.
template <typename R, int LGM, int LGC, int E, bool STG>
__global__ void __launch_bounds__((tile_threads_max<R, LGM, LGC>()))
__attribute__((amdgpu_waves_per_eu((STG && sizeof(R) == 4) ? 8 : 1)))
k_fft_fused(typename CT<R>::type* data, FftPass d, FftPass dD, FastDiag pre, int conjMid, int swapSh, int swapN,
            const typename CT<R>::type* __restrict__ tw, const typename CT<R>::type* __restrict__ tw4) {
    using C = typename CT<R>::type;
    constexpr int NST = Plan<LGM>::NST;
    constexpr int R0 = 1 << Plan<LGM>::lgr(0);
    constexpr int RL = 1 << Plan<LGM>::lgr(NST - 1);
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ PreLds P;
    C* lds = reinterpret_cast<C*>(smem_raw);
    const int tid = threadIdx.x, nt = blockDim.x;

    if (dD.usePre) {
        if (nt >= CVDV_FAST_TABS + CVDV_FAST_GEN) {
#pragma unroll
            for (int i = 0; i < CVDV_FAST_TABS; ++i)
                if (tid == i && i < pre.ntab) P.tab[i] = pre.tab[i];
#pragma unroll
            for (int i = 0; i < CVDV_FAST_GEN; ++i)
                if (tid == CVDV_FAST_TABS + i && i < pre.ngen) P.gen[i] = pre.gen[i];
        } else if (tid == 0) {
#pragma unroll
            for (int i = 0; i < CVDV_FAST_TABS; ++i)
                if (i < pre.ntab) P.tab[i] = pre.tab[i];
#pragma unroll
            for (int i = 0; i < CVDV_FAST_GEN; ++i)
                if (i < pre.ngen) P.gen[i] = pre.gen[i];
        }
        __syncthreads();
    }
    const R scale = (R)d.outScale;
    const R scaleLo = (R)(d.outScale - (double)scale);

    const uint64_t t = blockIdx.x;
    const uint64_t loBlk = t & ((1ull << d.lgNLoBlk) - 1);
    const uint64_t rr = t >> d.lgNLoBlk;
    const uint64_t h2Blk = rr & ((1ull << d.lgNH2Blk) - 1);
    const uint64_t h1 = rr >> d.lgNH2Blk;
    const uint64_t base = (h1 << d.lgInH1) + ((h2Blk << d.lgHb) << d.lgInH2) + (loBlk << LGC);
    uint64_t natBase = base;
    if (swapN > 0) {
        const uint64_t m = (1ull << swapN) - 1ull;
        const uint64_t lo = (base >> swapSh) & m, hi = (base >> (swapSh + swapN)) & m;
        natBase = (base & ~(((m << swapN) | m) << swapSh)) | (((lo << swapN) | hi) << swapSh);
    }

    C x[E];
    stage0_load<R0, E, LGM, LGC>(data, x, d, base, tid, nt);
    if (conjMid) {  // an inverse's half is conj . forward half . conj
#pragma unroll
        for (int i = 0; i < E; ++i) x[i].y = -x[i].y;
    }
    // pending half: the four-step twiddle (d carries no phases and no input conjugation)
    stage0_pre<R0, E, LGM, LGC, true, TM_NONE>(x, d, base, P, pre, tw4, tid, nt, lds);
    stage_compute<R0, E, LGM, LGC, 0, false>(x, tw, tid, nt);
    run_stages<LGM, LGC, 1, E, false>(x, lds, tw, tid, nt);
    if (conjMid) {
#pragma unroll
        for (int i = 0; i < E; ++i) x[i].y = -x[i].y;
    }
    // last stage's order -> stage 0's order
    stage_write<RL, E, LGM, LGC, NST - 1>(lds, x, tid, nt);
    lds_barrier();
    stage_read<R0, E, LGM, LGC>(lds, x, tid, nt);
    lds_barrier();
    // phases at the natural index, then the next transform's first half.  An
    // opaque copy of the thread id: index math derived from it (the store
    // addresses above all) would otherwise be computed early and held across
    // the phases in registers that the workgroup size leaves no room for.
    int t2 = tid;
    asm volatile("" : "+v"(t2));
    stage0_pre<R0, E, LGM, LGC, false, STG ? TM_STAGED : TM_ANY>(x, dD, natBase, P, pre, tw4, t2, nt, lds);
    stage_compute<R0, E, LGM, LGC, 0, false>(x, tw, t2, nt);
    run_stages<LGM, LGC, 1, E, false>(x, lds, tw, t2, nt);
    out_finish<E>(x, d, scale, scaleLo);
    last_store<RL, E, LGM, LGC>(data, x, d, base, scale, scaleLo, t2, nt);
}

// Direct DFT for very small states (fewer than 8 elements per tile).
template <typename R, bool INV>
__global__ void k_dft_naive(const typename CT<R>::type* in, typename CT<R>::type* out, uint64_t total,
                            int sh, int nb, FastDiag pre, int usePre, double scale,
                            const double2* __restrict__ twN) {
    using C = typename CT<R>::type;
    const uint64_t M = 1ull << nb, fmask = (M - 1) << sh;
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
         g += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t k = (g >> sh) & (M - 1);
        const uint64_t base = g & ~fmask;
        double ax = 0, ay = 0;
        for (uint64_t j = 0; j < M; ++j) {
            const uint64_t gi = base | (j << sh);
            C v = in[gi];
            if (usePre) v = apply_fast(pre, gi, v);
            const double2 p = mul_w<INV>(make_double2((double)v.x, (double)v.y), twN[(j * k) & (M - 1)]);
            ax += p.x;
            ay += p.y;
        }
        out[g] = cmk<C>(ax * scale, ay * scale);  // accumulated in FP64, rounded once
    }
}

// ---------------------------------------------------------------- planning
struct PassGeom {
    int M;
    uint64_t nLo, nH2, nH1;
    uint64_t inH1, inH2, inR, outH1, outH2, outR;
};

static int ilog2(uint64_t v) { return 63 - __builtin_clzll(v); }

// Index bits that vary within one tile of a pass: columns, rows, row blocks.
static uint64_t tile_bits(const FftPass& d) {
    auto field = [](int s, int n) { return n > 0 ? (((1ull << n) - 1ull) << s) : 0ull; };
    return field(0, d.lgC) | field(d.lgInR, d.lgM) | field(d.lgInH2, d.lgHb);
}
static size_t lds_bytes(uint64_t tile, size_t esz) { return (tile + tile / 32) * esz; }

// Tile limits: a strided tile stays within 64 KiB of LDS so several workgroups
// share a CU; a whole contiguous register may take one large workgroup.  cfg
// (cvdvFftAutotune) may narrow the rows and change how far a tile grows over
// row blocks.
static bool fit_pass(const PassGeom& g, int prec, FftPass& d, bool allowBig, const FftCfg& cfg) {
    const size_t esz = prec == CVDV_C128 ? 16 : 8;
    // 256 contiguous bytes per row: the measured HBM rate of strided tiles rises
    // from ~4.0 TB/s with 128-byte rows to ~5.1 TB/s with 256-byte rows
    const uint64_t Wmax = prec == CVDV_C128 ? 16 : 32;
    const uint64_t W = cfg.rowElems > 0 && (uint64_t)cfg.rowElems < Wmax ? (uint64_t)cfg.rowElems : Wmax;
    const uint64_t cap = cfg.tileCap > 0 ? (uint64_t)cfg.tileCap : 2048;
    uint64_t C = g.nLo < W ? g.nLo : W;
    const size_t small = 64 * 1024;
    // the passes of a two-step split narrow their rows before giving up (a tall
    // pass of a contiguous register above one tile has M up to 2^12 rows)
    if (!allowBig)
        while (C > 1 && lds_bytes(C * (uint64_t)g.M, esz) > small) C /= 2;
    const uint64_t tileCM = C * (uint64_t)g.M;
    const bool big = allowBig && C == 1 && lds_bytes(tileCM, esz) <= device_max_lds() && tileCM <= 16384;
    if (lds_bytes(tileCM, esz) > small && !big) return false;
    uint64_t Hb = 1;
    while (Hb * 2 <= g.nH2 && tileCM * Hb * 2 <= cap && lds_bytes(tileCM * Hb * 2, esz) <= small) Hb *= 2;
    if (tileCM * Hb < 8) return false;
    const uint64_t tile = tileCM * Hb;
    d.lgM = ilog2(g.M);
    d.lgC = ilog2(C);
    d.lgHb = ilog2(Hb);
    d.lgNLoBlk = ilog2(g.nLo / C);
    d.lgNH2Blk = ilog2(g.nH2 / Hb);
    auto lg0 = [](uint64_t v) { return v ? ilog2(v) : 0; };  // stride 0: that index is always 0
    d.lgInH1 = lg0(g.inH1); d.lgInH2 = lg0(g.inH2); d.lgInR = lg0(g.inR);
    d.lgOutH1 = lg0(g.outH1); d.lgOutH2 = lg0(g.outH2); d.lgOutR = lg0(g.outR);
    d.rowFastOut = g.outR <= g.outH2 ? 1 : 0;
    // 16-byte pairs along the fastest dimension when it is contiguous in memory
    auto vec_ok = [&](int rowFast, uint64_t sR, uint64_t sH) {
        if (esz != 8 || tile < 2) return 0;
        if (C >= 2) return 1;
        if (rowFast) return (sR == 1 && g.M >= 2) ? 1 : 0;
        return (sH == 1 && Hb >= 2) ? 1 : 0;
    };
    d.vecOut = vec_ok(d.rowFastOut, g.outR, g.outH2);
    // last stage writes rows q + m*M/R of column c: coalesced when a full
    // column group is one full-width run or when rows follow the columns directly
    d.directOut = (C == W || g.outR == C) ? 1 : 0;
    return true;
}

// Workgroups of a kernel instance resident on the whole device at once.
static uint64_t resident_blocks(const void* kfn, int threads, size_t lds) {
    static std::map<std::tuple<int, const void*, int, size_t>, uint64_t> cache;
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    const auto key = std::make_tuple(dev, kfn, threads, lds);
    {
        std::lock_guard<std::mutex> lk(g_devCacheMu);
        auto it = cache.find(key);
        if (it != cache.end()) return it->second;
    }
    int cus = 0, per = 0;
    HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kfn, threads, lds));
    const uint64_t n = (uint64_t)std::max(per, 1) * (uint64_t)std::max(cus, 1);
    std::lock_guard<std::mutex> lk(g_devCacheMu);
    cache[key] = n;
    return n;
}

// Kernel instances exist for the tile shapes fit_pass can produce: C = 2^LGC
// columns (up to 256 bytes per row) with C*M within 64 KiB of LDS, or one whole
// contiguous row (C = 1) of up to 2^14 complex64 / 2^13 complex128 points.
template <typename R, int LGM, int LGC>
constexpr bool tile_instance() {
    constexpr bool c64 = sizeof(R) == 4;
    if (LGC > (c64 ? 5 : 4)) return false;
    if (LGC == 0 && LGM <= (c64 ? 14 : 13)) return true;
    return LGM + LGC <= (c64 ? 13 : 12);
}

// Every table of the program is read by row in a one-row tile over the field
// [fs, fs + fn): its field holds the whole row field or none of it, and its
// control bit lies outside (the persistent instances compile only that path).
static bool table_by_row(const FastDiag& pre, int fs, int fn) {
    for (int i = 0; i < pre.ntab; ++i) {
        const DiagTab& T = pre.tab[i];
        const bool in = T.sh <= fs && fs + fn <= T.sh + T.nb;
        const bool out = T.sh >= fs + fn || fs >= T.sh + T.nb;
        if (!in && !out) return false;
        if (T.ctrl >= fs && T.ctrl < fs + fn) return false;
        for (int j = 0; j < T.nsel; ++j)
            if (tab_sel(T, j) >= fs && tab_sel(T, j) < fs + fn) return false;
    }
    return true;
}

static bool gen_affine_in_rows(const FastDiag& pre, int rowLo, int rowN);

template <typename R, int LGM, int LGC>
static void launch_geom(CVDVContext* ctx, const typename CT<R>::type* in, typename CT<R>::type* out,
                        const FftPass& d, const FastDiag& pre, const typename CT<R>::type* tw,
                        const typename CT<R>::type* tw4, uint64_t blocks, int tile, size_t lds) {
    if constexpr (tile_instance<R, LGM, LGC>()) {
        FftPass dg = d;
        dg.genSteps = d.usePre && pre.ngen > 0 && gen_affine_in_rows(pre, d.lgInR, LGM) ? 1 : 0;
        auto go = [&](auto kfn, int E, bool persistent) {
            if (tile / E > tile_threads_max<R, LGM, LGC>())
                throw CvdvError{CVDV_ERR_UNSUPPORTED, "FFT tile larger than its kernel's workgroup bound"};
            if (persistent && tile != (1 << LGM))
                throw CvdvError{CVDV_ERR_UNSUPPORTED, "persistent FFT tile of more than one row"};
            if (lds > 64 * 1024)
                HIPCHK(hipFuncSetAttribute((const void*)kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            // persistent instances: as many workgroups as the device holds at once
            const uint64_t grid =
                persistent ? std::min<uint64_t>(blocks, resident_blocks((const void*)kfn, tile / E, lds)) : blocks;
            hipLaunchKernelGGL(kfn, dim3((unsigned)grid), dim3(tile / E), lds, ctx->stream, in, out, dg, pre, tw,
                               tw4);
        };
        // one-row complex64 tiles that fill a CU's LDS alone prefetch the next row
        // and keep their (split) twiddles in LDS
        // (the second pass of a four-step split, which carries the four-step
        // twiddle, a pass that conjugates on load, and a program with a table not
        // read by row run one tile per workgroup)
        constexpr bool PF = LGC == 0 && LGM >= 13 && sizeof(R) == 4;
        const bool pf = PF && !d.tw4 && !d.conjLoad && (!d.usePre || table_by_row(pre, d.lgInR, LGM));
        if constexpr (one_row_layouts<R, LGM, LGC>()) {
            if (tile != (1 << LGM))
                throw CvdvError{CVDV_ERR_UNSUPPORTED, "one-row FFT tile of more than one row"};
            lds = (size_t)(tile + tile / 8) * sizeof(typename CT<R>::type);
        }
        if (pf) {
            tw = (const typename CT<R>::type*)split_table(ctx, LGM, sizeof(R) == 4 ? CVDV_MIXED : CVDV_C128);
            lds += (size_t)SplitPlan<LGM>::total() * sizeof(typename CT<R>::type);
        }
        if (lds > device_max_lds())
            throw CvdvError{CVDV_ERR_UNSUPPORTED, "FFT tile needs more LDS than the device has"};
        if constexpr (LGM == 14 && LGC == 0 && sizeof(R) == 4) {
            if (tile / 8 > 1024) {
                if (pf) go(k_fft_tile<R, LGM, LGC, 16, PF>, 16, PF);
                else go(k_fft_tile<R, LGM, LGC, 16, false>, 16, false);
                return;
            }
        }
        if (pf) go(k_fft_tile<R, LGM, LGC, 8, PF>, 8, PF);
        else go(k_fft_tile<R, LGM, LGC, 8, false>, 8, false);
    } else {
        throw CvdvError{CVDV_ERR_UNSUPPORTED, "FFT tile shape without a kernel instance"};
    }
}

template <typename R, int LGM>
static void launch_lgm(CVDVContext* ctx, const typename CT<R>::type* in, typename CT<R>::type* out,
                       const FftPass& d, const FastDiag& pre, const typename CT<R>::type* tw,
                       const typename CT<R>::type* tw4, uint64_t blocks, int tile, size_t lds) {
    switch (d.lgC) {
        case 0: launch_geom<R, LGM, 0>(ctx, in, out, d, pre, tw, tw4, blocks, tile, lds); break;
        case 1: launch_geom<R, LGM, 1>(ctx, in, out, d, pre, tw, tw4, blocks, tile, lds); break;
        case 2: launch_geom<R, LGM, 2>(ctx, in, out, d, pre, tw, tw4, blocks, tile, lds); break;
        case 3: launch_geom<R, LGM, 3>(ctx, in, out, d, pre, tw, tw4, blocks, tile, lds); break;
        case 4: launch_geom<R, LGM, 4>(ctx, in, out, d, pre, tw, tw4, blocks, tile, lds); break;
        case 5: launch_geom<R, LGM, 5>(ctx, in, out, d, pre, tw, tw4, blocks, tile, lds); break;
        default: throw CvdvError{CVDV_ERR_UNSUPPORTED, "FFT tile column count out of range"};
    }
}

template <typename R>
static void launch_pass(CVDVContext* ctx, const typename CT<R>::type* in, typename CT<R>::type* out,
                        const FftPass& d, const FastDiag& pre, int prec, uint64_t nH1) {
    using C = typename CT<R>::type;
    const C* tw = (const C*)stage_table(ctx, d.lgM, prec);
    const C* tw4 = d.tw4 ? (const C*)full_table(ctx, d.tw4LogN, prec) : nullptr;
    const int tile = 1 << (d.lgM + d.lgC + d.lgHb);
    if (tile / 16 > 1024 || (tile / 8 > 1024 && d.lgM != 14))
        throw CvdvError{CVDV_ERR_UNSUPPORTED, "FFT tile larger than one workgroup"};
    const uint64_t blocks = nH1 << (d.lgNH2Blk + d.lgNLoBlk);
    FftPass dd = d;
    dd.nTiles = blocks;
    const size_t lds = lds_bytes(tile, sizeof(C));
    const int pid = prof_begin(ctx, (d.lgC == 0 && d.lgHb == 0 && d.lgM >= 12) ? KK_FFT_ROW : KK_FFT);
    switch (d.lgM) {
#define CVDV_LGM_CASE(L) \
    case L: launch_lgm<R, L>(ctx, in, out, dd, pre, tw, tw4, blocks, tile, lds); break;
        CVDV_LGM_CASE(1) CVDV_LGM_CASE(2) CVDV_LGM_CASE(3) CVDV_LGM_CASE(4) CVDV_LGM_CASE(5)
        CVDV_LGM_CASE(6) CVDV_LGM_CASE(7) CVDV_LGM_CASE(8) CVDV_LGM_CASE(9) CVDV_LGM_CASE(10)
        CVDV_LGM_CASE(11) CVDV_LGM_CASE(12) CVDV_LGM_CASE(13) CVDV_LGM_CASE(14)
#undef CVDV_LGM_CASE
        default: throw CvdvError{CVDV_ERR_UNSUPPORTED, "FFT tile length out of range"};
    }
    HIPCHK(hipGetLastError());
    prof_end(ctx, pid);
}

// The four-step split of field [sh, sh+nb) (j = j1 + N1 j2, N1 = 2^(nb/2)):
//   d1: transform over j2 (stride N1 S), columns (j1, lo), in place;
//   d2: twiddle W_N^{j1 k2} on load, transform over j1 (stride S), row blocks
//       k2, written transposed (k2 + N2 k1) out of place;
// and, for equal digits, the in-place halves a pending transform completes with:
//   dL:  d2 stored in place (momentum left with its digits exchanged);
//   d1c: d1 with the twiddle W_N^{j1 j2} (j1 now the tile's) on load.
struct FourStep {
    FftPass d1, d2, dL, d1c;
    uint64_t H;
    int n1, n2;
    bool hasL;
};

static bool four_step_geom(int prec, uint64_t total, int sh, int nb, FourStep& f, const FftCfg& cfg) {
    const uint64_t N = 1ull << nb, S = 1ull << sh;
    const uint64_t H = total / (N * S);
    if (H == 0 || (H & (H - 1)) != 0 || H * N * S != total) return false;
    const int n1 = nb / 2, n2 = nb - n1;
    const uint64_t N1 = 1ull << n1, N2 = 1ull << n2;
    f = FourStep{};
    f.H = H;
    f.n1 = n1;
    f.n2 = n2;
    auto uni = [](const FftPass& d, int s, int n) {
        const uint64_t m = n > 0 ? (((1ull << n) - 1ull) << s) : 0ull;
        return (tile_bits(d) & m) ? 0 : 1;
    };
    PassGeom p1{(int)N2, N1 * S, 1, H, N * S, 0, N1 * S, N * S, 0, N1 * S};
    PassGeom p2{(int)N1, S, N2, H, N * S, N1 * S, S, N * S, S, N2 * S};
    if (!fit_pass(p1, prec, f.d1, false, cfg) || !fit_pass(p2, prec, f.d2, false, cfg)) return false;
    f.d1.tw4 = 0; f.d1.tw4Shift = 0; f.d1.tw4HiBits = 0; f.d1.tw4LogN = 0; f.d1.tw4Uni = 1;
    f.d2.tw4 = 1; f.d2.tw4Shift = sh + n1; f.d2.tw4HiBits = n2; f.d2.tw4LogN = nb;
    f.d2.tw4Uni = uni(f.d2, sh + n1, n2);
    f.d1c = f.d1;
    f.d1c.tw4 = 1; f.d1c.tw4Shift = sh; f.d1c.tw4HiBits = n1; f.d1c.tw4LogN = nb;
    f.d1c.tw4Uni = uni(f.d1c, sh, n1);
    if (n1 == n2) {
        PassGeom pL{(int)N1, S, N2, H, N * S, N1 * S, S, N * S, N1 * S, S};
        f.hasL = fit_pass(pL, prec, f.dL, false, cfg);
        f.dL.tw4 = 1; f.dL.tw4Shift = sh + n1; f.dL.tw4HiBits = n2; f.dL.tw4LogN = nb;
        f.dL.tw4Uni = uni(f.dL, sh + n1, n2);
    }
    return true;
}

// Tile choices for field [sh, sh+nb) of a buffer of `total` elements: the
// context's tuned ones for its own state, the defaults for any other buffer.
static FftCfg field_cfg(const CVDVContext* ctx, uint64_t total, int sh, int nb) {
    if (total != ctx->totalSize) return FftCfg{};
    const auto it = ctx->fftCfg.find(fft_cfg_key(sh, nb));
    return it == ctx->fftCfg.end() ? FftCfg{} : it->second;
}

// Runs the FFT over field [sh, sh+nb) of a buffer of `total` elements.
// Returns true when the result landed in the buffer returned by tmpFn.
template <typename R, typename TmpFn>
static bool fft_run(CVDVContext* ctx, typename CT<R>::type* data, TmpFn tmpFn, uint64_t total,
                    int sh, int nb, bool inverse, const DiagProg& preProg, double outScale, int prec) {
    using C = typename CT<R>::type;
    const uint64_t N = 1ull << nb, S = 1ull << sh;
    const uint64_t H = total / (N * S);
    if (H == 0 || (H & (H - 1)) != 0 || H * N * S != total)
        throw CvdvError{CVDV_ERR_ARG, "FFT over a buffer whose size is not a power-of-two multiple of the field"};
    bool act = false, folded = false;
    const int inv = inverse ? 1 : 0;
    const FftCfg cfg = field_cfg(ctx, total, sh, nb);
    // The pending phases are compiled once the tiling of the pass that applies
    // them is known: their tables take no choice bit that varies within a tile.
    auto compile_pre = [&](uint64_t tileBits, const DiagProg& prog) {
        const FastDiag f = compile_diag(ctx, prog, act, outScale, &folded, tileBits);
        return f;
    };

    // single pass: rows j (stride S), lo in [0,S), hi blocks in h2 (stride N*S)
    PassGeom g1{(int)N, S, H, 1, 0, N * S, S, 0, N * S, S};
    FftPass d{};
    if (fit_pass(g1, prec, d, true, cfg)) {
        // A persistent complex64 one-row pass folds the linear terms over its row
        // field per butterfly (one sin/cos each): a table would cost a global
        // load per element, queued behind the previous row's stores.
        const bool foldRow = prec == CVDV_MIXED && d.lgC == 0 && d.lgHb == 0 && d.lgM >= 13;
        const FastDiag pre =
            compile_pre(tile_bits(d), foldRow ? diag_fold_linear(preProg, d.lgInR, d.lgM) : preProg);
        const int usePre = act ? 1 : 0;
        d.tw4 = 0; d.tw4Shift = 0; d.tw4HiBits = 0; d.tw4LogN = 0;
        d.usePre = usePre;
        d.useScale = folded ? 0 : 1;
        d.conjIn = inv;
        d.conjOut = inv;
        d.outScale = outScale;
        launch_pass<R>(ctx, data, data, d, pre, prec, 1);
        return false;
    }
    if (total < 4096 || N <= 64) {
        // tiny state: direct DFT, out of place, FP64 accumulation rounded once
        // (per-element table reads: any choice bits)
        const FastDiag pre = compile_pre(0, preProg);
        const int usePre = act ? 1 : 0;
        const double sc = folded ? 1.0 : outScale;
        const double2* twN = (const double2*)full_table(ctx, nb);
        C* tmp = tmpFn();
        uint64_t blocks = (total + 255) / 256;
        const int pid = prof_begin(ctx, KK_FFT);
        if (inverse)
            hipLaunchKernelGGL((k_dft_naive<R, true>), dim3((unsigned)blocks), dim3(256), 0, ctx->stream,
                               data, tmp, total, sh, nb, pre, usePre, sc, twN);
        else
            hipLaunchKernelGGL((k_dft_naive<R, false>), dim3((unsigned)blocks), dim3(256), 0, ctx->stream,
                               data, tmp, total, sh, nb, pre, usePre, sc, twN);
        HIPCHK(hipGetLastError());
        prof_end(ctx, pid);
        return true;
    }
    FourStep f;
    if (!four_step_geom(prec, total, sh, nb, f, cfg))
        throw CvdvError{CVDV_ERR_UNSUPPORTED, "register too large for the two-pass FFT"};
    FftPass d1 = f.d1, d2 = f.d2;
    const FastDiag pre = compile_pre(tile_bits(d1), preProg);
    const int usePre = act ? 1 : 0;
    // inverse: conj(P2(T P1(conj x))) - pass 1 conjugates its input, pass 2 its
    // output; the twiddle T = W_N^{j1 k2} is applied by pass 2 on load, where the
    // pass has memory time to spare (pass 1 also applies the pending phases)
    d1.usePre = usePre;
    d1.useScale = 0;
    d1.conjIn = inv;
    d1.conjOut = 0;
    d1.outScale = 1.0;
    d2.usePre = 0;
    d2.useScale = folded ? 0 : 1;
    d2.conjIn = 0;
    d2.conjOut = inv;
    d2.outScale = outScale;
    C* tmp = tmpFn();
    launch_pass<R>(ctx, data, data, d1, pre, prec, f.H);
    launch_pass<R>(ctx, data, tmp, d2, FastDiag{}, prec, f.H);
    return true;
}

// ---------------------------------------------------------------- half transforms kept pending
// Fused instances exist for digits of 2^FUSE_LGM_MIN .. 2^FUSE_LGM_MAX points
// (registers of 2^10 .. 2^16 points).
#define FUSE_LGM_MIN 5
#define FUSE_LGM_MAX 8

// Every table of the program is read by row in tiles of geometry d (the test
// stage0_pre makes per table, made here on the host): its field misses the
// tile's column and row-block bits and holds all or none of the row bits, and
// no control or choice bit of it varies inside the tile.
static bool tile_tables_by_row(const FastDiag& pre, const FftPass& d) {
    const int rowLo = d.lgInR, rowHi = d.lgInR + d.lgM;
    auto hits = [](int sh, int nb, int lo, int n) { return n > 0 && sh < lo + n && lo < sh + nb; };
    auto varies = [&](int b) {
        return b >= 0 && (b < d.lgC || (b >= rowLo && b < rowHi) ||
                          (d.lgHb > 0 && b >= d.lgInH2 && b < d.lgInH2 + d.lgHb));
    };
    for (int i = 0; i < pre.ntab; ++i) {
        const DiagTab& T = pre.tab[i];
        const bool rowIn = T.sh <= rowLo && rowHi <= T.sh + T.nb;
        bool ctrlVar = varies(T.ctrl);
        for (int j = 0; j < T.nsel; ++j) ctrlVar = ctrlVar || varies(tab_sel(T, j));
        const bool fast = !hits(T.sh, T.nb, 0, d.lgC) && !hits(T.sh, T.nb, d.lgInH2, d.lgHb) &&
                          (rowIn || !hits(T.sh, T.nb, rowLo, d.lgM)) && !ctrlVar;
        if (!fast) return false;
    }
    return true;
}

// Every generic term is affine in the row index of a tile whose rows are the
// bits [rowLo, rowLo + rowN) (column and row-block bits fixed): each field holds
// all of the row bits or none (a bilinear term at most one field with them, a
// polynomial with them only its linear coefficient), and no control or Z bit
// is a row bit.
static bool gen_affine_in_rows(const FastDiag& pre, int rowLo, int rowN) {
    auto rel = [&](int sh, int nb) {  // 0: misses the rows, 1: holds them all, 2: part of them
        if (!(sh < rowLo + rowN && rowLo < sh + nb)) return 0;
        return (sh <= rowLo && rowLo + rowN <= sh + nb) ? 1 : 2;
    };
    for (int i = 0; i < pre.ngen; ++i) {
        const DiagTerm& T = pre.gen[i];
        if (T.ctrl >= rowLo && T.ctrl < rowLo + rowN) return false;
        if (T.kind == DT_POLY || T.kind == DT_LINF) {
            const int r = rel(T.sh1, T.nb1);
            if (r == 2 || (r == 1 && (T.c2 != 0.0 || T.c3 != 0.0))) return false;
        } else if (T.kind == DT_BILIN) {
            const int r1 = rel(T.sh1, T.nb1), r2 = rel(T.sh2, T.nb2);
            if (r1 == 2 || r2 == 2 || (r1 == 1 && r2 == 1)) return false;
        } else if (T.sh1 >= rowLo && T.sh1 < rowLo + rowN) {
            return false;
        }
    }
    return true;
}

template <typename R, int LGM, int LGC>
static void launch_fused_geom(CVDVContext* ctx, typename CT<R>::type* data, const FftPass& d, const FftPass& dD,
                              const FastDiag& pre, int conjMid, int swapSh, int swapN,
                              const typename CT<R>::type* tw, const typename CT<R>::type* tw4, uint64_t blocks,
                              int tile, size_t lds) {
    if constexpr (tile_instance<R, LGM, LGC>()) {
        if (tile / 8 > tile_threads_max<R, LGM, LGC>())
            throw CvdvError{CVDV_ERR_UNSUPPORTED, "fused FFT tile larger than its kernel's workgroup bound"};
        // tiles of several columns whose tables are all read by row (and, in
        // complex64, whose generic terms are affine in the row index) take the
        // instance that stages them through LDS: fewer registers, more
        // workgroups per CU
        FftPass dDg = dD;
        dDg.genSteps = dD.usePre && pre.ngen > 0 && gen_affine_in_rows(pre, dD.lgInR, LGM) ? 1 : 0;
        if constexpr (LGC > 0) {
            const bool gen = sizeof(R) == 8 || pre.ngen == 0 || dDg.genSteps;
            if (!dD.usePre || (tile_tables_by_row(pre, dD) && gen)) {
                hipLaunchKernelGGL((k_fft_fused<R, LGM, LGC, 8, true>), dim3((unsigned)blocks), dim3(tile / 8), lds,
                                   ctx->stream, data, d, dDg, pre, conjMid, swapSh, swapN, tw, tw4);
                return;
            }
        }
        hipLaunchKernelGGL((k_fft_fused<R, LGM, LGC, 8, false>), dim3((unsigned)blocks), dim3(tile / 8), lds,
                           ctx->stream, data, d, dDg, pre, conjMid, swapSh, swapN, tw, tw4);
    } else {
        throw CvdvError{CVDV_ERR_UNSUPPORTED, "fused FFT tile shape without a kernel instance"};
    }
}

template <typename R, int LGM>
static void launch_fused_lgm(CVDVContext* ctx, typename CT<R>::type* data, const FftPass& d, const FftPass& dD,
                             const FastDiag& pre, int conjMid, int swapSh, int swapN,
                             const typename CT<R>::type* tw, const typename CT<R>::type* tw4, uint64_t blocks,
                             int tile, size_t lds) {
    switch (d.lgC) {
#define CVDV_FUSE_C(L) \
    case L: launch_fused_geom<R, LGM, L>(ctx, data, d, dD, pre, conjMid, swapSh, swapN, tw, tw4, blocks, tile, lds); break;
        CVDV_FUSE_C(0) CVDV_FUSE_C(1) CVDV_FUSE_C(2) CVDV_FUSE_C(3) CVDV_FUSE_C(4) CVDV_FUSE_C(5)
#undef CVDV_FUSE_C
        default: throw CvdvError{CVDV_ERR_UNSUPPORTED, "fused FFT tile column count out of range"};
    }
}

template <typename R>
static void launch_fused(CVDVContext* ctx, const FftPass& d, const FftPass& dD, const FastDiag& pre, int conjMid,
                         int swapSh, int swapN, uint64_t nH1) {
    using C = typename CT<R>::type;
    const int prec = ctx->prec;
    const C* tw = (const C*)stage_table(ctx, d.lgM, prec);
    const C* tw4 = (const C*)full_table(ctx, d.tw4LogN, prec);
    const int tile = 1 << (d.lgM + d.lgC + d.lgHb);
    const uint64_t blocks = nH1 << (d.lgNH2Blk + d.lgNLoBlk);
    FftPass dd = d;
    dd.nTiles = blocks;
    const size_t lds = lds_bytes(tile, sizeof(C));
    C* data = (C*)ctx->state();
    const int pid = prof_begin(ctx, KK_FFT_FUSED);
    switch (d.lgM) {
#define CVDV_FUSE_M(L) \
    case L: launch_fused_lgm<R, L>(ctx, data, dd, dD, pre, conjMid, swapSh, swapN, tw, tw4, blocks, tile, lds); break;
        CVDV_FUSE_M(5) CVDV_FUSE_M(6) CVDV_FUSE_M(7) CVDV_FUSE_M(8)
#undef CVDV_FUSE_M
        default: throw CvdvError{CVDV_ERR_UNSUPPORTED, "fused FFT digit length out of range"};
    }
    HIPCHK(hipGetLastError());
    prof_end(ctx, pid);
}

static uint64_t bit_field(int sh, int n) { return n > 0 ? (((1ull << n) - 1ull) << sh) : 0ull; }

// The four-step of this register can leave halves pending: equal digits, fused
// instances for the digit length, and tiles that store straight to global memory.
static bool fusable(const CVDVContext* ctx, const FourStep& f, int nb, const FftCfg& cfg) {
    return ctx->fuseFt && cfg.fuse != 0 && !ctx->rec && f.n1 == f.n2 && f.n1 >= FUSE_LGM_MIN && f.n1 <= FUSE_LGM_MAX && f.hasL &&
           f.d1.directOut && f.dL.directOut && nb == 2 * f.n1;
}

// Complete the half-done transform D (its record already cleared).
template <typename R>
static void settle_impl(CVDVContext* ctx, const FtDefer D) {
    using C = typename CT<R>::type;
    FourStep f;
    if (!four_step_geom(ctx->prec, ctx->totalSize, D.sh, D.nb, f, field_cfg(ctx, ctx->totalSize, D.sh, D.nb)))
        throw CvdvError{CVDV_ERR_STATE, "half-done register transform without its four-step geometry"};
    C* data = (C*)ctx->state();
    if (D.high) {
        // twiddle W_N^{lo hi} and the transform over hi, in place: natural order
        FftPass d = f.d1c;
        d.usePre = 0;
        d.conjLoad = D.inv;
        d.conjIn = 0;
        d.conjOut = D.inv;
        d.useScale = D.scale != 1.0 ? 1 : 0;
        d.outScale = D.scale;
        launch_pass<R>(ctx, data, data, d, FastDiag{}, ctx->prec, f.H);
    } else {
        // twiddle and the transform over lo, stored transposed into the other buffer
        FftPass d = f.d2;
        d.usePre = 0;
        d.conjLoad = D.inv;
        d.conjIn = 0;
        d.conjOut = D.inv;
        d.useScale = D.scale != 1.0 ? 1 : 0;
        d.outScale = D.scale;
        launch_pass<R>(ctx, data, (C*)ctx->other(), d, FastDiag{}, ctx->prec, f.H);
        ctx->cur = 1 - ctx->cur;
    }
}

static void settle_one(CVDVContext* ctx, FtDefer& rec) {
    const FtDefer D = rec;
    rec = FtDefer{};
    if (ctx->prec == CVDV_C128) settle_impl<double>(ctx, D);
    else settle_impl<float>(ctx, D);
}

void fft_settle(CVDVContext* ctx, bool mayAllocate) {
    if (!ctx->defer.any()) return;
    if (!mayAllocate && !ctx->buf[1 - ctx->cur])
        for (const FtDefer& d : ctx->defer.f)
            if (d.on && !d.high)
                throw CvdvError{CVDV_ERR_STATE,
                                "a register transform is half done and completing it needs the second state "
                                "buffer: call cvdvFlush first"};
    for (FtDefer& d : ctx->defer.f)
        if (d.on) settle_one(ctx, d);
}

void fft_settle_bits(CVDVContext* ctx, uint64_t mask) {
    for (FtDefer& d : ctx->defer.f)
        if (d.on && (mask & bit_field(d.sh, d.nb))) settle_one(ctx, d);
}

// A transform of field [sh, sh+nb) that can keep a half pending.  Returns false
// when the ordinary path must run (the caller has settled nothing yet).
template <typename R>
static bool fft_deferred(CVDVContext* ctx, int sh, int nb, bool inverse, const DiagProg& preProg, double outScale) {
    // only where the ordinary path takes the four-step split (fft_run)
    const uint64_t N = 1ull << nb, S = 1ull << sh, total = ctx->totalSize;
    if (total < 4096 || N <= 64 || (total % (N * S)) != 0) return false;
    const FftCfg cfg = field_cfg(ctx, total, sh, nb);
    {
        PassGeom g1{(int)N, S, total / (N * S), 1, 0, N * S, S, 0, N * S, S};
        FftPass d{};
        if (fit_pass(g1, ctx->prec, d, true, cfg)) return false;
    }
    FourStep f;
    if (!four_step_geom(ctx->prec, total, sh, nb, f, cfg) || !fusable(ctx, f, nb, cfg)) return false;
    const int n = f.n1;
    bool act = false, folded = false;
    if (FtDefer* rec = ctx->defer.find(sh)) {
        FtDefer& D = *rec;
        // fuse the pending half with this transform's first half
        const bool high = D.high != 0;
        FftPass d = high ? f.d1c : f.dL;
        FftPass dD = d;  // the tile in the natural index of the completed transform
        if (!high) {
            dD.lgInR = sh + n;   // rows: the lo digit here is the momentum's hi digit
            dD.lgInH2 = sh;      // row blocks: the hi digit here is the momentum's lo digit
        }
        const uint64_t natBits = tile_bits(dD);
        const double scale = outScale * D.scale;
        if (diag_fits(preProg, natBits)) {
            const FastDiag pre = compile_diag(ctx, preProg, act, scale, &folded, natBits);
            d.usePre = 0;
            d.conjLoad = 0;  // the kernel's conjMid conjugates around the pending half
            d.conjIn = 0;
            d.conjOut = inverse ? 1 : 0;
            d.useScale = folded ? 0 : 1;
            d.outScale = scale;
            dD.tw4 = 0;
            dD.usePre = act ? 1 : 0;
            dD.conjIn = inverse ? 1 : 0;
            launch_fused<R>(ctx, d, dD, pre, D.inv, sh, high ? 0 : n, f.H);
            D = FtDefer{1, sh, nb, high ? 0 : 1, inverse ? 1 : 0, 1.0};
            return true;
        }
        settle_one(ctx, D);
    }
    // first half from the natural order: pending phases on load, transform over hi
    FftPass d = f.d1;
    const FastDiag pre = compile_diag(ctx, preProg, act, outScale, &folded, tile_bits(d));
    d.usePre = act ? 1 : 0;
    d.conjIn = inverse ? 1 : 0;
    d.conjOut = inverse ? 1 : 0;  // a half-done inverse stays a linear map of the state
    d.useScale = folded ? 0 : 1;
    d.outScale = outScale;
    FtDefer* slot = ctx->defer.free_slot();
    if (!slot) throw CvdvError{CVDV_ERR_STATE, "no free record for a half-done register transform"};
    launch_pass<R>(ctx, (typename CT<R>::type*)ctx->state(), (typename CT<R>::type*)ctx->state(), d, pre, ctx->prec,
                   f.H);
    *slot = FtDefer{1, sh, nb, 0, inverse ? 1 : 0, 1.0};
    return true;
}

bool fft_defers(const CVDVContext* ctx, int sh, int nb) {
    const uint64_t N = 1ull << nb, S = 1ull << sh, total = ctx->totalSize;
    if (total < 4096 || N <= 64 || (total % (N * S)) != 0) return false;
    const FftCfg cfg = field_cfg(ctx, total, sh, nb);
    PassGeom g1{(int)N, S, total / (N * S), 1, 0, N * S, S, 0, N * S, S};
    FftPass d{};
    if (fit_pass(g1, ctx->prec, d, true, cfg)) return false;
    FourStep f;
    return four_step_geom(ctx->prec, total, sh, nb, f, cfg) && fusable(ctx, f, nb, cfg);
}

uint64_t fft_plan_signature(const CVDVContext* ctx, int sh, int nb, const FftCfg& cfg) {
    const uint64_t N = 1ull << nb, S = 1ull << sh, total = ctx->totalSize;
    if (total % (N * S) != 0) return 0;
    auto sig = [](const FftPass& d) { return (uint64_t)(d.lgC * 64 + d.lgHb * 4 + d.lgM / 8 + 1); };
    PassGeom g1{(int)N, S, total / (N * S), 1, 0, N * S, S, 0, N * S, S};
    FftPass d{};
    if (fit_pass(g1, ctx->prec, d, true, cfg)) return 1000000 + sig(d);
    if (total < 4096 || N <= 64) return 0;
    FourStep f;
    if (!four_step_geom(ctx->prec, total, sh, nb, f, cfg)) return 0;
    const uint64_t fz = fusable(ctx, f, nb, cfg) ? 1 : 0;
    return 2000000 + ((sig(f.d1) * 4096 + sig(f.d2)) * 4096 + (f.hasL ? sig(f.dL) : 0)) * 2 + fz;
}

// Every twiddle table an FFT of any register of the context can ask for.
void fft_prepare(CVDVContext* ctx) {
    for (int r = 0; r < ctx->numReg; ++r) {
        const int nb = ctx->nq[r];
        if (nb == 0 || nb > 28) continue;  // above 28 qubits the two-pass FFT refuses anyway
        for (int lg = 1; lg <= nb && lg <= 14; ++lg) (void)stage_table(ctx, lg, ctx->prec);
        if (ctx->prec == CVDV_MIXED)
            for (int lg = 13; lg <= nb && lg <= 14; ++lg) (void)split_table(ctx, lg, ctx->prec);
        (void)full_table(ctx, nb, ctx->prec);                   // four-step twiddles
        if (nb <= 12) (void)full_table(ctx, nb, CVDV_C128);     // direct DFT of tiny states
    }
}

void fft_register(CVDVContext* ctx, int pshift, int nbits, bool inverse, const DiagProg& pre,
                  double outScale) {
    const bool done = ctx->prec == CVDV_C128 ? fft_deferred<double>(ctx, pshift, nbits, inverse, pre, outScale)
                                             : fft_deferred<float>(ctx, pshift, nbits, inverse, pre, outScale);
    if (done) return;
    if (FtDefer* rec = ctx->defer.find(pshift)) settle_one(ctx, *rec);
    bool swapped;
    if (ctx->prec == CVDV_C128)
        swapped = fft_run<double>(ctx, (double2*)ctx->state(), [ctx] { return (double2*)ctx->other(); },
                                  ctx->totalSize, pshift, nbits, inverse, pre, outScale, CVDV_C128);
    else
        swapped = fft_run<float>(ctx, (float2*)ctx->state(), [ctx] { return (float2*)ctx->other(); },
                                 ctx->totalSize, pshift, nbits, inverse, pre, outScale, CVDV_MIXED);
    if (swapped) ctx->cur = 1 - ctx->cur;
}

void fft_generic_c128(CVDVContext* ctx, double2* data, double2* tmp, uint64_t total, int shift,
                      int nbits, bool inverse, double outScale) {
    if (fft_run<double>(ctx, data, [tmp] { return tmp; }, total, shift, nbits, inverse, DiagProg{},
                        outScale, CVDV_C128))
        HIPCHK(hipMemcpyAsync(data, tmp, total * sizeof(double2), hipMemcpyDeviceToDevice, ctx->stream));
}
