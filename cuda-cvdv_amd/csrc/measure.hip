// measure.hip - marginals, norm, overlaps and photon number.
//
// Reference interfaces replaced: src/cuda/measure.cu:15-123 (measureMultiple,
// inner products, expectX2) and :276-366 (getNorm, getFidelity,
// getPhotonNumber, fidelityStatevectors); kernels measure.cuh:33-252.
//
// Marginal P[o] = sum_{g -> o} |psi_g|^2.  The flat index bits split into
// selected (S) and unselected (U) bits.  A workgroup of 2^TB threads takes the
// lowest TB index bits from its thread id (so every wave reads contiguous
// amplitudes), one combination of the S bits above TB from its group id, and
// a contiguous share of the U-bit combinations above TB, which each thread
// walks in increasing order.  Every thread therefore accumulates a single bin
// (Neumaier-compensated FP64), a fixed-order tree joins the threads that share
// a bin, and a second pass adds the per-split partials in fixed order.  No
// atomics: results are bitwise reproducible run to run.

#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "cvdv_internal.h"
#include "../../include/cvdv_hip.h"

struct KSum {
    double s, c;
};

__host__ __device__ inline void ksum_add(double& s, double& c, double v) {
    const double t = s + v;
    if (fabs(s) >= fabs(v)) c += (s - t) + v;
    else c += (v - t) + s;
    s = t;
}

struct MargArgs {
    uint64_t wgSmask;    // selected bits at or above TB
    uint64_t loopUmask;  // unselected bits at or above TB
    uint64_t lowUmask;   // unselected bits below TB
    int TB;
    int usplit;
    uint64_t loopPerSplit;
    uint64_t outSize;
    int nsel;
    int sh[CVDV_MAX_REG];
    int nb[CVDV_MAX_REG];
    uint64_t flip[CVDV_MAX_REG];
    uint64_t stride[CVDV_MAX_REG];
};

__device__ __forceinline__ uint64_t deposit(uint64_t v, uint64_t mask) {
    uint64_t r = 0;
    while (mask) {
        const uint64_t low = mask & (~mask + 1ull);
        if (v & 1ull) r |= low;
        v >>= 1;
        mask &= mask - 1ull;
    }
    return r;
}

template <typename C>
__device__ __forceinline__ double abs2d(C v) {
    const double x = (double)v.x, y = (double)v.y;
    return x * x + y * y;
}

template <typename C>
__global__ void __launch_bounds__(256) k_marginal(const C* __restrict__ st, MargArgs a, double* __restrict__ part) {
    __shared__ double ss[256];
    __shared__ double cc[256];
    const int tid = threadIdx.x;
    const uint64_t group = blockIdx.x / a.usplit;
    const uint64_t split = blockIdx.x % a.usplit;
    const uint64_t gbase = deposit(group, a.wgSmask) | (uint64_t)tid;
    uint64_t lp = deposit(split * a.loopPerSplit, a.loopUmask);
    double s = 0.0, c = 0.0;
    uint64_t it = 0;
    const uint64_t L = a.loopPerSplit;
    for (; it + 4 <= L; it += 4) {
        const uint64_t l0 = lp;
        const uint64_t l1 = (l0 - a.loopUmask) & a.loopUmask;
        const uint64_t l2 = (l1 - a.loopUmask) & a.loopUmask;
        const uint64_t l3 = (l2 - a.loopUmask) & a.loopUmask;
        lp = (l3 - a.loopUmask) & a.loopUmask;
        const C v0 = st[gbase | l0], v1 = st[gbase | l1], v2 = st[gbase | l2], v3 = st[gbase | l3];
        ksum_add(s, c, abs2d(v0));
        ksum_add(s, c, abs2d(v1));
        ksum_add(s, c, abs2d(v2));
        ksum_add(s, c, abs2d(v3));
    }
    for (; it < L; ++it) {
        ksum_add(s, c, abs2d(st[gbase | lp]));
        lp = (lp - a.loopUmask) & a.loopUmask;
    }
    ss[tid] = s;
    cc[tid] = c;
    // join threads that differ only in unselected low bits, lowest bit first
    for (int b = 0; b < a.TB; ++b) {
        __syncthreads();
        if (((a.lowUmask >> b) & 1ull) && !((tid >> b) & 1)) {
            const int o = tid | (1 << b);
            if (!((uint64_t)tid & (a.lowUmask & ((1ull << b) - 1ull)))) {
                double s2 = ss[tid], c2 = cc[tid];
                ksum_add(s2, c2, ss[o]);
                c2 += cc[o];
                ss[tid] = s2;
                cc[tid] = c2;
            }
        }
    }
    __syncthreads();
    if ((uint64_t)tid & a.lowUmask) return;
    uint64_t o = 0;
    for (int r = 0; r < a.nsel; ++r) {
        const uint64_t j = ((gbase >> a.sh[r]) & ((1ull << a.nb[r]) - 1ull)) ^ a.flip[r];
        o += j * a.stride[r];
    }
    part[split * a.outSize + o] = ss[tid] + cc[tid];
}

__global__ void k_merge_splits(const double* __restrict__ part, double* __restrict__ out,
                               uint64_t outSize, int usplit) {
    for (uint64_t o = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; o < outSize;
         o += (uint64_t)gridDim.x * blockDim.x) {
        double s = 0.0, c = 0.0;
        for (int k = 0; k < usplit; ++k) ksum_add(s, c, part[(uint64_t)k * outSize + o]);
        out[o] = s + c;
    }
}

// Many splits and few bins (a qubit's marginal: 2 bins, thousands of splits):
// one workgroup per bin, each thread summing a strided share of the partials,
// then a fixed-order tree over the threads (compensated; deterministic).
__global__ void __launch_bounds__(256) k_merge_splits_bin(const double* __restrict__ part,
                                                          double* __restrict__ out, uint64_t outSize,
                                                          int usplit) {
    __shared__ double ss[256];
    __shared__ double cc[256];
    const uint64_t o = blockIdx.x;
    const int tid = threadIdx.x;
    double s = 0.0, c = 0.0;
    for (int k = tid; k < usplit; k += 256) ksum_add(s, c, part[(uint64_t)k * outSize + o]);
    ss[tid] = s;
    cc[tid] = c;
    for (int w = 128; w > 0; w >>= 1) {
        __syncthreads();
        if (tid < w) {
            double s2 = ss[tid], c2 = cc[tid] + cc[tid + w];
            ksum_add(s2, c2, ss[tid + w]);
            ss[tid] = s2;
            cc[tid] = c2;
        }
    }
    if (tid == 0) out[o] = ss[0] + cc[0];
}

static int popc(uint64_t m) { return __builtin_popcountll(m); }

// Marginal over `regs` (logical registers; output column-major, first fastest).
// The per-split partials go to `fixedScratch` (capElems doubles; the split count
// is capped to fit, and a marginal whose bins alone do not fit is refused) or,
// when it is null, to the context's scratch buffer (grown on demand).
static void marginal_core(CVDVContext* ctx, const int* regs, int nsel, double* hostOut, double* fixedScratch,
                          size_t capElems) {
    // a half-done register transform is completed first (without allocating on
    // the fixed-scratch path)
    fft_settle(ctx, fixedScratch == nullptr);
    const int T = ctx->totalQbt;
    const int TB = T < 8 ? T : 8;
    const uint64_t all = (T >= 64) ? ~0ull : ((1ull << T) - 1ull);
    const uint64_t low = (1ull << TB) - 1ull;
    MargArgs a{};
    uint64_t smask = 0, outSize = 1;
    a.nsel = nsel;
    for (int i = 0; i < nsel; ++i) {
        const int r = regs[i];
        const uint64_t fm = (1ull << ctx->nq[r]) - 1ull;
        smask |= fm << ctx->pshift[r];
        a.sh[i] = ctx->pshift[r];
        a.nb[i] = ctx->nq[r];
        a.flip[i] = ctx->refl[r] ? fm : 0;
        a.stride[i] = outSize;
        outSize <<= ctx->nq[r];
    }
    a.TB = TB;
    a.wgSmask = smask & ~low & all;
    a.loopUmask = ~smask & ~low & all;
    a.lowUmask = ~smask & low;
    a.outSize = outSize;
    const uint64_t nGroups = 1ull << popc(a.wgSmask);
    const uint64_t nLoop = 1ull << popc(a.loopUmask);
    if (fixedScratch && outSize > capElems)
        throw CvdvError{CVDV_ERR_ARG, "marginal of " + std::to_string(outSize) + " bins exceeds the fixed scratch of " +
                                          std::to_string(capElems) + " doubles"};
    // partials + merged output: usplit * outSize (+ outSize when usplit > 1)
    auto fits = [&](uint64_t us) { return !fixedScratch || (us > 1 ? us + 1 : 1) * outSize <= capElems; };
    uint64_t usplit = 1;
    while (usplit * 2 <= nLoop && nGroups * usplit < 4096 && usplit * 2 * outSize <= (32ull << 20) &&
           fits(usplit * 2))
        usplit *= 2;
    a.usplit = (int)usplit;
    a.loopPerSplit = nLoop / usplit;

    double* part = fixedScratch ? fixedScratch
                                : (double*)ctx_scratch(ctx, (usplit + 1) * outSize * sizeof(double));
    double* fin = part + usplit * outSize;
    const unsigned blocks = (unsigned)(nGroups * usplit);
    const unsigned threads = 1u << TB;
    // event timing creates events: not on the fixed-scratch path
    const int pid = fixedScratch ? -1 : prof_begin(ctx, KK_REDUCE);
    if (ctx->prec == CVDV_C128)
        hipLaunchKernelGGL(k_marginal<double2>, dim3(blocks), dim3(threads), 0, ctx->stream,
                           (const double2*)ctx->state(), a, part);
    else
        hipLaunchKernelGGL(k_marginal<float2>, dim3(blocks), dim3(threads), 0, ctx->stream,
                           (const float2*)ctx->state(), a, part);
    HIPCHK(hipGetLastError());
    if (pid >= 0) prof_end(ctx, pid);
    const double* res = part;
    if (usplit > 1) {
        if (usplit >= 256 && outSize <= 4096) {
            hipLaunchKernelGGL(k_merge_splits_bin, dim3((unsigned)outSize), dim3(256), 0, ctx->stream, part,
                               fin, outSize, (int)usplit);
        } else {
            uint64_t b = (outSize + 255) / 256;
            if (b > 4096) b = 4096;
            hipLaunchKernelGGL(k_merge_splits, dim3((unsigned)b), dim3(256), 0, ctx->stream, part, fin,
                               outSize, (int)usplit);
        }
        HIPCHK(hipGetLastError());
        res = fin;
    }
    HIPCHK(hipMemcpyAsync(hostOut, res, outSize * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
}

void marginal_probs(CVDVContext* ctx, const int* regs, int nsel, double* hostOut) {
    marginal_core(ctx, regs, nsel, hostOut, nullptr, 0);
}

// Fixed scratch of the allocation-free marginal (cvdvMeasureMultipleNoAlloc):
// every split partial and the merged bins of one call.  Each GPU the library
// runs on holds its own copy until the process ends.  Every entry a call reads
// is written by that call first (each split writes every bin), so nothing needs
// clearing between calls.
#define CVDV_NOALLOC_SCRATCH_DOUBLES (1u << 20)  // 8 MiB
static __device__ __align__(256) double scratch[CVDV_NOALLOC_SCRATCH_DOUBLES];

// Device address of `scratch` on the current device, looked up once per device
// (the host-side name of a __device__ array is not a device address).
static double* noalloc_scratch() {
    static std::mutex mu;
    static std::map<int, double*> ptrs;
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    auto it = ptrs.find(dev);
    if (it != ptrs.end()) return it->second;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        throw CvdvError{CVDV_ERR_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e)};
    ptrs[dev] = (double*)ptr;
    return (double*)ptr;
}

// ---------------------------------------------------------------- overlaps
struct SepDot {
    const double2* arr[CVDV_MAX_REG];
    int sh[CVDV_MAX_REG];
    int nb[CVDV_MAX_REG];
    int refl[CVDV_MAX_REG];
    int nreg;
};

struct CSum {
    double sx, cx, sy, cy;
};

__device__ void block_reduce_store(CSum v, CSum* out) {
    __shared__ double sh[4][256];
    const int t = threadIdx.x;
    sh[0][t] = v.sx; sh[1][t] = v.cx; sh[2][t] = v.sy; sh[3][t] = v.cy;
    for (int w = 128; w > 0; w >>= 1) {
        __syncthreads();
        if (t < w) {
            double a = sh[0][t], b = sh[1][t] + sh[1][t + w];
            ksum_add(a, b, sh[0][t + w]);
            sh[0][t] = a; sh[1][t] = b;
            a = sh[2][t]; b = sh[3][t] + sh[3][t + w];
            ksum_add(a, b, sh[2][t + w]);
            sh[2][t] = a; sh[3][t] = b;
        }
    }
    if (t == 0) out[blockIdx.x] = CSum{sh[0][0], sh[1][0], sh[2][0], sh[3][0]};
}

// sum_g conj(psi_g) * prod_r phi_r[logical local index]   (measure.cuh:182-224)
template <typename C>
__global__ void __launch_bounds__(256) k_dot_sep(const C* __restrict__ st, uint64_t total, DiagProg p, int usep,
                                                 SepDot sd, CSum* out) {
    // the register loop indexes the descriptor at run time: staged in LDS, not
    // copied to per-lane scratch memory
    __shared__ SepDot s;
    if (threadIdx.x == 0) s = sd;
    __syncthreads();
    CSum acc{0, 0, 0, 0};
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
         g += (uint64_t)gridDim.x * blockDim.x) {
        C v = st[g];
        if (usep) v = apply_diag(p, g, v);
        double2 t = make_double2(1.0, 0.0);
        for (int r = 0; r < s.nreg; ++r) {
            const uint64_t m = (1ull << s.nb[r]) - 1ull;
            uint64_t j = (g >> s.sh[r]) & m;
            if (s.refl[r]) j = m - j;
            t = cmul(t, s.arr[r][j]);
        }
        const double2 pr = cmulc(t, make_double2((double)v.x, (double)v.y));  // conj(psi) * t
        ksum_add(acc.sx, acc.cx, pr.x);
        ksum_add(acc.sy, acc.cy, pr.y);
    }
    block_reduce_store(acc, out);
}

// apply_diag for a program in LDS: a plain loop over its terms (the unrolled
// form would hold every term of both programs in registers across the loop).
template <typename C>
__device__ __forceinline__ C apply_diag_lds(const DiagProg& p, uint64_t g, C v) {
    using R = decltype(v.x);
    double r = p.grev;
#pragma unroll 1
    for (int i = 0; i < p.n; ++i) r += cvdv_frac(diag_term_rev(p.t[i], g));
    return cmul(v, unit_rev(cvdv_frac(r), R(0)));
}

// sum_g conj(a_g) b_g   (measure.cuh:227-252)
template <typename CA, typename CB>
__global__ void __launch_bounds__(256) k_dot_states(const CA* __restrict__ a, const CB* __restrict__ b, uint64_t total,
                                                    DiagProg pa, int usea, DiagProg pb, int useb, CSum* out) {
    // two programs held in registers would spill: they are read from LDS
    __shared__ DiagProg spa, spb;
    if (threadIdx.x == 0) spa = pa;
    if (threadIdx.x == 1) spb = pb;
    __syncthreads();
    CSum acc{0, 0, 0, 0};
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
         g += (uint64_t)gridDim.x * blockDim.x) {
        CA va = a[g];
        CB vb = b[g];
        if (usea) va = apply_diag_lds(spa, g, va);
        if (useb) vb = apply_diag_lds(spb, g, vb);
        const double2 pr = cmulc(make_double2((double)vb.x, (double)vb.y),
                                 make_double2((double)va.x, (double)va.y));
        ksum_add(acc.sx, acc.cx, pr.x);
        ksum_add(acc.sy, acc.cy, pr.y);
    }
    block_reduce_store(acc, out);
}

static unsigned dot_grid(uint64_t total) {
    uint64_t b = (total + 255) / 256;
    return (unsigned)(b > 2048 ? 2048 : (b ? b : 1));
}

static void finish_dot(CVDVContext* ctx, CSum* d, unsigned nb, double& re, double& im) {
    std::vector<CSum> h(nb);
    HIPCHK(hipMemcpyAsync(h.data(), d, nb * sizeof(CSum), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    double sx = 0, cx = 0, sy = 0, cy = 0;
    for (unsigned i = 0; i < nb; ++i) {
        ksum_add(sx, cx, h[i].sx);
        cx += h[i].cx;
        ksum_add(sy, cy, h[i].sy);
        cy += h[i].cy;
    }
    re = sx + cx;
    im = sy + cy;
}

static inline bool prog_on(const DiagProg& p) { return p.n > 0 || p.grev != 0.0; }

// ---------------------------------------------------------------- C ABI
#define CVDV_TRY try {
#define CVDV_CATCH(ret)                                                      \
    }                                                                        \
    catch (const CvdvError& e) {                                             \
        cvdv_set_error(e.code, e.msg);                                       \
        return ret;                                                          \
    }                                                                        \
    catch (...) {                                                            \
        cvdv_set_error(CVDV_ERR_HIP, "unexpected internal error");           \
        return ret;                                                          \
    }

extern "C" {

void cvdvMeasureMultiple(CVDVContext* ctx, const int* regIdxs, int numRegs, double* probsOut) {
    if (!ctx || !regIdxs || numRegs <= 0 || !probsOut || refuse_recording(ctx)) return;
    if (!ctx->initialized) {
        cvdv_set_error(CVDV_ERR_STATE, "State not initialized");
        return;
    }
    uint64_t seen = 0;
    for (int i = 0; i < numRegs; i++) {
        if (regIdxs[i] < 0 || regIdxs[i] >= ctx->numReg) {
            cvdv_set_error(CVDV_ERR_ARG, "Invalid register index: " + std::to_string(regIdxs[i]));
            return;
        }
        if (seen & (1ull << regIdxs[i])) {
            cvdv_set_error(CVDV_ERR_ARG, "Register listed twice: " + std::to_string(regIdxs[i]));
            return;
        }
        seen |= 1ull << regIdxs[i];
    }
    CVDV_TRY marginal_probs(ctx, regIdxs, numRegs, probsOut); CVDV_CATCH()
}

void cvdvMeasureMultipleNoAlloc(CVDVContext* ctx, const int* regIdxs, int numRegs, double* probsOut) {
    if (!ctx || !regIdxs || numRegs <= 0 || !probsOut || refuse_recording(ctx)) return;
    if (!ctx->initialized) {
        cvdv_set_error(CVDV_ERR_STATE, "State not initialized");
        return;
    }
    uint64_t seen = 0;
    for (int i = 0; i < numRegs; i++) {
        if (regIdxs[i] < 0 || regIdxs[i] >= ctx->numReg) {
            cvdv_set_error(CVDV_ERR_ARG, "Invalid register index: " + std::to_string(regIdxs[i]));
            return;
        }
        if (seen & (1ull << regIdxs[i])) {
            cvdv_set_error(CVDV_ERR_ARG, "Register listed twice: " + std::to_string(regIdxs[i]));
            return;
        }
        seen |= 1ull << regIdxs[i];
    }
    CVDV_TRY
    double* ptr = noalloc_scratch();
    marginal_core(ctx, regIdxs, numRegs, probsOut, ptr, CVDV_NOALLOC_SCRATCH_DOUBLES);
    CVDV_CATCH()
}

double cvdvGetNorm(CVDVContext* ctx) {
    // measure.cu:276-300: sum |psi|^2 (not its square root)
    if (!ctx || refuse_recording(ctx)) return 0.0;
    if (!ctx->initialized) {
        cvdv_set_error(CVDV_ERR_STATE, "State not initialized");
        return 0.0;
    }
    CVDV_TRY
    double r = 0.0;
    marginal_probs(ctx, nullptr, 0, &r);
    return r;
    CVDV_CATCH(0.0)
}

void cvdvGetFidelity(CVDVContext* ctx, void** devicePtrs, int numReg, double* fidOut) {
    // measure.cu:302-318: |<psi|sep>|^2
    if (!ctx || !fidOut || refuse_recording(ctx)) return;
    if (!ctx->initialized || ctx->numReg == 0) {
        cvdv_set_error(CVDV_ERR_STATE, "State not initialized");
        *fidOut = 0.0;
        return;
    }
    if (numReg != ctx->numReg || !devicePtrs) {
        cvdv_set_error(CVDV_ERR_ARG, "numReg mismatch: context has " + std::to_string(ctx->numReg) +
                                         ", got " + std::to_string(numReg));
        *fidOut = 0.0;
        return;
    }
    CVDV_TRY
    SepDot sd{};
    for (int r = 0; r < numReg; ++r) {
        sd.arr[r] = (const double2*)devicePtrs[r];
        sd.sh[r] = ctx->pshift[r];
        sd.nb[r] = ctx->nq[r];
        sd.refl[r] = ctx->refl[r];
    }
    sd.nreg = numReg;
    fft_settle(ctx);
    HIPCHK(hipDeviceSynchronize());  // register arrays come from the caller's stream
    const uint64_t n = ctx->totalSize;
    const unsigned nb = dot_grid(n);
    CSum* d = (CSum*)ctx_scratch(ctx, nb * sizeof(CSum));
    const DiagProg& p = ctx->pending;
    if (ctx->prec == CVDV_C128)
        hipLaunchKernelGGL(k_dot_sep<double2>, dim3(nb), dim3(256), 0, ctx->stream,
                           (const double2*)ctx->state(), n, p, prog_on(p) ? 1 : 0, sd, d);
    else
        hipLaunchKernelGGL(k_dot_sep<float2>, dim3(nb), dim3(256), 0, ctx->stream,
                           (const float2*)ctx->state(), n, p, prog_on(p) ? 1 : 0, sd, d);
    HIPCHK(hipGetLastError());
    double re, im;
    finish_dot(ctx, d, nb, re, im);
    *fidOut = re * re + im * im;
    CVDV_CATCH()
}

double cvdvGetPhotonNumber(CVDVContext* ctx, int regIdx) {
    // measure.cu:320-331: (<x^2> + <p^2> - 1) / 2 through an ftQ2P / ftP2Q round trip
    if (!ctx || !ctx->initialized || refuse_recording(ctx)) return -1.0;
    if (regIdx < 0 || regIdx >= ctx->numReg) return -1.0;
    CVDV_TRY
    const int N = 1 << ctx->nq[regIdx];
    const double dx = ctx->dx[regIdx];
    std::vector<double> P(N);
    auto expect = [&]() {
        marginal_probs(ctx, &regIdx, 1, P.data());
        double s = 0.0, c = 0.0;
        for (int j = 0; j < N; ++j) {
            const double x = (j - (N - 1) * 0.5) * dx;
            ksum_add(s, c, P[j] * x * x);
        }
        return s + c;
    };
    const double q2 = expect();
    ctx_ft(ctx, regIdx, false);
    const double p2 = expect();
    ctx_ft(ctx, regIdx, true);
    return (q2 + p2 - 1.0) / 2.0;
    CVDV_CATCH(-1.0)
}

void cvdvFidelityStatevectors(CVDVContext* ctx1, CVDVContext* ctx2, double* fidOut) {
    // measure.cu:333-366: |<psi1|psi2>|^2
    if (!fidOut) return;
    *fidOut = 0.0;
    if (!ctx1 || !ctx2 || refuse_recording(ctx1) || refuse_recording(ctx2)) return;
    if (!ctx1->initialized || !ctx2->initialized) {
        cvdv_set_error(CVDV_ERR_STATE, "cvdvFidelityStatevectors: state not initialized");
        return;
    }
    if (ctx1->totalQbt != ctx2->totalQbt) {
        cvdv_set_error(CVDV_ERR_ARG, "cvdvFidelityStatevectors: total qubit count mismatch (" +
                                         std::to_string(ctx1->totalQbt) + " vs " +
                                         std::to_string(ctx2->totalQbt) + ")");
        return;
    }
    CVDV_TRY
    fft_settle(ctx1);
    fft_settle(ctx2);
    bool same = ctx1->numReg == ctx2->numReg;
    for (int r = 0; same && r < ctx1->numReg; ++r)
        same = ctx1->nq[r] == ctx2->nq[r] && ctx1->pshift[r] == ctx2->pshift[r] &&
               ctx1->refl[r] == ctx2->refl[r];
    if (!same) {
        ctx_materialize(ctx1);
        ctx_materialize(ctx2);
    }
    if (ctx1->stream != ctx2->stream) HIPCHK(hipStreamSynchronize(ctx2->stream));
    const uint64_t n = ctx1->totalSize;
    const unsigned nb = dot_grid(n);
    CSum* d = (CSum*)ctx_scratch(ctx1, nb * sizeof(CSum));
    const DiagProg &pa = ctx1->pending, &pb = ctx2->pending;
    const int ua = prog_on(pa) ? 1 : 0, ub = prog_on(pb) ? 1 : 0;
    const bool a64 = ctx1->prec == CVDV_C128, b64 = ctx2->prec == CVDV_C128;
    if (a64 && b64)
        hipLaunchKernelGGL((k_dot_states<double2, double2>), dim3(nb), dim3(256), 0, ctx1->stream,
                           (const double2*)ctx1->state(), (const double2*)ctx2->state(), n, pa, ua, pb, ub, d);
    else if (a64)
        hipLaunchKernelGGL((k_dot_states<double2, float2>), dim3(nb), dim3(256), 0, ctx1->stream,
                           (const double2*)ctx1->state(), (const float2*)ctx2->state(), n, pa, ua, pb, ub, d);
    else if (b64)
        hipLaunchKernelGGL((k_dot_states<float2, double2>), dim3(nb), dim3(256), 0, ctx1->stream,
                           (const float2*)ctx1->state(), (const double2*)ctx2->state(), n, pa, ua, pb, ub, d);
    else
        hipLaunchKernelGGL((k_dot_states<float2, float2>), dim3(nb), dim3(256), 0, ctx1->stream,
                           (const float2*)ctx1->state(), (const float2*)ctx2->state(), n, pa, ua, pb, ub, d);
    HIPCHK(hipGetLastError());
    double re, im;
    finish_dot(ctx1, d, nb, re, im);
    *fidOut = re * re + im * im;
    CVDV_CATCH()
}

}  // extern "C"
