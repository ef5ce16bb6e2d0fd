// ctx.hip - context lifecycle, errors, separable initialisation, state I/O and
// layout materialisation.
//
// Reference interfaces replaced: src/cuda/context.cu:7-165 (create, destroy,
// free, initFromSeparable, setStateFromDevicePtr, getters) and
// src/cuda/measure.cu:258-274 (getState).

#include <cmath>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <new>

#include "cvdv_internal.h"
#include "diag_fast.h"
#include "../../include/cvdv_hip.h"

// ---------------------------------------------------------------- errors
static thread_local CvdvError g_last{CVDV_OK, ""};

void cvdv_set_error(int code, const std::string& msg, bool print) {
    g_last.code = code;
    g_last.msg = msg;
    if (print) fprintf(stderr, "%s\n", msg.c_str());
}

void cvdv_check_hip(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        std::string m = std::string("HIP error '") + hipGetErrorString(e) + "' in " + what;
        throw CvdvError{e == hipErrorOutOfMemory ? CVDV_ERR_MEMORY : CVDV_ERR_HIP, m};
    }
}

#define CVDV_TRY try {
#define CVDV_CATCH(ret)                                                      \
    }                                                                        \
    catch (const CvdvError& e) {                                             \
        cvdv_set_error(e.code, e.msg);                                       \
        return ret;                                                          \
    }                                                                        \
    catch (const std::bad_alloc&) {                                          \
        cvdv_set_error(CVDV_ERR_MEMORY, "host allocation failed");           \
        return ret;                                                          \
    }                                                                        \
    catch (...) {                                                            \
        cvdv_set_error(CVDV_ERR_HIP, "unexpected internal error");           \
        return ret;                                                          \
    }

// ---------------------------------------------------------------- context helpers
void* CVDVContext::other() {
    const int o = 1 - cur;
    if (!buf[o]) {
        if (capturing) throw CvdvError{CVDV_ERR_UNSUPPORTED, "buffer allocation inside a graph capture"};
        HIPCHK(hipMalloc(&buf[o], totalSize * elemBytes()));
    }
    return buf[o];
}

bool CVDVContext::canonical() const {
    for (int r = 0; r < numReg; ++r)
        if (pshift[r] != flw[r] || refl[r]) return false;
    return true;
}

void* ctx_scratch(CVDVContext* ctx, size_t bytes) {
    if (bytes > ctx->scratchBytes) {
        if (ctx->capturing) throw CvdvError{CVDV_ERR_UNSUPPORTED, "scratch allocation inside a graph capture"};
        if (ctx->scratch) HIPCHK(hipFree(ctx->scratch));
        ctx->scratch = nullptr;
        ctx->scratchBytes = 0;
        HIPCHK(hipMalloc(&ctx->scratch, bytes));
        ctx->scratchBytes = bytes;
    }
    return ctx->scratch;
}

// ---------------------------------------------------------------- profiling
static hipEvent_t take_event(CVDVContext* ctx) {
    if (!ctx->evPool.empty()) {
        hipEvent_t e = ctx->evPool.back();
        ctx->evPool.pop_back();
        return e;
    }
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    return e;
}

int prof_begin(CVDVContext* ctx, int kind) {
    if (!ctx->profOn || ctx->capturing) return -1;
    ProfRec r{kind, take_event(ctx), take_event(ctx)};
    HIPCHK(hipEventRecord(r.a, ctx->stream));
    ctx->prof.push_back(r);
    return (int)ctx->prof.size() - 1;
}

void prof_end(CVDVContext* ctx, int idx) {
    if (idx < 0) return;
    HIPCHK(hipEventRecord(ctx->prof[idx].b, ctx->stream));
}

static void prof_clear(CVDVContext* ctx) {
    for (auto& r : ctx->prof) {
        ctx->evPool.push_back(r.a);
        ctx->evPool.push_back(r.b);
    }
    ctx->prof.clear();
}

static void ctx_release(CVDVContext* ctx) {
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    prof_clear(ctx);
    for (hipEvent_t e : ctx->evPool) (void)hipEventDestroy(e);
    ctx->evPool.clear();
    for (int i = 0; i < 2; ++i)
        if (ctx->buf[i]) (void)hipFree(ctx->buf[i]);
    ctx->buf[0] = ctx->buf[1] = nullptr;
    if (ctx->scratch) (void)hipFree(ctx->scratch);
    ctx->scratch = nullptr;
    ctx->scratchBytes = 0;
    if (ctx->arena) (void)hipFree(ctx->arena);
    ctx->arena = nullptr;
    ctx->arenaBytes = 0;
    for (void* p : ctx->retired) (void)hipFree(p);
    ctx->retired.clear();
    if (ctx->tw) twiddle_cache_free(ctx->tw);
    ctx->tw = nullptr;
    blas_release(ctx);
    ctx->generation++;  // captured graphs of the released buffers become invalid
    ctx->defer = FtDefers{};
    ctx->initialized = false;
}

static void reset_layout(CVDVContext* ctx) {
    for (int r = 0; r < ctx->numReg; ++r) {
        ctx->pshift[r] = ctx->flw[r];
        ctx->refl[r] = 0;
    }
    ctx->pending = DiagProg{};
    ctx->defer = FtDefers{};  // a new state: nothing half done
    ctx->cur = 0;
}

static void ensure_state(CVDVContext* ctx) {
    if (!ctx->buf[0]) HIPCHK(hipMalloc(&ctx->buf[0], ctx->totalSize * ctx->elemBytes()));
}

// ---------------------------------------------------------------- kernels
struct SepArgs {
    const double2* arr[CVDV_MAX_REG];
    int sh[CVDV_MAX_REG];
    int nb[CVDV_MAX_REG];
    int nreg;
};

// psi[g] = prod_r arr_r[local_r(g)].  The state is cut into chunks of 2^k
// amplitudes (k <= INIT_LO_BITS).  Registers wholly below bit k form a product
// table in LDS, built once per workgroup; registers wholly above it give one
// factor per chunk; the register (at most one) that straddles bit k is read per
// amplitude.  An amplitude then costs one LDS read, at most one gather from a
// register array (cache-resident), two FP64 complex products and one store,
// instead of one gather and one product per register.
#define INIT_LO_BITS 11
template <typename C>
__global__ void __launch_bounds__(256) k_init_product(C* __restrict__ out, uint64_t nchunks, int k,
                                                      SepArgs a) {
    __shared__ double2 lo[1 << INIT_LO_BITS];
    const int L = 1 << k;
    int s = -1;
    for (int r = 0; r < a.nreg; ++r)
        if (a.sh[r] < k && a.sh[r] + a.nb[r] > k) s = r;
    for (int l = threadIdx.x; l < L; l += blockDim.x) {
        double2 p = make_double2(1.0, 0.0);
        for (int r = 0; r < a.nreg; ++r)
            if (a.sh[r] + a.nb[r] <= k) p = cmul(p, a.arr[r][(l >> a.sh[r]) & ((1 << a.nb[r]) - 1)]);
        lo[l] = p;
    }
    __syncthreads();
    const double2* sa = s >= 0 ? a.arr[s] : nullptr;
    const int ssh = s >= 0 ? a.sh[s] : 0;
    const uint64_t smask = s >= 0 ? (1ull << a.nb[s]) - 1 : 0;
    for (uint64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const uint64_t g0 = c << k;
        double2 hp = make_double2(1.0, 0.0);
        // a 0-qubit register at bit k is in the LDS table already (sh + nb <= k)
        for (int r = 0; r < a.nreg; ++r)
            if (a.sh[r] >= k && a.nb[r] > 0) hp = cmul(hp, a.arr[r][(g0 >> a.sh[r]) & ((1ull << a.nb[r]) - 1)]);
#pragma unroll 4
        for (int l = threadIdx.x; l < L; l += blockDim.x) {
            double2 p = cmul(hp, lo[l]);
            if (sa) p = cmul(p, sa[((g0 | (uint64_t)l) >> ssh) & smask]);
            out[g0 + l] = cmk<C>(p.x, p.y);
        }
    }
}

template <typename C>
__global__ void k_from_c128(C* __restrict__ out, const double2* __restrict__ in, uint64_t total) {
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
         g += (uint64_t)gridDim.x * blockDim.x)
        out[g] = cmk<C>(in[g].x, in[g].y);
}

struct LayoutArgs {
    int pshift[CVDV_MAX_REG];
    int flw[CVDV_MAX_REG];
    int nq[CVDV_MAX_REG];
    int refl[CVDV_MAX_REG];
    int nreg;
};

__device__ __forceinline__ uint64_t canon_index(const LayoutArgs& L, uint64_t g) {
    uint64_t o = 0;
    for (int r = 0; r < L.nreg; ++r) {
        const uint64_t mask = (1ull << L.nq[r]) - 1ull;
        uint64_t j = (g >> L.pshift[r]) & mask;
        if (L.refl[r]) j = mask - j;
        o |= j << L.flw[r];
    }
    return o;
}

// out[canonical(g)] = pending(g) * in[g]
template <typename C>
__global__ void k_materialize(C* __restrict__ out, const C* __restrict__ in, uint64_t total,
                              LayoutArgs L, DiagProg p) {
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
         g += (uint64_t)gridDim.x * blockDim.x) {
        C v = in[g];
        if (p.n || p.grev != 0.0) v = apply_diag(p, g, v);
        out[canon_index(L, g)] = v;
    }
}

template <typename C>
__global__ void k_split(double* __restrict__ re, double* __restrict__ im, const C* __restrict__ in,
                        uint64_t total) {
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
         g += (uint64_t)gridDim.x * blockDim.x) {
        const C v = in[g];
        re[g] = (double)v.x;
        im[g] = (double)v.y;
    }
}

template <typename C>
__global__ void k_to_c128(double2* __restrict__ out, const C* __restrict__ in, uint64_t total) {
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
         g += (uint64_t)gridDim.x * blockDim.x)
        out[g] = make_double2((double)in[g].x, (double)in[g].y);
}

template <typename C>
__global__ void k_copy(C* __restrict__ out, const C* __restrict__ in, uint64_t total) {
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
         g += (uint64_t)gridDim.x * blockDim.x)
        out[g] = in[g];
}

static unsigned grid_for(uint64_t total, unsigned block = 256) {
    uint64_t b = (total + block - 1) / block;
    if (b > 65536ull * 8) b = 65536ull * 8;
    return (unsigned)(b ? b : 1);
}

void ctx_materialize(CVDVContext* ctx) {
    fft_settle(ctx);
    if (ctx->canonical()) {
        ctx_flush(ctx);
        return;
    }
    LayoutArgs L{};
    for (int r = 0; r < ctx->numReg; ++r) {
        L.pshift[r] = ctx->pshift[r];
        L.flw[r] = ctx->flw[r];
        L.nq[r] = ctx->nq[r];
        L.refl[r] = ctx->refl[r];
    }
    L.nreg = ctx->numReg;
    DiagProg p = ctx_take_pending(ctx);
    void* dst = ctx->other();
    const uint64_t n = ctx->totalSize;
    const int pid = prof_begin(ctx, KK_OTHER);
    if (ctx->prec == CVDV_C128)
        hipLaunchKernelGGL(k_materialize<double2>, dim3(grid_for(n)), dim3(256), 0, ctx->stream,
                           (double2*)dst, (const double2*)ctx->state(), n, L, p);
    else
        hipLaunchKernelGGL(k_materialize<float2>, dim3(grid_for(n)), dim3(256), 0, ctx->stream,
                           (float2*)dst, (const float2*)ctx->state(), n, L, p);
    HIPCHK(hipGetLastError());
    prof_end(ctx, pid);
    ctx->cur = 1 - ctx->cur;
    for (int r = 0; r < ctx->numReg; ++r) {
        ctx->pshift[r] = ctx->flw[r];
        ctx->refl[r] = 0;
    }
}

// ---------------------------------------------------------------- graph capture
// A captured gate sequence.  Capture starts from the canonical layout with no
// pending phases in buffer `cur`, and ends the same way (pending phases
// flushed, layout materialised, the result copied back to `cur` if the last
// pass left it in the other buffer), so a replay maps that state to the state
// after the sequence and replays compose.
struct CVDVGraph {
    CVDVContext* ctx;
    hipGraph_t graph;
    hipGraphExec_t exec;
    int cur;
    int generation;
};


// Copy the state between the two buffers (a kernel, so it can sit in a graph).
static void copy_to(CVDVContext* ctx, int dst) {
    const uint64_t n = ctx->totalSize;
    if (ctx->prec == CVDV_C128)
        hipLaunchKernelGGL(k_copy<double2>, dim3(grid_for(n)), dim3(256), 0, ctx->stream,
                           (double2*)ctx->buf[dst], (const double2*)ctx->buf[ctx->cur], n);
    else
        hipLaunchKernelGGL(k_copy<float2>, dim3(grid_for(n)), dim3(256), 0, ctx->stream,
                           (float2*)ctx->buf[dst], (const float2*)ctx->buf[ctx->cur], n);
    HIPCHK(hipGetLastError());
    ctx->cur = dst;
}

// ---------------------------------------------------------------- C ABI
static CVDVContext* create_impl(int numReg, const int* numQubits, int precision, bool recording = false) {
    CVDVContext* ctx = new CVDVContext{};
    if (numReg <= 0 || numQubits == nullptr) return ctx;  // context.cu:11-13
    if (numReg > CVDV_MAX_REG) {
        cvdv_set_error(CVDV_ERR_ARG, "cvdvCreate: at most 32 registers are supported");
        return ctx;
    }
    if (precision != CVDV_C128 && precision != CVDV_MIXED) {
        cvdv_set_error(CVDV_ERR_ARG, "cvdvCreateEx: precision must be 0 (complex128) or 1 (mixed)");
        return ctx;
    }
    int total = 0;
    for (int i = 0; i < numReg; ++i) {
        if (numQubits[i] < 0 || numQubits[i] > 30) {
            cvdv_set_error(CVDV_ERR_ARG, "cvdvCreate: register qubit count out of range [0, 30]");
            return ctx;
        }
        total += numQubits[i];
    }
    if (total > 36) {
        cvdv_set_error(CVDV_ERR_ARG, "cvdvCreate: more than 36 total qubits");
        return ctx;
    }
    ctx->numReg = numReg;
    ctx->totalQbt = total;
    ctx->prec = precision;
    ctx->totalSize = (size_t)1 << total;
    for (int i = 0; i < numReg; ++i) {
        ctx->nq[i] = numQubits[i];
        ctx->dx[i] = std::sqrt(2.0 * M_PI / (double)((size_t)1 << numQubits[i]));  // context.cu:31
        int f = 0;
        for (int j = i + 1; j < numReg; ++j) f += numQubits[j];
        ctx->flw[i] = f;  // context.cu:33-38
    }
    reset_layout(ctx);
    if (recording) {
        ctx->rec = new BatchRecording();
        ctx->initialized = true;  // gates record from the canonical layout
        return ctx;
    }
    HIPCHK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    return ctx;
}

CVDVContext* ctx_create_like(const CVDVContext* src) {
    CVDVContext* ctx = create_impl(src->numReg, src->nq, src->prec);
    if (ctx->totalSize != src->totalSize) {
        cvdvDestroy(ctx);
        throw CvdvError{CVDV_ERR_STATE, "cannot create a context of the same shape"};
    }
    HIPCHK(hipMalloc(&ctx->buf[0], ctx->totalSize * ctx->elemBytes()));
    HIPCHK(hipMemsetAsync(ctx->buf[0], 0, ctx->totalSize * ctx->elemBytes(), ctx->stream));
    ctx->initialized = true;
    return ctx;
}

CVDVContext* ctx_create_recording(int numReg, const int* numQubits, int precision) {
    CVDVContext* ctx = create_impl(numReg, numQubits, precision, true);
    if (!ctx->rec) {  // invalid shape: the error is recorded
        delete ctx;
        return nullptr;
    }
    return ctx;
}

extern "C" {

CVDVContext* cvdvCreate(int numReg, int* numQubits) {
    CVDV_TRY
    return create_impl(numReg, numQubits, CVDV_C128);
    CVDV_CATCH(nullptr)
}

CVDVContext* cvdvCreateEx(int numReg, int* numQubits, int precision) {
    CVDV_TRY
    return create_impl(numReg, numQubits, precision);
    CVDV_CATCH(nullptr)
}

int cvdvGetPrecision(CVDVContext* ctx) { return ctx ? ctx->prec : -1; }

void cvdvDestroy(CVDVContext* ctx) {
    if (!ctx || refuse_recording(ctx)) return;
    ctx_release(ctx);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

void cvdvFree(CVDVContext* ctx) {
    if (!ctx || refuse_recording(ctx)) return;
    ctx_release(ctx);  // context.cu:120-129: drop state and metadata, keep the handle
    ctx->numReg = 0;
    ctx->totalQbt = 0;
    ctx->totalSize = 0;
}

int cvdvGetLastError(char* buf, size_t len) {
    const int c = g_last.code;
    if (buf && len) {
        strncpy(buf, g_last.msg.c_str(), len - 1);
        buf[len - 1] = 0;
    }
    g_last.code = CVDV_OK;
    g_last.msg.clear();
    return c;
}

void cvdvSynchronize(CVDVContext* ctx) {
    if (ctx && ctx->capturing) {
        cvdv_set_error(CVDV_ERR_STATE, "the context is capturing a graph: end the capture first");
        return;
    }
    CVDV_TRY
    if (ctx && ctx->stream) HIPCHK(hipStreamSynchronize(ctx->stream));
    CVDV_CATCH()
}

void cvdvInitFromSeparable(CVDVContext* ctx, void** devicePtrs, int numReg) {
    if (!ctx || refuse_recording(ctx)) return;
    if (ctx->numReg == 0) {
        cvdv_set_error(CVDV_ERR_STATE,
                       "Error: Must call cvdvCreate with registers before cvdvInitFromSeparable");
        return;
    }
    if (numReg != ctx->numReg) {
        cvdv_set_error(CVDV_ERR_ARG, "Error: numReg mismatch: context has " +
                                         std::to_string(ctx->numReg) + ", got " + std::to_string(numReg));
        return;
    }
    if (!devicePtrs) {
        cvdv_set_error(CVDV_ERR_ARG, "Error: null device pointer array");
        return;
    }
    CVDV_TRY
    SepArgs a{};
    for (int r = 0; r < numReg; ++r) {
        if (!devicePtrs[r]) {
            cvdv_set_error(CVDV_ERR_ARG, "Error: null register array");
            return;
        }
        a.arr[r] = (const double2*)devicePtrs[r];
        a.sh[r] = ctx->flw[r];
        a.nb[r] = ctx->nq[r];
    }
    a.nreg = numReg;
    ensure_state(ctx);
    reset_layout(ctx);
    // The register arrays were written on another stream (the caller's); the
    // copy below must see them, and they may be freed as soon as we return.
    HIPCHK(hipDeviceSynchronize());
    const int k = std::min(INIT_LO_BITS, ctx->totalQbt);
    const uint64_t nchunks = ctx->totalSize >> k;
    const dim3 grid((unsigned)std::min<uint64_t>(nchunks, 2048));
    if (ctx->prec == CVDV_C128)
        hipLaunchKernelGGL(k_init_product<double2>, grid, dim3(256), 0, ctx->stream,
                           (double2*)ctx->state(), nchunks, k, a);
    else
        hipLaunchKernelGGL(k_init_product<float2>, grid, dim3(256), 0, ctx->stream,
                           (float2*)ctx->state(), nchunks, k, a);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->initialized = true;
    CVDV_CATCH()
}

void cvdvSetStateFromDevicePtr(CVDVContext* ctx, void* d_src) {
    if (!ctx || !d_src || refuse_recording(ctx)) return;
    if (ctx->numReg == 0) {
        cvdv_set_error(CVDV_ERR_STATE, "Error: context has no registers");
        return;
    }
    CVDV_TRY
    ensure_state(ctx);
    reset_layout(ctx);
    HIPCHK(hipDeviceSynchronize());
    const uint64_t n = ctx->totalSize;
    if (ctx->prec == CVDV_C128) {
        HIPCHK(hipMemcpyAsync(ctx->state(), d_src, n * sizeof(double2), hipMemcpyDeviceToDevice,
                              ctx->stream));
    } else {
        hipLaunchKernelGGL(k_from_c128<float2>, dim3(grid_for(n)), dim3(256), 0, ctx->stream,
                           (float2*)ctx->state(), (const double2*)d_src, n);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->initialized = true;
    CVDV_CATCH()
}

int cvdvGetNumRegisters(CVDVContext* ctx) { return ctx ? ctx->numReg : 0; }

size_t cvdvGetTotalSize(CVDVContext* ctx) { return ctx ? ctx->totalSize : 0; }

void cvdvGetRegisterInfo(CVDVContext* ctx, int* qubitCountsOut, double* gridStepsOut) {
    if (!ctx || ctx->numReg == 0) return;
    for (int r = 0; r < ctx->numReg; ++r) {
        if (qubitCountsOut) qubitCountsOut[r] = ctx->nq[r];
        if (gridStepsOut) gridStepsOut[r] = ctx->dx[r];
    }
}

int cvdvGetRegisterDim(CVDVContext* ctx, int regIdx) {
    if (!ctx || regIdx < 0 || regIdx >= ctx->numReg) return -1;
    return 1 << ctx->nq[regIdx];
}

double cvdvGetRegisterDx(CVDVContext* ctx, int regIdx) {
    if (!ctx || regIdx < 0 || regIdx >= ctx->numReg) return -1.0;
    return ctx->dx[regIdx];
}

void cvdvProfileEnable(CVDVContext* ctx, int on) {
    if (!ctx) return;
    CVDV_TRY
    if (ctx->stream) HIPCHK(hipStreamSynchronize(ctx->stream));
    prof_clear(ctx);
    ctx->profOn = on != 0;
    CVDV_CATCH()
}

int cvdvProfileRead(CVDVContext* ctx, int kind, double* totalMs, double* maxMs) {
    if (!ctx) return 0;
    CVDV_TRY
    if (ctx->stream) HIPCHK(hipStreamSynchronize(ctx->stream));
    int n = 0;
    double tot = 0.0, mx = 0.0;
    for (auto& r : ctx->prof) {
        const bool fft = r.kind == KK_FFT || r.kind == KK_FFT_FUSED || r.kind == KK_FFT_ROW;
        if (r.kind != kind && !(kind == KK_FFT && fft)) continue;
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, r.a, r.b));
        tot += ms;
        mx = ms > mx ? ms : mx;
        ++n;
    }
    if (totalMs) *totalMs = tot;
    if (maxMs) *maxMs = mx;
    return n;
    CVDV_CATCH(0)
}

void cvdvProfileReset(CVDVContext* ctx) {
    if (!ctx) return;
    CVDV_TRY
    if (ctx->stream) HIPCHK(hipStreamSynchronize(ctx->stream));
    prof_clear(ctx);
    CVDV_CATCH()
}

void cvdvSetTransformFusion(CVDVContext* ctx, int on) {
    if (!ctx || refuse_recording(ctx)) return;
    CVDV_TRY
    if (!on && ctx->initialized) fft_settle(ctx);
    ctx->fuseFt = on ? 1 : 0;
    CVDV_CATCH()
}

void cvdvFlush(CVDVContext* ctx) {
    if (!ctx || !ctx->initialized || refuse_recording(ctx)) return;
    CVDV_TRY
    ctx_materialize(ctx);
    CVDV_CATCH()
}

void cvdvGetState(CVDVContext* ctx, double* realOut, double* imagOut) {
    if (!ctx || refuse_recording(ctx)) return;
    if (!ctx->initialized) {
        cvdv_set_error(CVDV_ERR_STATE, "Error: State not initialized. Call cvdvInitStateVector first.");
        return;
    }
    CVDV_TRY
    ctx_materialize(ctx);
    const uint64_t n = ctx->totalSize;
    double* d = (double*)ctx_scratch(ctx, 2 * n * sizeof(double));
    if (ctx->prec == CVDV_C128)
        hipLaunchKernelGGL(k_split<double2>, dim3(grid_for(n)), dim3(256), 0, ctx->stream, d, d + n,
                           (const double2*)ctx->state(), n);
    else
        hipLaunchKernelGGL(k_split<float2>, dim3(grid_for(n)), dim3(256), 0, ctx->stream, d, d + n,
                           (const float2*)ctx->state(), n);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(realOut, d, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(imagOut, d + n, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    CVDV_CATCH()
}

void cvdvGraphBegin(CVDVContext* ctx) {
    if (!ctx || refuse_recording(ctx)) return;
    if (!ctx->initialized) {
        cvdv_set_error(CVDV_ERR_STATE, "State not initialized");
        return;
    }
    CVDV_TRY
    ctx_materialize(ctx);
    (void)ctx->other();  // both buffers exist before the capture
    int maxNb = 0;
    for (int r = 0; r < ctx->numReg; ++r) maxNb = std::max(maxNb, ctx->nq[r]);
    arena_reserve(ctx, diag_arena_bound(maxNb, ctx->elemBytes()));
    fft_prepare(ctx);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (int r = 0; r < ctx->numReg; ++r) {
        ctx->capSave.pshift[r] = ctx->pshift[r];
        ctx->capSave.refl[r] = ctx->refl[r];
    }
    ctx->capSave.pending = ctx->pending;
    ctx->capSave.cur = ctx->cur;
    HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    ctx->capturing = true;
    CVDV_CATCH()
}

CVDVGraph* cvdvGraphEnd(CVDVContext* ctx) {
    if (!ctx) return nullptr;
    if (!ctx->capturing) {
        cvdv_set_error(CVDV_ERR_STATE, "cvdvGraphEnd: no capture in progress");
        return nullptr;
    }
    hipGraph_t g = nullptr;
    bool ok = true;
    std::string why;
    try {
        ctx_materialize(ctx);
        if (ctx->cur != ctx->capSave.cur) copy_to(ctx, ctx->capSave.cur);
    } catch (const CvdvError& e) {
        ok = false;
        why = e.msg;
    }
    ctx->capturing = false;
    const hipError_t ce = hipStreamEndCapture(ctx->stream, &g);
    // the host-side layout returns to the capture start: nothing has executed
    for (int r = 0; r < ctx->numReg; ++r) {
        ctx->pshift[r] = ctx->capSave.pshift[r];
        ctx->refl[r] = ctx->capSave.refl[r];
    }
    ctx->pending = ctx->capSave.pending;
    ctx->defer = FtDefers{};  // cvdvGraphBegin settled it
    ctx->cur = ctx->capSave.cur;
    if (!ok || ce != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        cvdv_set_error(CVDV_ERR_HIP, "cvdvGraphEnd: capture failed" + (why.empty() ? std::string() : ": " + why));
        return nullptr;
    }
    hipGraphExec_t e = nullptr;
    if (hipGraphInstantiate(&e, g, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGraphDestroy(g);
        cvdv_set_error(CVDV_ERR_HIP, "cvdvGraphEnd: hipGraphInstantiate failed");
        return nullptr;
    }
    ctx->graphs++;
    return new CVDVGraph{ctx, g, e, ctx->capSave.cur, ctx->generation};
}

void cvdvGraphLaunch(CVDVGraph* gr) {
    if (!gr) return;
    CVDVContext* ctx = gr->ctx;
    if (ctx->capturing || refuse_recording(ctx)) {
        cvdv_set_error(CVDV_ERR_STATE, "cvdvGraphLaunch: the context is capturing");
        return;
    }
    if (!ctx->initialized || gr->generation != ctx->generation) {
        cvdv_set_error(CVDV_ERR_STATE, "cvdvGraphLaunch: the context's state was released after the capture");
        return;
    }
    CVDV_TRY
    ctx_materialize(ctx);
    if (ctx->cur != gr->cur) copy_to(ctx, gr->cur);
    HIPCHK(hipGraphLaunch(gr->exec, ctx->stream));
    CVDV_CATCH()
}

void cvdvGraphDestroy(CVDVGraph* gr) {
    if (!gr) return;
    if (gr->ctx && gr->ctx->stream) (void)hipStreamSynchronize(gr->ctx->stream);
    if (gr->exec) (void)hipGraphExecDestroy(gr->exec);
    if (gr->graph) (void)hipGraphDestroy(gr->graph);
    delete gr;
}

void cvdvGetStateDevice(CVDVContext* ctx, void* d_dst) {
    if (!ctx || !d_dst || refuse_recording(ctx)) return;
    if (!ctx->initialized) {
        cvdv_set_error(CVDV_ERR_STATE, "Error: State not initialized");
        return;
    }
    CVDV_TRY
    ctx_materialize(ctx);
    const uint64_t n = ctx->totalSize;
    if (ctx->prec == CVDV_C128)
        HIPCHK(hipMemcpyAsync(d_dst, ctx->state(), n * sizeof(double2), hipMemcpyDeviceToDevice,
                              ctx->stream));
    else {
        hipLaunchKernelGGL(k_to_c128<float2>, dim3(grid_for(n)), dim3(256), 0, ctx->stream,
                           (double2*)d_dst, (const float2*)ctx->state(), n);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    CVDV_CATCH()
}

}  // extern "C"
