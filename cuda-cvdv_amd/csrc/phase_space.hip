// phase_space.hip - native-grid Wigner and Husimi-Q readouts.
//
// Reference interfaces replaced: src/cuda/bosonic.cu:10-152 and the kernels of
// src/cuda/kernels/bosonic.cuh:7-186; numerical semantics follow
// src/torchCvdv.py:344-404.  Readouts always compute in FP64 (amplitudes of a
// mixed-precision state are widened on load).
//
// Wigner:  rho = X X^H over the slices X[s][j] of register r (a complex
//          Hermitian rank-S product), integrand G[i][k] = rho[i-kd][i+kd] with
//          kd = k - (N-1)/2, an unnormalised inverse FFT along k, and
//          W[jc][i] = Re(e^{-i (N-1) dx p_jc} G[i][(jc+N/2) mod N]) dx / pi,
//          p_jc = (jc - N/2) pi / (N dx)  (bosonic.cuh:39-52, torchCvdv.py:361-368).
// Husimi:  Q[jc][q] = sum_s |FFT_x(psi_s[x] w(q,x))[(jc+N/2) mod N]|^2 / pi with
//          the non-periodic window w(q,x) = pi^-1/4 exp(-((x-q) dx)^2/2) sqrt(dx)
//          (torchCvdv.py:380-397).

#include <cmath>
#include <string>
#include <vector>

#include "cvdv_internal.h"
#include "../../include/cvdv_hip.h"

static const double PI = 3.14159265358979323846;

static unsigned grid_for(uint64_t total, unsigned block = 256) {
    uint64_t b = (total + block - 1) / block;
    if (b > 65536ull * 8) b = 65536ull * 8;
    return (unsigned)(b ? b : 1);
}

// X[s][j] = psi(slice s, logical j) in FP64
template <typename C>
__global__ void k_extract(double2* __restrict__ X, const C* __restrict__ st, uint64_t S, int sh, int n,
                          int refl) {
    const uint64_t N = 1ull << n, tot = S * N, slo = (1ull << sh) - 1ull;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t s = e >> n, j = e & (N - 1);
        const uint64_t jp = refl ? (N - 1 - j) : j;
        const uint64_t g = ((s & ~slo) << n) | (jp << sh) | (s & slo);
        const C v = st[g];
        X[e] = make_double2((double)v.x, (double)v.y);
    }
}

// rho[a][b] = sum_s X[s][a] conj(X[s][b]); 32x32 output tile per workgroup.
__global__ void __launch_bounds__(256) k_rho(const double2* __restrict__ X, uint64_t S, int N,
                                             double2* __restrict__ rho) {
    __shared__ double2 xa[16][33];
    __shared__ double2 xb[16][33];
    const int tid = threadIdx.x;
    const int a0 = blockIdx.y * 32, b0 = blockIdx.x * 32;
    const int ta = (tid / 16) * 2, tb = (tid % 16) * 2;
    double2 acc[2][2] = {{{0, 0}, {0, 0}}, {{0, 0}, {0, 0}}};
    for (uint64_t s0 = 0; s0 < S; s0 += 16) {
        for (int e = tid; e < 16 * 32; e += 256) {
            const int ks = e / 32, c = e % 32;
            const uint64_t s = s0 + ks;
            const bool ok = s < S;
            xa[ks][c] = (ok && a0 + c < N) ? X[s * N + a0 + c] : make_double2(0, 0);
            xb[ks][c] = (ok && b0 + c < N) ? X[s * N + b0 + c] : make_double2(0, 0);
        }
        __syncthreads();
#pragma unroll 4
        for (int ks = 0; ks < 16; ++ks) {
            const double2 u0 = xa[ks][ta], u1 = xa[ks][ta + 1];
            const double2 v0 = xb[ks][tb], v1 = xb[ks][tb + 1];
            acc[0][0].x += u0.x * v0.x + u0.y * v0.y; acc[0][0].y += u0.y * v0.x - u0.x * v0.y;
            acc[0][1].x += u0.x * v1.x + u0.y * v1.y; acc[0][1].y += u0.y * v1.x - u0.x * v1.y;
            acc[1][0].x += u1.x * v0.x + u1.y * v0.y; acc[1][0].y += u1.y * v0.x - u1.x * v0.y;
            acc[1][1].x += u1.x * v1.x + u1.y * v1.y; acc[1][1].y += u1.y * v1.x - u1.x * v1.y;
        }
        __syncthreads();
    }
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
            const int a = a0 + ta + i, b = b0 + tb + j;
            if (a < N && b < N) rho[(uint64_t)a * N + b] = acc[i][j];
        }
}

__global__ void k_wigner_integrand(double2* __restrict__ G, const double2* __restrict__ rho, int N) {
    const uint64_t tot = (uint64_t)N * N;
    const int h = (N - 1) / 2;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const int i = (int)(e / N), k = (int)(e % N);
        const int kd = k - h, ipy = i + kd, imy = i - kd;
        G[e] = (ipy >= 0 && ipy < N && imy >= 0 && imy < N) ? rho[(uint64_t)imy * N + ipy]
                                                             : make_double2(0, 0);
    }
}

// W[jc][i] = Re(e^{-i (N-1) dx p_jc} G[i][kf]) dx/pi; optionally written as complex
template <bool COMPLEX_OUT>
__global__ void k_wigner_finalize(void* __restrict__ outv, const double2* __restrict__ G, int N, double dx) {
    const uint64_t tot = (uint64_t)N * N;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const int jc = (int)(e / N), i = (int)(e % N);
        const int kf = (jc + N / 2) % N;
        const double2 v = G[(uint64_t)i * N + kf];
        // -(N-1) dx p_jc = -(N-1)(jc - N/2) pi / N  (radians) -> sincospi
        const double a = -(double)(N - 1) * ((double)jc - 0.5 * N) / (double)N;
        const double ar = a - 2.0 * rint(0.5 * a);
        double s, c;
        sincospi(ar, &s, &c);
        const double w = (v.x * c - v.y * s) * dx / PI;
        if (COMPLEX_OUT) ((double2*)outv)[e] = make_double2(w, 0.0);
        else ((double*)outv)[e] = w;
    }
}

// B[q][s][x] = X[s][x] w(q, x) for q in [q0, q0+Qc)
__global__ void k_husimi_window(double2* __restrict__ B, const double2* __restrict__ X, uint64_t S, int n,
                                int q0, int Qc, double dx) {
    const uint64_t N = 1ull << n, tot = (uint64_t)Qc * S * N;
    const double norm = pow(PI, -0.25) * sqrt(dx);
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t x = e & (N - 1);
        const uint64_t sq = e >> n;
        const uint64_t s = sq % S;
        const int q = q0 + (int)(sq / S);
        const double d = ((double)x - (double)q) * dx;
        const double w = norm * exp(-0.5 * d * d);
        const double2 v = X[s * N + x];
        B[e] = make_double2(v.x * w, v.y * w);
    }
}

// acc[q][m] += sum_s |B[q][s][m]|^2 (fixed order over s)
__global__ void k_husimi_accum(double* __restrict__ acc, const double2* __restrict__ B, uint64_t S, int N,
                               int q0, int Qc) {
    const uint64_t tot = (uint64_t)Qc * N;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t ql = e / N, m = e % N;
        double sum = 0.0;
        for (uint64_t s = 0; s < S; ++s) {
            const double2 v = B[(ql * S + s) * N + m];
            sum += v.x * v.x + v.y * v.y;
        }
        acc[(uint64_t)(q0 + ql) * N + m] = sum;
    }
}

__global__ void k_husimi_finalize(double* __restrict__ Q, const double* __restrict__ acc, int N) {
    const uint64_t tot = (uint64_t)N * N;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const int jc = (int)(e / N), q = (int)(e % N);
        Q[e] = acc[(uint64_t)q * N + (jc + N / 2) % N] / PI;
    }
}

// bosonic.cuh:158-178: spectrum *= exp(-(wq^2 + wp^2)/4)
__global__ void k_husimi_gauss2d(double2* __restrict__ spec, int N, double dx) {
    const uint64_t tot = (uint64_t)N * N;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const int pIdx = (int)(e / N), qIdx = (int)(e % N);
        const int kq = (qIdx <= N / 2) ? qIdx : (qIdx - N);
        const int kp = (pIdx <= N / 2) ? pIdx : (pIdx - N);
        const double Lq = (double)N * dx;
        const double dp = PI / ((double)N * dx);
        const double Lp = (double)N * dp;
        const double wq = 2.0 * PI * kq / Lq, wp = 2.0 * PI * kp / Lp;
        const double g = exp(-0.25 * (wq * wq + wp * wp));
        spec[e] = make_double2(spec[e].x * g, spec[e].y * g);
    }
}

// bosonic.cuh:180-186
__global__ void k_real_clamp(double* __restrict__ out, const double2* __restrict__ in, uint64_t n2,
                             double norm) {
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n2;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const double v = in[e].x / norm;
        out[e] = v < 0.0 ? 0.0 : v;
    }
}

// ---------------------------------------------------------------- drivers
struct Bufs {
    double2* X;
    uint64_t S;
    int N, n, sh;
};

static Bufs extract_slices(CVDVContext* ctx, int r, double2* X) {
    Bufs b{};
    b.n = ctx->nq[r];
    b.N = 1 << b.n;
    b.sh = ctx->pshift[r];
    b.S = ctx->totalSize >> b.n;
    b.X = X;
    const uint64_t tot = ctx->totalSize;
    if (ctx->prec == CVDV_C128)
        hipLaunchKernelGGL(k_extract<double2>, dim3(grid_for(tot)), dim3(256), 0, ctx->stream, X,
                           (const double2*)ctx->state(), b.S, b.sh, b.n, ctx->refl[r]);
    else
        hipLaunchKernelGGL(k_extract<float2>, dim3(grid_for(tot)), dim3(256), 0, ctx->stream, X,
                           (const float2*)ctx->state(), b.S, b.sh, b.n, ctx->refl[r]);
    HIPCHK(hipGetLastError());
    return b;
}

// Wigner into dev (N*N doubles or complex when asComplex); scratch layout below.
static void wigner_device(CVDVContext* ctx, int r, void* dev, bool asComplex, double2* X, double2* rho,
                          double2* G, double2* tmp) {
    ctx_flush(ctx);
    Bufs b = extract_slices(ctx, r, X);
    const int N = b.N;
    const uint64_t N2 = (uint64_t)N * N;
    dim3 gr((N + 31) / 32, (N + 31) / 32);
    hipLaunchKernelGGL(k_rho, gr, dim3(256), 0, ctx->stream, X, b.S, N, rho);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_wigner_integrand, dim3(grid_for(N2)), dim3(256), 0, ctx->stream, G, rho, N);
    HIPCHK(hipGetLastError());
    // unnormalised inverse FFT along k (rows of G)
    if (b.n > 0) fft_generic_c128(ctx, G, tmp, N2, 0, b.n, true, 1.0);
    if (asComplex)
        hipLaunchKernelGGL(k_wigner_finalize<true>, dim3(grid_for(N2)), dim3(256), 0, ctx->stream, dev, G, N,
                           ctx->dx[r]);
    else
        hipLaunchKernelGGL(k_wigner_finalize<false>, dim3(grid_for(N2)), dim3(256), 0, ctx->stream, dev, G,
                           N, ctx->dx[r]);
    HIPCHK(hipGetLastError());
}

// The readouts below write their N x N result into `out`: host memory (through
// a device scratch copy), or, with outOnDevice, caller device memory directly.
static void finish_out(CVDVContext* ctx, double* out, const double* dev, uint64_t n, bool outOnDevice) {
    if (!outOnDevice) HIPCHK(hipMemcpyAsync(out, dev, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
}

static void get_wigner(CVDVContext* ctx, int r, double* out, bool outOnDevice) {
    const int N = 1 << ctx->nq[r];
    const uint64_t N2 = (uint64_t)N * N, A = ctx->totalSize;
    // scratch: X (A c128) | rho (N2) | G (N2) | tmp (N2) | W (N2 doubles)
    char* base = (char*)ctx_scratch(ctx, A * 16 + 3 * N2 * 16 + (outOnDevice ? 0 : N2 * 8));
    double2* X = (double2*)base;
    double2* rho = X + A;
    double2* G = rho + N2;
    double2* tmp = G + N2;
    double* W = outOnDevice ? out : (double*)(tmp + N2);
    const int pid = prof_begin(ctx, KK_OTHER);  // the readout's device time, first to last kernel
    wigner_device(ctx, r, W, false, X, rho, G, tmp);
    prof_end(ctx, pid);
    finish_out(ctx, out, W, N2, outOnDevice);
}

static void get_husimi_overlap(CVDVContext* ctx, int r, double* out, bool outOnDevice) {
    ctx_flush(ctx);
    const int n = ctx->nq[r], N = 1 << n;
    const uint64_t N2 = (uint64_t)N * N, A = ctx->totalSize, S = A >> n;
    // q chunk so that the windowed batch stays below 2^26 elements (1 GiB)
    uint64_t Qc = 1;
    while (Qc * 2 <= (uint64_t)N && Qc * 2 * S * N <= (1ull << 26)) Qc *= 2;
    const uint64_t batch = Qc * S * N;
    // scratch: X (A) | B (batch) | tmp (batch) | acc (N2 doubles) | Q (N2 doubles)
    char* base = (char*)ctx_scratch(ctx, A * 16 + 2 * batch * 16 + (outOnDevice ? 1 : 2) * N2 * 8);
    double2* X = (double2*)base;
    double2* B = X + A;
    double2* tmp = B + batch;
    double* acc = (double*)(tmp + batch);
    double* Q = outOnDevice ? out : acc + N2;
    const int pid = prof_begin(ctx, KK_OTHER);
    Bufs b = extract_slices(ctx, r, X);
    for (uint64_t q0 = 0; q0 < (uint64_t)N; q0 += Qc) {
        hipLaunchKernelGGL(k_husimi_window, dim3(grid_for(batch)), dim3(256), 0, ctx->stream, B, X, b.S, n,
                           (int)q0, (int)Qc, ctx->dx[r]);
        HIPCHK(hipGetLastError());
        if (n > 0) fft_generic_c128(ctx, B, tmp, batch, 0, n, false, 1.0);
        hipLaunchKernelGGL(k_husimi_accum, dim3(grid_for(Qc * N)), dim3(256), 0, ctx->stream, acc, B, b.S, N,
                           (int)q0, (int)Qc);
        HIPCHK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_husimi_finalize, dim3(grid_for(N2)), dim3(256), 0, ctx->stream, Q, acc, N);
    HIPCHK(hipGetLastError());
    prof_end(ctx, pid);
    finish_out(ctx, out, Q, N2, outOnDevice);
}

// bosonic.cu:63-101: Wigner -> 2D FFT -> Gaussian -> inverse 2D FFT -> Re/N^2, clamped
static void get_husimi_wigner(CVDVContext* ctx, int r, double* out, bool outOnDevice) {
    const int n = ctx->nq[r], N = 1 << n;
    const uint64_t N2 = (uint64_t)N * N, A = ctx->totalSize;
    char* base = (char*)ctx_scratch(ctx, A * 16 + 4 * N2 * 16 + (outOnDevice ? 0 : N2 * 8));
    double2* X = (double2*)base;
    double2* rho = X + A;
    double2* G = rho + N2;
    double2* tmp = G + N2;
    double2* Wc = tmp + N2;
    double* Q = outOnDevice ? out : (double*)(Wc + N2);
    const int pid = prof_begin(ctx, KK_OTHER);
    wigner_device(ctx, r, Wc, true, X, rho, G, tmp);
    if (n > 0) {
        fft_generic_c128(ctx, Wc, tmp, N2, 0, n, false, 1.0);  // along q (columns index)
        fft_generic_c128(ctx, Wc, tmp, N2, n, n, false, 1.0);  // along p (rows index)
    }
    hipLaunchKernelGGL(k_husimi_gauss2d, dim3(grid_for(N2)), dim3(256), 0, ctx->stream, Wc, N, ctx->dx[r]);
    HIPCHK(hipGetLastError());
    if (n > 0) {
        fft_generic_c128(ctx, Wc, tmp, N2, 0, n, true, 1.0);
        fft_generic_c128(ctx, Wc, tmp, N2, n, n, true, 1.0);
    }
    hipLaunchKernelGGL(k_real_clamp, dim3(grid_for(N2)), dim3(256), 0, ctx->stream, Q, Wc, N2, (double)N2);
    HIPCHK(hipGetLastError());
    prof_end(ctx, pid);
    finish_out(ctx, out, Q, N2, outOnDevice);
}

// bosonic.cu:13-25: the overlap route holds one N-point window, two N x N double
// arrays, every slice and an N x N complex block per slice; above 80% of the
// free device memory (or when the device cannot say) the FFT2 route runs.  Both
// are FP64; they differ in grid and normalisation (getHusimiQWigner's p spacing
// is pi/(N dx) and its scale that of the Wigner), as in the reference.  The
// overlap route here is chunked and needs far less than that estimate; the
// estimate is kept so the same states take the same route.
static void get_husimi(CVDVContext* ctx, int r, double* out, bool outOnDevice) {
    const uint64_t N = 1ull << ctx->nq[r], N2 = N * N, S = ctx->totalSize >> ctx->nq[r];
    const double need = (double)(N * 16 + 2 * N2 * 8) + (double)S * (double)N * 16.0 + (double)S * (double)N2 * 16.0;
    size_t freeB = ctx->freeMemOverride, totalB = 0;
    bool known = true;
    if (!freeB) known = hipMemGetInfo(&freeB, &totalB) == hipSuccess;
    if (!known || need > 0.8 * (double)freeB) get_husimi_wigner(ctx, r, out, outOnDevice);
    else get_husimi_overlap(ctx, r, out, outOnDevice);
}

#define CVDV_TRY try {
#define CVDV_CATCH()                                                         \
    }                                                                        \
    catch (const CvdvError& e) {                                             \
        cvdv_set_error(e.code, e.msg);                                       \
    }                                                                        \
    catch (...) {                                                            \
        cvdv_set_error(CVDV_ERR_HIP, "unexpected internal error");           \
    }

static bool ps_ok(CVDVContext* ctx, int r, double* out) {
    if (!ctx || !out || refuse_recording(ctx)) return false;
    if (!ctx->initialized) {
        cvdv_set_error(CVDV_ERR_STATE, "State not initialized");
        return false;
    }
    if (r < 0 || r >= ctx->numReg) {
        cvdv_set_error(CVDV_ERR_ARG, "Invalid register index: " + std::to_string(r));
        return false;
    }
    return true;
}

extern "C" {

void cvdvGetWigner(CVDVContext* ctx, int regIdx, double* wignerOut) {
    if (!ps_ok(ctx, regIdx, wignerOut)) return;
    CVDV_TRY get_wigner(ctx, regIdx, wignerOut, false); CVDV_CATCH()
}

void cvdvGetHusimiQ(CVDVContext* ctx, int regIdx, double* husimiOut) {
    if (!ps_ok(ctx, regIdx, husimiOut)) return;
    CVDV_TRY get_husimi(ctx, regIdx, husimiOut, false); CVDV_CATCH()
}

void cvdvGetHusimiQOverlap(CVDVContext* ctx, int regIdx, double* husimiOut) {
    if (!ps_ok(ctx, regIdx, husimiOut)) return;
    CVDV_TRY get_husimi_overlap(ctx, regIdx, husimiOut, false); CVDV_CATCH()
}

void cvdvGetHusimiQWigner(CVDVContext* ctx, int regIdx, double* husimiOut) {
    if (!ps_ok(ctx, regIdx, husimiOut)) return;
    CVDV_TRY get_husimi_wigner(ctx, regIdx, husimiOut, false); CVDV_CATCH()
}

void cvdvGetWignerDevice(CVDVContext* ctx, int regIdx, double* d_out) {
    if (!ps_ok(ctx, regIdx, d_out)) return;
    CVDV_TRY get_wigner(ctx, regIdx, d_out, true); CVDV_CATCH()
}

void cvdvGetHusimiQDevice(CVDVContext* ctx, int regIdx, double* d_out) {
    if (!ps_ok(ctx, regIdx, d_out)) return;
    CVDV_TRY get_husimi(ctx, regIdx, d_out, true); CVDV_CATCH()
}

void cvdvSetFreeMemoryOverride(CVDVContext* ctx, size_t bytes) {
    if (!ctx) return;
    ctx->freeMemOverride = bytes;
}

}  // extern "C"
