// phase_grid.hip - Wigner and Husimi-Q of one register at arbitrary phase-space
// points, traced over the other registers or one block per slice (per branch of
// the other registers' values).  An addition: the reference evaluates both only
// on its native grid (src/cuda/bosonic.cu:10-152); on that grid these equal
// cvdvGetWigner / cvdvGetHusimiQ.
//
// Wigner.  With x = x_i + t dx (t in [0,1), t = 0 on grid points) and phi_s the
// trigonometric interpolant of slice s sampled at x_j + t dx (one FFT per slice,
// a phase ramp and one inverse FFT per x point), A(d) = sum_s phi_s(i-d)
// conj(phi_s(i+d)) for |d| <= N/2 (zero outside the grid), and
//     W(x, p) = dx/pi Re sum_d A(d) e^{i dx p m(d)},
// m(d) = 2(d + (N-1)/2) - (N-1) (native: the e^{-i (N-1) dx p} finalisation of
// bosonic.cuh:39-52) or m(d) = 2d (the symmetric sum).  The sum over d for all
// points at once is a complex GEMM (rocBLAS zgemm): Z[p][x] = sum_d E[p][d] A[x][d].
// Husimi.  Q(q, p) = dx/pi^(3/2) sum_s |H_s(q,p)|^2 with
//     H_s(q, p) = sum_j e^{-i p x_j} psi_s(x_j) e^{-(x_j-q)^2/2},
// again one GEMM per slice over j.  Angles are FP64 with range reduction
// (revolutions, frac, sincospi); all arithmetic is FP64.

#include <rocblas/rocblas.h>

#include <cmath>
#include <string>
#include <vector>

#include "cvdv_internal.h"
#include "../../include/cvdv_hip.h"

static const double PI = 3.14159265358979323846;
static const uint64_t GRID_CHUNK = 1ull << 26;  // complex128 elements per intermediate buffer (1 GiB)

static unsigned grid_for(uint64_t total, unsigned block = 256) {
    uint64_t b = (total + block - 1) / block;
    if (b > 65536ull * 8) b = 65536ull * 8;
    return (unsigned)(b ? b : 1);
}

static rocblas_handle blas(CVDVContext* ctx) {
    if (!ctx->blas) {
        rocblas_handle h = nullptr;
        if (rocblas_create_handle(&h) != rocblas_status_success)
            throw CvdvError{CVDV_ERR_HIP, "rocblas_create_handle failed"};
        ctx->blas = h;
    }
    rocblas_handle h = (rocblas_handle)ctx->blas;
    if (rocblas_set_stream(h, ctx->stream) != rocblas_status_success)
        throw CvdvError{CVDV_ERR_HIP, "rocblas_set_stream failed"};
    return h;
}

void blas_release(CVDVContext* ctx) {
    if (ctx->blas) (void)rocblas_destroy_handle((rocblas_handle)ctx->blas);
    ctx->blas = nullptr;
}

static void blas_check(rocblas_status s, const char* what) {
    if (s != rocblas_status_success) throw CvdvError{CVDV_ERR_HIP, std::string(what) + " failed"};
}

// Zcm(m x n, ldc) = Acm^T Bcm for `batch` slices (column-major; A: k x m, B: k x n).
static void zgemm_tn(CVDVContext* ctx, int m, int n, int k, const double2* A, int lda, uint64_t strideA,
                     const double2* B, int ldb, uint64_t strideB, double2* C, int ldc, uint64_t strideC,
                     int batch) {
    const rocblas_double_complex one{1.0, 0.0}, zero{0.0, 0.0};
    blas_check(rocblas_zgemm_strided_batched(blas(ctx), rocblas_operation_transpose, rocblas_operation_none, m,
                                             n, k, &one, (const rocblas_double_complex*)A, lda,
                                             (rocblas_stride)strideA, (const rocblas_double_complex*)B, ldb,
                                             (rocblas_stride)strideB, &zero, (rocblas_double_complex*)C, ldc,
                                             (rocblas_stride)strideC, batch),
               "rocblas_zgemm_strided_batched");
}

// ---------------------------------------------------------------- kernels
// X[s][j]: slices of register r in the canonical layout (field at sh), FP64.
template <typename C>
__global__ void k_grid_slices(double2* __restrict__ X, const C* __restrict__ st, uint64_t S, int sh, int n) {
    const uint64_t N = 1ull << n, tot = S * N, slo = (1ull << sh) - 1ull;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t s = e >> n, j = e & (N - 1);
        const C v = st[((s & ~slo) << n) | (j << sh) | (s & slo)];
        X[e] = make_double2((double)v.x, (double)v.y);
    }
}

// Y[x][s][k] = spec[s][k] e^{2 pi i kk t_x / N} / N (kk the signed frequency);
// an inverse FFT along k then samples the interpolant at x_j + t_x dx.
__global__ void k_grid_shift(double2* __restrict__ Y, const double2* __restrict__ spec, uint64_t S, int n,
                             const double* __restrict__ tx, int nxc) {
    const uint64_t N = 1ull << n, per = S * N, tot = per * (uint64_t)nxc;
    const double invN = 1.0 / (double)N;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t x = e / per, sk = e - x * per, k = sk & (N - 1);
        const int64_t kk = k < N / 2 ? (int64_t)k : (int64_t)k - (int64_t)N;
        const double2 u = unit_rev(cvdv_frac((double)kk * tx[x] * invN), 0.0);
        const double2 c = spec[sk];
        Y[e] = make_double2((c.x * u.x - c.y * u.y) * invN, (c.x * u.y + c.y * u.x) * invN);
    }
}

// A[(s,) x][d + D] = sum_s (or slice s) phi[x][s][i_x - d] conj(phi[x][s][i_x + d]).
__global__ void k_grid_pairs(double2* __restrict__ A, const double2* __restrict__ phi, uint64_t S, int n,
                             const int* __restrict__ ix, int nxc, int perSlice) {
    const int N = 1 << n, D = N / 2, nd = 2 * D + 1;
    const uint64_t blocks = perSlice ? S * (uint64_t)nxc : (uint64_t)nxc;
    const uint64_t tot = blocks * (uint64_t)nd;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t row = e / nd;
        const int d = (int)(e - row * nd) - D;
        const uint64_t x = perSlice ? row % nxc : row;
        const uint64_t s0 = perSlice ? row / nxc : 0, s1 = perSlice ? s0 + 1 : S;
        const int i = ix[x], a = i - d, b = i + d;
        double re = 0.0, im = 0.0;
        if (a >= 0 && a < N && b >= 0 && b < N) {
            for (uint64_t s = s0; s < s1; ++s) {
                const double2 u = phi[(x * S + s) * N + a], v = phi[(x * S + s) * N + b];
                re += u.x * v.x + u.y * v.y;  // u conj(v)
                im += u.y * v.x - u.x * v.y;
            }
        }
        A[e] = make_double2(re, im);
    }
}

// E[p][d + D] = e^{i dx p m(d)}, m = 2d + 2o - (N-1) (native) or 2d.
__global__ void k_grid_wigner_E(double2* __restrict__ E, const double* __restrict__ ps, int np, int n, double dx,
                                int native) {
    const int N = 1 << n, D = N / 2, nd = 2 * D + 1, o = (N - 1) / 2;
    const uint64_t tot = (uint64_t)np * nd;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const int p = (int)(e / nd), d = (int)(e % nd) - D;
        const double m = native ? (double)(2 * (d + o) - (N - 1)) : (double)(2 * d);
        E[e] = unit_rev(cvdv_frac(cvdv_frac(dx * ps[p] / (2.0 * PI)) * m), 0.0);
    }
}

// B[s][q][j] = X[s][j] e^{-(x_j - q)^2 / 2} for the q chunk.
__global__ void k_grid_husimi_B(double2* __restrict__ B, const double2* __restrict__ X, uint64_t S, int n,
                                const double* __restrict__ qs, int nqc, double dx) {
    const uint64_t N = 1ull << n, tot = S * (uint64_t)nqc * N;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t j = e & (N - 1), sq = e >> n, q = sq % nqc, s = sq / nqc;
        const double x = ((double)j - 0.5 * (double)(N - 1)) * dx - qs[q];
        const double w = exp(-0.5 * x * x);
        const double2 v = X[s * N + j];
        B[e] = make_double2(v.x * w, v.y * w);
    }
}

// E2[p][j] = e^{-i p x_j}, x_j = (2j - N + 1) dx / 2.
__global__ void k_grid_husimi_E(double2* __restrict__ E, const double* __restrict__ ps, int np, int n, double dx) {
    const uint64_t N = 1ull << n, tot = (uint64_t)np * N;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t p = e >> n, j = e & (N - 1);
        const double m = (double)(2 * (int64_t)j - (int64_t)N + 1);
        E[e] = unit_rev(cvdv_frac(cvdv_frac(-dx * ps[p] / (4.0 * PI)) * m), 0.0);
    }
}

__global__ void k_grid_real(double* __restrict__ out, const double2* __restrict__ Z, uint64_t tot, double scale) {
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot;
         e += (uint64_t)gridDim.x * blockDim.x)
        out[e] = Z[e].x * scale;
}

// out[(s,) p][q] = scale * (|H_s|^2, or summed over s in slice order).
__global__ void k_grid_abs2(double* __restrict__ out, const double2* __restrict__ H, uint64_t S, uint64_t per,
                            int perSlice, double scale) {
    const uint64_t tot = perSlice ? S * per : per;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot;
         e += (uint64_t)gridDim.x * blockDim.x) {
        double acc = 0.0;
        if (perSlice) {
            const double2 h = H[e];
            acc = h.x * h.x + h.y * h.y;
        } else {
            for (uint64_t s = 0; s < S; ++s) {
                const double2 h = H[s * per + e];
                acc += h.x * h.x + h.y * h.y;
            }
        }
        out[e] = acc * scale;
    }
}

// ---------------------------------------------------------------- drivers
// x -> (i, t) with x = x_i + t dx; within 1e-9 of a grid point: t = 0.
static void grid_split(double x, int N, double dx, int& i, double& t) {
    const double u = (x + 0.5 * (double)(N - 1) * dx) / dx;
    const double r = std::nearbyint(u);
    if (std::fabs(u - r) < 1e-9) {
        i = (int)r;
        t = 0.0;
    } else {
        const double f = std::floor(u);
        i = (int)f;
        t = u - f;
    }
    if (i < 0 || !std::isfinite(u)) {
        i = 0;
        t = 0.0;
    }
    if (i > N - 1) {
        i = N - 1;
        t = 0.0;
    }
}

// Canonical slices of register r (pending phases applied, lazy layout undone).
static double2* grid_slices(CVDVContext* ctx, int r, double2* X) {
    ctx_materialize(ctx);
    const int n = ctx->nq[r];
    const uint64_t S = ctx->totalSize >> n, tot = ctx->totalSize;
    if (ctx->prec == CVDV_C128)
        hipLaunchKernelGGL(k_grid_slices<double2>, dim3(grid_for(tot)), dim3(256), 0, ctx->stream, X,
                           (const double2*)ctx->state(), S, ctx->flw[r], n);
    else
        hipLaunchKernelGGL(k_grid_slices<float2>, dim3(grid_for(tot)), dim3(256), 0, ctx->stream, X,
                           (const float2*)ctx->state(), S, ctx->flw[r], n);
    HIPCHK(hipGetLastError());
    return X;
}

template <typename T>
static T* upload(CVDVContext* ctx, T* dst, const std::vector<T>& v) {
    HIPCHK(hipMemcpyAsync(dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
    return dst;
}

static void get_wigner_grid(CVDVContext* ctx, int r, const double* xs, int nx, const double* ps, int np,
                            int perSlice, int native, double* out) {
    const int n = ctx->nq[r], N = 1 << n, D = N / 2, nd = 2 * D + 1;
    const double dx = ctx->dx[r];
    const uint64_t A = ctx->totalSize, S = A >> n;
    const uint64_t outBlocks = perSlice ? S : 1;
    // x chunk: a power of two (the batched FFT over the chunk's S rows per point
    // needs a power-of-two row count); the last chunk is padded with t = 0 rows
    int nxc = 1;
    while (nxc < nx && (uint64_t)nxc * 2 * S * N <= GRID_CHUNK && (uint64_t)nxc * 2 * nd * outBlocks <= GRID_CHUNK)
        nxc *= 2;
    const int nChunks = (nx + nxc - 1) / nxc;
    const int nxPad = nChunks * nxc;
    std::vector<int> hi(nxPad, 0);
    std::vector<double> ht(nxPad, 0.0);
    for (int x = 0; x < nx; ++x) grid_split(xs[x], N, dx, hi[x], ht[x]);
    // scratch: X | spec | Y (nxc S N) | tmp (nxc S N) | Apairs (outBlocks nxc nd) | E (np nd) | Z (outBlocks np nx)
    //          | ix | tx | ps | out
    const uint64_t nY = (uint64_t)nxc * S * N, nA = outBlocks * nxc * nd, nE = (uint64_t)np * nd;
    const uint64_t nZ = outBlocks * (uint64_t)np * nx;
    const size_t bytes = (2 * A + 2 * nY + nA + nE + nZ) * 16 + (size_t)nxPad * 12 + (size_t)np * 8 + nZ * 8 + 256;
    char* base = (char*)ctx_scratch(ctx, bytes);
    double2* X = (double2*)base;
    double2* spec = X + A;
    double2* Y = spec + A;
    double2* tmp = Y + nY;
    double2* Ap = tmp + nY;
    double2* E = Ap + nA;
    double2* Z = E + nE;
    double* dps = (double*)(Z + nZ);
    double* dtx = dps + np;
    double* dout = dtx + nxPad;
    int* dix = (int*)(dout + nZ);
    grid_slices(ctx, r, X);
    upload(ctx, dps, std::vector<double>(ps, ps + np));
    upload(ctx, dtx, ht);
    upload(ctx, dix, hi);
    // spectra of the slices (unnormalised forward FFT)
    HIPCHK(hipMemcpyAsync(spec, X, A * 16, hipMemcpyDeviceToDevice, ctx->stream));
    if (n > 0) fft_generic_c128(ctx, spec, tmp, A, 0, n, false, 1.0);
    hipLaunchKernelGGL(k_grid_wigner_E, dim3(grid_for(nE)), dim3(256), 0, ctx->stream, E, dps, np, n, dx, native);
    HIPCHK(hipGetLastError());
    for (int x0 = 0; x0 < nx; x0 += nxc) {
        const int c = nx - x0 < nxc ? nx - x0 : nxc;
        const uint64_t ny = (uint64_t)nxc * S * N;  // padded rows keep the FFT's row count a power of two
        hipLaunchKernelGGL(k_grid_shift, dim3(grid_for(ny)), dim3(256), 0, ctx->stream, Y, spec, S, n, dtx + x0, nxc);
        HIPCHK(hipGetLastError());
        if (n > 0) fft_generic_c128(ctx, Y, tmp, ny, 0, n, true, 1.0);
        const uint64_t na = (perSlice ? S : 1) * (uint64_t)c * nd;
        hipLaunchKernelGGL(k_grid_pairs, dim3(grid_for(na)), dim3(256), 0, ctx->stream, Ap, Y, S, n, dix + x0, c,
                           perSlice);
        HIPCHK(hipGetLastError());
        // Z[(s,) p][x0 + x] = sum_d E[p][d] Ap[(s,) x][d]
        zgemm_tn(ctx, c, np, nd, Ap, nd, (uint64_t)c * nd, E, nd, 0, Z + x0, nx, (uint64_t)np * nx,
                 (int)outBlocks);
    }
    hipLaunchKernelGGL(k_grid_real, dim3(grid_for(nZ)), dim3(256), 0, ctx->stream, dout, Z, nZ, dx / PI);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, dout, nZ * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
}

static void get_husimi_grid(CVDVContext* ctx, int r, const double* qs, int nq, const double* ps, int np,
                            int perSlice, double* out) {
    const int n = ctx->nq[r], N = 1 << n;
    const double dx = ctx->dx[r];
    const uint64_t A = ctx->totalSize, S = A >> n;
    int nqc = nq;
    while (nqc > 1 && (uint64_t)nqc * S * N > GRID_CHUNK) nqc = (nqc + 1) / 2;
    const uint64_t nB = (uint64_t)nqc * S * N, nE = (uint64_t)np * N, nH = S * (uint64_t)np * nq;
    const uint64_t nOut = (perSlice ? S : 1) * (uint64_t)np * nq;
    const size_t bytes = (A + nB + nE + nH) * 16 + ((size_t)nq + np + nOut) * 8 + 256;
    char* base = (char*)ctx_scratch(ctx, bytes);
    double2* X = (double2*)base;
    double2* B = X + A;
    double2* E = B + nB;
    double2* H = E + nE;
    double* dqs = (double*)(H + nH);
    double* dps = dqs + nq;
    double* dout = dps + np;
    grid_slices(ctx, r, X);
    upload(ctx, dqs, std::vector<double>(qs, qs + nq));
    upload(ctx, dps, std::vector<double>(ps, ps + np));
    hipLaunchKernelGGL(k_grid_husimi_E, dim3(grid_for(nE)), dim3(256), 0, ctx->stream, E, dps, np, n, dx);
    HIPCHK(hipGetLastError());
    for (int q0 = 0; q0 < nq; q0 += nqc) {
        const int c = nq - q0 < nqc ? nq - q0 : nqc;
        const uint64_t nb = (uint64_t)c * S * N;
        hipLaunchKernelGGL(k_grid_husimi_B, dim3(grid_for(nb)), dim3(256), 0, ctx->stream, B, X, S, n, dqs + q0, c,
                           dx);
        HIPCHK(hipGetLastError());
        // H_s[p][q0 + q] = sum_j E[p][j] B[s][q][j]
        zgemm_tn(ctx, c, np, N, B, N, (uint64_t)c * N, E, N, 0, H + q0, nq, (uint64_t)np * nq, (int)S);
    }
    hipLaunchKernelGGL(k_grid_abs2, dim3(grid_for(nOut)), dim3(256), 0, ctx->stream, dout, H, S,
                       (uint64_t)np * nq, perSlice, dx / (PI * std::sqrt(PI)));
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, dout, nOut * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
}

static bool grid_ok(CVDVContext* ctx, int r, const double* a, int na, const double* b, int nb, double* out) {
    if (!ctx || !out || refuse_recording(ctx)) return false;
    if (!ctx->initialized) {
        cvdv_set_error(CVDV_ERR_STATE, "State not initialized");
        return false;
    }
    if (r < 0 || r >= ctx->numReg) {
        cvdv_set_error(CVDV_ERR_ARG, "Invalid register index: " + std::to_string(r));
        return false;
    }
    if (!a || !b || na <= 0 || nb <= 0) {
        cvdv_set_error(CVDV_ERR_ARG, "grid readout: empty point list");
        return false;
    }
    for (int i = 0; i < na; ++i)
        if (!std::isfinite(a[i])) {
            cvdv_set_error(CVDV_ERR_ARG, "grid readout: non-finite coordinate");
            return false;
        }
    for (int i = 0; i < nb; ++i)
        if (!std::isfinite(b[i])) {
            cvdv_set_error(CVDV_ERR_ARG, "grid readout: non-finite coordinate");
            return false;
        }
    return true;
}

#define GRID_TRY try {
#define GRID_CATCH()                                                         \
    }                                                                        \
    catch (const CvdvError& e) {                                             \
        cvdv_set_error(e.code, e.msg);                                       \
    }                                                                        \
    catch (...) {                                                            \
        cvdv_set_error(CVDV_ERR_HIP, "unexpected internal error");           \
    }

extern "C" {

void cvdvGetWignerGrid(CVDVContext* ctx, int regIdx, const double* xs, int nx, const double* ps, int np,
                       int perSlice, int native, double* out) {
    if (!grid_ok(ctx, regIdx, xs, nx, ps, np, out)) return;
    GRID_TRY get_wigner_grid(ctx, regIdx, xs, nx, ps, np, perSlice ? 1 : 0, native ? 1 : 0, out); GRID_CATCH()
}

void cvdvGetHusimiQGrid(CVDVContext* ctx, int regIdx, const double* qs, int nq, const double* ps, int np,
                        int perSlice, double* out) {
    if (!grid_ok(ctx, regIdx, qs, nq, ps, np, out)) return;
    GRID_TRY get_husimi_grid(ctx, regIdx, qs, nq, ps, np, perSlice ? 1 : 0, out); GRID_CATCH()
}

}  // extern "C"
