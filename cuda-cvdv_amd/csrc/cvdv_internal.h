// cvdv_internal.h - shared host/device definitions of libcvdv_hip.
//
// State layout.  The state is one flat vector of 2^T amplitudes (T = sum of
// register qubit counts), register 0 most significant, exactly as in the
// reference (src/cuda/context.cu:27-38, src/cuda/kernels/utils.cuh:34-37).
// Amplitudes are complex128 (double2) or, in the mixed mode, complex64 (float2)
// with every phase angle, twiddle and probability sum kept in FP64.
//
// Lazy layout.  Each logical register r lives in a physical bit field
// [pshift[r], pshift[r]+nq[r]) of the flat index and may be stored reflected
// (physical j = N-1-logical j).  SWAP of equal registers and PARITY only update
// this map; every kernel receives physical shifts.  Readouts that expose the
// flat order (getState, state-state fidelity) first materialise the canonical
// layout.
//
// Deferred diagonal phases.  Diagonal gates e^{i f(x)} are not launched on
// their own: they accumulate in a DiagProg (phase terms in revolutions over
// physical indices) that the next state-touching kernel applies to each
// element as it is loaded.  A register's grid point x_j = (j-(N-1)/2) dx is
// written x = m dx / 2 with the exact odd integer m = 2j-N+1, so every term is
// a coefficient times an exact integer monomial (m, m^2, m^3, m1*m2).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#define CVDV_MAX_REG 32
#define CVDV_MAX_TERMS 12

enum CvdvPrecision { CVDV_C128 = 0, CVDV_MIXED = 1 };

enum CvdvErr {
    CVDV_OK = 0,
    CVDV_ERR_ARG = 1,      // invalid argument (the reference prints and no-ops)
    CVDV_ERR_STATE = 2,    // state not initialised
    CVDV_ERR_HIP = 3,      // HIP runtime failure
    CVDV_ERR_MEMORY = 4,   // allocation failure
    CVDV_ERR_UNSUPPORTED = 5,
};

// ---------------------------------------------------------------- diag program
// DT_LINF: a linear term c1 m over one field that its consumer evaluates in
// place instead of reading a table (see diag_fold_linear); evaluated as DT_POLY.
enum DiagKind { DT_POLY = 1, DT_BILIN = 2, DT_ZBIT = 3, DT_LINF = 4 };

struct DiagTerm {
    int kind;
    int sh1, nb1;  // field 1: m1 = 2*((g>>sh1)&(2^nb1-1)) + 1 - 2^nb1 ; ZBIT: bit sh1
    int sh2, nb2;  // field 2 (BILIN)
    int ctrl;      // control bit position (sign -1 when set), -1 = unconditional
    double c1, c2, c3;  // POLY: c1 m + c2 m^2 + c3 m^3 ; BILIN: c1 m1 m2 ; ZBIT: +-c1
};

struct DiagProg {
    int n;          // number of terms
    int pad;
    double grev;    // global phase, revolutions
    DiagTerm t[CVDV_MAX_TERMS];
};

__host__ __device__ inline double cvdv_frac(double r) { return r - rint(r); }

__device__ __forceinline__ double diag_field_m(uint64_t g, int sh, int nb) {
    const uint64_t j = (g >> sh) & ((1ull << nb) - 1ull);
    return (double)((int64_t)(2 * j) + 1 - ((int64_t)1 << nb));
}

// One term's phase of element g in revolutions (not yet reduced).
__device__ __forceinline__ double diag_term_rev(const DiagTerm& t, uint64_t g) {
    double v;
    if (t.kind == DT_POLY || t.kind == DT_LINF) {
        const double m = diag_field_m(g, t.sh1, t.nb1);
        v = fma(fma(t.c3, m, t.c2), m, t.c1) * m;
    } else if (t.kind == DT_BILIN) {
        const double m1 = diag_field_m(g, t.sh1, t.nb1);
        const double m2 = diag_field_m(g, t.sh2, t.nb2);
        v = t.c1 * (m1 * m2);  // m1*m2 is an exact integer below 2^53
    } else {
        v = ((g >> t.sh1) & 1ull) ? t.c1 : -t.c1;
    }
    if (t.ctrl >= 0 && ((g >> t.ctrl) & 1ull)) v = -v;
    return v;
}

// Phase of element g in revolutions, reduced to [-1/2, 1/2] term by term.
// The loop is unrolled to constant term indices so a program passed as a
// kernel argument is read in place (a runtime index would make the compiler
// copy the whole argument into per-lane scratch memory).
__device__ __forceinline__ double diag_rev(const DiagProg& p, uint64_t g) {
    double r = p.grev;
#pragma unroll
    for (int i = 0; i < CVDV_MAX_TERMS; ++i)
        if (i < p.n) r += cvdv_frac(diag_term_rev(p.t[i], g));
    return cvdv_frac(r);
}

// ---------------------------------------------------------------- complex helpers
template <typename R> struct CT;
template <> struct CT<float> { using type = float2; };
template <> struct CT<double> { using type = double2; };

template <typename C> __host__ __device__ inline C cmk(double x, double y) {
    C r; r.x = (decltype(r.x))x; r.y = (decltype(r.y))y; return r;
}
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {  // a * conj(b)
    return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {
    return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}

// e^{2 pi i f} for |f| <= 1/2.  Complex128: FP64 sincospi.  Mixed: the angle
// is reduced in FP64 (diag_rev) and only the final sin/cos runs in FP32.
__device__ __forceinline__ double2 unit_rev(double f, double) {
    double s, c;
    sincospi(2.0 * f, &s, &c);
    return make_double2(c, s);
}
__device__ __forceinline__ float2 unit_rev(double f, float) {
    float s, c;
    sincospif((float)(2.0 * f), &s, &c);
    // Pull |e|^2 back to 1 (to first order, exactly in FP64) before the single
    // rounding to FP32, so repeated phase passes leave the norm unbiased.
    const double r2 = (double)c * c + (double)s * s;
    const double k = 1.5 - 0.5 * r2;
    return make_float2((float)(c * k), (float)(s * k));
}

template <typename C>
__device__ __forceinline__ C apply_diag(const DiagProg& p, uint64_t g, C v) {
    using R = decltype(v.x);
    return cmul(v, unit_rev(diag_rev(p, g), R(0)));
}

// ---------------------------------------------------------------- batch recording
// A recording context (one per circuit of a batch, batch.hip) runs the same gate
// compiler but appends these ops instead of launching kernels.
enum BatchOpKind { BOP_FFT = 1, BOP_QUBIT = 2, BOP_CPARITY = 3, BOP_DIAG = 4 };

struct BatchOp {
    int kind;
    int sh, nb, inv;     // register field (FFT, CPARITY target); inv: inverse FFT
    int bit;             // QUBIT: qubit bit; CPARITY: control bit
    int ctrlFlip, tgtRefl;
    int prog;            // index of the phase program applied on load, -1 = none
    double outScale;     // FFT: 1/sqrt(N)
    double2 m[4];        // QUBIT: 2x2 matrix, FP64
};

struct BatchRecording {
    std::vector<BatchOp> ops;
    std::vector<DiagProg> progs;
};

// ---------------------------------------------------------------- context
struct TwiddleCache;  // fft.hip

// Kernel kinds for the optional per-launch event timing (cvdvProfileEnable).
// KK_FFT counts every FFT pass when read; launches are recorded as KK_FFT_FUSED
// (two half transforms in one pass), KK_FFT_ROW (one whole contiguous row of
// 2^12 points or more per tile) or KK_FFT (any other FFT pass).
enum CvdvKernelKind { KK_FFT = 0, KK_DIAG = 1, KK_QUBIT = 2, KK_CPARITY = 3, KK_REDUCE = 4,
                      KK_OTHER = 5, KK_FFT_FUSED = 6, KK_FFT_ROW = 7, KK_COUNT = 8 };

struct ProfRec {
    int kind;
    hipEvent_t a, b;
};

// A register transform left half done (fft.hip).  The four-step FFT of a
// register field of 2n bits (digits lo = bits [sh, sh+n), hi = [sh+n, sh+2n))
// is two in-place half passes: H (transform over hi) and L (twiddle W_N^{lo*hi},
// transform over lo).  After H the second half is pending (!high: L, whose
// in-place result is the momentum register with its digits exchanged); after L
// of a second transform, the twiddle and H of that one (high).  The next
// transform of the same field fuses the pending half with its own first half
// in one pass (pending phases applied between them); anything else that needs
// the field in natural order completes it first.
struct FtDefer {
    int on = 0;
    int sh = 0, nb = 0;   // the register field
    int high = 0;         // pending half: 0 = L (+ digit exchange), 1 = twiddle + H
    int inv = 0;          // the pending transform is an inverse
    double scale = 1.0;   // its normalisation, not yet applied
};

// Tile choices of the FFT of one register field (fft.hip), as cvdvFftAutotune
// sets them; zero / negative members keep the planner's defaults.
struct FftCfg {
    int rowElems = 0;  // columns of a strided tile (default: 256 bytes of them)
    int tileCap = 0;   // elements a tile grows to over row blocks (default 2048)
    int fuse = -1;     // 0: this field's transforms run whole (no pending halves)
};
inline uint64_t fft_cfg_key(int sh, int nb) { return ((uint64_t)(uint32_t)sh << 8) | (uint32_t)nb; }

// Every register field with a half-done transform (at most one record per field).
struct FtDefers {
    FtDefer f[CVDV_MAX_REG];
    FtDefer* find(int sh) {
        for (FtDefer& d : f)
            if (d.on && d.sh == sh) return &d;
        return nullptr;
    }
    FtDefer* free_slot() {
        for (FtDefer& d : f)
            if (!d.on) return &d;
        return nullptr;
    }
    bool any() const {
        for (const FtDefer& d : f)
            if (d.on) return true;
        return false;
    }
};

// Host-side state at the start of a graph capture (restored at its end: nothing
// has executed).  Kept per context, so captures on several contexts may nest.
struct CaptureSave {
    int pshift[CVDV_MAX_REG], refl[CVDV_MAX_REG];
    DiagProg pending;
    int cur;
};

struct CVDVContext {
    int numReg = 0;
    int totalQbt = 0;
    int prec = CVDV_C128;
    size_t totalSize = 0;
    int nq[CVDV_MAX_REG] = {0};
    int flw[CVDV_MAX_REG] = {0};      // canonical shift (reference flwQbts)
    double dx[CVDV_MAX_REG] = {0};
    int pshift[CVDV_MAX_REG] = {0};   // physical shift of each logical register
    int refl[CVDV_MAX_REG] = {0};     // 1 = stored reflected
    void* buf[2] = {nullptr, nullptr};
    int cur = 0;
    bool initialized = false;
    hipStream_t stream = nullptr;
    DiagProg pending{};
    FtDefers defer{};                 // half-done register transforms (fft.hip)
    int fuseFt = 1;                   // 0: every transform runs whole (no deferred halves)
    std::map<uint64_t, FftCfg> fftCfg;  // tuned tile choices per register field (fft_cfg_key)
    size_t freeMemOverride = 0;       // readout route choice: assumed free device bytes (0 = ask)
    TwiddleCache* tw = nullptr;
    void* scratch = nullptr;          // readout scratch (grown on demand)
    size_t scratchBytes = 0;
    void* arena = nullptr;            // phase tables of the compiled pending program
    size_t arenaBytes = 0;
    BatchRecording* rec = nullptr;    // non-null: record ops, launch nothing
    void* blas = nullptr;             // rocBLAS handle of the grid readouts (created on first use)
    bool capturing = false;           // between cvdvGraphBegin and cvdvGraphEnd
    CaptureSave capSave{};            // host-side state at cvdvGraphBegin
    int graphs = 0;                   // graphs captured (they hold arena pointers)
    std::vector<void*> retired;       // arenas outgrown while a graph may still read them
    int generation = 0;               // bumped when the state buffers are released
    bool profOn = false;
    std::vector<ProfRec> prof;        // recorded launches
    std::vector<hipEvent_t> evPool;   // reusable events

    size_t elemBytes() const { return prec == CVDV_C128 ? 16 : 8; }
    void* state() const { return buf[cur]; }
    void* other();                    // allocates the ping-pong buffer on first use
    bool canonical() const;
};

// ---------------------------------------------------------------- errors
struct CvdvError {
    int code;
    std::string msg;
};

void cvdv_set_error(int code, const std::string& msg, bool print = true);
void cvdv_check_hip(hipError_t e, const char* what);  // throws CvdvError

#define HIPCHK(x) cvdv_check_hip((x), #x)

// ---------------------------------------------------------------- internal API
// gates.hip
void ctx_push_term(CVDVContext* ctx, const DiagTerm& t);
void ctx_flush(CVDVContext* ctx);              // apply pending diagonal in place
void ctx_materialize(CVDVContext* ctx);        // flush + canonical layout
void ctx_ft(CVDVContext* ctx, int reg, bool inverse);
DiagProg ctx_take_pending(CVDVContext* ctx);
void launch_diag(CVDVContext* ctx, const DiagProg& p);

// fft.hip
void fft_register(CVDVContext* ctx, int pshift, int nbits, bool inverse, const DiagProg& pre,
                  double outScale);
// largest dynamic LDS a workgroup may request on the current device
size_t device_max_lds();
void fft_generic_c128(CVDVContext* ctx, double2* data, double2* tmp, uint64_t total, int shift,
                      int nbits, bool inverse, double outScale);
void twiddle_cache_free(TwiddleCache* tc);
// Complete a half-done transform (no-op when none is pending).  mayAllocate:
// completing an L half writes the other state buffer; when it does not exist
// yet and allocation is not allowed, throws CVDV_ERR_STATE.
void fft_settle(CVDVContext* ctx, bool mayAllocate = true);
// A transform of field [sh, sh+nb) runs as halves and keeps one pending.
bool fft_defers(const CVDVContext* ctx, int sh, int nb);
// Tile geometry the planner picks for field [sh, sh+nb) of the context's state
// under cfg, folded into one number (equal numbers: the same passes); 0 when
// the FFT of that field takes no tile pass (direct DFT of a tiny state).
uint64_t fft_plan_signature(const CVDVContext* ctx, int sh, int nb, const FftCfg& cfg);
// A context of src's shape and precision in the canonical layout, its state
// zeroed, nothing pending (the FFT autotuner times candidates on it).
CVDVContext* ctx_create_like(const CVDVContext* src);
// Complete the half-done transforms whose fields the program reads, outside the
// bits `except`.
void fft_settle_for(CVDVContext* ctx, const DiagProg& p, uint64_t except = 0);
// Complete a half-done transform when the bits `mask` overlap its field.
void fft_settle_bits(CVDVContext* ctx, uint64_t mask);

// measure.hip
void marginal_probs(CVDVContext* ctx, const int* regs, int nsel, double* hostOut);

// scratch
void* ctx_scratch(CVDVContext* ctx, size_t bytes);

// batch recording (gates.hip): append an op whose phase program is `p`
void rec_op(CVDVContext* ctx, BatchOp op, const DiagProg& p);
// Everything a gate may allocate or upload, done before a graph capture (fft.hip)
void fft_prepare(CVDVContext* ctx);
void arena_reserve(CVDVContext* ctx, size_t bytes);  // gates.hip
// rocBLAS handle of the grid readouts (phase_grid.hip)
void blas_release(CVDVContext* ctx);
// readouts refuse recording contexts (no state of their own) and contexts that
// are capturing a graph (nothing has executed yet)
bool refuse_recording(CVDVContext* ctx);
// Stockham stage twiddles of length 2^logM in the given precision (fft.hip):
// device table (cached in tc) and a host copy for tables built elsewhere.
const void* fft_stage_table(TwiddleCache*& tc, int logM, int prec);
std::vector<uint8_t> fft_stage_host(int logM, int prec);
// A context for one circuit of a batch: records gates, owns no state (ctx.hip).
CVDVContext* ctx_create_recording(int numReg, const int* numQubits, int precision);

// profiling (ctx.hip): bracket a launch on ctx->stream when profiling is on
int prof_begin(CVDVContext* ctx, int kind);
void prof_end(CVDVContext* ctx, int idx);

inline uint64_t ctx_ctrl_bit(const CVDVContext* c, int reg, int qubit) {
    return (uint64_t)(c->pshift[reg] + c->nq[reg] - 1 - qubit);
}
