/* cvdv_hip.h - C ABI of libcvdv_hip.so, the MI355X (gfx950) hybrid CV-DV
 * state-vector engine.
 *
 * Every function named cvdv* below except those marked "addition" has the
 * same name, argument list and meaning as the reference's C ABI in
 * src/cuda/api.h (line cited per entry), so a caller of libcvdv.so (e.g.
 * src/cudaCvdv.py:35-73) can load this library instead without change.
 *
 * Conventions (identical to the reference): register 0 is the most
 * significant part of the flat index; qubit k of register r is bit
 * (n_r - 1 - k) of that register's local index; controlled gates apply +phase
 * when the control bit is 0 and -phase when it is 1; grid x_j = (j-(N-1)/2)dx,
 * dx = sqrt(2 pi / N).  Device pointers passed in hold complex128 data.
 *
 * Errors: invalid arguments print a message to stderr and leave the state
 * unchanged, as the reference does; in addition every failure (including HIP
 * runtime failures, which the reference turns into C++ exceptions) is recorded
 * and can be read with cvdvGetLastError.  No C++ exception crosses this ABI.
 */
#ifndef CVDV_HIP_H
#define CVDV_HIP_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct CVDVContext CVDVContext;

/* ---- lifecycle and state (api.h:14-23) ---- */
CVDVContext* cvdvCreate(int numReg, int* numQubits);                 /* api.h:14 */
void   cvdvDestroy(CVDVContext* ctx);                                /* api.h:15 */
void   cvdvFree(CVDVContext* ctx);                                   /* api.h:16 */
void   cvdvInitFromSeparable(CVDVContext* ctx, void** devicePtrs, int numReg); /* api.h:17 */
void   cvdvSetStateFromDevicePtr(CVDVContext* ctx, void* d_src);     /* api.h:18 */
int    cvdvGetNumRegisters(CVDVContext* ctx);                        /* api.h:19 */
size_t cvdvGetTotalSize(CVDVContext* ctx);                           /* api.h:20 */
void   cvdvGetRegisterInfo(CVDVContext* ctx, int* qubitCountsOut, double* gridStepsOut); /* api.h:21 */
int    cvdvGetRegisterDim(CVDVContext* ctx, int regIdx);             /* api.h:22 */
double cvdvGetRegisterDx(CVDVContext* ctx, int regIdx);              /* api.h:23 */

/* ---- gates (api.h:25-46) ---- */
void cvdvFtQ2P(CVDVContext* ctx, int regIdx);                        /* api.h:25 */
void cvdvFtP2Q(CVDVContext* ctx, int regIdx);                        /* api.h:26 */
void cvdvDisplacement(CVDVContext* ctx, int regIdx, double betaRe, double betaIm); /* api.h:27 */
void cvdvConditionalDisplacement(CVDVContext* ctx, int targetReg, int ctrlReg, int ctrlQubit,
                                 double betaRe, double betaIm);      /* api.h:28-29 */
void cvdvPauliRotation(CVDVContext* ctx, int regIdx, int qubitIdx, int axis, double theta); /* api.h:30 */
void cvdvHadamard(CVDVContext* ctx, int regIdx, int qubitIdx);       /* api.h:31 */
void cvdvParity(CVDVContext* ctx, int regIdx);                       /* api.h:32 */
void cvdvConditionalParity(CVDVContext* ctx, int targetReg, int ctrlReg, int ctrlQubit); /* api.h:33 */
void cvdvSwapRegisters(CVDVContext* ctx, int reg1, int reg2);        /* api.h:34 */
void cvdvPhaseSquare(CVDVContext* ctx, int regIdx, double t);        /* api.h:35 */
void cvdvPhaseCubic(CVDVContext* ctx, int regIdx, double t);         /* api.h:36 */
void cvdvRotation(CVDVContext* ctx, int regIdx, double theta);       /* api.h:37 */
void cvdvConditionalRotation(CVDVContext* ctx, int targetReg, int ctrlReg, int ctrlQubit,
                             double theta);                          /* api.h:38-39 */
void cvdvSqueeze(CVDVContext* ctx, int regIdx, double r);            /* api.h:40 */
void cvdvConditionalSqueeze(CVDVContext* ctx, int targetReg, int ctrlReg, int ctrlQubit,
                            double r);                               /* api.h:41-42 */
void cvdvBeamSplitter(CVDVContext* ctx, int reg1, int reg2, double theta); /* api.h:43 */
void cvdvConditionalBeamSplitter(CVDVContext* ctx, int reg1, int reg2, int ctrlReg,
                                 int ctrlQubit, double theta);       /* api.h:44-45 */
void cvdvQ1Q2Gate(CVDVContext* ctx, int reg1, int reg2, double coeff); /* api.h:46 */

/* ---- phase-space readout (api.h:48-51); outputs are row-major N x N doubles,
 *      rows = momentum, columns = position ---- */
void cvdvGetWigner(CVDVContext* ctx, int regIdx, double* wignerOut);   /* api.h:48 */
void cvdvGetHusimiQ(CVDVContext* ctx, int regIdx, double* husimiOut);  /* api.h:49 */
void cvdvGetHusimiQOverlap(CVDVContext* ctx, int regIdx, double* husimiOut); /* api.h:50 */
void cvdvGetHusimiQWigner(CVDVContext* ctx, int regIdx, double* husimiOut);  /* api.h:51 */

/* ---- measurement and state readout (api.h:53-61) ---- */
/* probsOut holds prod(dim) doubles, column-major: first selected register fastest. */
void   cvdvMeasureMultiple(CVDVContext* ctx, const int* regIdxs, int numRegs, double* probsOut); /* api.h:53-54 */
/* The library-reduction comparator (api.h:55-56; the reference's cuTENSOR
 * route): |psi|^2 of the canonical state reduced by one hipTensor reduction,
 * same output as cvdvMeasureMultiple to FP64 rounding (the library's summation
 * order, not deterministic).  For comparisons and timing; selections beyond
 * what hipTensor supports fail with CVDV_ERR_UNSUPPORTED. */
void   cvdvMeasureMultipleCT(CVDVContext* ctx, const int* regIdxs, int numRegs, double* probsOut); /* api.h:55-56 */
void   cvdvGetState(CVDVContext* ctx, double* realOut, double* imagOut);    /* api.h:57 */
double cvdvGetNorm(CVDVContext* ctx);                                      /* api.h:58 */
void   cvdvGetFidelity(CVDVContext* ctx, void** devicePtrs, int numReg, double* fidOut); /* api.h:59 */
double cvdvGetPhotonNumber(CVDVContext* ctx, int regIdx);                  /* api.h:60 */
void   cvdvFidelityStatevectors(CVDVContext* ctx1, CVDVContext* ctx2, double* fidOut); /* api.h:61 */

/* ---- additions (no counterpart in api.h) ---- */
/* Precision mode: 0 = complex128 (as cvdvCreate), 1 = mixed (complex64
 * amplitudes; FP64 phase angles with range reduction, FP64 twiddle
 * generation, FP64 compensated probability/norm sums). */
CVDVContext* cvdvCreateEx(int numReg, int* numQubits, int precision);
int    cvdvGetPrecision(CVDVContext* ctx);
/* Returns the last error code (0 = none) of this thread and copies its message
 * (truncated to len) into buf when buf is not NULL; clears it. */
int    cvdvGetLastError(char* buf, size_t len);
/* Blocks until every operation issued on the context has finished. */
void   cvdvSynchronize(CVDVContext* ctx);
/* Applies any deferred diagonal phases and restores the canonical register
 * layout (both are otherwise applied lazily by the next kernel). */
void   cvdvFlush(CVDVContext* ctx);
/* Consecutive Fourier transforms of one register that the four-step split
 * serves (registers of 2^10 .. 2^16 points whose tile does not hold them whole)
 * share HBM passes: each transform runs as two half passes, and the second half
 * of one is fused with the first half of the next, the phases between them
 * applied in the same pass (m transforms: m + 1 passes instead of 2m).  on = 0
 * runs every transform whole (two passes each), the reference's pass structure;
 * results agree to FFT rounding.  Default on. */
void   cvdvSetTransformFusion(CVDVContext* ctx, int on);
/* FFT autotuner (the reference leaves plan choices to cuFFT, context.cu:48-65).
 * For every register field of the context it times the distinct plans of the
 * planner's free choices (columns per strided tile, how far a tile grows over
 * row blocks, transform fusion on or off) as chains of forward and inverse
 * transforms on a private context of the same shape (its own state buffers,
 * allocated for the call: the caller's state is not touched), and keeps the
 * fastest per field in this context (another plan replaces the default only
 * when it is at least 2% faster).  Results are kept per (precision, state
 * size, field) for the process: a later call for the same shape times nothing.
 * Completes a pending half transform first.  Returns the number of plans
 * timed, or -1 on error. */
int    cvdvFftAutotune(CVDVContext* ctx);
/* The choices in force for register regIdx's field: 1 and the tuned values
 * (0 / -1 members: the planner's default) with the timed milliseconds per
 * transform, 0 when the field was not tuned, -1 on error. */
int    cvdvFftGetTuning(CVDVContext* ctx, int regIdx, int* rowElems, int* tileCap, int* fuse,
                        double* msPerTransform);
/* Copies the state (canonical layout, complex128) to a caller device buffer. */
void   cvdvGetStateDevice(CVDVContext* ctx, void* d_dst);
/* Per-launch timing with HIP events recorded on the context's stream.
 * kind: 0 FFT passes (all of them), 1 diagonal passes, 2 qubit-pair passes,
 * 3 conditional parity, 4 reductions, 5 layout/other, and among the FFT
 * passes: 6 fused half-transform passes (cvdvSetTransformFusion), 7 one-row
 * tiles (a whole contiguous register of 2^12 points or more per tile).  Read
 * returns the launch count and the summed (and largest) elapsed milliseconds
 * since the last enable/reset. */
void   cvdvProfileEnable(CVDVContext* ctx, int on);
int    cvdvProfileRead(CVDVContext* ctx, int kind, double* totalMs, double* maxMs);
void   cvdvProfileReset(CVDVContext* ctx);

/* ---- phase-space readouts into device memory (addition) ----
 * cvdvGetWigner / cvdvGetHusimiQ (same arrays, same route choice) written to
 * caller device memory (N x N doubles); they return when the result is
 * complete.  The host-output forms spend most of their time copying an N x N
 * result into pageable host memory (128 MiB at N = 4096). */
void   cvdvGetWignerDevice(CVDVContext* ctx, int regIdx, double* d_out);
void   cvdvGetHusimiQDevice(CVDVContext* ctx, int regIdx, double* d_out);
/* cvdvGetHusimiQ picks its route as the reference does (bosonic.cu:13-25): the
 * overlap route unless its memory estimate exceeds 80% of the free device
 * memory, then the Wigner -> 2D FFT route (cvdvGetHusimiQWigner: other grid and
 * scale).  bytes > 0 replaces the device's free-memory figure in that choice;
 * 0 asks the device (the default). */
void   cvdvSetFreeMemoryOverride(CVDVContext* ctx, size_t bytes);

/* ---- graph capture of a gate sequence (addition) ----
 * cvdvGraphBegin brings the context to its canonical layout with no pending
 * phases, allocates everything a gate may need and starts capturing the
 * context's stream; the gate calls that follow are recorded, not executed, and
 * readouts are refused until cvdvGraphEnd, which returns the sequence as an
 * instantiated hipGraph (NULL on failure; the state is untouched either way).
 * cvdvGraphLaunch applies the sequence to the context's current state
 * (asynchronously, replays compose).  Destroy graphs before their context; a
 * graph is refused after cvdvFree of its context. */
typedef struct CVDVGraph CVDVGraph;
void   cvdvGraphBegin(CVDVContext* ctx);
CVDVGraph* cvdvGraphEnd(CVDVContext* ctx);
void   cvdvGraphLaunch(CVDVGraph* graph);
void   cvdvGraphDestroy(CVDVGraph* graph);

/* ---- phase space on arbitrary points (addition) ----
 * Wigner / Husimi-Q of register regIdx at the points (xs[i], ps[k]) (host
 * arrays), written row-major out[k * nx + i] (rows momentum, columns position,
 * as cvdvGetWigner).  perSlice = 0: traced over the other registers;
 * perSlice = 1: one np x nx block per slice, slices in the canonical order of
 * the other registers' flat index (one block per branch).  The Wigner is the
 * discrete sum of cvdvGetWigner with each slice's trigonometric interpolant
 * between grid points; native = 1 keeps cvdvGetWigner's e^{-i (N-1) dx p}
 * finalisation (equal to cvdvGetWigner on its grid), native = 0 drops it.
 * The Husimi is cvdvGetHusimiQ's overlap sum at arbitrary (q, p) (equal to it
 * on its grid).  FP64 throughout. */
void   cvdvGetWignerGrid(CVDVContext* ctx, int regIdx, const double* xs, int nx, const double* ps, int np,
                         int perSlice, int native, double* out);
void   cvdvGetHusimiQGrid(CVDVContext* ctx, int regIdx, const double* qs, int nq, const double* ps, int np,
                          int perSlice, double* out);

/* ---- batched circuits (addition) ----
 * Many small independent circuits executed by one kernel launch, one
 * workgroup per circuit.  numQubitsFlat lists every circuit's register sizes
 * in order (numRegs[i] entries for circuit i); at most 8 registers of at most
 * 12 qubits and 20 qubits in total per circuit.  cvdvBatchCircuit returns a
 * recording context: every gate entry point above applies to it and records
 * the gate (same decompositions); readouts and state setters on it fail with
 * CVDV_ERR_STATE, and it is destroyed with the batch.  The initial state of
 * each circuit is the product of its register arrays (hostC128: interleaved
 * re/im doubles, circuits in order, registers in order).  cvdvBatchRun starts
 * every circuit from that initial state and runs its recorded gates
 * (asynchronous; recordings may be extended between runs).  Measure returns,
 * for every circuit, the marginal of register regIdx (canonical order,
 * 2^n entries per circuit, concatenated); Norms one sum of |psi|^2 per
 * circuit; both are deterministic compensated FP64 sums. */
typedef struct CVDVBatch CVDVBatch;
CVDVBatch* cvdvBatchCreate(int numCircuits, const int* numRegs, const int* numQubitsFlat, int precision);
void   cvdvBatchDestroy(CVDVBatch* batch);
int    cvdvBatchNumCircuits(CVDVBatch* batch);
CVDVContext* cvdvBatchCircuit(CVDVBatch* batch, int i);
size_t cvdvBatchStateSize(CVDVBatch* batch, int i);
size_t cvdvBatchNumOps(CVDVBatch* batch);
void   cvdvBatchInitFromSeparable(CVDVBatch* batch, const double* hostC128);
/* Launch shape: threads per workgroup (256 or 512) and the LDS budget in
 * bytes; a state held in LDS takes 9/8 of its size (one pad slot per 8
 * amplitudes), and circuits whose padded state exceeds the budget run from a
 * global work area.
 * Autotune times every shape (reps runs each), keeps the fastest and returns
 * its time per run in microseconds. */
void   cvdvBatchSetLaunch(CVDVBatch* batch, int threads, int ldsBytes);
void   cvdvBatchGetLaunch(CVDVBatch* batch, int* threads, int* ldsBytes);
double cvdvBatchAutotune(CVDVBatch* batch, int reps);
void   cvdvBatchRun(CVDVBatch* batch);
void   cvdvBatchSynchronize(CVDVBatch* batch);
void   cvdvBatchMeasure(CVDVBatch* batch, int regIdx, double* probsOut);
void   cvdvBatchNorms(CVDVBatch* batch, double* normsOut);
void   cvdvBatchGetState(CVDVBatch* batch, int i, double* realOut, double* imagOut);

/* ---- allocation-free marginal (addition; no counterpart in the reference) ----
 * For callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * output and errors as cvdvMeasureMultiple, and the same deterministic
 * compensated reduction, but the per-split partials live in a fixed 8 MiB
 * device array of the library (one per GPU, held until the process ends):
 * no hipMalloc, hipFree or allocating library call.  Marginals of more than
 * 2^20 bins are refused (CVDV_ERR_ARG); fewer splits may be used than by
 * cvdvMeasureMultiple, so results can differ from it by FP64 rounding.  Calls
 * that overlap on one GPU (several host threads, several contexts) share the
 * array and race: serialise them. */
void   cvdvMeasureMultipleNoAlloc(CVDVContext* ctx, const int* regIdxs, int numRegs, double* probsOut);

/* ---- copy-bandwidth probe (addition; measurement only) ----
 * Times a plain copy of one 2^28-amplitude complex64 state (2 GiB read + 2 GiB
 * written, buffers allocated and freed inside) with the access pattern of one
 * FFT pass of a [14,14] state and no arithmetic; writes the mean milliseconds
 * per copy over reps launches (after one warmup) to msOut.  Returns 0 or an
 * error code.  Patterns: */
enum {
    CVDV_PROBE_COPY = 0,      /* float4 grid-stride copy (the counter calibration kernel) */
    CVDV_PROBE_MEMCPY = 1,    /* hipMemcpyAsync device to device */
    CVDV_PROBE_STRIDED = 2,   /* 128 rows x 256-byte runs at stride 16 MiB, in place */
    CVDV_PROBE_TRANSPOSE = 3, /* rows at stride 128 KiB written at stride 16 MiB, out of place */
    CVDV_PROBE_MID = 4,       /* rows at stride 128 KiB, in place */
    CVDV_PROBE_ROW = 5,       /* one contiguous 128 KiB row per workgroup, one workgroup per CU */
    CVDV_PROBE_COUNT = 6
};
int    cvdvCopyProbe(int pattern, int reps, double* msOut);

#ifdef __cplusplus
}
#endif
#endif /* CVDV_HIP_H */
