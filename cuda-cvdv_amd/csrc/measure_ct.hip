// measure_ct.hip - the library-reduction comparator for marginals.
//
// Replaces the reference's cuTENSOR route (src/cuda/measure.cu:125-256):
// |psi|^2 of the canonical state into an FP64 buffer, then one hipTensor
// reduction (opReduce = add, FP64 compute) over the unselected registers,
// output column-major with the first selected register fastest, as
// cvdvMeasureMultiple.  It exists to compare and time the own deterministic
// reduction (measure.hip) against a library one; the summation order is the
// library's, so results agree to FP64 rounding, not bit for bit.
//
// The reduction writes the kept registers in the state's own order and the
// host permutes the (small) result into the requested order, so registers that
// are adjacent in the state merge into one tensor mode whenever both are kept
// or both are reduced: the tensor rank is what the selection needs and no more
// (a joint marginal of 14 one-qubit registers next to an oscillator is a rank-2
// reduction, whatever order it is requested in).
//
// hipTensor is opened with dlopen on the first call, not linked: its load-time
// initialisation ends the process on a machine without a GPU, and the library
// must load there (builds, ABI checks).

#include <dlfcn.h>
#include <hiptensor/hiptensor.h>

#include <mutex>
#include <string>
#include <vector>

#include "cvdv_internal.h"
#include "../../include/cvdv_hip.h"

template <typename C>
__global__ void k_abs2(double* __restrict__ out, const C* __restrict__ st, uint64_t n) {
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += (uint64_t)gridDim.x * blockDim.x) {
        const C v = st[g];
        out[g] = (double)v.x * (double)v.x + (double)v.y * (double)v.y;
    }
}

namespace {

// The hipTensor entry points this file calls, resolved once per process.
struct HtApi {
#define CVDV_HT_FN(name) decltype(&::hiptensor##name) name = nullptr;
    CVDV_HT_FN(Create) CVDV_HT_FN(Destroy) CVDV_HT_FN(CreateTensorDescriptor) CVDV_HT_FN(DestroyTensorDescriptor)
    CVDV_HT_FN(CreateReduction) CVDV_HT_FN(DestroyOperationDescriptor) CVDV_HT_FN(CreatePlanPreference)
    CVDV_HT_FN(DestroyPlanPreference) CVDV_HT_FN(EstimateWorkspaceSize) CVDV_HT_FN(CreatePlan)
    CVDV_HT_FN(DestroyPlan) CVDV_HT_FN(Reduce) CVDV_HT_FN(GetErrorString)
#undef CVDV_HT_FN
};

const HtApi& ht() {
    static std::once_flag once;
    static HtApi api;
    static std::string err;
    std::call_once(once, [] {
        void* lib = dlopen("libhiptensor.so.0", RTLD_NOW | RTLD_LOCAL);
        if (!lib) lib = dlopen("/opt/rocm/lib/libhiptensor.so.0", RTLD_NOW | RTLD_LOCAL);
        if (!lib) {
            const char* e = dlerror();
            err = std::string("cannot open libhiptensor: ") + (e ? e : "unknown error");
            return;
        }
        HtApi a;
        bool ok = true;
#define CVDV_HT_SYM(name)                                                          \
    a.name = reinterpret_cast<decltype(a.name)>(dlsym(lib, "hiptensor" #name)); \
    ok = ok && a.name != nullptr;
        CVDV_HT_SYM(Create) CVDV_HT_SYM(Destroy) CVDV_HT_SYM(CreateTensorDescriptor)
        CVDV_HT_SYM(DestroyTensorDescriptor) CVDV_HT_SYM(CreateReduction) CVDV_HT_SYM(DestroyOperationDescriptor)
        CVDV_HT_SYM(CreatePlanPreference) CVDV_HT_SYM(DestroyPlanPreference) CVDV_HT_SYM(EstimateWorkspaceSize)
        CVDV_HT_SYM(CreatePlan) CVDV_HT_SYM(DestroyPlan) CVDV_HT_SYM(Reduce) CVDV_HT_SYM(GetErrorString)
#undef CVDV_HT_SYM
        if (!ok) {
            err = "libhiptensor lacks an entry point the comparator needs";
            return;
        }
        api = a;
    });
    if (!api.Reduce) throw CvdvError{CVDV_ERR_UNSUPPORTED, err};
    return api;
}

// One register of the canonical state: extent, input stride and the stride of
// its index in the requested (column-major) output, -1 when it is reduced.
struct RegMode {
    int64_t len, inStride, outReq;
};

// A tensor mode: consecutive registers that are all kept or all reduced.
struct TensorMode {
    int64_t len, inStride;
    bool kept;
    int32_t id;
};

void ht_check(hiptensorStatus_t s, const char* what) {
    if (s != HIPTENSOR_STATUS_SUCCESS)
        throw CvdvError{s == HIPTENSOR_STATUS_NOT_SUPPORTED ? CVDV_ERR_UNSUPPORTED : CVDV_ERR_HIP,
                        std::string("hipTensor: ") + what + " failed (" + ht().GetErrorString(s) + ")"};
}

// Owns the descriptors of one reduction; released on every path.
struct Reduction {
    hiptensorHandle_t h = nullptr;
    hiptensorTensorDescriptor_t in = nullptr, out = nullptr;
    hiptensorOperationDescriptor_t op = nullptr;
    hiptensorPlanPreference_t pref = nullptr;
    hiptensorPlan_t plan = nullptr;
    ~Reduction() {
        if (plan) ht().DestroyPlan(plan);
        if (pref) ht().DestroyPlanPreference(pref);
        if (op) ht().DestroyOperationDescriptor(op);
        if (in) ht().DestroyTensorDescriptor(in);
        if (out) ht().DestroyTensorDescriptor(out);
        if (h) ht().Destroy(h);
    }
};

// Registers of the canonical state for a selection, least significant first.
std::vector<RegMode> reg_modes(const CVDVContext* ctx, const int* regs, int nsel, uint64_t& outSize) {
    std::vector<int64_t> ostride(ctx->numReg, -1);
    outSize = 1;
    for (int i = 0; i < nsel; ++i) {
        ostride[regs[i]] = (int64_t)outSize;
        outSize <<= ctx->nq[regs[i]];
    }
    std::vector<RegMode> out;
    for (int r = ctx->numReg - 1; r >= 0; --r)
        if (ctx->nq[r] > 0) out.push_back(RegMode{1ll << ctx->nq[r], 1ll << ctx->flw[r], ostride[r]});
    return out;
}

// The reduction's modes: the output keeps the kept registers in the state's own
// order (compact strides; the host permutes afterwards), so consecutive
// registers merge whenever both are kept or both are reduced and their input
// strides continue.  The rank is then at most the number of alternations
// between kept and reduced registers, plus one.
std::vector<TensorMode> tensor_modes(const std::vector<RegMode>& regs) {
    std::vector<TensorMode> modes;
    bool anyReduced = false;
    for (const RegMode& r : regs) anyReduced = anyReduced || r.outReq < 0;
    if (!anyReduced) {
        // nothing to sum (every register kept): the library refuses a rank-1
        // tensor with no reduced mode, so keep one mode per register, and split a
        // lone register in two
        for (const RegMode& r : regs) modes.push_back(TensorMode{r.len, r.inStride, true, (int32_t)modes.size()});
        if (modes.size() == 1 && modes[0].len >= 4) {
            int64_t lo = 1;
            while (lo * lo < modes[0].len) lo *= 2;
            const int64_t hi = modes[0].len / lo;
            modes = {TensorMode{lo, modes[0].inStride, true, 0}, TensorMode{hi, modes[0].inStride * lo, true, 1}};
        }
        return modes;
    }
    for (const RegMode& r : regs) {
        const bool kept = r.outReq >= 0;
        if (!modes.empty()) {
            TensorMode& m = modes.back();
            if (m.kept == kept && r.inStride == m.inStride * m.len) {
                m.len *= r.len;
                continue;
            }
        }
        modes.push_back(TensorMode{r.len, r.inStride, kept, (int32_t)modes.size()});
    }
    return modes;
}

}  // namespace

extern "C" {

void cvdvMeasureMultipleCT(CVDVContext* ctx, const int* regIdxs, int numRegs, double* probsOut) {
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
    try {
        ctx_materialize(ctx);  // canonical layout, pending phases applied
        uint64_t outSize = 0;
        const std::vector<RegMode> regs = reg_modes(ctx, regIdxs, numRegs, outSize);
        const std::vector<TensorMode> modes = tensor_modes(regs);
        std::vector<int64_t> inLen, inStr, outLen, outStr;
        std::vector<int32_t> inId, outId;
        int64_t compact = 1;
        for (const TensorMode& m : modes) {
            inLen.push_back(m.len);
            inStr.push_back(m.inStride);
            inId.push_back(m.id);
            if (m.kept) {
                outLen.push_back(m.len);
                outStr.push_back(compact);
                outId.push_back(m.id);
                compact *= m.len;
            }
        }
        const uint64_t n = ctx->totalSize;
        double* sq = (double*)ctx_scratch(ctx, (n + outSize) * sizeof(double));
        double* res = sq + n;
        Reduction R;
        ht_check(ht().Create(&R.h), "hiptensorCreate");
        ht_check(ht().CreateTensorDescriptor(R.h, &R.in, (uint32_t)inLen.size(), inLen.data(), inStr.data(),
                                                 HIPTENSOR_R_64F, 256),
                 "input descriptor");
        ht_check(ht().CreateTensorDescriptor(R.h, &R.out, (uint32_t)outLen.size(), outLen.data(),
                                                 outStr.data(), HIPTENSOR_R_64F, 256),
                 "output descriptor");
        ht_check(ht().CreateReduction(R.h, &R.op, R.in, inId.data(), HIPTENSOR_OP_IDENTITY, R.out, outId.data(),
                                          HIPTENSOR_OP_IDENTITY, R.out, outId.data(), HIPTENSOR_OP_ADD,
                                          HIPTENSOR_COMPUTE_DESC_64F),
                 "reduction descriptor");
        ht_check(ht().CreatePlanPreference(R.h, &R.pref, HIPTENSOR_ALGO_DEFAULT, HIPTENSOR_JIT_MODE_NONE),
                 "plan preference");
        uint64_t ws = 0;
        ht_check(ht().EstimateWorkspaceSize(R.h, R.op, R.pref, HIPTENSOR_WORKSPACE_DEFAULT, &ws), "workspace size");
        ht_check(ht().CreatePlan(R.h, &R.plan, R.op, R.pref, ws), "plan");
        void* wsp = nullptr;
        if (ws > 0) HIPCHK(hipMalloc(&wsp, ws));
        HIPCHK(hipMemsetAsync(res, 0, outSize * sizeof(double), ctx->stream));  // C of D = a A + b C
        const int pid = prof_begin(ctx, KK_REDUCE);
        if (ctx->prec == CVDV_C128)
            hipLaunchKernelGGL(k_abs2<double2>, dim3(2048), dim3(256), 0, ctx->stream, sq, (const double2*)ctx->state(), n);
        else
            hipLaunchKernelGGL(k_abs2<float2>, dim3(2048), dim3(256), 0, ctx->stream, sq, (const float2*)ctx->state(), n);
        hipError_t le = hipGetLastError();
        const double one = 1.0, zero = 0.0;
        hiptensorStatus_t rs = HIPTENSOR_STATUS_SUCCESS;
        if (le == hipSuccess) rs = ht().Reduce(R.h, R.plan, &one, sq, &zero, res, res, wsp, ws, ctx->stream);
        prof_end(ctx, pid);
        hipError_t se = hipStreamSynchronize(ctx->stream);
        if (wsp) (void)hipFree(wsp);
        HIPCHK(le);
        ht_check(rs, "hiptensorReduce");
        HIPCHK(se);
        std::vector<double> natural(outSize);
        HIPCHK(hipMemcpy(natural.data(), res, outSize * sizeof(double), hipMemcpyDeviceToHost));
        // natural index (kept registers, least significant first) -> requested
        // column-major index
        for (uint64_t e = 0; e < outSize; ++e) {
            uint64_t rem = e, o = 0;
            for (const RegMode& r : regs) {
                if (r.outReq < 0) continue;
                o += (rem % (uint64_t)r.len) * (uint64_t)r.outReq;
                rem /= (uint64_t)r.len;
            }
            probsOut[o] = natural[e];
        }
    } catch (const CvdvError& e) {
        cvdv_set_error(e.code, e.msg);
    } catch (...) {
        cvdv_set_error(CVDV_ERR_HIP, "unexpected internal error");
    }
}

}  // extern "C"
