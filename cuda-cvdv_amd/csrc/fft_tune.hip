// fft_tune.hip - autotuner of the register FFT's tile choices.
//
// The reference leaves plan choices to cuFFT (src/cuda/context.cu:48-65).  Here
// the planner (fft.hip) has a few free choices per register field: how many
// columns a strided tile takes (its contiguous bytes per row), how far a tile
// grows over row blocks, and whether consecutive transforms of the field share
// half passes (transform fusion).  cvdvFftAutotune times every distinct plan
// these choices give, for every register field of a context, and keeps the
// fastest per field in the context.
//
// The timing runs on a private context of the same shape and precision (its
// own zeroed state, its own stream): the caller's state, pending phases and
// half-done transforms are not touched.  The measure is the time per transform
// of a chain of alternating forward and inverse transforms with nothing
// between them, completed at the end, best of a few chains.  Results are kept
// per (precision, state size, field) for the process, so contexts of one
// shape time their candidates once.

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "cvdv_internal.h"
#include "../../include/cvdv_hip.h"

namespace {

struct Tuned {
    FftCfg cfg;
    double ms = 0.0;  // per transform
};

std::mutex g_tuneMu;
std::map<std::tuple<int, uint64_t, int, int>, Tuned> g_tuned;

constexpr int CHAIN = 8;  // transforms per timed chain
constexpr int REPS = 3;   // timed chains per candidate (the best counts)

// Milliseconds per transform of a chain over field [sh, sh+nb) of c.
double time_chain(CVDVContext* c, int sh, int nb, hipEvent_t e0, hipEvent_t e1) {
    const double scale = 1.0 / std::sqrt((double)(1ull << nb));
    auto chain = [&] {
        for (int i = 0; i < CHAIN; ++i) fft_register(c, sh, nb, (i & 1) != 0, DiagProg{}, scale);
        fft_settle(c);
    };
    chain();  // warm: code objects, twiddle tables, the second buffer
    double best = 1e30;
    for (int r = 0; r < REPS; ++r) {
        HIPCHK(hipEventRecord(e0, c->stream));
        chain();
        HIPCHK(hipEventRecord(e1, c->stream));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        best = std::min(best, (double)ms / CHAIN);
    }
    return best;
}

// The candidates of one field: row widths and tile caps around the defaults,
// fusion on and off; the defaults first, so a tie keeps them.
std::vector<FftCfg> candidates(int prec) {
    const int wmax = prec == CVDV_C128 ? 16 : 32;
    std::vector<FftCfg> out;
    out.push_back(FftCfg{});
    for (int fuse : {1, 0})
        for (int w : {wmax, wmax / 2, wmax / 4})
            for (int cap : {2048, 1024, 4096}) {
                FftCfg c;
                c.rowElems = w;
                c.tileCap = cap;
                c.fuse = fuse;
                out.push_back(c);
            }
    return out;
}

// Tune field [sh, sh+nb) on the private context; returns the candidates timed.
int tune_field(CVDVContext* c, int sh, int nb, Tuned& best, hipEvent_t e0, hipEvent_t e1) {
    const uint64_t key = fft_cfg_key(sh, nb);
    std::vector<uint64_t> seen;
    int timed = 0;
    best = Tuned{FftCfg{}, 1e30};
    for (const FftCfg& cand : candidates(c->prec)) {
        const uint64_t sig = fft_plan_signature(c, sh, nb, cand);
        if (std::find(seen.begin(), seen.end(), sig) != seen.end()) continue;  // the same passes
        seen.push_back(sig);
        c->fftCfg[key] = cand;
        double ms;
        try {
            ms = time_chain(c, sh, nb, e0, e1);
        } catch (const CvdvError&) {
            // a plan without a kernel instance: not a candidate
            HIPCHK(hipStreamSynchronize(c->stream));
            c->defer = FtDefers{};
            continue;
        }
        ++timed;
        // another plan replaces the defaults only when it is at least 2% faster
        if (timed == 1 || ms < 0.98 * best.ms) best = Tuned{cand, ms};
    }
    c->fftCfg.erase(key);
    return timed;
}

}  // namespace

extern "C" {

int cvdvFftAutotune(CVDVContext* ctx) {
    if (!ctx || refuse_recording(ctx)) return -1;
    if (!ctx->initialized) {
        cvdv_set_error(CVDV_ERR_STATE, "State not initialized");
        return -1;
    }
    if (ctx->capturing) {
        cvdv_set_error(CVDV_ERR_STATE, "cvdvFftAutotune inside a graph capture");
        return -1;
    }
    int timed = 0;
    CVDVContext* c = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    try {
        // the register fields as laid out now (swaps exchange fields of equal size)
        std::vector<std::pair<int, int>> fields;
        for (int r = 0; r < ctx->numReg; ++r)
            if (ctx->nq[r] > 0) fields.emplace_back(ctx->pshift[r], ctx->nq[r]);
        std::sort(fields.begin(), fields.end());
        fields.erase(std::unique(fields.begin(), fields.end()), fields.end());
        // a half left pending completes under the plan that started it
        fft_settle(ctx);
        for (const auto& f : fields) {
            const auto gkey = std::make_tuple(ctx->prec, (uint64_t)ctx->totalSize, f.first, f.second);
            Tuned t;
            bool have = false;
            {
                std::lock_guard<std::mutex> lk(g_tuneMu);
                const auto it = g_tuned.find(gkey);
                if (it != g_tuned.end()) {
                    t = it->second;
                    have = true;
                }
            }
            if (!have) {
                if (!c) {
                    c = ctx_create_like(ctx);
                    HIPCHK(hipEventCreate(&e0));
                    HIPCHK(hipEventCreate(&e1));
                }
                timed += tune_field(c, f.first, f.second, t, e0, e1);
                std::lock_guard<std::mutex> lk(g_tuneMu);
                g_tuned[gkey] = t;
            }
            ctx->fftCfg[fft_cfg_key(f.first, f.second)] = t.cfg;
        }
    } catch (const CvdvError& e) {
        cvdv_set_error(e.code, e.msg);
        timed = -1;
    } catch (...) {
        cvdv_set_error(CVDV_ERR_HIP, "unexpected internal error");
        timed = -1;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (c) cvdvDestroy(c);
    return timed;
}

int cvdvFftGetTuning(CVDVContext* ctx, int regIdx, int* rowElems, int* tileCap, int* fuse, double* msPerTransform) {
    if (!ctx || regIdx < 0 || regIdx >= ctx->numReg) {
        cvdv_set_error(CVDV_ERR_ARG, "Invalid register index: " + std::to_string(regIdx));
        return -1;
    }
    const auto it = ctx->fftCfg.find(fft_cfg_key(ctx->pshift[regIdx], ctx->nq[regIdx]));
    if (it == ctx->fftCfg.end()) return 0;
    if (rowElems) *rowElems = it->second.rowElems;
    if (tileCap) *tileCap = it->second.tileCap;
    if (fuse) *fuse = it->second.fuse;
    if (msPerTransform) {
        std::lock_guard<std::mutex> lk(g_tuneMu);
        const auto t = g_tuned.find(std::make_tuple(ctx->prec, (uint64_t)ctx->totalSize, ctx->pshift[regIdx],
                                                    ctx->nq[regIdx]));
        *msPerTransform = t == g_tuned.end() ? 0.0 : t->second.ms;
    }
    return 1;
}

}  // extern "C"
