// ppf_batch.hip -- batched solves of many independent problems.
//
// The reference solves one problem per plan (a module singleton,
// 01_Fortran/src/fft_3d_poisson.f90:43-55) and one per call.  A batch binds
// N problems of any extents and BCs to their device buffers, tunes the z-pass
// shape once per distinct problem shape, and after a first direct solve
// (which also learns how many refinement sweeps each FP32_IR problem needs)
// captures every problem as an independent branch of one hipGraph.  Branches
// run concurrently on the device, so small problems that cannot fill 256 CUs
// on their own overlap instead of queueing, and the whole batch costs one
// graph launch on the host.  An FP32_IR branch runs the learned sweeps and
// then tests its residual on the device: a new right-hand side that needs one
// more sweep gets it within the same replay, and one that needs more is
// reported (PPF_ERR_NOT_CONVERGED from a solve with stats).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <map>
#include <new>
#include <tuple>
#include <vector>

#include "../../include/pascal_poisson_fft.h"
#include "ppf_plan.hpp"

namespace {
constexpr int kBranchStreams = 8;  // capture streams (default priority)
// The fused launch (ppf_fused.hpp: one persistent kernel over a device work
// queue of (stage, problem, tile) items) is switched off: its GPU runs did not
// complete (a hang whose cause has not been shown), so every batch runs one
// plan per problem.  The code is kept, unreachable, for that investigation.
constexpr bool kFusedBatch = false;
// guarded refinement sweeps an FP32_IR branch carries beyond the learned
// count: each is ~9 graph nodes per problem even when its gate is closed
constexpr int kGuardedSweeps = 1;
}

struct ppf_batch {
    int n = 0;
    int flags = 0;
    std::vector<ppf_batch_io> io;
    std::vector<ppf_plan *> plan;
    std::vector<int> fp32;
    std::vector<int> sweeps;
    std::vector<hipStream_t> streams;
    std::vector<hipEvent_t> events;  // [0] fork, [1..] joins
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool calibrated = false;
    ppf_fused *fused = nullptr;  // the fused launch (FP64 batches), else plans
};

static void batch_free(ppf_batch *b) {
    if (!b) return;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (hipEvent_t e : b->events) (void)hipEventDestroy(e);
    for (hipStream_t s : b->streams) (void)hipStreamDestroy(s);
    if (b->cap) (void)hipStreamDestroy(b->cap);
    for (ppf_plan *p : b->plan)
        if (p) (void)ppf_plan_destroy(p);
    ppf::fused_destroy(b->fused);
    delete b;
}

static bool io_ok(const ppf_batch_io &q) {
    return q.prhs && q.p && q.dmx1 && q.dmx2 && q.dmx3 && q.dx3;
}

// Capture every problem as its own branch (round-robin over the capture
// streams) into one graph.
static int batch_capture(ppf_batch *b) {
    if (hipStreamBeginCapture(b->cap, hipStreamCaptureModeRelaxed) != hipSuccess) return PPF_ERR_HIP;
    int rc = PPF_OK;
    const int ns = (int)b->streams.size();
    if (hipEventRecord(b->events[0], b->cap) != hipSuccess) rc = PPF_ERR_HIP;
    for (int s = 0; s < ns && !rc; ++s)
        if (hipStreamWaitEvent(b->streams[s], b->events[0], 0) != hipSuccess) rc = PPF_ERR_HIP;
    if (b->fused) {
        rc = ppf::fused_launch(b->fused, b->streams[0]);
    } else {
        for (int i = 0; i < b->n && !rc; ++i) {
            const ppf_batch_io &q = b->io[i];
            rc = ppf_solve_async(b->plan[i], q.prhs, q.p, q.dmx1, q.dmx2, q.dmx3, q.dx3,
                                 b->fp32[i] ? b->sweeps[i] : 0, kGuardedSweeps, b->streams[i % ns]);
        }
    }
    for (int s = 0; s < ns && !rc; ++s) {
        if (hipEventRecord(b->events[1 + s], b->streams[s]) != hipSuccess ||
            hipStreamWaitEvent(b->cap, b->events[1 + s], 0) != hipSuccess)
            rc = PPF_ERR_HIP;
    }
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(b->cap, &g);
    if (!rc && ec != hipSuccess) rc = PPF_ERR_HIP;
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    b->graph = g;
    if (hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess) {
        b->exec = nullptr;
        return PPF_ERR_HIP;
    }
    return PPF_OK;
}

extern "C" {

int ppf_batch_create(int nprob, const ppf_desc *descs, const ppf_batch_io *io, int flags, ppf_batch **out) {
    if (!out) return PPF_ERR_INVALID_ARG;
    *out = nullptr;
    if (nprob <= 0 || !descs || !io) return PPF_ERR_INVALID_ARG;
    for (int i = 0; i < nprob; ++i)
        if (!io_ok(io[i])) return PPF_ERR_INVALID_ARG;
    ppf_batch *b = new (std::nothrow) ppf_batch();
    if (!b) return PPF_ERR_ALLOC;
    b->n = nprob;
    b->flags = flags;
    b->io.assign(io, io + nprob);
    b->plan.assign(nprob, nullptr);
    b->fp32.assign(nprob, 0);
    b->sweeps.assign(nprob, 0);
    int rc = PPF_OK;
    bool fusable = kFusedBatch;
    for (int i = 0; i < nprob && fusable; ++i) fusable = ppf::fused_supported(descs[i]);
    if (fusable) {
        rc = ppf::fused_create(nprob, descs, io, &b->fused);
        if (rc) {
            batch_free(b);
            return rc;
        }
    }
    for (int i = 0; i < nprob && !rc && !b->fused; ++i) {
        rc = ppf_plan_create(&descs[i], &b->plan[i]);
        b->fp32[i] = descs[i].precision == PPF_PREC_FP32_IR;
    }
    // one autotuner run per distinct (extents, BCs, precision)
    if (!rc && !b->fused && !(flags & PPF_BATCH_NO_TUNE)) {
        // every problem of a shape gets the x, y and z variants tuned on its first
        std::map<std::tuple<int, int, int, int, int, int>, std::array<int, 3>> tuned;
        for (int i = 0; i < nprob && !rc; ++i) {
            const ppf_desc &d = descs[i];
            const auto key = std::make_tuple(d.n1m, d.n2m, d.n3m, d.pbc1 != 0, d.pbc2 != 0, d.precision);
            auto it = tuned.find(key);
            if (it == tuned.end()) {
                int v = 0;
                std::array<int, 3> vs{};
                const ppf_batch_io &q = io[i];
                rc = ppf_plan_tune(b->plan[i], q.prhs, q.p, q.dmx1, q.dmx2, q.dmx3, q.dx3, nullptr, &v);
                if (!rc) rc = ppf_plan_variants(b->plan[i], vs.data());
                tuned[key] = vs;
            } else {
                rc = ppf_plan_set_variants(b->plan[i], it->second[0], it->second[1], it->second[2]);
            }
        }
    }
    if (!rc && !(flags & PPF_BATCH_NO_GRAPH)) {
        const int ns = b->fused ? 1 : std::min(nprob, kBranchStreams);
        if (hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking) != hipSuccess) rc = PPF_ERR_HIP;
        for (int s = 0; s < ns && !rc; ++s) {
            hipStream_t st;
            if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) rc = PPF_ERR_HIP;
            else b->streams.push_back(st);
        }
        for (int e = 0; e < ns + 1 && !rc; ++e) {
            hipEvent_t ev;
            if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) rc = PPF_ERR_HIP;
            else b->events.push_back(ev);
        }
    }
    if (rc) {
        batch_free(b);
        return rc;
    }
    *out = b;
    return PPF_OK;
}

int ppf_batch_solve(ppf_batch *b, void *stream, ppf_stats *stats) {
    if (!b) return PPF_ERR_INVALID_ARG;
    hipStream_t s = static_cast<hipStream_t>(stream);
    int rc = PPF_OK;
    bool replayed = false;
    if (!b->calibrated && b->fused) {
        if ((rc = ppf::fused_launch(b->fused, s))) return rc;
        b->calibrated = true;
        if (!(b->flags & PPF_BATCH_NO_GRAPH)) {
            if (hipStreamSynchronize(s) != hipSuccess) return PPF_ERR_HIP;
            if ((rc = batch_capture(b))) return rc;
        }
    } else if (b->fused && !b->exec) {
        if ((rc = ppf::fused_launch(b->fused, s))) return rc;
    } else if (!b->calibrated) {
        // direct solves; FP32_IR problems run their convergence test and the
        // sweeps they needed are kept for the captured graph
        for (int i = 0; i < b->n && !rc; ++i) {
            const ppf_batch_io &q = b->io[i];
            ppf_stats st{0, -1.0};
            rc = ppf_solve(b->plan[i], q.prhs, q.p, q.dmx1, q.dmx2, q.dmx3, q.dx3, stream, &st);
            b->sweeps[i] = st.sweeps;
        }
        if (rc) return rc;
        b->calibrated = true;
        if (!(b->flags & PPF_BATCH_NO_GRAPH)) {
            if (hipStreamSynchronize(s) != hipSuccess) return PPF_ERR_HIP;
            if ((rc = batch_capture(b))) return rc;
        }
    } else if (b->exec) {
        if (hipGraphLaunch(b->exec, s) != hipSuccess) return PPF_ERR_HIP;
        replayed = true;
    } else {
        for (int i = 0; i < b->n && !rc; ++i) {
            const ppf_batch_io &q = b->io[i];
            rc = ppf_solve_async(b->plan[i], q.prhs, q.p, q.dmx1, q.dmx2, q.dmx3, q.dx3, b->fp32[i] ? b->sweeps[i] : 0,
                                 kGuardedSweeps, stream);
        }
        if (rc) return rc;
        replayed = true;
    }
    if (stats && b->fused) {
        // a dependency wait that timed out leaves the flag set: report it
        if ((rc = ppf::fused_status(b->fused, s))) return rc;
        for (int i = 0; i < b->n && !rc; ++i) {
            stats[i].sweeps = 0;
            rc = ppf::fused_residual(b->fused, i, s, &stats[i].rel_residual);
        }
        return rc;
    }
    if (stats) {
        // FP32_IR problems of a replay report the sweeps and the residual their
        // device-side test measured (PPF_ERR_NOT_CONVERGED if that stayed above
        // tol); FP64 problems (and the first, eager solve) a fresh FP64 residual
        int miss = 0;
        for (int i = 0; i < b->n && !rc; ++i) {
            const ppf_batch_io &q = b->io[i];
            if (replayed && b->fp32[i]) {
                rc = ppf_solve_async_result(b->plan[i], stream, &stats[i]);
                if (rc == PPF_ERR_NOT_CONVERGED) { miss = 1; rc = PPF_OK; }
            } else {
                stats[i].sweeps = b->sweeps[i];
                rc = ppf_residual(b->plan[i], q.prhs, q.p, q.dmx3, q.dx3, stream, &stats[i].rel_residual);
            }
        }
        if (!rc && miss) rc = PPF_ERR_NOT_CONVERGED;
    }
    return rc;
}

int ppf_batch_destroy(ppf_batch *b) {
    if (!b) return PPF_ERR_INVALID_ARG;
    batch_free(b);
    return PPF_OK;
}

}  // extern "C"
