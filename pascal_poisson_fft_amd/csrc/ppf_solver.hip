// ppf_solver.hip -- plan, tables, launch sequence and C ABI of the solve path.
//
// Replaces module fft_poisson of 01_Fortran/src/fft_3d_poisson.f90 (plan
// create :70-108, initial :110-121, solve :353-626, clean :123-130) for one
// GPU.  The reference's ~57 field passes (transposes, DCT pre/post kernels,
// cuFFT calls, coefficient arrays, dcopy/dscal, mean and BC sweeps) become the
// five passes of ppf_kernels.hpp.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <array>
#include <cstring>
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <new>
#include <type_traits>
#include <vector>

#include "../../include/pascal_poisson_fft.h"
#include "ppf_aux.hpp"
#include "ppf_plan.hpp"

using namespace ppf;


namespace {
// Autotuner results per problem shape, for the life of the process: a second
// plan of a tuned shape takes them without timing anything.
struct TuneKey {
    int n1m, n2m, n3m, pbc1, pbc2, pbc3, precision;
    bool operator<(const TuneKey &o) const {
        return std::tie(n1m, n2m, n3m, pbc1, pbc2, pbc3, precision) <
               std::tie(o.n1m, o.n2m, o.n3m, o.pbc1, o.pbc2, o.pbc3, o.precision);
    }
};
std::mutex g_tune_mu;
std::map<TuneKey, std::array<int, 3>> g_tuned;
}  // namespace

template <typename T>
static int dev_alloc(ppf_plan *p, T **ptr, size_t count) {
    void *q = nullptr;
    if (count == 0) count = 1;
    if (hipMalloc(&q, count * sizeof(T)) != hipSuccess) return PPF_ERR_ALLOC;
    *ptr = static_cast<T *>(q);
    p->bytes += count * sizeof(T);
    return PPF_OK;
}

template <typename T>
static int upload_tables(ppf_plan *p, AxisTables<T> &tb, int N, int kind) {
    std::vector<cx<T>> tw, wk, qk;
    host_axis_tables<T>(N, kind, tw, wk, qk);
    const int M = (int)tw.size();
    int rc;
    if ((rc = dev_alloc(p, &tb.tw, M)) || (rc = dev_alloc(p, &tb.wk, M + 1)) || (rc = dev_alloc(p, &tb.qk, M + 1)))
        return rc;
    PPF_HIP(hipMemcpy(tb.tw, tw.data(), sizeof(cx<T>) * M, hipMemcpyHostToDevice));
    PPF_HIP(hipMemcpy(tb.wk, wk.data(), sizeof(cx<T>) * (M + 1), hipMemcpyHostToDevice));
    PPF_HIP(hipMemcpy(tb.qk, qk.data(), sizeof(cx<T>) * (M + 1), hipMemcpyHostToDevice));
    return PPF_OK;
}

// Checks of a problem description shared by plans and the allocation-free solve.
static int desc_check(const ppf_desc &d) {
    if (d.n1m <= 0 || d.n2m <= 0 || d.n3m <= 0 || !(d.L1 > 0) || !(d.L2 > 0) || !(d.L3 > 0)) return PPF_ERR_INVALID_ARG;
    if (d.precision != PPF_PREC_FP64 && d.precision != PPF_PREC_FP32_IR) return PPF_ERR_INVALID_ARG;
    // every n must be even (01_Fortran/examples/global.f90:114-120)
    // z rows live in LDS in the z pass (2 doubles per row): n3m <= 4096
    if (!size_supported(d.n1m) || !size_supported(d.n2m) || d.n3m < 2 || (d.n3m & 1) || d.n3m > 4096)
        return PPF_ERR_SIZE;
    return PPF_OK;
}

static void free_plan(ppf_plan *p) {
    if (!p) return;
    void *ptrs[] = {p->W,        p->W32,      p->cs,       p->zcoef,    p->lamx,     p->lamy,     p->zx,
                    p->zy,       p->mean,     p->partial,  p->red,      p->tx64.tw,  p->tx64.wk,  p->tx64.qk,
                    p->ty64.tw,  p->ty64.wk,  p->ty64.qk,  p->tx32.tw,  p->tx32.wk,  p->tx32.qk,  p->ty32.tw,
                    p->ty32.wk,  p->ty32.qk,  p->dgate,    p->dres};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    if (p->hred) (void)hipHostFree(p->hred);
    delete p;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char *ppf_error_string(int code) {
    switch (code) {
        case PPF_OK: return "success";
        case PPF_ERR_INVALID_ARG: return "invalid argument";
        case PPF_ERR_SIZE: return "unsupported extent (x/y need even 8..1024 of the form 2^a times 1, 3, 5, 7 or 15; z even, 2..4096)";
        case PPF_ERR_HIP: return "HIP runtime error";
        case PPF_ERR_ALLOC: return "device allocation failed";
        case PPF_ERR_UNSUPPORTED: return "unsupported configuration";
        case PPF_ERR_NOT_CONVERGED: return "refinement did not reach the tolerance";
        case PPF_ERR_SCRATCH:
            return "problem too large for the static device scratch of ppf_solve_noalloc (64 MiB: workspace "
                   "(n1m or n1m+2)(n2m+2)(n3m+2) doubles, rows padded to 128 bytes, plus tables)";
    }
    return "unknown error";
}

int ppf_plan_create(const ppf_desc *desc, ppf_plan **out) {
    if (!desc || !out) return PPF_ERR_INVALID_ARG;
    *out = nullptr;
    const ppf_desc &d = *desc;
    int rc = desc_check(d);
    if (rc) return rc;
    ppf_plan *p = new (std::nothrow) ppf_plan();
    if (!p) return PPF_ERR_ALLOC;
    plan_geometry(d, p);
    const bool fp32 = d.precision == PPF_PREC_FP32_IR;
    const size_t wsz = (size_t)p->Lxs * p->Ly * p->Lz;
    // Thomas scratch only where the one-thread-per-line kernel runs by
    // default (n3m > 2048); ppf_plan_set_zvariant allocates it when that
    // kernel is selected explicitly
    const bool need_cs = d.n3m > 2048;
    const size_t csz = need_cs ? (size_t)p->Lx * p->znjr * d.n3m : 1;
    p->cs_elems = csz;
    if ((rc = fp32 ? dev_alloc(p, &p->W32, wsz) : dev_alloc(p, &p->W, wsz)) || (rc = dev_alloc(p, &p->cs, csz)) || (rc = dev_alloc(p, &p->lamx, p->Lx)) ||
        (rc = dev_alloc(p, &p->lamy, p->Ly)) || (rc = dev_alloc(p, &p->zx, p->Lx)) || (rc = dev_alloc(p, &p->zy, p->Ly)) ||
        (rc = dev_alloc(p, &p->mean, 1)) || (rc = dev_alloc(p, &p->red, 3)) ||
        (rc = dev_alloc(p, &p->zcoef, 2 * (size_t)d.n3m + 4))) {
        free_plan(p);
        return rc;
    }
    // partial sums: one pair per workgroup of the residual kernels (persistent
    // grids are capped at this count)
    p->npartial = 4096;
    if ((rc = dev_alloc(p, &p->partial, 3 * (size_t)p->npartial)) || (rc = dev_alloc(p, &p->dgate, 2)) ||
        (rc = dev_alloc(p, &p->dres, 1))) {
        free_plan(p);
        return rc;
    }
    if (hipHostMalloc((void **)&p->hred, 3 * sizeof(double), hipHostMallocDefault) != hipSuccess) {
        p->hred = nullptr;
        free_plan(p);
        return PPF_ERR_ALLOC;
    }

    // eigenvalue and zero-mode tables per scalar column / row
    const EigGeom eg = eig_geom(p);
    std::vector<double> lx(p->Lx), ly(p->Ly, 0.0);
    std::vector<unsigned char> zx(p->Lx, 0), zy(p->Ly, 0);
    for (int kx = 0; kx < p->Lx; ++kx) eig_col(eg, kx, lx[kx], zx[kx]);
    for (int jp = 0; jp < p->Ly; ++jp) eig_row(eg, jp, ly[jp], zy[jp]);
    if (hipMemcpy(p->lamx, lx.data(), sizeof(double) * p->Lx, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->lamy, ly.data(), sizeof(double) * p->Ly, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->zx, zx.data(), p->Lx, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->zy, zy.data(), p->Ly, hipMemcpyHostToDevice) != hipSuccess) {
        free_plan(p);
        return PPF_ERR_HIP;
    }
    if (fp32) rc = upload_tables(p, p->tx32, d.n1m, p->xkind) ? PPF_ERR_ALLOC : upload_tables(p, p->ty32, d.n2m, p->ykind);
    else rc = upload_tables(p, p->tx64, d.n1m, p->xkind) ? PPF_ERR_ALLOC : upload_tables(p, p->ty64, d.n2m, p->ykind);
    if (rc) {
        free_plan(p);
        return rc;
    }
    *out = p;
    return PPF_OK;
}

int ppf_plan_destroy(ppf_plan *plan) {
    if (!plan) return PPF_ERR_INVALID_ARG;
    free_plan(plan);
    return PPF_OK;
}

size_t ppf_plan_workspace_bytes(const ppf_plan *plan) { return plan ? plan->bytes : 0; }

}  // extern "C"

// The five passes of one solve of the plan's precision: in -> P.  ev (6 events
// or nullptr) brackets the passes for ppf_solve_timed.
template <typename T>
static int run_passes(ppf_plan *p, T *W, const AxisTables<T> &tx, const AxisTables<T> &ty, const double *in,
                      double *P, const double *dmx1, const double *dmx2, const double *dmx3, const double *dx3,
                      hipStream_t s, hipEvent_t *ev) {
    int rc;
    if (ev) PPF_HIP(hipEventRecord(ev[0], s));
    // pass 1 also fills the z row table of this solve
    if ((rc = dispatch_xfwd<T>(p, in, W, tx, false, s, -1, z_rows(p, dx3, dmx3)))) return rc;
    if (ev) PPF_HIP(hipEventRecord(ev[1], s));
    if ((rc = dispatch_y<T>(p, W, ty, false, dmx2, s))) return rc;
    if (ev) PPF_HIP(hipEventRecord(ev[2], s));
    if ((rc = launch_z<T>(p, W, dx3, dmx3, s))) return rc;
    if (ev) PPF_HIP(hipEventRecord(ev[3], s));
    if ((rc = dispatch_y<T>(p, W, ty, true, dmx2, s))) return rc;
    if (ev) PPF_HIP(hipEventRecord(ev[4], s));
    if ((rc = dispatch_xinv<T>(p, W, P, tx, dmx1, false, s))) return rc;
    if (ev) PPF_HIP(hipEventRecord(ev[5], s));
    PPF_HIP(hipGetLastError());
    return PPF_OK;
}

static XRes make_xres(const ppf_plan *p, const double *P) {
    XRes r{};
    r.P = P;
    r.zam = p->zcoef;
    r.zap = p->zcoef + p->d.n3m;
    r.partial = p->partial;
    const double h1 = p->d.L1 / p->d.n1m, h2 = p->d.L2 / p->d.n2m;
    r.ih1 = 1.0 / (h1 * h1);
    r.ih2 = 1.0 / (h2 * h2);
    r.pbc1 = p->d.pbc1 != 0;
    r.pbc2 = p->d.pbc2 != 0;
    return r;
}

// FP64 residual of P against f into p->red (sum r^2, sum f^2, sum p), from
// the z row table already in p->zcoef; gated by p->gate.  out (FP32 plans):
// r is also stored in W32, where the refinement form of pass 1 reads it.
// reduce = false skips the totals (a store right after a test of the same P,
// whose totals are already in p->red); test folds the device-side
// refinement test into the reduction.
static int launch_residual(ppf_plan *p, const double *f, const double *P, hipStream_t s, bool out = false,
                           bool reduce = true, const GateTest &test = GateTest{}) {
    const int n3m = p->d.n3m;
    XRes xr = make_xres(p, P);
    if (out) {
        xr.rout = p->W32;
        xr.Ly = p->Ly;
        xr.Lxs = p->Lxs;
    }
    const int ty = residual_rows(p->d.n1m, p->d.n2m, n3m, 4L * num_cus());
    const long ntiles = residual_tiles(p->d.n1m, p->d.n2m, n3m, ty);
    const int g = (int)std::min<long>(std::min<long>(p->npartial, 8L * num_cus()), ntiles);
    launch_k_residual(ty, g, s, f, xr, p->d.n1m, p->d.n2m, n3m, p->gate);
    if (reduce) hipLaunchKernelGGL(k_reduce3, dim3(1), dim3(256), 0, s, p->partial, g, p->red, p->gate, test);
    PPF_HIP(hipGetLastError());
    return PPF_OK;
}

// Ghost layers of P re-derived from its interior, in the reference's order
// (walls x, y, z, then wraps x, y, z; fft_3d_poisson.f90:620-622).
static int refresh_ghosts(const ppf_plan *p, double *P, const double *dmx1, const double *dmx2, const double *dmx3,
                          hipStream_t s) {
    const int n[3] = {p->d.n1m, p->d.n2m, p->d.n3m};
    const int per[3] = {p->d.pbc1 != 0, p->d.pbc2 != 0, p->d.pbc3 != 0};
    const double *dmx[3] = {dmx1, dmx2, dmx3};
    for (int wrap = 0; wrap < 2; ++wrap)
        for (int dir = 0; dir < 3; ++dir) {
            if (per[dir] != wrap) continue;
            const size_t nb = (size_t)(dir == 0 ? n[1] : n[0]) + 2, nc = (size_t)(dir == 2 ? n[1] : n[2]) + 2;
            const unsigned grid = (unsigned)((nb * nc + 255) / 256);
            hipLaunchKernelGGL(k_ghost_face, dim3(grid), dim3(256), 0, s, P, n[0], n[1], n[2], dir, wrap, dmx[dir]);
        }
    PPF_HIP(hipGetLastError());
    return PPF_OK;
}

// The correction of one refinement sweep of an FP32_IR plan: pass 1
// transforms the residual launch_residual(out) left in W32, passes 2-4 in
// FP32, pass 5 adds the correction to P.  Gated by p->gate.
static int correct32(ppf_plan *p, double *P, const double *dmx1, const double *dmx2, const double *dmx3,
                     const double *dx3, hipStream_t s) {
    int rc;
    if ((rc = dispatch_xfwd<float>(p, nullptr, p->W32, p->tx32, true, s)) ||
        (rc = dispatch_y<float>(p, p->W32, p->ty32, false, dmx2, s)) || (rc = launch_z<float>(p, p->W32, dx3, dmx3, s)) ||
        (rc = dispatch_y<float>(p, p->W32, p->ty32, true, dmx2, s)) ||
        (rc = dispatch_xinv<float>(p, p->W32, P, p->tx32, dmx1, true, s)))
        return rc;
    PPF_HIP(hipGetLastError());
    return PPF_OK;
}

// FP32 transforms and Thomas storage with FP64 refinement: x_0 = solve32(f);
// then r = f - L x (FP64, fused into pass 1), x += solve32(r) (fused into
// pass 5) until ||r||/||f|| <= tol or max_sweeps corrections.  The z pass
// computes in FP64 on FP32 data, so the pinned (0,0) system and the mean stay
// FP64.  tol < 0: exactly max_sweeps corrections, no convergence test.
// fixed >= 0 overrides the plan: exactly `fixed` corrections, no test.
// The ghost layers are accumulated with each correction, which carries the
// FP32 rounding of the first solve's ghosts along (the residual never sees
// them), so they are re-derived from the refined interior at the end.
static int solve_fp32_ir(ppf_plan *p, const double *f, double *P, const double *dmx1, const double *dmx2,
                         const double *dmx3, const double *dx3, hipStream_t s, ppf_stats *stats, int fixed = -1) {
    const double tol = fixed >= 0 ? -1.0 : (p->d.tol == 0.0 ? 1e-10 : p->d.tol);
    const int maxs = fixed >= 0 ? fixed : (p->d.max_sweeps > 0 ? p->d.max_sweeps : 8);
    int rc;
    if ((rc = run_passes<float>(p, p->W32, p->tx32, p->ty32, f, P, dmx1, dmx2, dmx3, dx3, s, nullptr))) return rc;
    // the residuals below use the z row table pass 1 of this solve filled
    int sweeps = 0;
    double rel = -1.0;
    bool converged = false;
    if (tol < 0) {
        for (; sweeps < maxs; ++sweeps)
            if ((rc = launch_residual(p, f, P, s, true)) || (rc = correct32(p, P, dmx1, dmx2, dmx3, dx3, s))) return rc;
        if (sweeps > 0 && (rc = refresh_ghosts(p, P, dmx1, dmx2, dmx3, s))) return rc;
        if (stats) {
            if ((rc = ppf_residual(p, f, P, dmx3, dx3, s, &rel))) return rc;
        }
        converged = true;
    } else {
        // Each test computes the FP64 residual.  The FP32 copy of r that pass
        // 1 of a correction reads is stored with the test only after the
        // FP32 solve (which never meets a tight tol); after a correction the
        // test goes without it (the store made the residual 0.72 instead of
        // 0.49 ms at 512^3) and a miss recomputes r with the store.
        for (;;) {
            const bool store = sweeps == 0 && sweeps < maxs;
            if ((rc = launch_residual(p, f, P, s, store))) return rc;
            PPF_HIP(hipMemcpyAsync(p->hred, p->red, 3 * sizeof(double), hipMemcpyDeviceToHost, s));
            PPF_HIP(hipStreamSynchronize(s));
            rel = p->hred[1] > 0 ? std::sqrt(p->hred[0] / p->hred[1]) : std::sqrt(p->hred[0]);
            if (rel <= tol) { converged = true; break; }
            if (sweeps == maxs || !(rel == rel)) break;
            if (!store && (rc = launch_residual(p, f, P, s, true, false))) return rc;
            if ((rc = correct32(p, P, dmx1, dmx2, dmx3, dx3, s))) return rc;
            ++sweeps;
        }
        if (sweeps > 0 && (rc = refresh_ghosts(p, P, dmx1, dmx2, dmx3, s))) return rc;
    }
    if (stats) {
        stats->sweeps = sweeps;
        stats->rel_residual = rel;
    }
    return converged ? PPF_OK : PPF_ERR_NOT_CONVERGED;
}

// FP32_IR solve with the convergence test on the device, for callers that
// cannot synchronise (graph capture, graph replays of new right-hand sides).
// min_sweeps corrections always run; then a residual test (k_residual and
// the gate test in its reduction) decides on the device whether each of up to
// extra_sweeps gated corrections runs (extra_sweeps < 0: up to max_sweeps in all).  Once
// the residual is <= tol every later gated kernel returns at once -- but is
// still a launch, which is why a captured graph keeps extra_sweeps small.
// Only kernel launches: no host synchronisation, no memset.
static int solve_fp32_async(ppf_plan *p, const double *f, double *P, const double *dmx1, const double *dmx2,
                            const double *dmx3, const double *dx3, int min_sweeps, int extra_sweeps, hipStream_t s) {
    const double tol = p->d.tol == 0.0 ? 1e-10 : p->d.tol;  // tol < 0 never closes the gate
    const int maxs = p->d.max_sweeps > 0 ? p->d.max_sweeps : 8;
    const int nmin = std::min(min_sweeps, maxs);
    const int nmax = extra_sweeps < 0 ? maxs : std::min(maxs, nmin + extra_sweeps);
    struct GateReset {  // p->gate is back to nullptr on every return path
        ppf_plan *p;
        ~GateReset() { p->gate = nullptr; }
    } reset{p};
    p->gate = nullptr;
    hipLaunchKernelGGL(k_gate_init, dim3(1), dim3(1), 0, s, p->dgate, nmin);
    int rc;
    if ((rc = run_passes<float>(p, p->W32, p->tx32, p->ty32, f, P, dmx1, dmx2, dmx3, dx3, s, nullptr))) return rc;
    // the z row table of this solve (filled by pass 1) serves the residuals
    for (int i = 0; i < nmin; ++i)
        if ((rc = launch_residual(p, f, P, s, true)) || (rc = correct32(p, P, dmx1, dmx2, dmx3, dx3, s))) return rc;
    // each test without the FP32 store of r; a gated residual with the store
    // feeds the gated correction (both return at once once the gate closed)
    for (int i = nmin;; ++i) {
        p->gate = p->dgate;  // open at the first test
        if ((rc = launch_residual(p, f, P, s, false, true, GateTest{p->dgate, tol, i > nmin ? 1 : 0, p->dres})))
            return rc;
        if (i >= nmax) break;
        if ((rc = launch_residual(p, f, P, s, true, false)) || (rc = correct32(p, P, dmx1, dmx2, dmx3, dx3, s)))
            return rc;
    }
    p->gate = nullptr;
    if ((rc = refresh_ghosts(p, P, dmx1, dmx2, dmx3, s))) return rc;
    PPF_HIP(hipGetLastError());
    return PPF_OK;
}

extern "C" {

int ppf_solve(ppf_plan *p, const double *prhs, double *P, const double *dmx1, const double *dmx2, const double *dmx3,
              const double *dx3, void *stream, ppf_stats *stats) {
    if (!p || !prhs || !P || !dmx1 || !dmx2 || !dmx3 || !dx3) return PPF_ERR_INVALID_ARG;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (p->d.precision == PPF_PREC_FP32_IR) return solve_fp32_ir(p, prhs, P, dmx1, dmx2, dmx3, dx3, s, stats);
    int rc;
    if ((rc = run_passes<double>(p, p->W, p->tx64, p->ty64, prhs, P, dmx1, dmx2, dmx3, dx3, s, nullptr)))
        return rc;
    if (stats) { stats->sweeps = 0; stats->rel_residual = -1.0; }
    return PPF_OK;
}

int ppf_solve_sweeps(ppf_plan *p, const double *prhs, double *P, const double *dmx1, const double *dmx2,
                     const double *dmx3, const double *dx3, int sweeps, void *stream) {
    if (!p || !prhs || !P || !dmx1 || !dmx2 || !dmx3 || !dx3 || sweeps < 0) return PPF_ERR_INVALID_ARG;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (p->d.precision == PPF_PREC_FP32_IR) return solve_fp32_ir(p, prhs, P, dmx1, dmx2, dmx3, dx3, s, nullptr, sweeps);
    if (sweeps != 0) return PPF_ERR_INVALID_ARG;
    return run_passes<double>(p, p->W, p->tx64, p->ty64, prhs, P, dmx1, dmx2, dmx3, dx3, s, nullptr);
}

int ppf_solve_async(ppf_plan *p, const double *prhs, double *P, const double *dmx1, const double *dmx2,
                    const double *dmx3, const double *dx3, int min_sweeps, int extra_sweeps, void *stream) {
    if (!p || !prhs || !P || !dmx1 || !dmx2 || !dmx3 || !dx3 || min_sweeps < 0) return PPF_ERR_INVALID_ARG;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (p->d.precision == PPF_PREC_FP32_IR)
        return solve_fp32_async(p, prhs, P, dmx1, dmx2, dmx3, dx3, min_sweeps, extra_sweeps, s);
    return run_passes<double>(p, p->W, p->tx64, p->ty64, prhs, P, dmx1, dmx2, dmx3, dx3, s, nullptr);
}

int ppf_solve_async_result(ppf_plan *p, void *stream, ppf_stats *stats) {
    if (!p || !stats) return PPF_ERR_INVALID_ARG;
    PPF_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    if (p->d.precision != PPF_PREC_FP32_IR) {
        stats->sweeps = 0;
        stats->rel_residual = -1.0;
        return PPF_OK;
    }
    int g[2] = {0, 0};
    double rel = -1.0;
    PPF_HIP(hipMemcpy(g, p->dgate, sizeof(g), hipMemcpyDeviceToHost));
    PPF_HIP(hipMemcpy(&rel, p->dres, sizeof(double), hipMemcpyDeviceToHost));
    stats->sweeps = g[1];
    stats->rel_residual = rel;
    const double tol = p->d.tol == 0.0 ? 1e-10 : p->d.tol;
    return (tol < 0 || rel <= tol) ? PPF_OK : PPF_ERR_NOT_CONVERGED;
}

int ppf_plan_tune(ppf_plan *p, const double *prhs, double *P, const double *dmx1, const double *dmx2,
                  const double *dmx3, const double *dx3, void *stream, int *zvariant) {
    if (!p || !prhs || !P || !dmx1 || !dmx2 || !dmx3 || !dx3) return PPF_ERR_INVALID_ARG;
    const TuneKey key{p->d.n1m, p->d.n2m, p->d.n3m, p->d.pbc1 != 0, p->d.pbc2 != 0, p->d.pbc3 != 0, p->d.precision};
    {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        auto it = g_tuned.find(key);
        if (it != g_tuned.end()) {
            (void)x_variant_decode(it->second[0], p->xvariant, p->xvinv);
            p->yvariant = it->second[1];
            const int rc = ppf_plan_set_zvariant(p, it->second[2]);
            if (zvariant) *zvariant = p->zvariant;
            return rc;
        }
    }
    const int keep[4] = {p->xvariant, p->yvariant, p->zvariant, p->xvinv};
    // per-stage minimum over 4 timed solves after 1 untimed (variants differ
    // by a few percent, about the run-to-run spread of one solve);
    // UNSUPPORTED if the current shapes do not fit the extents
    auto time_stages = [&](float out[5]) -> int {
        for (int rep = 0; rep < 5; ++rep) {
            float st[5];
            const int rc = ppf_solve_timed(p, prhs, P, dmx1, dmx2, dmx3, dx3, stream, st);
            if (rc) return rc;
            for (int i = 0; i < 5; ++i) out[i] = rep == 1 ? st[i] : (rep > 1 && st[i] < out[i] ? st[i] : out[i]);
        }
        return PPF_OK;
    };
    auto restore = [&](int rc) {
        p->xvariant = keep[0];
        p->yvariant = keep[1];
        p->zvariant = keep[2];
        p->xvinv = keep[3];
        return rc;
    };
    // x and y passes: one run per variant times all four line passes
    // (independent kernels); the x-forward, x-inverse and y passes each keep
    // their faster variant
    float bxf = 0.f, bxi = 0.f, by = 0.f;
    int vxf = 0, vxi = 0, vy = 0;
    p->zvariant = keep[2];
    for (int v = 0; v < std::max(kXVariants, kYVariants); ++v) {
        p->xvariant = p->xvinv = v < kXVariants ? v : 0;
        p->yvariant = v < kYVariants ? v : 0;
        float st[5];
        const int rc = time_stages(st);
        if (rc) return restore(rc);
        if (v < kXVariants && (v == 0 || st[0] < bxf)) { bxf = st[0]; vxf = v; }
        if (v < kXVariants && (v == 0 || st[4] < bxi)) { bxi = st[4]; vxi = v; }
        if (v < kYVariants && (v == 0 || st[1] + st[3] < by)) { by = st[1] + st[3]; vy = v; }
    }
    p->xvariant = vxf;
    p->xvinv = vxi;
    p->yvariant = vy;
    // z pass: the partitioned (K,S) shapes and the pipelined wavefront forms
    static const int cand[] = {3, 1, 2, 4, 5, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17};
    int best = 0;
    float best_ms = 0.f;
    for (int v : cand) {
        p->zvariant = v;
        float st[5];
        const int rc = time_stages(st);
        if (rc == PPF_ERR_UNSUPPORTED) continue;
        if (rc) return restore(rc);
        if (best == 0 || st[2] < best_ms) { best = v; best_ms = st[2]; }
    }
    p->zvariant = best ? best : keep[2];
    {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        g_tuned[key] = {x_variant_code(p->xvariant, p->xvinv), p->yvariant, p->zvariant};
    }
    if (zvariant) *zvariant = p->zvariant;
    return PPF_OK;
}

int ppf_plan_variants(const ppf_plan *p, int variants[3]) {
    if (!p || !variants) return PPF_ERR_INVALID_ARG;
    variants[0] = x_variant_code(p->xvariant, p->xvinv);
    variants[1] = p->yvariant;
    variants[2] = p->zvariant;
    return PPF_OK;
}

int ppf_plan_set_variants(ppf_plan *p, int xvariant, int yvariant, int zvariant) {
    int xf = 0, xi = 0;
    if (!p || !x_variant_decode(xvariant, xf, xi) || yvariant < 0 || yvariant >= kYVariants)
        return PPF_ERR_INVALID_ARG;
    const int rc = ppf_plan_set_zvariant(p, zvariant);
    if (rc) return rc;
    p->xvariant = xf;
    p->xvinv = xi;
    p->yvariant = yvariant;
    return PPF_OK;
}

int ppf_plan_set_zvariant(ppf_plan *p, int zvariant) {
    if (!p || zvariant < 0 || zvariant > kZVariants) return PPF_ERR_INVALID_ARG;
    if (zvariant == 6) {
        // the one-thread-per-line kernel keeps its forward multipliers in cs
        const size_t need = (size_t)p->Lx * p->znjr * p->d.n3m;
        if (p->cs_elems < need) {
            double *cs = nullptr;
            if (hipMalloc(&cs, need * sizeof(double)) != hipSuccess) return PPF_ERR_ALLOC;
            if (p->cs) (void)hipFree(p->cs);
            p->bytes += (need - p->cs_elems) * sizeof(double);
            p->cs = cs;
            p->cs_elems = need;
        }
    }
    p->zvariant = zvariant;
    return PPF_OK;
}

int ppf_solve_timed(ppf_plan *p, const double *prhs, double *P, const double *dmx1, const double *dmx2,
                    const double *dmx3, const double *dx3, void *stream, float stage_ms[5]) {
    if (!p || !prhs || !P || !dmx1 || !dmx2 || !dmx3 || !dx3 || !stage_ms) return PPF_ERR_INVALID_ARG;
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipEvent_t ev[6];
    int made = 0;
    int rc = PPF_OK;
    for (; made < 6; ++made)
        if (hipEventCreate(&ev[made]) != hipSuccess) { rc = PPF_ERR_HIP; break; }
    if (!rc) {
        if (p->d.precision == PPF_PREC_FP32_IR)
            rc = run_passes<float>(p, p->W32, p->tx32, p->ty32, prhs, P, dmx1, dmx2, dmx3, dx3, s, ev);
        else
            rc = run_passes<double>(p, p->W, p->tx64, p->ty64, prhs, P, dmx1, dmx2, dmx3, dx3, s, ev);
    }
    if (!rc && hipEventSynchronize(ev[5]) != hipSuccess) rc = PPF_ERR_HIP;
    for (int i = 0; i < 5 && !rc; ++i) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ev[i], ev[i + 1]) != hipSuccess) rc = PPF_ERR_HIP;
        stage_ms[i] = ms;
    }
    for (int i = 0; i < made; ++i) (void)hipEventDestroy(ev[i]);
    return rc;
}

int ppf_residual(ppf_plan *p, const double *f, const double *P, const double *dmx3, const double *dx3, void *stream,
                 double *rel) {
    if (!p || !f || !P || !dmx3 || !dx3 || !rel) return PPF_ERR_INVALID_ARG;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int n3m = p->d.n3m;
    hipLaunchKernelGGL(k_zcoef, dim3((n3m + 255) / 256), dim3(256), 0, s, dx3, dmx3, n3m, p->zcoef, p->zcoef + n3m);
    int rc;
    if ((rc = launch_residual(p, f, P, s))) return rc;
    PPF_HIP(hipMemcpyAsync(p->hred, p->red, 3 * sizeof(double), hipMemcpyDeviceToHost, s));
    PPF_HIP(hipStreamSynchronize(s));
    *rel = p->hred[1] > 0 ? std::sqrt(p->hred[0] / p->hred[1]) : std::sqrt(p->hred[0]);
    return PPF_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Allocation-free solve (ppf_solve_noalloc), for callers that cannot allocate
// device memory: the workspace, axis tables, z row table and scalars of one
// FP64 solve live in `scratch`, a file-scope device array every GPU holds once
// for the life of the process.  Calls on one GPU share it (and keep the last
// call's contents), so calls that overlap on one GPU race.  Nothing in it has
// to start at zero: every pass writes what the next one reads.
// ---------------------------------------------------------------------------
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch[kMaxDevices] = {};  // device address of `scratch`, per GPU

// The device address of the current GPU's scratch array, looked up once per
// GPU (the array's name is a host-side stand-in, not a device address).
int scratch_base(double **out) {
    int dev = 0;
    PPF_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return PPF_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return PPF_ERR_HIP;
        g_scratch[dev] = static_cast<double *>(ptr);
    }
    *out = g_scratch[dev];
    return PPF_OK;
}

// 256-byte aligned slices of the scratch array
struct ScratchCarver {
    char *base;
    size_t used = 0;
    template <typename T>
    T *take(size_t count) {
        T *q = reinterpret_cast<T *>(base + used);
        used += (count * sizeof(T) + 255) / 256 * 256;
        return q;
    }
};

struct NoAllocTables {
    cx<double> *tw[2], *wk[2], *qk[2];
    int N[2], kind[2];
    double *lamx, *lamy;
    unsigned char *zx, *zy;
    int Lx, Ly;
    EigGeom eg;
};

// One axis's FFT tables on the device (the layout upload_tables writes on the
// host: per-stage twiddles, real-FFT split and DCT quarter-wave factors), from
// sincospi of the exact rational angle.
__device__ void fill_axis_tables(cx<double> *tw, cx<double> *wk, cx<double> *qk, int N, int kind, int tid, int nt) {
    const int M = (kind == AX_C2C) ? N : N / 2;
    for (int e = tid; e < M; e += nt) {
        cx<double> v{1.0, 0.0};  // entry M-1 is unused
        int LS = 1;
        while (LS < M) {
            const int R = pick_radix(M / LS);
            if (e < LS - 1 + (R - 1) * LS) {
                const int q = e - (LS - 1), r = q / LS + 1, j = q - (r - 1) * LS;
                double sv, cv;
                sincospi(-2.0 * (double)(j * r) / (double)(LS * R), &sv, &cv);
                v = cx<double>{cv, sv};
                break;
            }
            LS *= R;
        }
        tw[e] = v;
    }
    for (int k = tid; k <= M; k += nt) {
        double sv, cv;
        sincospi(-2.0 * (double)k / (double)N, &sv, &cv);
        wk[k] = cx<double>{cv, sv};
        sincospi(-(double)k / (2.0 * (double)N), &sv, &cv);
        qk[k] = cx<double>{cv, sv};
    }
}

__global__ void __launch_bounds__(256) k_noalloc_tables(NoAllocTables t) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x, nt = gridDim.x * blockDim.x;
    for (int a = 0; a < 2; ++a) fill_axis_tables(t.tw[a], t.wk[a], t.qk[a], t.N[a], t.kind[a], tid, nt);
    for (int kx = tid; kx < t.Lx; kx += nt) eig_col(t.eg, kx, t.lamx[kx], t.zx[kx]);
    for (int jp = tid; jp < t.Ly; jp += nt) eig_row(t.eg, jp, t.lamy[jp], t.zy[jp]);
}
}  // namespace

extern "C" int ppf_solve_noalloc(const ppf_desc *desc, const double *prhs, double *P, const double *dmx1,
                                 const double *dmx2, const double *dmx3, const double *dx3, void *stream) {
    if (!desc || !prhs || !P || !dmx1 || !dmx2 || !dmx3 || !dx3) return PPF_ERR_INVALID_ARG;
    int rc = desc_check(*desc);
    if (rc) return rc;
    if (desc->precision != PPF_PREC_FP64) return PPF_ERR_UNSUPPORTED;
    ppf_plan p;  // geometry and device pointers only; owns nothing
    plan_geometry(*desc, &p);
    const int Mx = (p.xkind == AX_C2C) ? desc->n1m : desc->n1m / 2;
    const int My = (p.ykind == AX_C2C) ? desc->n2m : desc->n2m / 2;
    const int n3m = desc->n3m;
    const bool need_cs = n3m > 2048;  // the one-thread-per-line z kernel's multipliers
    // bytes the slices below take (each rounded up to 256)
    auto sz = [](size_t bytes) { return (bytes + 255) / 256 * 256; };
    const size_t need = sz(sizeof(double) * (size_t)p.Lxs * p.Ly * p.Lz) + sz(sizeof(cx<double>) * Mx) +
                        2 * sz(sizeof(cx<double>) * (Mx + 1)) + sz(sizeof(cx<double>) * My) +
                        2 * sz(sizeof(cx<double>) * (My + 1)) + sz(sizeof(double) * (2 * (size_t)n3m + 4)) +
                        sz(sizeof(double) * p.Lx) + sz(sizeof(double) * p.Ly) + sz(p.Lx) + sz(p.Ly) +
                        sz(sizeof(double)) +
                        (need_cs ? sz(sizeof(double) * (size_t)p.Lx * p.znjr * n3m) : 0);
    if (need > kScratchBytes) return PPF_ERR_SCRATCH;
    double *base = nullptr;
    if ((rc = scratch_base(&base))) return rc;
    ScratchCarver c{reinterpret_cast<char *>(base)};
    p.W = c.take<double>((size_t)p.Lxs * p.Ly * p.Lz);
    p.tx64.tw = c.take<cx<double>>(Mx);
    p.tx64.wk = c.take<cx<double>>(Mx + 1);
    p.tx64.qk = c.take<cx<double>>(Mx + 1);
    p.ty64.tw = c.take<cx<double>>(My);
    p.ty64.wk = c.take<cx<double>>(My + 1);
    p.ty64.qk = c.take<cx<double>>(My + 1);
    p.zcoef = c.take<double>(2 * (size_t)n3m + 4);
    p.lamx = c.take<double>(p.Lx);
    p.lamy = c.take<double>(p.Ly);
    p.zx = c.take<unsigned char>(p.Lx);
    p.zy = c.take<unsigned char>(p.Ly);
    p.mean = c.take<double>(1);
    if (need_cs) {
        p.cs_elems = (size_t)p.Lx * p.znjr * n3m;
        p.cs = c.take<double>(p.cs_elems);
    }
    if (c.used != need) return PPF_ERR_UNSUPPORTED;  // the size check and the slices disagree
    hipStream_t s = static_cast<hipStream_t>(stream);
    NoAllocTables t{};
    t.tw[0] = p.tx64.tw; t.wk[0] = p.tx64.wk; t.qk[0] = p.tx64.qk; t.N[0] = desc->n1m; t.kind[0] = p.xkind;
    t.tw[1] = p.ty64.tw; t.wk[1] = p.ty64.wk; t.qk[1] = p.ty64.qk; t.N[1] = desc->n2m; t.kind[1] = p.ykind;
    t.lamx = p.lamx; t.lamy = p.lamy; t.zx = p.zx; t.zy = p.zy;
    t.Lx = p.Lx; t.Ly = p.Ly;
    t.eg = eig_geom(&p);
    hipLaunchKernelGGL(k_noalloc_tables, dim3(4), dim3(256), 0, s, t);
    PPF_HIP(hipGetLastError());
    return run_passes<double>(&p, p.W, p.tx64, p.ty64, prhs, P, dmx1, dmx2, dmx3, dx3, s, nullptr);
}

extern "C" {

// ---- device memory shim ------------------------------------------------------
int ppf_malloc(void **ptr, size_t bytes) {
    if (!ptr) return PPF_ERR_INVALID_ARG;
    return hipMalloc(ptr, bytes ? bytes : 1) == hipSuccess ? PPF_OK : PPF_ERR_ALLOC;
}
int ppf_free(void *ptr) { return hipFree(ptr) == hipSuccess ? PPF_OK : PPF_ERR_HIP; }
int ppf_memcpy(void *dst, const void *src, size_t bytes, int kind) {
    hipMemcpyKind k;
    switch (kind) {
        case PPF_MEMCPY_HOST_TO_DEVICE: k = hipMemcpyHostToDevice; break;
        case PPF_MEMCPY_DEVICE_TO_HOST: k = hipMemcpyDeviceToHost; break;
        case PPF_MEMCPY_DEVICE_TO_DEVICE: k = hipMemcpyDeviceToDevice; break;
        default: return PPF_ERR_INVALID_ARG;
    }
    return hipMemcpy(dst, src, bytes, k) == hipSuccess ? PPF_OK : PPF_ERR_HIP;
}
int ppf_memset(void *ptr, int value, size_t bytes) { return hipMemset(ptr, value, bytes) == hipSuccess ? PPF_OK : PPF_ERR_HIP; }
int ppf_sync(void *stream) {
    return hipStreamSynchronize(static_cast<hipStream_t>(stream)) == hipSuccess ? PPF_OK : PPF_ERR_HIP;
}
int ppf_device_count(int *count) {
    if (!count) return PPF_ERR_INVALID_ARG;
    return hipGetDeviceCount(count) == hipSuccess ? PPF_OK : PPF_ERR_HIP;
}
int ppf_set_device(int device) { return hipSetDevice(device) == hipSuccess ? PPF_OK : PPF_ERR_HIP; }

int ppf_copy_timed(double *dst, const double *src, size_t count, int reps, void *stream, float *ms) {
    if (!dst || !src || reps <= 0 || !ms || count == 0 || (count & 1)) return PPF_ERR_INVALID_ARG;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const size_t n2 = count / 2;
    if ((n2 + 255) / 256 > 0x7fffffffu) return PPF_ERR_SIZE;
    const unsigned grid = (unsigned)((n2 + 255) / 256);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess) return PPF_ERR_HIP;
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return PPF_ERR_HIP;
    }
    int rc = PPF_OK;
    if (hipEventRecord(e0, s) != hipSuccess) rc = PPF_ERR_HIP;
    for (int r = 0; r < reps && rc == PPF_OK; ++r) {
        hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, s, reinterpret_cast<D2v *>(dst),
                           reinterpret_cast<const D2v *>(src), n2);
        if (hipGetLastError() != hipSuccess) rc = PPF_ERR_HIP;
    }
    if (rc == PPF_OK && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                         hipEventElapsedTime(ms, e0, e1) != hipSuccess))
        rc = PPF_ERR_HIP;
    if (rc == PPF_OK) *ms /= (float)reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

}  // extern "C"
