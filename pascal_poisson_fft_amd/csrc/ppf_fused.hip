// ppf_fused.hip -- host side of the fused batch (ppf_fused.hpp): one pooled
// allocation holds every problem's workspace, axis tables, eigenvalue tables,
// z row table and mean; the item schedule (tiles per stage and problem) is
// built once; a solve is a counter reset plus one persistent launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "ppf_aux.hpp"
#include "ppf_fused.hpp"
#include "ppf_plan.hpp"

using namespace ppf;

struct ppf_fused {
    int nprob = 0, nitems = 0;
    std::vector<FusedProb> h;
    std::vector<ppf_desc> desc;
    FusedProb *dprobs = nullptr;
    int *gstart = nullptr;   // 5 nprob + 1
    int *ctr = nullptr;      // 2 + 5 nprob
    char *pool = nullptr;
    double *partial = nullptr, *red = nullptr;  // residual reductions (stats)
    size_t shm = 0;
    unsigned grid = 0;
};

namespace {

constexpr int kResPartials = 2048;

// geometry of the tile functions for line length N (host side)
template <int N>
struct FusedGeom {
    static int x_lpb(bool per) { return per ? XGeom<double, N, true>::LPB : XGeom<double, N, false>::LPB; }
    static size_t x_lds(bool per) {
        return sizeof(double) * (per ? XGeom<double, N, true>::L::SIZE : XGeom<double, N, false>::L::SIZE);
    }
    static int y_tx(int mode) {
        return mode == 0 ? YGeom<double, N, 0, 0, kFusedNT>::TX
                         : (mode == 1 ? YGeom<double, N, 1, 0, kFusedNT>::TX : YGeom<double, N, 2, 0, kFusedNT>::TX);
    }
    static size_t y_lds(int mode) {
        return sizeof(double) * (mode == 0 ? YGeom<double, N, 0, 0, kFusedNT>::L::SIZE
                                           : (mode == 1 ? YGeom<double, N, 1, 0, kFusedNT>::L::SIZE
                                                        : YGeom<double, N, 2, 0, kFusedNT>::L::SIZE));
    }
};

bool fused_size(int n) {
    switch (n) {
#define PPF_FC(v) case v:
        PPF_FUSED_SIZES(PPF_FC)
#undef PPF_FC
        return true;
    }
    return false;
}

template <typename F>
auto by_size(int n, F f) {
    switch (n) {
#define PPF_FC(v) case v: return f(FusedGeom<v>{});
        PPF_FUSED_SIZES(PPF_FC)
#undef PPF_FC
    }
    return f(FusedGeom<32>{});  // unreachable: sizes are checked first
}

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

}  // namespace

namespace ppf {

bool fused_supported(const ppf_desc &d) {
    return d.precision == PPF_PREC_FP64 && fused_size(d.n1m) && fused_size(d.n2m) && d.n3m >= 2 &&
           (d.n3m & 1) == 0 && d.n3m <= kFusedK * kFusedMaxP;
}

void fused_destroy(ppf_fused *f) {
    if (!f) return;
    void *ptrs[] = {f->dprobs, f->gstart, f->ctr, f->pool, f->partial, f->red};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    delete f;
}

int fused_create(int n, const ppf_desc *descs, const ppf_batch_io *io, ppf_fused **out) {
    *out = nullptr;
    for (int i = 0; i < n; ++i)
        if (!fused_supported(descs[i])) return PPF_ERR_UNSUPPORTED;
    ppf_fused *f = new (std::nothrow) ppf_fused();
    if (!f) return PPF_ERR_ALLOC;
    f->nprob = n;
    f->h.resize(n);
    f->desc.assign(descs, descs + n);
    // pool layout and the host image of everything but the workspaces
    struct Slices {
        size_t W, xtw, xwk, xqk, ytw, ywk, yqk, lamx, lamy, zx, zy, zcoef, mean;
    };
    std::vector<Slices> sl(n);
    std::vector<ppf_plan> geo(n);
    size_t used = 0;
    auto take = [&](size_t bytes) {
        const size_t at = used;
        used += align256(bytes);
        return at;
    };
    size_t shm = 0;
    for (int i = 0; i < n; ++i) {
        ppf_plan &p = geo[i];
        plan_geometry(descs[i], &p);
        const int Mx = p.xkind == AX_C2C ? descs[i].n1m : descs[i].n1m / 2;
        const int My = p.ykind == AX_C2C ? descs[i].n2m : descs[i].n2m / 2;
        Slices &s = sl[i];
        s.W = take(sizeof(double) * (size_t)p.Lxs * p.Ly * p.Lz);
        s.xtw = take(sizeof(cx<double>) * Mx);
        s.xwk = take(sizeof(cx<double>) * (Mx + 1));
        s.xqk = take(sizeof(cx<double>) * (Mx + 1));
        s.ytw = take(sizeof(cx<double>) * My);
        s.ywk = take(sizeof(cx<double>) * (My + 1));
        s.yqk = take(sizeof(cx<double>) * (My + 1));
        s.lamx = take(sizeof(double) * p.Lx);
        s.lamy = take(sizeof(double) * p.Ly);
        s.zx = take(p.Lx);
        s.zy = take(p.Ly);
        s.zcoef = take(sizeof(double) * (2 * (size_t)descs[i].n3m + 4));
        s.mean = take(sizeof(double));
        const bool per = p.xkind == AX_R2C;
        const int ymode = p.ymode;
        const size_t zlds = sizeof(double) * (2 * (size_t)((kFusedK + kFusedS - 1) / kFusedS) * kFusedNT +
                                              7 * (size_t)kFusedMaxP * kFusedS);
        shm = std::max({shm, by_size(descs[i].n1m, [&](auto g) { return g.x_lds(per); }),
                        by_size(descs[i].n2m, [&](auto g) { return g.y_lds(ymode); }), zlds});
    }
    f->shm = shm;
    int rc = PPF_OK;
    if (hipMalloc(&f->pool, used ? used : 1) != hipSuccess || hipMalloc(&f->dprobs, sizeof(FusedProb) * n) != hipSuccess ||
        hipMalloc(&f->gstart, sizeof(int) * (5 * (size_t)n + 1)) != hipSuccess ||
        hipMalloc(&f->ctr, sizeof(int) * (2 + 5 * (size_t)n)) != hipSuccess ||
        hipMalloc(&f->partial, sizeof(double) * 3 * kResPartials) != hipSuccess ||
        hipMalloc(&f->red, sizeof(double) * 3) != hipSuccess) {
        fused_destroy(f);
        return PPF_ERR_ALLOC;
    }
    std::vector<char> img(used, 0);
    std::vector<int> gstart(5 * (size_t)n + 1, 0);
    for (int i = 0; i < n; ++i) {
        const ppf_plan &p = geo[i];
        const ppf_desc &d = descs[i];
        const Slices &s = sl[i];
        std::vector<cx<double>> tw, wk, qk;
        auto put = [&](size_t at, const void *src, size_t bytes) { std::memcpy(img.data() + at, src, bytes); };
        host_axis_tables<double>(d.n1m, p.xkind, tw, wk, qk);
        put(s.xtw, tw.data(), sizeof(cx<double>) * tw.size());
        put(s.xwk, wk.data(), sizeof(cx<double>) * wk.size());
        put(s.xqk, qk.data(), sizeof(cx<double>) * qk.size());
        host_axis_tables<double>(d.n2m, p.ykind, tw, wk, qk);
        put(s.ytw, tw.data(), sizeof(cx<double>) * tw.size());
        put(s.ywk, wk.data(), sizeof(cx<double>) * wk.size());
        put(s.yqk, qk.data(), sizeof(cx<double>) * qk.size());
        const EigGeom eg = eig_geom(&p);
        for (int kx = 0; kx < p.Lx; ++kx)
            eig_col(eg, kx, reinterpret_cast<double *>(img.data() + s.lamx)[kx],
                    reinterpret_cast<unsigned char *>(img.data() + s.zx)[kx]);
        for (int jp = 0; jp < p.Ly; ++jp)
            eig_row(eg, jp, reinterpret_cast<double *>(img.data() + s.lamy)[jp],
                    reinterpret_cast<unsigned char *>(img.data() + s.zy)[jp]);
        FusedProb &q = f->h[i];
        char *base = f->pool;
        q.in = io[i].prhs;
        q.P = io[i].p;
        q.dmx1 = io[i].dmx1;
        q.dmx2 = io[i].dmx2;
        q.dmx3 = io[i].dmx3;
        q.dx3 = io[i].dx3;
        q.W = reinterpret_cast<double *>(base + s.W);
        q.xtw = reinterpret_cast<const cx<double> *>(base + s.xtw);
        q.xwk = reinterpret_cast<const cx<double> *>(base + s.xwk);
        q.xqk = reinterpret_cast<const cx<double> *>(base + s.xqk);
        q.ytw = reinterpret_cast<const cx<double> *>(base + s.ytw);
        q.ywk = reinterpret_cast<const cx<double> *>(base + s.ywk);
        q.yqk = reinterpret_cast<const cx<double> *>(base + s.yqk);
        q.lamx = reinterpret_cast<const double *>(base + s.lamx);
        q.lamy = reinterpret_cast<const double *>(base + s.lamy);
        q.zx = reinterpret_cast<const unsigned char *>(base + s.zx);
        q.zy = reinterpret_cast<const unsigned char *>(base + s.zy);
        q.zcoef = reinterpret_cast<double *>(base + s.zcoef);
        q.mean = reinterpret_cast<double *>(base + s.mean);
        q.n1m = d.n1m;
        q.n2m = d.n2m;
        q.n3m = d.n3m;
        q.Lx = p.Lx;
        q.Lxs = p.Lxs;
        q.Ly = p.Ly;
        q.Lz = p.Lz;
        q.per_x = p.xkind == AX_R2C;
        q.ymode = p.ymode;
        q.zper = d.pbc3 != 0;
        q.zjlo = p.zjlo;
        q.znjr = p.znjr;
        q.mean_kx = p.mean_kx;
        q.mean_jp = p.mean_jp;
        q.mean_scale = p.mean_scale;
        q.nzp = (d.n3m + kFusedK - 1) / kFusedK;
        const int lpb = by_size(d.n1m, [&](auto g) { return g.x_lpb(q.per_x != 0); });
        const int tx = by_size(d.n2m, [&](auto g) { return g.y_tx(q.ymode); });
        const int nct = (p.Lx + tx - 1) / tx;
        q.ntiles[0] = (d.n2m * d.n3m + lpb - 1) / lpb;
        q.ntiles[1] = nct * d.n3m;
        q.ntiles[2] = ((p.Lx + kFusedS - 1) / kFusedS) * p.znjr;
        q.ntiles[3] = nct * p.Lz;
        q.ntiles[4] = (p.Ly * p.Lz + lpb - 1) / lpb;
    }
    // groups stage-major: all problems' stage-0 tiles, then stage 1, ...
    int items = 0;
    for (int st = 0; st < 5; ++st)
        for (int i = 0; i < n; ++i) {
            gstart[(size_t)st * n + i] = items;
            items += f->h[i].ntiles[st];
        }
    gstart[5 * (size_t)n] = items;
    f->nitems = items;
    if (hipMemcpy(f->pool, img.data(), used, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(f->dprobs, f->h.data(), sizeof(FusedProb) * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(f->gstart, gstart.data(), sizeof(int) * gstart.size(), hipMemcpyHostToDevice) != hipSuccess)
        rc = PPF_ERR_HIP;
    if (!rc && f->shm > 65536 &&
        hipFuncSetAttribute(reinterpret_cast<const void *>(&k_fused_batch), hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)f->shm) != hipSuccess)
        rc = PPF_ERR_HIP;
    if (!rc) f->grid = persistent_grid(k_fused_batch, kFusedNT, f->nitems, f->shm);
    if (rc) {
        fused_destroy(f);
        return rc;
    }
    *out = f;
    return PPF_OK;
}

int fused_launch(ppf_fused *f, hipStream_t s) {
    const int nctr = 2 + 5 * f->nprob;
    hipLaunchKernelGGL(k_fused_reset, dim3((nctr + 255) / 256), dim3(256), 0, s, f->ctr, nctr);
    FusedArgs a{f->dprobs, f->nprob, f->gstart, f->nitems, f->ctr};
    hipLaunchKernelGGL(k_fused_batch, dim3(f->grid), dim3(kFusedNT), f->shm, s, a);
    PPF_HIP(hipGetLastError());
    return PPF_OK;
}

int fused_status(ppf_fused *f, hipStream_t s) {
    int flag = 0;
    PPF_HIP(hipStreamSynchronize(s));
    PPF_HIP(hipMemcpy(&flag, f->ctr + 1, sizeof(int), hipMemcpyDeviceToHost));
    return flag ? PPF_ERR_HIP : PPF_OK;
}

int fused_residual(ppf_fused *f, int i, hipStream_t s, double *rel) {
    const FusedProb &q = f->h[i];
    const ppf_desc &d = f->desc[i];
    XRes xr{};
    xr.P = q.P;
    xr.zam = q.zcoef;
    xr.zap = q.zcoef + d.n3m;
    xr.partial = f->partial;
    const double h1 = d.L1 / d.n1m, h2 = d.L2 / d.n2m;
    xr.ih1 = 1.0 / (h1 * h1);
    xr.ih2 = 1.0 / (h2 * h2);
    xr.pbc1 = d.pbc1 != 0;
    xr.pbc2 = d.pbc2 != 0;
    const int ty = residual_rows(d.n1m, d.n2m, d.n3m, 4L * num_cus());
    const long ntiles = residual_tiles(d.n1m, d.n2m, d.n3m, ty);
    const int g = (int)std::min<long>(std::min<long>(kResPartials, 8L * num_cus()), ntiles);
    launch_k_residual(ty, g, s, q.in, xr, d.n1m, d.n2m, d.n3m, nullptr);
    hipLaunchKernelGGL(k_reduce3, dim3(1), dim3(256), 0, s, f->partial, g, f->red, nullptr);
    PPF_HIP(hipGetLastError());
    double hr[3];
    PPF_HIP(hipStreamSynchronize(s));
    PPF_HIP(hipMemcpy(hr, f->red, sizeof(hr), hipMemcpyDeviceToHost));
    *rel = hr[1] > 0 ? std::sqrt(hr[0] / hr[1]) : std::sqrt(hr[0]);
    return PPF_OK;
}

}  // namespace ppf
