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
#include <cstring>
#include <new>
#include <type_traits>
#include <vector>

#include "../../include/pascal_poisson_fft.h"
#include "ppf_kernels.hpp"

using namespace ppf;

namespace {

const double kPi = 3.14159265358979323846264338327950288;

// Transform kind along one axis.
enum AxisKind { AX_DCT = 0, AX_R2C = 1, AX_C2C = 2 };

template <typename T>
struct AxisTables {
    cx<T> *tw = nullptr;  // FFT twiddles, M entries
    cx<T> *wk = nullptr;  // exp(-2 pi i k / N), M+1 entries
    cx<T> *qk = nullptr;  // exp(-i pi k / 2N), M+1 entries
};

bool size_supported(int n) {
    if (n < 8 || n > 1024 || (n & 1)) return false;
    int m = n;
    if (m % 3 == 0) m /= 3;
    while (m % 2 == 0) m /= 2;
    return m == 1;
}

}  // namespace

constexpr int kMaxSlabs = 16;

struct ppf_plan {
    ppf_desc d;
    int Lx, Ly, Lz;        // workspace extents (scalars)
    int Lxs;               // workspace row stride (Lx rounded up to 128 bytes)
    int xkind, ykind;      // AxisKind
    int ymode;             // y pass MODE (0 DCT, 1 C2C, 2 R2C)
    int zjlo, znjr;        // rows solved by the z pass
    int mean_kx, mean_jp;
    double mean_scale;
    int zvariant;          // z-pass kernel shape (0 = chosen by n3m; PPF_ZVARIANT overrides)
    // slab pipeline: x-forward -> y-forward and y-inverse -> x-inverse run
    // slab by slab on two streams, the second pass reading what the first
    // just wrote while it is still cache resident
    int nslab = 1;
    hipStream_t s2 = nullptr;
    hipEvent_t pev[2 * kMaxSlabs + 4] = {};
    // device memory
    double *W = nullptr;   // spectral workspace (FP64 plans)
    float *W32 = nullptr;  // spectral workspace (FP32_IR plans)
    double *cs = nullptr;  // Thomas scratch of the one-thread-per-line z kernel
    size_t cs_elems = 0;
    double *zcoef = nullptr;  // z row coefficients of the current solve (2*n3m+4)
    double *lamx = nullptr, *lamy = nullptr;
    unsigned char *zx = nullptr, *zy = nullptr;
    double *mean = nullptr;     // device scalar
    double *partial = nullptr;  // residual partial sums
    double *red = nullptr;      // residual totals: sum r^2, sum f^2, sum p (3)
    int npartial = 0;
    double *hred = nullptr;     // pinned host copy of red (refinement convergence test)
    AxisTables<double> tx64, ty64;
    AxisTables<float> tx32, ty32;
    size_t bytes = 0;
};

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------
#define PPF_HIP(call)                                  \
    do {                                               \
        hipError_t e_ = (call);                        \
        if (e_ != hipSuccess) return PPF_ERR_HIP;      \
    } while (0)

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
    const int M = (kind == AX_C2C) ? N : N / 2;
    std::vector<cx<T>> tw(M, cx<T>{T(1), T(0)}), wk(M + 1), qk(M + 1);
    // per-stage FFT twiddles: stage of span LS, radix R at LS-1 + (r-1) LS + j
    for (int LS = 1; LS < M;) {
        const int R = pick_radix(M / LS);
        for (int r = 1; r < R; ++r)
            for (int j = 0; j < LS; ++j) {
                long double a = -2.0L * (long double)kPi * (long double)(j * r) / (long double)(LS * R);
                tw[(size_t)(LS - 1) + (size_t)(r - 1) * LS + j] = {(T)cosl(a), (T)sinl(a)};
            }
        LS *= R;
    }
    for (int k = 0; k <= M; ++k) {
        long double a = -2.0L * (long double)kPi * (long double)k / (long double)N;
        wk[k] = {(T)cosl(a), (T)sinl(a)};
        long double b = -(long double)kPi * (long double)k / (2.0L * (long double)N);
        qk[k] = {(T)cosl(b), (T)sinl(b)};
    }
    int rc;
    if ((rc = dev_alloc(p, &tb.tw, M)) || (rc = dev_alloc(p, &tb.wk, M + 1)) || (rc = dev_alloc(p, &tb.qk, M + 1)))
        return rc;
    PPF_HIP(hipMemcpy(tb.tw, tw.data(), sizeof(cx<T>) * M, hipMemcpyHostToDevice));
    PPF_HIP(hipMemcpy(tb.wk, wk.data(), sizeof(cx<T>) * (M + 1), hipMemcpyHostToDevice));
    PPF_HIP(hipMemcpy(tb.qk, qk.data(), sizeof(cx<T>) * (M + 1), hipMemcpyHostToDevice));
    return PPF_OK;
}

// modified wavenumbers, fft_3d_poisson.f90:288-351
static double eig_neumann(int m, int n, double h) {
    double s = std::sin(0.5 * (double)m * kPi / (double)n);
    return 4.0 * s * s / (h * h);
}
static double eig_periodic(int m, int n, double h) {
    return 2.0 * (1.0 - std::cos(2.0 * (double)m * kPi / (double)n)) / (h * h);
}

// ---------------------------------------------------------------------------
// kernel dispatch over the supported line lengths
// ---------------------------------------------------------------------------
#define PPF_FOR_SIZES(X) X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768) X(1024)

static int num_cus() {
    static int n = 0;
    if (n <= 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
            n = 256;
    }
    return n;
}

// grid of a persistent kernel: as many workgroups as can be resident at once
template <typename K>
static unsigned persistent_grid(K kernel, int nthreads, int ntiles, size_t dyn_lds = 0) {
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, nthreads, dyn_lds) != hipSuccess || occ < 1) occ = 1;
    const long g = (long)occ * num_cus();
    return (unsigned)std::max(1L, std::min<long>(g, ntiles));
}

// the workgroup count of a persistent launch that writes per-workgroup partial
// sums is capped by the partial buffer
template <typename T, int N, bool PER, bool RES>
static unsigned launch_xfwd_k(const ppf_plan *p, const double *in, T *W, const AxisTables<T> &tb, const XRes &xr,
                              const ZRows &zr, hipStream_t s, int nplanes) {
    using X = XGeom<T, N, PER>;
    const int nlines = p->d.n2m * nplanes;
    unsigned g = persistent_grid(k_xfwd<T, N, PER, RES>, X::NT, (nlines + X::LPB - 1) / X::LPB);
    if (RES) g = std::min<unsigned>(g, (unsigned)p->npartial);
    hipLaunchKernelGGL((k_xfwd<T, N, PER, RES>), dim3(g), dim3(X::NT), 0, s, in, W, nlines, p->d.n2m, p->Ly, tb.tw,
                       tb.wk, tb.qk, xr, zr);
    return g;
}

template <typename T, int N>
static unsigned launch_xfwd(const ppf_plan *p, const double *in, T *W, const AxisTables<T> &tb, const XRes *xr,
                            const ZRows &zr, hipStream_t s, int nplanes) {
    XRes none{};
    if (p->xkind == AX_R2C)
        return xr ? launch_xfwd_k<T, N, true, true>(p, in, W, tb, *xr, zr, s, nplanes)
                  : launch_xfwd_k<T, N, true, false>(p, in, W, tb, none, zr, s, nplanes);
    return xr ? launch_xfwd_k<T, N, false, true>(p, in, W, tb, *xr, zr, s, nplanes)
              : launch_xfwd_k<T, N, false, false>(p, in, W, tb, none, zr, s, nplanes);
}

template <typename T, int N, bool PER, bool ACC>
static void launch_xinv_k(const ppf_plan *p, const T *W, double *P, const AxisTables<T> &tb, const double *dmx1,
                          hipStream_t s, int nplanes) {
    using X = XGeom<T, N, PER>;
    const int nlines = p->Ly * nplanes;
    const unsigned g = persistent_grid(k_xinv<T, N, PER, ACC>, X::NT, (nlines + X::LPB - 1) / X::LPB);
    const double inv_cells = 1.0 / ((double)p->d.n1m * (double)p->d.n2m * (double)p->d.n3m);
    hipLaunchKernelGGL((k_xinv<T, N, PER, ACC>), dim3(g), dim3(X::NT), 0, s, W, P, nlines, tb.tw, tb.wk, tb.qk, dmx1,
                       p->mean, p->red, inv_cells);
}

template <typename T, int N>
static void launch_xinv(const ppf_plan *p, const T *W, double *P, const AxisTables<T> &tb, const double *dmx1,
                        bool acc, hipStream_t s, int nplanes) {
    if (p->xkind == AX_R2C) {
        if (acc) launch_xinv_k<T, N, true, true>(p, W, P, tb, dmx1, s, nplanes);
        else launch_xinv_k<T, N, true, false>(p, W, P, tb, dmx1, s, nplanes);
    } else {
        if (acc) launch_xinv_k<T, N, false, true>(p, W, P, tb, dmx1, s, nplanes);
        else launch_xinv_k<T, N, false, false>(p, W, P, tb, dmx1, s, nplanes);
    }
}

// nplanes: planes of the forward pass (interior planes after the base) or of
// the inverse pass (all planes from the base, ghost planes included)
template <typename T, int N, int MODE>
static void launch_y_mode(const ppf_plan *p, T *W, const AxisTables<T> &tb, bool inverse, const double *dmx2,
                          hipStream_t s, int nplanes) {
    using Y = YGeom<T, N, MODE>;
    const int nb = (p->Lx + Y::TX - 1) / Y::TX;
    if (!inverse) {
        const unsigned g = persistent_grid(k_yfwd<T, N, MODE>, Y::NT, nb * nplanes);
        hipLaunchKernelGGL((k_yfwd<T, N, MODE>), dim3(g), dim3(Y::NT), 0, s, W, p->Lx, p->Lxs, p->Ly, nplanes, tb.tw, tb.wk,
                           tb.qk);
    } else {
        const unsigned g = persistent_grid(k_yinv<T, N, MODE>, Y::NT, nb * nplanes);
        hipLaunchKernelGGL((k_yinv<T, N, MODE>), dim3(g), dim3(Y::NT), 0, s, W, p->Lx, p->Lxs, p->Ly, nplanes, tb.tw, tb.wk,
                           tb.qk, dmx2);
    }
}

template <typename T, int N>
static void launch_y(const ppf_plan *p, T *W, const AxisTables<T> &tb, bool inverse, const double *dmx2,
                     hipStream_t s, int nplanes) {
    if (p->ymode == 0) launch_y_mode<T, N, 0>(p, W, tb, inverse, dmx2, s, nplanes);
    else if (p->ymode == 1) launch_y_mode<T, N, 1>(p, W, tb, inverse, dmx2, s, nplanes);
    else launch_y_mode<T, N, 2>(p, W, tb, inverse, dmx2, s, nplanes);
}

static ZRows z_rows(const ppf_plan *p, const double *dx3, const double *dmx3) {
    return ZRows{dx3, dmx3, p->zcoef, p->zcoef + p->d.n3m, p->d.n3m};
}

// zr: when zr.zam is set, workgroup 0 also fills the plan's z row table
template <typename T>
static int dispatch_xfwd(const ppf_plan *p, const double *in, T *W, const AxisTables<T> &tb, const XRes *xr,
                         hipStream_t s, unsigned *grid = nullptr, int nplanes = -1, ZRows zr = ZRows{}) {
    unsigned g = 0;
    if (nplanes < 0) nplanes = p->d.n3m;
    switch (p->d.n1m) {
#define X_CASE(n) \
    case n:                                                         \
        g = launch_xfwd<T, n>(p, in, W, tb, xr, zr, s, nplanes);   \
        if (grid) *grid = g;                                        \
        return PPF_OK;
        PPF_FOR_SIZES(X_CASE)
#undef X_CASE
    }
    return PPF_ERR_SIZE;
}
template <typename T>
static int dispatch_xinv(const ppf_plan *p, const T *W, double *P, const AxisTables<T> &tb, const double *dmx1,
                         bool acc, hipStream_t s, int nplanes = -1) {
    if (nplanes < 0) nplanes = p->Lz;
    switch (p->d.n1m) {
#define X_CASE(n) \
    case n: launch_xinv<T, n>(p, W, P, tb, dmx1, acc, s, nplanes); return PPF_OK;
        PPF_FOR_SIZES(X_CASE)
#undef X_CASE
    }
    return PPF_ERR_SIZE;
}
template <typename T>
static int dispatch_y(const ppf_plan *p, T *W, const AxisTables<T> &tb, bool inverse, const double *dmx2,
                      hipStream_t s, int nplanes = -1) {
    if (nplanes < 0) nplanes = inverse ? p->Lz : p->d.n3m;
    switch (p->d.n2m) {
#define Y_CASE(n) \
    case n: launch_y<T, n>(p, W, tb, inverse, dmx2, s, nplanes); return PPF_OK;
        PPF_FOR_SIZES(Y_CASE)
#undef Y_CASE
    }
    return PPF_ERR_SIZE;
}

template <typename T>
static int launch_z(ppf_plan *p, T *W, const double *dx3, const double *dmx3, hipStream_t s) {
    // k_zpart and k_zpipe read the z row table that pass 1 of this solve filled
    ZArgs z;
    z.Lx = p->Lx;
    z.Lxs = p->Lxs;
    z.Ly = p->Ly;
    z.n3m = p->d.n3m;
    z.jlo = p->zjlo;
    z.njr = p->znjr;
    z.lamx = p->lamx;
    z.lamy = p->lamy;
    z.zx = p->zx;
    z.zy = p->zy;
    z.mean_kx = p->mean_kx;
    z.mean_jp = p->mean_jp;
    z.mean_scale = p->mean_scale;
    z.zper = p->d.pbc3 != 0;
    const int n3m = p->d.n3m;
    // partitioned register-resident Thomas, shape (K rows per thread, S
    // columns per workgroup); launched only when S*P <= NTMAX and the LDS
    // footprint stays within 64 KiB, else the one-thread-per-line kernel
    bool done = false;
    auto part = [&](auto kk, auto ss, auto nt, auto wpe) {
        constexpr int K = decltype(kk)::value, S = decltype(ss)::value;
        constexpr int NTMAX = decltype(nt)::value, WPE = decltype(wpe)::value;
        const int P = (n3m + K - 1) / K;
        constexpr int KS = (K + S - 1) / S;  // z table rows per thread
        const size_t shm = sizeof(double) * (2 * (size_t)KS * S * P + 7 * (size_t)P * S);
        if (P < 1 || S * P > NTMAX || shm > 65536) return;
        const int tiles = (p->Lx + S - 1) / S;
        const int ntiles = tiles * p->znjr;
        // FP64 arithmetic for FP32 storage too: FP32 arithmetic measured no
        // faster at 512^3 (0.509 vs 0.515 ms) and costs the refinement a sweep
        hipLaunchKernelGGL((k_zpart<T, K, S, NTMAX, WPE>), dim3((unsigned)ntiles), dim3(S * P), shm, s, W, z,
                           p->zcoef, p->zcoef + n3m, p->mean, P);
        done = true;
    };
    // pipelined variant: persistent, prefetching, wavefront cyclic reduction
    auto pipe = [&](auto kk, auto ss, auto wpe) {
        constexpr int K = decltype(kk)::value, S = decltype(ss)::value, WPE = decltype(wpe)::value;
        const int P = (n3m + K - 1) / K;
        const size_t shm = sizeof(double) * (2 * (size_t)P * K + 3 * (size_t)S * 65 + (size_t)P * S);
        if (P < 1 || P > 32 || K < 2 || shm > 65536) return;
        const double *zam = p->zcoef, *zap = p->zcoef + n3m;
        const int ntiles = ((p->Lx + S - 1) / S) * p->znjr;
        const unsigned g = persistent_grid(k_zpipe<T, K, S, WPE>, S * P, ntiles, shm);
        hipLaunchKernelGGL((k_zpipe<T, K, S, WPE>), dim3(g), dim3(S * P), shm, s, W, z, zam, zap, p->mean, P);
        done = true;
    };
    using std::integral_constant;
    int v = p->zvariant;
    // measured on MI355X at 512^3: (K=32, S=16) 0.54 ms, pipelined K=16 0.81 ms
    if (v == 0) v = (n3m <= 1024) ? 3 : 4;
    // the pipelined form writes wall ghost planes only
    if (z.zper && v >= 7) v = (n3m <= 1024) ? 3 : 4;
    switch (v) {
        case 7: pipe(integral_constant<int, 16>{}, integral_constant<int, 16>{}, integral_constant<int, 2>{}); break;
        case 8: pipe(integral_constant<int, 16>{}, integral_constant<int, 8>{}, integral_constant<int, 2>{}); break;
        case 9: pipe(integral_constant<int, 32>{}, integral_constant<int, 8>{}, integral_constant<int, 2>{}); break;
        case 10: pipe(integral_constant<int, 32>{}, integral_constant<int, 16>{}, integral_constant<int, 2>{}); break;
        case 11: pipe(integral_constant<int, 8>{}, integral_constant<int, 16>{}, integral_constant<int, 4>{}); break;
        default: break;
    }
    switch (v) {
        case 1: part(integral_constant<int, 16>{}, integral_constant<int, 16>{}, integral_constant<int, 512>{}, integral_constant<int, 3>{}); break;
        case 2: part(integral_constant<int, 16>{}, integral_constant<int, 8>{}, integral_constant<int, 512>{}, integral_constant<int, 3>{}); break;
        case 3: part(integral_constant<int, 32>{}, integral_constant<int, 16>{}, integral_constant<int, 512>{}, integral_constant<int, 2>{}); break;
        case 4: part(integral_constant<int, 32>{}, integral_constant<int, 8>{}, integral_constant<int, 512>{}, integral_constant<int, 2>{}); break;
        case 5: part(integral_constant<int, 8>{}, integral_constant<int, 8>{}, integral_constant<int, 512>{}, integral_constant<int, 4>{}); break;
        case 12: part(integral_constant<int, 24>{}, integral_constant<int, 16>{}, integral_constant<int, 512>{}, integral_constant<int, 2>{}); break;
        case 13: part(integral_constant<int, 24>{}, integral_constant<int, 8>{}, integral_constant<int, 512>{}, integral_constant<int, 2>{}); break;
        case 14: part(integral_constant<int, 48>{}, integral_constant<int, 8>{}, integral_constant<int, 512>{}, integral_constant<int, 1>{}); break;
        default: break;
    }
    if (!done) {
        const long ncol = (long)p->Lx * p->znjr;
        // the one-thread-per-line kernel keeps its forward multipliers in cs
        if (p->cs_elems < (size_t)ncol * (size_t)n3m) return PPF_ERR_UNSUPPORTED;
        const int nt = 256;
        const size_t shm = 2 * sizeof(double) * (size_t)n3m;
        hipLaunchKernelGGL((k_zthomas<T>), dim3((unsigned)((ncol + nt - 1) / nt)), dim3(nt), shm, s, W,
                           reinterpret_cast<T *>(p->cs), z,
                           dx3, dmx3, p->mean);
    }
    return PPF_OK;
}

static void free_plan(ppf_plan *p) {
    if (!p) return;
    void *ptrs[] = {p->W,        p->W32,      p->cs,       p->zcoef,    p->lamx,     p->lamy,     p->zx,
                    p->zy,       p->mean,     p->partial,  p->red,      p->tx64.tw,  p->tx64.wk,  p->tx64.qk,
                    p->ty64.tw,  p->ty64.wk,  p->ty64.qk,  p->tx32.tw,  p->tx32.wk,  p->tx32.qk,  p->ty32.tw,
                    p->ty32.wk,  p->ty32.qk};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    if (p->hred) (void)hipHostFree(p->hred);
    for (hipEvent_t e : p->pev)
        if (e) (void)hipEventDestroy(e);
    if (p->s2) (void)hipStreamDestroy(p->s2);
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
        case PPF_ERR_SIZE: return "unsupported extent (x/y need even 8..1024 of the form 2^a or 3*2^a; z >= 2 even)";
        case PPF_ERR_HIP: return "HIP runtime error";
        case PPF_ERR_ALLOC: return "device allocation failed";
        case PPF_ERR_UNSUPPORTED: return "unsupported configuration";
        case PPF_ERR_NOT_CONVERGED: return "refinement did not reach the tolerance";
    }
    return "unknown error";
}

int ppf_plan_create(const ppf_desc *desc, ppf_plan **out) {
    if (!desc || !out) return PPF_ERR_INVALID_ARG;
    *out = nullptr;
    const ppf_desc &d = *desc;
    if (d.n1m <= 0 || d.n2m <= 0 || d.n3m <= 0 || !(d.L1 > 0) || !(d.L2 > 0) || !(d.L3 > 0)) return PPF_ERR_INVALID_ARG;
    if (d.precision != PPF_PREC_FP64 && d.precision != PPF_PREC_FP32_IR) return PPF_ERR_INVALID_ARG;
    // every n must be even (01_Fortran/examples/global.f90:114-120)
    // z rows live in LDS in the z pass (2 doubles per row): n3m <= 4096
    if (!size_supported(d.n1m) || !size_supported(d.n2m) || d.n3m < 2 || (d.n3m & 1) || d.n3m > 4096)
        return PPF_ERR_SIZE;

    ppf_plan *p = new (std::nothrow) ppf_plan();
    if (!p) return PPF_ERR_ALLOC;
    p->d = d;
    const bool px = d.pbc1 != 0, py = d.pbc2 != 0;
    p->xkind = px ? AX_R2C : AX_DCT;
    if (!py) { p->ykind = AX_DCT; p->ymode = 0; }
    else if (px) { p->ykind = AX_C2C; p->ymode = 1; }
    else { p->ykind = AX_R2C; p->ymode = 2; }
    p->Lx = px ? d.n1m + 2 : d.n1m;
    p->Ly = d.n2m + 2;
    p->Lz = d.n3m + 2;
    if (p->ymode == 2) { p->zjlo = 0; p->znjr = d.n2m + 2; }
    else { p->zjlo = 1; p->znjr = d.n2m; }
    p->mean_kx = 0;
    p->mean_jp = (p->ymode == 2) ? 0 : 1;
    p->zvariant = 0;
    if (const char *zv = std::getenv("PPF_ZVARIANT")) p->zvariant = std::atoi(zv);
    const double fx = px ? 1.0 : 2.0, fy = (p->ymode == 0) ? 2.0 : 1.0;
    p->mean_scale = 1.0 / (fx * fy * (double)d.n1m * (double)d.n2m * (double)d.n3m);

    int rc = PPF_OK;
    const bool fp32 = d.precision == PPF_PREC_FP32_IR;
    p->Lxs = w_stride(p->Lx, fp32 ? (int)sizeof(float) : (int)sizeof(double));
    const size_t wsz = (size_t)p->Lxs * p->Ly * p->Lz;
    // Thomas scratch only where the one-thread-per-line kernel can run: long
    // lines (n3m > 2048) or when it is selected explicitly
    const bool need_cs = d.n3m > 2048 || p->zvariant == 6;
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
    if ((rc = dev_alloc(p, &p->partial, 3 * (size_t)p->npartial))) { free_plan(p); return rc; }
    if (hipHostMalloc((void **)&p->hred, 3 * sizeof(double), hipHostMallocDefault) != hipSuccess) {
        p->hred = nullptr;
        free_plan(p);
        return PPF_ERR_ALLOC;
    }

    // eigenvalue and zero-mode tables per scalar column / row
    const double h1 = d.L1 / d.n1m, h2 = d.L2 / d.n2m;
    std::vector<double> lx(p->Lx), ly(p->Ly, 0.0);
    std::vector<unsigned char> zx(p->Lx, 0), zy(p->Ly, 0);
    for (int kx = 0; kx < p->Lx; ++kx) {
        const int m = px ? kx / 2 : kx;
        lx[kx] = px ? eig_periodic(m, d.n1m, h1) : eig_neumann(m, d.n1m, h1);
        zx[kx] = (m == 0);
    }
    for (int jp = 0; jp < p->Ly; ++jp) {
        int m;
        if (p->ymode == 0) m = jp - 1;
        else if (p->ymode == 1) { const int j = jp - 1; m = (j <= d.n2m / 2) ? j : d.n2m - j; }
        else m = jp / 2;
        if (m < 0 || m > d.n2m) { ly[jp] = 0.0; zy[jp] = 0; continue; }  // unused ghost rows
        if (p->ymode == 0 && jp > d.n2m) { ly[jp] = 0.0; zy[jp] = 0; continue; }
        if (p->ymode == 1 && (jp < 1 || jp > d.n2m)) { ly[jp] = 0.0; zy[jp] = 0; continue; }
        ly[jp] = (p->ymode == 0) ? eig_neumann(m, d.n2m, h2) : eig_periodic(m, d.n2m, h2);
        zy[jp] = (m == 0);
    }
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
    // slab pipeline for fields large enough to keep every stream busy
    // (PPF_SLABS overrides; 1 disables)
    {
        // measured at 512^3 on MI355X: 1 slab 2.29 ms, 4 slabs 2.39, 8 slabs 2.40, 16 slabs 2.53
        int ns = 1;
        if (const char *e = std::getenv("PPF_SLABS")) ns = std::atoi(e);
        ns = std::max(1, std::min(ns, std::min(kMaxSlabs, d.n3m / 2)));
        if (ns > 1) {
            if (hipStreamCreateWithFlags(&p->s2, hipStreamNonBlocking) != hipSuccess) {
                p->s2 = nullptr;
                free_plan(p);
                return PPF_ERR_HIP;
            }
            for (hipEvent_t &e : p->pev)
                if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
                    e = nullptr;
                    free_plan(p);
                    return PPF_ERR_HIP;
                }
        }
        p->nslab = ns;
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

// The five passes with the slab pipeline: slab g of x-forward (stream s),
// then slab g of y-forward (stream s2, after x-forward's slab g); z on s
// after both; then slab g of y-inverse (s) and slab g of x-inverse (s2).
// Slabs are whole z-planes, so every pass of a slab only touches its planes.
template <typename T>
static int run_passes_slabs(ppf_plan *p, T *W, const AxisTables<T> &tx, const AxisTables<T> &ty, const double *in,
                            double *P, bool acc, const double *dmx1, const double *dmx2, const double *dmx3,
                            const double *dx3, hipStream_t s) {
    const int ns = p->nslab, n3m = p->d.n3m;
    const size_t wplane = (size_t)p->Lxs * p->Ly, pplane = (size_t)(p->d.n1m + 2) * p->Ly;
    const size_t fplane = (size_t)p->d.n1m * p->d.n2m;
    hipEvent_t *e = p->pev;
    int rc;
    // s2 starts after everything already queued on s
    PPF_HIP(hipEventRecord(e[0], s));
    PPF_HIP(hipStreamWaitEvent(p->s2, e[0], 0));
    for (int g = 0; g < ns; ++g) {
        const int k0 = (int)((long)g * n3m / ns), k1 = (int)((long)(g + 1) * n3m / ns);
        if ((rc = dispatch_xfwd<T>(p, in + (size_t)k0 * fplane, W + (size_t)k0 * wplane, tx, nullptr, s, nullptr,
                                   k1 - k0, g == 0 ? z_rows(p, dx3, dmx3) : ZRows{})))
            return rc;
        PPF_HIP(hipEventRecord(e[1 + g], s));
        PPF_HIP(hipStreamWaitEvent(p->s2, e[1 + g], 0));
        if ((rc = dispatch_y<T>(p, W + (size_t)k0 * wplane, ty, false, dmx2, p->s2, k1 - k0))) return rc;
    }
    PPF_HIP(hipEventRecord(e[1 + ns], p->s2));
    PPF_HIP(hipStreamWaitEvent(s, e[1 + ns], 0));
    if ((rc = launch_z<T>(p, W, dx3, dmx3, s))) return rc;
    PPF_HIP(hipEventRecord(e[2 + ns], s));
    PPF_HIP(hipStreamWaitEvent(p->s2, e[2 + ns], 0));
    const int Lz = p->Lz;
    for (int g = 0; g < ns; ++g) {
        const int q0 = (int)((long)g * Lz / ns), q1 = (int)((long)(g + 1) * Lz / ns);
        if ((rc = dispatch_y<T>(p, W + (size_t)q0 * wplane, ty, true, dmx2, s, q1 - q0))) return rc;
        PPF_HIP(hipEventRecord(e[3 + ns + g], s));
        PPF_HIP(hipStreamWaitEvent(p->s2, e[3 + ns + g], 0));
        if ((rc = dispatch_xinv<T>(p, W + (size_t)q0 * wplane, P + (size_t)q0 * pplane, tx, dmx1, acc, p->s2,
                                   q1 - q0)))
            return rc;
    }
    PPF_HIP(hipEventRecord(e[3 + 2 * ns], p->s2));
    PPF_HIP(hipStreamWaitEvent(s, e[3 + 2 * ns], 0));
    PPF_HIP(hipGetLastError());
    return PPF_OK;
}

// The five passes of one solve of the plan's precision.  xr != nullptr: pass 1
// transforms the residual f - L P (refinement form); acc: pass 5 adds to P.
// ev (7 events or nullptr) brackets the passes for ppf_solve_timed.
template <typename T>
static int run_passes(ppf_plan *p, T *W, const AxisTables<T> &tx, const AxisTables<T> &ty, const double *in,
                      const XRes *xr, double *P, bool acc, const double *dmx1, const double *dmx2, const double *dmx3,
                      const double *dx3, hipStream_t s, hipEvent_t *ev, unsigned *res_grid) {
    int rc;
    if (p->nslab > 1 && !ev && !xr) return run_passes_slabs<T>(p, W, tx, ty, in, P, acc, dmx1, dmx2, dmx3, dx3, s);
    if (ev) PPF_HIP(hipEventRecord(ev[0], s));
    // pass 1 also fills the z row table of this solve (not in the residual
    // form, which reads it: the table of the solve that made the iterate)
    if ((rc = dispatch_xfwd<T>(p, in, W, tx, xr, s, res_grid, -1, xr ? ZRows{} : z_rows(p, dx3, dmx3)))) return rc;
    if (ev) PPF_HIP(hipEventRecord(ev[1], s));
    if ((rc = dispatch_y<T>(p, W, ty, false, dmx2, s))) return rc;
    if (ev) PPF_HIP(hipEventRecord(ev[2], s));
    if ((rc = launch_z<T>(p, W, dx3, dmx3, s))) return rc;
    if (ev) PPF_HIP(hipEventRecord(ev[3], s));
    if ((rc = dispatch_y<T>(p, W, ty, true, dmx2, s))) return rc;
    if (ev) PPF_HIP(hipEventRecord(ev[4], s));
    if ((rc = dispatch_xinv<T>(p, W, P, tx, dmx1, acc, s))) return rc;
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

// FP32 transforms and Thomas storage with FP64 refinement: x_0 = solve32(f);
// then r = f - L x (FP64, fused into pass 1), x += solve32(r) (fused into
// pass 5) until ||r||/||f|| <= tol or max_sweeps corrections.  The z pass
// computes in FP64 on FP32 data, so the pinned (0,0) system and the mean stay
// FP64.  tol < 0: exactly max_sweeps corrections, no convergence test.
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

// fixed >= 0 overrides the plan: exactly `fixed` corrections, no test.
// The ghost layers are accumulated with each correction, which carries the
// FP32 rounding of the first solve's ghosts along (the residual never sees
// them), so they are re-derived from the refined interior at the end.
static int solve_fp32_ir(ppf_plan *p, const double *f, double *P, const double *dmx1, const double *dmx2,
                         const double *dmx3, const double *dx3, hipStream_t s, ppf_stats *stats, int fixed = -1) {
    const double tol = fixed >= 0 ? -1.0 : (p->d.tol == 0.0 ? 1e-10 : p->d.tol);
    const int maxs = fixed >= 0 ? fixed : (p->d.max_sweeps > 0 ? p->d.max_sweeps : 8);
    int rc;
    if ((rc = run_passes<float>(p, p->W32, p->tx32, p->ty32, f, nullptr, P, false, dmx1, dmx2, dmx3, dx3, s, nullptr,
                                nullptr)))
        return rc;
    const XRes xr = make_xres(p, P);
    int sweeps = 0;
    double rel = -1.0;
    bool converged = false;
    if (tol < 0) {
        for (; sweeps < maxs; ++sweeps) {
            unsigned g = 0;
            if ((rc = dispatch_xfwd<float>(p, f, p->W32, p->tx32, &xr, s, &g))) return rc;
            hipLaunchKernelGGL(k_reduce3, dim3(1), dim3(256), 0, s, p->partial, (int)g, p->red);
            if ((rc = dispatch_y<float>(p, p->W32, p->ty32, false, dmx2, s)) ||
                (rc = launch_z<float>(p, p->W32, dx3, dmx3, s)) ||
                (rc = dispatch_y<float>(p, p->W32, p->ty32, true, dmx2, s)) ||
                (rc = dispatch_xinv<float>(p, p->W32, P, p->tx32, dmx1, true, s)))
                return rc;
            PPF_HIP(hipGetLastError());
        }
        if (sweeps > 0 && (rc = refresh_ghosts(p, P, dmx1, dmx2, dmx3, s))) return rc;
        if (stats) {
            if ((rc = ppf_residual(p, f, P, dmx3, dx3, s, &rel))) return rc;
        }
        converged = true;
    } else {
        // The residual form of pass 1 computes ||r|| and transforms r for the
        // next sweep; after a sweep the plain residual kernel (about half its
        // time) checks convergence, so a converged solve does not pay for a
        // transform it discards.  Both evaluate the same FP64 operator.
        for (;;) {
            // residual of the current iterate, transformed by pass 1
            unsigned g = 0;
            if ((rc = dispatch_xfwd<float>(p, f, p->W32, p->tx32, &xr, s, &g))) return rc;
            hipLaunchKernelGGL(k_reduce3, dim3(1), dim3(256), 0, s, p->partial, (int)g, p->red);
            PPF_HIP(hipMemcpyAsync(p->hred, p->red, 3 * sizeof(double), hipMemcpyDeviceToHost, s));
            PPF_HIP(hipStreamSynchronize(s));
            rel = p->hred[1] > 0 ? std::sqrt(p->hred[0] / p->hred[1]) : std::sqrt(p->hred[0]);
            if (rel <= tol) { converged = true; break; }
            if (sweeps == maxs || !(rel == rel)) break;
            if ((rc = dispatch_y<float>(p, p->W32, p->ty32, false, dmx2, s)) ||
                (rc = launch_z<float>(p, p->W32, dx3, dmx3, s)) ||
                (rc = dispatch_y<float>(p, p->W32, p->ty32, true, dmx2, s)) ||
                (rc = dispatch_xinv<float>(p, p->W32, P, p->tx32, dmx1, true, s)))
                return rc;
            PPF_HIP(hipGetLastError());
            ++sweeps;
            if ((rc = ppf_residual(p, f, P, dmx3, dx3, s, &rel))) return rc;
            if (rel <= tol) { converged = true; break; }
            if (sweeps == maxs || !(rel == rel)) break;
        }
        if (sweeps > 0 && (rc = refresh_ghosts(p, P, dmx1, dmx2, dmx3, s))) return rc;
    }
    if (stats) {
        stats->sweeps = sweeps;
        stats->rel_residual = rel;
    }
    return converged ? PPF_OK : PPF_ERR_NOT_CONVERGED;
}

extern "C" {

int ppf_solve(ppf_plan *p, const double *prhs, double *P, const double *dmx1, const double *dmx2, const double *dmx3,
              const double *dx3, void *stream, ppf_stats *stats) {
    if (!p || !prhs || !P || !dmx1 || !dmx2 || !dmx3 || !dx3) return PPF_ERR_INVALID_ARG;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (p->d.precision == PPF_PREC_FP32_IR) return solve_fp32_ir(p, prhs, P, dmx1, dmx2, dmx3, dx3, s, stats);
    int rc;
    if ((rc = run_passes<double>(p, p->W, p->tx64, p->ty64, prhs, nullptr, P, false, dmx1, dmx2, dmx3, dx3, s,
                                 nullptr, nullptr)))
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
    return run_passes<double>(p, p->W, p->tx64, p->ty64, prhs, nullptr, P, false, dmx1, dmx2, dmx3, dx3, s, nullptr,
                              nullptr);
}

int ppf_plan_tune(ppf_plan *p, const double *prhs, double *P, const double *dmx1, const double *dmx2,
                  const double *dmx3, const double *dx3, void *stream, int *zvariant) {
    if (!p || !prhs || !P || !dmx1 || !dmx2 || !dmx3 || !dx3) return PPF_ERR_INVALID_ARG;
    // candidates: partitioned (K,S) shapes and the pipelined wavefront-PCR form
    static const int cand[] = {3, 1, 2, 4, 5, 7, 8, 12, 13, 14};
    const int keep = p->zvariant;
    int best = 0;
    float best_ms = 0.f;
    for (int v : cand) {
        p->zvariant = v;
        float st[5];
        float z = 0.f;
        bool ok = true;
        for (int rep = 0; rep < 3 && ok; ++rep) {
            const int rc = ppf_solve_timed(p, prhs, P, dmx1, dmx2, dmx3, dx3, stream, st);
            if (rc == PPF_ERR_UNSUPPORTED) ok = false;
            else if (rc) { p->zvariant = keep; return rc; }
            else if (rep > 0) z += st[2];
        }
        if (ok && (best == 0 || z < best_ms)) { best = v; best_ms = z; }
    }
    p->zvariant = best ? best : keep;
    if (zvariant) *zvariant = p->zvariant;
    return PPF_OK;
}

int ppf_plan_set_zvariant(ppf_plan *p, int zvariant) {
    if (!p || zvariant < 0 || zvariant > 14) return PPF_ERR_INVALID_ARG;
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
            rc = run_passes<float>(p, p->W32, p->tx32, p->ty32, prhs, nullptr, P, false, dmx1, dmx2, dmx3, dx3, s, ev,
                                   nullptr);
        else
            rc = run_passes<double>(p, p->W, p->tx64, p->ty64, prhs, nullptr, P, false, dmx1, dmx2, dmx3, dx3, s, ev,
                                    nullptr);
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
    const XRes xr = make_xres(p, P);
#ifndef PPF_RES_KZ
#define PPF_RES_KZ 32
#endif
#ifndef PPF_RES_WGS
#define PPF_RES_WGS 8
#endif
    constexpr int KZ = PPF_RES_KZ;  // planes per tile
    const long ntiles = (long)((p->d.n1m + 255) / 256) * p->d.n2m * ((n3m + KZ - 1) / KZ);
    const int g = (int)std::min<long>(std::min<long>(p->npartial, (long)PPF_RES_WGS * num_cus()), ntiles);
    hipLaunchKernelGGL(k_residual<KZ>, dim3(g), dim3(256), 0, s, f, xr, p->d.n1m, p->d.n2m, n3m);
    hipLaunchKernelGGL(k_reduce3, dim3(1), dim3(256), 0, s, p->partial, g, p->red);
    PPF_HIP(hipGetLastError());
    PPF_HIP(hipMemcpyAsync(p->hred, p->red, 3 * sizeof(double), hipMemcpyDeviceToHost, s));
    PPF_HIP(hipStreamSynchronize(s));
    *rel = p->hred[1] > 0 ? std::sqrt(p->hred[0] / p->hred[1]) : std::sqrt(p->hred[0]);
    return PPF_OK;
}

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
    if (!dst || !src || reps <= 0 || !ms || (count & 1)) return PPF_ERR_INVALID_ARG;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const size_t n2 = count / 2;
    const unsigned grid = 8u * (unsigned)num_cus();
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess) return PPF_ERR_HIP;
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return PPF_ERR_HIP;
    }
    int rc = PPF_OK;
    if (hipEventRecord(e0, s) != hipSuccess) rc = PPF_ERR_HIP;
    for (int r = 0; r < reps && rc == PPF_OK; ++r) {
        hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, s, reinterpret_cast<double2 *>(dst),
                           reinterpret_cast<const double2 *>(src), n2);
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
