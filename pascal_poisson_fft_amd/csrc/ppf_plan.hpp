// ppf_plan.hpp -- the plan object of the solve path and the host-side launch
// interface shared by the translation units that instantiate the passes
// (ppf_xfwd*.hip, ppf_xinv*.hip, ppf_yfwd*.hip, ppf_yinv*.hip, ppf_zpass.hip)
// and the one that sequences them (ppf_solver.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/pascal_poisson_fft.h"
#include "ppf_kernels.hpp"

namespace ppf {

// Transform kind along one axis.
enum AxisKind { AX_DCT = 0, AX_R2C = 1, AX_C2C = 2 };

template <typename T>
struct AxisTables {
    cx<T> *tw = nullptr;  // FFT twiddles, M entries
    cx<T> *wk = nullptr;  // exp(-2 pi i k / N), M+1 entries
    cx<T> *qk = nullptr;  // exp(-i pi k / 2N), M+1 entries
};

// Line lengths with compiled transform kernels (x and y extents): the even
// lengths 8..1024 of the form 2^a * {1, 3, 5, 7, 15}, in three groups that
// compile as separate translation units.
#define PPF_SIZES_A(X) X(8) X(10) X(12) X(14) X(16) X(20) X(24) X(28) X(30) X(32) X(40) X(48) X(56) X(60) X(64) X(80) X(96)
#define PPF_SIZES_B(X) X(112) X(120) X(128) X(160) X(192) X(224) X(240) X(256) X(320) X(384)
#define PPF_SIZES_C(X) X(448) X(480) X(512) X(640) X(768) X(896) X(960) X(1024)
#define PPF_FOR_SIZES(X) PPF_SIZES_A(X) PPF_SIZES_B(X) PPF_SIZES_C(X)

inline bool size_supported(int n) {
    switch (n) {
#define PPF_SIZE_CASE(v) case v:
        PPF_FOR_SIZES(PPF_SIZE_CASE)
#undef PPF_SIZE_CASE
        return true;
    }
    return false;
}

#define PPF_HIP(call)                                  \
    do {                                               \
        hipError_t e_ = (call);                        \
        if (e_ != hipSuccess) return PPF_ERR_HIP;      \
    } while (0)

inline int num_cus() {
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
unsigned persistent_grid(K kernel, int nthreads, int ntiles, size_t dyn_lds = 0) {
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, nthreads, dyn_lds) != hipSuccess || occ < 1) occ = 1;
    const long g = (long)occ * num_cus();
    return (unsigned)std::max(1L, std::min<long>(g, ntiles));
}

}  // namespace ppf

struct ppf_plan {
    ppf_desc d;
    int Lx, Ly, Lz;        // workspace extents (scalars)
    int Lxs;               // workspace scalars per row: Lx rounded up to 128 bytes (whole column tiles)
    int xkind, ykind;      // ppf::AxisKind
    int ymode;             // y pass MODE (0 DCT, 1 C2C, 2 R2C)
    int zjlo, znjr;        // rows solved by the z pass
    int mean_kx, mean_jp;
    double mean_scale;
    int zvariant;          // z-pass kernel shape (0 = chosen by n3m)
    int xvariant = 0;      // x-forward pass variant (kXVariants forms)
    int xvinv = 0;         // x-inverse pass variant (the tuner picks it apart)
    int yvariant = 0;      // y-pass threads per line: 0 one butterfly per stage, 1 two
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
    int *dgate = nullptr;       // device refinement test: [0] more sweeps needed, [1] sweeps run
    double *dres = nullptr;     // device refinement test: last relative residual
    const int *gate = nullptr;  // gate the launch helpers pass (nullptr: kernels always run)
    ppf::AxisTables<double> tx64, ty64;
    ppf::AxisTables<float> tx32, ty32;
    size_t bytes = 0;
};

namespace ppf {

constexpr double kPi = 3.14159265358979323846264338327950288;

// Host tables of one axis (uploaded by plans and by the fused batch): the
// per-stage FFT twiddles -- the stage of span LS and radix R at LS-1 + (r-1) LS
// + j -- the real-FFT split factors wk and the DCT quarter-wave factors qk.
template <typename T>
void host_axis_tables(int N, int kind, std::vector<cx<T>> &tw, std::vector<cx<T>> &wk, std::vector<cx<T>> &qk) {
    const int M = (kind == AX_C2C) ? N : N / 2;
    tw.assign(M, cx<T>{T(1), T(0)});
    wk.resize(M + 1);
    qk.resize(M + 1);
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
}

// modified wavenumbers, fft_3d_poisson.f90:288-351 (host for plans, device
// for the allocation-free solve: one formula)
__host__ __device__ inline double eig_neumann(int m, int n, double h) {
    const double s = sin(0.5 * (double)m * kPi / (double)n);
    return 4.0 * s * s / (h * h);
}
__host__ __device__ inline double eig_periodic(int m, int n, double h) {
    return 2.0 * (1.0 - cos(2.0 * (double)m * kPi / (double)n)) / (h * h);
}

// Eigenvalue and zero-mode flag of workspace scalar column kx (x) and row jp
// (y).  Periodic x stores the half spectrum as re/im pairs (column kx is mode
// kx/2); y rows follow the y pass: DCT rows 1..n2m are modes 0..n2m-1, C2C
// rows 1..n2m are the wrapped modes of the complex FFT, R2C rows 0..n2m+1 are
// re/im pairs of modes 0..n2m/2.  Unused ghost rows get 0.
struct EigGeom {
    int px, ymode, n1m, n2m;
    double h1, h2;
};
__host__ __device__ inline void eig_col(const EigGeom &g, int kx, double &lam, unsigned char &zero) {
    const int m = g.px ? kx / 2 : kx;
    lam = g.px ? eig_periodic(m, g.n1m, g.h1) : eig_neumann(m, g.n1m, g.h1);
    zero = (m == 0);
}
__host__ __device__ inline void eig_row(const EigGeom &g, int jp, double &lam, unsigned char &zero) {
    int m;
    lam = 0.0;
    zero = 0;
    if (g.ymode == 0) m = jp - 1;
    else if (g.ymode == 1) { const int j = jp - 1; m = (j <= g.n2m / 2) ? j : g.n2m - j; }
    else m = jp / 2;
    if (m < 0 || m > g.n2m) return;
    if (g.ymode == 0 && jp > g.n2m) return;
    if (g.ymode == 1 && (jp < 1 || jp > g.n2m)) return;
    lam = (g.ymode == 0) ? eig_neumann(m, g.n2m, g.h2) : eig_periodic(m, g.n2m, g.h2);
    zero = (m == 0);
}

// Workspace geometry, transform kinds and mean bookkeeping of a problem.
inline void plan_geometry(const ppf_desc &d, ppf_plan *p) {
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
    const double fx = px ? 1.0 : 2.0, fy = (p->ymode == 0) ? 2.0 : 1.0;
    p->mean_scale = 1.0 / (fx * fy * (double)d.n1m * (double)d.n2m * (double)d.n3m);
    const bool fp32 = d.precision == PPF_PREC_FP32_IR;
    p->Lxs = w_stride(p->Lx, fp32 ? (int)sizeof(float) : (int)sizeof(double));
}

inline EigGeom eig_geom(const ppf_plan *p) {
    return EigGeom{p->d.pbc1 != 0, p->ymode, p->d.n1m, p->d.n2m, p->d.L1 / p->d.n1m, p->d.L2 / p->d.n2m};
}

// z-pass shapes: 1-5 and 12-16 partitioned (k_zpart), 6 one thread per line
// (k_zthomas), 7-11 and 17 pipelined (k_zpipe); 0 = chosen by n3m.
constexpr int kZVariants = 17;
// x-pass variants (XGeom): 0 and 1 threads per line, 2 without the register
// prefetch, 3 the same kernel as 2 with one workgroup per tile (not
// persistent); y-pass variants (YGeom): 0 and 1.
constexpr int kXVariants = 4, kYVariants = 2;
// The x value of ppf_plan_variants / ppf_plan_set_variants: f when both x
// passes use form f, f + 16 i when the forward pass uses f and the inverse i.
inline int x_variant_code(int f, int i) { return f == i ? f : f + 16 * i; }
inline bool x_variant_decode(int v, int &f, int &i) {
    if (v >= 0 && v < kXVariants) { f = i = v; return true; }
    f = v % 16;
    i = v / 16;
    return v >= 16 && f < kXVariants && i < kXVariants && f != i;
}

// The z row table of a plan (filled by pass 1 when passed to dispatch_xfwd).
inline ZRows z_rows(const ppf_plan *p, const double *dx3, const double *dmx3) {
    return ZRows{dx3, dmx3, p->zcoef, p->zcoef + p->d.n3m, p->d.n3m};
}

// Pass launchers (T = double or float).  nplanes < 0: the whole field.
// res (FP32 plans only): the refinement form of pass 1, transforming the
// residual k_residual stored in the W rows (in unused).  zr.zam set:
// workgroup 0 of pass 1 also fills the z row table.  acc (FP32 plans only):
// pass 5 adds to P.
template <typename T>
int dispatch_xfwd(const ppf_plan *p, const double *in, T *W, const AxisTables<T> &tb, bool res, hipStream_t s,
                  int nplanes = -1, ZRows zr = ZRows{});
template <typename T>
int dispatch_xinv(const ppf_plan *p, const T *W, double *P, const AxisTables<T> &tb, const double *dmx1, bool acc,
                  hipStream_t s, int nplanes = -1);
template <typename T>
int dispatch_yfwd(const ppf_plan *p, T *W, const AxisTables<T> &tb, hipStream_t s, int nplanes = -1);
template <typename T>
int dispatch_yinv(const ppf_plan *p, T *W, const AxisTables<T> &tb, const double *dmx2, hipStream_t s,
                  int nplanes = -1);
template <typename T>
inline int dispatch_y(const ppf_plan *p, T *W, const AxisTables<T> &tb, bool inverse, const double *dmx2,
                      hipStream_t s, int nplanes = -1) {
    return inverse ? dispatch_yinv<T>(p, W, tb, dmx2, s, nplanes) : dispatch_yfwd<T>(p, W, tb, s, nplanes);
}
template <typename T>
int launch_z(ppf_plan *p, T *W, const double *dx3, const double *dmx3, hipStream_t s);

// The fused batch (ppf_fused.hpp / ppf_fused.hip): every FP64 problem of a
// batch in one persistent launch over a device work queue.
}  // namespace ppf
struct ppf_fused;
namespace ppf {
bool fused_supported(const ppf_desc &d);
int fused_create(int n, const ppf_desc *descs, const ppf_batch_io *io, ppf_fused **out);
int fused_launch(ppf_fused *f, hipStream_t s);
int fused_status(ppf_fused *f, hipStream_t s);  // synchronises; PPF_ERR_HIP if a dependency wait timed out
int fused_residual(ppf_fused *f, int i, hipStream_t s, double *rel);
void fused_destroy(ppf_fused *f);

// Launchers for one line length N, defined by the size-group translation
// units (ppf_xpass.inc, ppf_ypass.inc) and called by the dispatchers above.
template <typename T, int N>
int xfwd_n(const ppf_plan *p, const double *in, T *W, const AxisTables<T> &tb, bool res, const ZRows &zr,
           hipStream_t s, int nplanes);
template <typename T, int N>
int xinv_n(const ppf_plan *p, const T *W, double *P, const AxisTables<T> &tb, const double *dmx1, bool acc,
           hipStream_t s, int nplanes);
template <typename T, int N>
int y_n(const ppf_plan *p, T *W, const AxisTables<T> &tb, bool inverse, const double *dmx2, hipStream_t s,
        int nplanes);

}  // namespace ppf
