// ppf_plan.hpp -- the plan object of the solve path and the host-side launch
// interface shared by the translation units that instantiate the passes
// (ppf_xfwd*.hip, ppf_xinv*.hip, ppf_yfwd*.hip, ppf_yinv*.hip, ppf_zpass.hip)
// and the one that sequences them (ppf_solver.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>

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
    int Lxs;               // workspace row stride (Lx rounded up to 128 bytes)
    int xkind, ykind;      // ppf::AxisKind
    int ymode;             // y pass MODE (0 DCT, 1 C2C, 2 R2C)
    int zjlo, znjr;        // rows solved by the z pass
    int mean_kx, mean_jp;
    double mean_scale;
    int zvariant;          // z-pass kernel shape (0 = chosen by n3m)
    int xvariant = 0;      // x-pass threads per line: 0 one butterfly per stage, 1 two
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

// z-pass shapes: 1-5 and 12-16 partitioned (k_zpart), 6 one thread per line
// (k_zthomas), 7-11 and 17 pipelined (k_zpipe); 0 = chosen by n3m.
constexpr int kZVariants = 17;
// x- and y-pass variants (threads per line, see XGeom / YGeom): 0 and 1.
constexpr int kXYVariants = 2;

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
