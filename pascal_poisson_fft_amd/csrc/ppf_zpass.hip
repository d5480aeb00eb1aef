// ppf_zpass.hip -- host launcher of pass 3 (the z Thomas solves) for FP64 and
// FP32 workspaces: the z-pass shapes (partitioned register Thomas k_zpart, its
// pipelined wavefront form k_zpipe, the one-thread-per-line k_zthomas) and the
// choice among them.
#include <type_traits>

#include "ppf_plan.hpp"

namespace ppf {

template <typename T>
int launch_z(ppf_plan *p, T *W, const double *dx3, const double *dmx3, hipStream_t s) {
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
    z.gate = p->gate;
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
    if (z.zper && ((v >= 7 && v <= 11) || v == 17)) v = (n3m <= 1024) ? 3 : 4;
    switch (v) {
        case 7: pipe(integral_constant<int, 16>{}, integral_constant<int, 16>{}, integral_constant<int, 2>{}); break;
        case 8: pipe(integral_constant<int, 16>{}, integral_constant<int, 8>{}, integral_constant<int, 2>{}); break;
        case 9: pipe(integral_constant<int, 32>{}, integral_constant<int, 8>{}, integral_constant<int, 2>{}); break;
        case 10: pipe(integral_constant<int, 32>{}, integral_constant<int, 16>{}, integral_constant<int, 2>{}); break;
        case 11: pipe(integral_constant<int, 8>{}, integral_constant<int, 16>{}, integral_constant<int, 4>{}); break;
        case 17: pipe(integral_constant<int, 24>{}, integral_constant<int, 16>{}, integral_constant<int, 2>{}); break;
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
        // 32 columns: 128-byte row segments for FP32 workspaces
        case 15: part(integral_constant<int, 32>{}, integral_constant<int, 32>{}, integral_constant<int, 512>{}, integral_constant<int, 2>{}); break;
        case 16: part(integral_constant<int, 16>{}, integral_constant<int, 32>{}, integral_constant<int, 512>{}, integral_constant<int, 3>{}); break;
        default: break;
    }
    // a pinned shape that does not fit the extent is reported, never replaced:
    // only the default (v chosen above) and shape 6 fall to one thread per line
    if (!done && p->zvariant != 0 && p->zvariant != 6) return PPF_ERR_UNSUPPORTED;
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


template int launch_z<double>(ppf_plan *, double *, const double *, const double *, hipStream_t);
template int launch_z<float>(ppf_plan *, float *, const double *, const double *, hipStream_t);

}  // namespace ppf
