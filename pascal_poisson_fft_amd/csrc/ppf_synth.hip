// ppf_synth.hip -- synthetic right-hand sides for measurement (no counterpart
// in the reference): the discrete divergence of a staggered velocity field,
// as the pressure step of a projection method would hand to the solver.
//
// Velocity components live on cell faces: u on x-faces i (periodic in x), v on
// y-faces j (zero on the two walls j = 0, n2m), w on z-faces k (zero on the
// walls k = 0, n3m).  Each component is a sum of nmodes separable modes
// (per-mode 1-D factors supplied by the caller, amplitudes folded in) plus
// hashed uniform face noise.  The divergence
//   f = (u_{i+1} - u_i)/h1 + (v_{j+1} - v_j)/h2 + (w_{k+1} - w_k)/dx3_k
// sums to zero against the cell volumes h1 h2 dx3_k (telescoping with periodic
// x and zero wall fluxes), so f is compatible with the zero-flux operator by
// construction.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "../../include/pascal_poisson_fft.h"

namespace {

// splitmix64 finaliser: counter-based, stateless noise in [-1, 1)
__device__ __forceinline__ double hash_unit(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    x ^= x >> 31;
    return (double)(x >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

__device__ __forceinline__ uint64_t face_key(uint64_t seed, int comp, int i, int j, int k) {
    return seed * 0x100000001B3ull ^ ((uint64_t)comp << 62) ^ ((uint64_t)(uint32_t)k << 40) ^
           ((uint64_t)(uint32_t)j << 20) ^ (uint64_t)(uint32_t)i;
}

// tables: 9 blocks of nmodes x stride doubles, in this order:
//   Fu_x (x-faces, n1m, periodic) Cu_y (n2m) Cu_z (n3m)
//   Cv_x (n1m) Fv_y (y-faces, n2m+1, zero at 0 and n2m) Cv_z (n3m)
//   Cw_x (n1m) Cw_y (n2m) Fw_z (z-faces, n3m+1, zero at 0 and n3m)
__global__ void __launch_bounds__(256)
k_synth_div(double *__restrict__ f, int n1m, int n2m, int n3m, double ih1, double ih2, const double *__restrict__ dx3,
            const double *__restrict__ tab, int nmodes, int stride, uint64_t seed, double noise) {
    const size_t N = (size_t)n1m * n2m * n3m;
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= N) return;
    const int i = (int)(q % n1m);
    const size_t rest = q / n1m;
    const int j = (int)(rest % n2m), k = (int)(rest / n2m);
    const int ip = i + 1 == n1m ? 0 : i + 1;
    const size_t blk = (size_t)nmodes * stride;
    const double *Fux = tab, *Cuy = tab + blk, *Cuz = tab + 2 * blk;
    const double *Cvx = tab + 3 * blk, *Fvy = tab + 4 * blk, *Cvz = tab + 5 * blk;
    const double *Cwx = tab + 6 * blk, *Cwy = tab + 7 * blk, *Fwz = tab + 8 * blk;
    double du = 0.0, dv = 0.0, dw = 0.0;
    for (int m = 0; m < nmodes; ++m) {
        const size_t o = (size_t)m * stride;
        du += (Fux[o + ip] - Fux[o + i]) * Cuy[o + j] * Cuz[o + k];
        dv += Cvx[o + i] * (Fvy[o + j + 1] - Fvy[o + j]) * Cvz[o + k];
        dw += Cwx[o + i] * Cwy[o + j] * (Fwz[o + k + 1] - Fwz[o + k]);
    }
    // face noise: u periodic in x, v and w zero on the walls
    const double nu0 = hash_unit(face_key(seed, 0, i, j, k)), nu1 = hash_unit(face_key(seed, 0, ip, j, k));
    const double nv0 = j == 0 ? 0.0 : hash_unit(face_key(seed, 1, i, j, k));
    const double nv1 = j + 1 == n2m ? 0.0 : hash_unit(face_key(seed, 1, i, j + 1, k));
    const double nw0 = k == 0 ? 0.0 : hash_unit(face_key(seed, 2, i, j, k));
    const double nw1 = k + 1 == n3m ? 0.0 : hash_unit(face_key(seed, 2, i, j, k + 1));
    du += noise * (nu1 - nu0);
    dv += noise * (nv1 - nv0);
    dw += noise * (nw1 - nw0);
    f[q] = du * ih1 + dv * ih2 + dw / dx3[k + 1];
}

}  // namespace

extern "C" int ppf_synth_divergence_rhs(double *f, int n1m, int n2m, int n3m, double L1, double L2,
                                        const double *dx3, const double *tables, int nmodes, int stride,
                                        unsigned long long seed, double noise, void *stream) {
    if (!f || !dx3 || (nmodes > 0 && !tables) || n1m <= 0 || n2m <= 0 || n3m <= 0 || nmodes < 0 ||
        stride < (n1m > n2m ? n1m : n2m) + 1 || stride < n3m + 1 || !(L1 > 0) || !(L2 > 0))
        return PPF_ERR_INVALID_ARG;
    const size_t N = (size_t)n1m * n2m * n3m;
    const unsigned grid = (unsigned)((N + 255) / 256);
    hipLaunchKernelGGL(k_synth_div, dim3(grid), dim3(256), 0, static_cast<hipStream_t>(stream), f, n1m, n2m, n3m,
                       (double)n1m / L1, (double)n2m / L2, dx3, tables, nmodes, stride, (uint64_t)seed, noise);
    return hipGetLastError() == hipSuccess ? PPF_OK : PPF_ERR_HIP;
}

// ---------------------------------------------------------------------------
// The reference example's right-hand side on the device
// (cuda_pressure_RHS_kernel, 01_Fortran/examples/cuda_pressure.f90:66-90):
// prhs(i,j,k) = -cos(x1(i) pi) cos(x2(j) pi) cos(x3(k) pi) 3 pi^2 at the grid
// nodes, in the reference's left-to-right product order.
// ---------------------------------------------------------------------------
namespace {
__global__ void __launch_bounds__(256)
k_example_rhs(double *__restrict__ f, const double *__restrict__ x1, const double *__restrict__ x2,
              const double *__restrict__ x3, int n1m, int n2m, int n3m) {
    const double pi = 3.14159265358979323846264338327950288;
    const size_t n = (size_t)n1m * n2m * n3m;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(q % n1m);
        const size_t r = q / n1m;
        const int j = (int)(r % n2m), k = (int)(r / n2m);
        // Fortran element i of x(0:) is x[i]; cells are 1-based
        const double v = cos(x1[i + 1] * pi) * cos(x2[j + 1] * pi) * cos(x3[k + 1] * pi) * 3.0 * pi * pi;
        f[q] = -v;
    }
}
}  // namespace

extern "C" int ppf_example_rhs(double *f, const double *x1, const double *x2, const double *x3, int n1m, int n2m,
                               int n3m, void *stream) {
    if (!f || !x1 || !x2 || !x3 || n1m <= 0 || n2m <= 0 || n3m <= 0) return PPF_ERR_INVALID_ARG;
    const size_t n = (size_t)n1m * n2m * n3m;
    const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_example_rhs, dim3(grid), dim3(256), 0, static_cast<hipStream_t>(stream), f, x1, x2, x3, n1m,
                       n2m, n3m);
    return hipGetLastError() == hipSuccess ? PPF_OK : PPF_ERR_HIP;
}
