// ppf_aux.hpp -- small non-template kernels of the solve path (z row table,
// reductions, the device-side refinement test, ghost faces, the copy
// yardstick), with internal linkage so several translation units may include
// them.
#pragma once

#include "ppf_kernels.hpp"

namespace ppf {

// z row coefficients for one solve as a table (read by the residual forms and
// the pipelined z kernel): zam[k], zap[k] (k < n3m), then zap[n3m..n3m+3] =
// the wall extrapolation coefficients.
static __global__ void __launch_bounds__(256)
k_zcoef(const double *__restrict__ dx3, const double *__restrict__ dmx3, int n3m, double *__restrict__ zam,
        double *__restrict__ zap) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n3m) z_row(dx3, dmx3, n3m, k, zam[k], zap[k]);
    if (k == 0) z_wall_bc(dmx3, n3m, zap + n3m);
}

// Totals of the per-workgroup triples (sum r^2, sum f^2, sum x) in a fixed
// order (deterministic).  With check set it is also the device-side
// refinement test (below; check.gate is the gate to test and close, the same
// one that gates this kernel), which saves a launch per test.
struct GateTest {
    int *gate = nullptr;
    double tol = 0.0;
    int count = 0;
    double *res_out = nullptr;
};
static __global__ void __launch_bounds__(256) k_reduce3(const double *__restrict__ partial, int n, double *__restrict__ out,
                                                 const int *__restrict__ gate, GateTest check = GateTest{}) {
    if (gated_off(gate)) return;
    __shared__ double red[256];
    double a = 0.0, b = 0.0, c = 0.0;
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
        a += partial[3 * q];
        b += partial[3 * q + 1];
        c += partial[3 * q + 2];
    }
    a = block_sum<256>(a, red);
    b = block_sum<256>(b, red);
    c = block_sum<256>(c, red);
    if (threadIdx.x == 0) {
        out[0] = a;
        out[1] = b;
        out[2] = c;
        if (check.gate && check.gate[0] != 0) {
            const double rel = b > 0 ? sqrt(a / b) : sqrt(a);
            *check.res_out = rel;
            check.gate[1] += check.count;
            check.gate[0] = (rel <= check.tol) ? 0 : 1;
        }
    }
}

// Device-side refinement test.  k_gate_init opens the gate (gate[0] = 1: more
// sweeps may run) and sets the sweep count gate[1]; the test (k_reduce3 with
// a GateTest, while the gate is open) stores ||r||/||f|| of the totals it
// just reduced in res_out, counts the sweep that produced them (count = 1
// after a gated sweep) and closes the gate once the residual is <= tol (a
// NaN residual keeps it open, so the sweeps run out and the caller sees the
// miss).
static __global__ void k_gate_init(int *__restrict__ gate, int sweeps) {
    gate[0] = 1;
    gate[1] = sweeps;
}

// Ghost layers of P from its interior, one face pair per launch:
// wall extrapolation p0 = BC_a p1 - BC_b p2 (cuda_neumann_BC /
// apply_bc_on_direction, fft_3d_poisson.f90:798-852) or periodic wrap
// (cuda_ghostcell_update on one rank, :666-796), over the full (ghosted)
// ranges of the other two axes.  The reference's order -- walls along x, y,
// z, then wraps along x, y, z -- is the caller's launch order.
static __global__ void __launch_bounds__(256)
k_ghost_face(double *__restrict__ P, int n1m, int n2m, int n3m, int dir, int wrap, const double *__restrict__ dmx) {
    const size_t L1 = (size_t)n1m + 2, L2 = (size_t)n2m + 2, L3 = (size_t)n3m + 2;
    const size_t nb = dir == 0 ? L2 : L1, nc = dir == 2 ? L2 : L3;
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nb * nc) return;
    const size_t b = q % nb, c = q / nb;
    const int nm = dir == 0 ? n1m : (dir == 1 ? n2m : n3m);
    // element (b, c, position along dir)
    size_t base, step;
    if (dir == 0) { base = (c * L2 + b) * L1; step = 1; }
    else if (dir == 1) { base = c * L1 * L2 + b; step = L1; }
    else { base = c * L1 + b; step = L1 * L2; }
    double *x = P + base;
    if (wrap) {
        const double lo = x[1 * step], hi = x[(size_t)nm * step];
        x[0] = hi;
        x[(size_t)(nm + 1) * step] = lo;
    } else {
        const double d1 = dmx[1], d2 = dmx[2], s0 = d1 + d2;
        const double e1 = dmx[nm + 1], e2 = dmx[nm], s1 = e1 + e2;
        x[0] = (s0 * s0 / (s0 * s0 - d1 * d1)) * x[step] - (d1 * d1 / (s0 * s0 - d1 * d1)) * x[2 * step];
        x[(size_t)(nm + 1) * step] = (s1 * s1 / (s1 * s1 - e1 * e1)) * x[(size_t)nm * step] -
                                     (e1 * e1 / (s1 * s1 - e1 * e1)) * x[(size_t)(nm - 1) * step];
    }
}

// Streaming copy, the measured-bandwidth yardstick for the pass kernels and
// the calibration of HBM byte counters: one nontemporal 16-byte load and
// store per thread, one workgroup per 4 KiB.  Of the copy shapes measured at
// 1 GiB on MI355X this one is the fastest (6.5 TB/s; a persistent grid-stride
// loop 4.8-6.1 TB/s, hipMemcpy 5.1 TB/s).
using D2v = double __attribute__((ext_vector_type(2)));
static __global__ void __launch_bounds__(256) k_copy(D2v *__restrict__ dst, const D2v *__restrict__ src, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) st_stream(&dst[i], ld_stream(&src[i]));
}

}  // namespace ppf
