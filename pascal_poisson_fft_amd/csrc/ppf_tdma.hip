// ppf_tdma.hip -- batched Thomas solver on explicit coefficient arrays.
//
// Replaces the one-rank branch of PaScaL_TDMA_many_solve_cuda
// (01_Fortran/PaScaL_TDMA/src/pascal_tdma_cuda.f90:230-269) which launches
// tdma_many_cuda (tdmas_cuda.f90:35-102).  One thread per system; system index
// fastest in memory, so every row access of a wavefront is one contiguous
// 512-byte segment.  The recurrence keeps the previous row in registers
// instead of the reference's shared-memory pipeline copies, and the
// arithmetic is evaluated without contraction in the reference's order
// (d/b, c/b; r = 1/(b - a c'); d' = r (d - a d'); c' = r c; d -= c' d_next), so
// results are bitwise equal to the reference Thomas on the same inputs.
#include <hip/hip_runtime.h>

#include <new>

#include "../../include/pascal_poisson_fft.h"

struct ppf_tdma_plan {
    int nx, ny, nz;
    double *e;  // correction vector of the cyclic solve (nx*ny*nz)
};

namespace {

__global__ void __launch_bounds__(256)
k_tdma_many(const double *__restrict__ a, const double *__restrict__ b, double *__restrict__ c, double *__restrict__ d,
            long nsys, int nrow) {
#pragma clang fp contract(off)
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsys) return;
    double dp = d[s] / b[s];
    double cp = c[s] / b[s];
    d[s] = dp;
    c[s] = cp;
    for (int k = 1; k < nrow; ++k) {
        const size_t q = (size_t)k * (size_t)nsys + (size_t)s;
        const double ak = a[q];
        const double r = 1.0 / (b[q] - ak * cp);
        dp = r * (d[q] - ak * dp);
        cp = r * c[q];
        d[q] = dp;
        c[q] = cp;
    }
    for (int k = nrow - 2; k >= 0; --k) {
        const size_t q = (size_t)k * (size_t)nsys + (size_t)s;
        dp = d[q] - c[q] * dp;
        d[q] = dp;
    }
}

// Cyclic Thomas, op order of tdma_cycl_many_cuda (tdmas_cuda.f90:114-217):
// rows 1..n-1 are reduced for d' and the correction vector e (e_1 = -a_1,
// e_{n-1} = -c_{n-1}), x_0 follows from row 0, then x_k = d'_k + x_0 e_k.
// The reference keeps d'_1, e'_1 for row 0 in shared-memory registers after
// the backward loop; they are re-read here (same values, and defined for
// n = 2, where the reference's backward loop never sets them).
__global__ void __launch_bounds__(256)
k_tdma_cycl_many(const double *__restrict__ a, const double *__restrict__ b, double *__restrict__ c,
                 double *__restrict__ d, double *__restrict__ e, long nsys, int nrow) {
#pragma clang fp contract(off)
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsys) return;
    const size_t ns = (size_t)nsys;
    for (int k = 0; k < nrow; ++k) e[(size_t)k * ns + s] = 0.0;
    e[ns + s] = -a[ns + s];
    e[(size_t)(nrow - 1) * ns + s] = -c[(size_t)(nrow - 1) * ns + s];
    double d1 = d[ns + s], b1 = b[ns + s], c1 = c[ns + s], e1 = e[ns + s];
    d1 = d1 / b1;
    c1 = c1 / b1;
    e1 = e1 / b1;
    d[ns + s] = d1;
    c[ns + s] = c1;
    e[ns + s] = e1;
    for (int k = 2; k < nrow; ++k) {
        const size_t q = (size_t)k * ns + s;
        const double c0 = c1, d0 = d1, e0 = e1;
        const double ak = a[q];
        b1 = b[q];
        c1 = c[q];
        d1 = d[q];
        e1 = e[q];
        const double r = 1.0 / (b1 - ak * c0);
        d1 = r * (d1 - ak * d0);
        e1 = r * (e1 - ak * e0);
        c1 = r * c1;
        e[q] = e1;
        d[q] = d1;
        c[q] = c1;
    }
    for (int k = nrow - 2; k >= 1; --k) {
        const size_t q = (size_t)k * ns + s;
        double c0 = c[q], d0 = d[q], e0 = e[q];
        d0 = d0 - c0 * d1;
        e0 = e0 - c0 * e1;
        d1 = d0;
        e1 = e0;
        d[q] = d0;
        e[q] = e0;
    }
    const size_t qn = (size_t)(nrow - 1) * ns + s;
    const double a0 = a[s], b0 = b[s], c0 = c[s];
    double x0 = d[s];
    x0 = (x0 - a0 * d[qn] - c0 * d[ns + s]) / (b0 + a0 * e[qn] + c0 * e[ns + s]);
    d[s] = x0;
    for (int k = 1; k < nrow; ++k) {
        const size_t q = (size_t)k * ns + s;
        d[q] = d[q] + x0 * e[q];
    }
}

}  // namespace

extern "C" {

int ppf_tdma_plan_many_create(int nx_sys, int ny_sys, int nz_row, ppf_tdma_plan **plan) {
    if (!plan) return PPF_ERR_INVALID_ARG;
    *plan = nullptr;
    if (nx_sys <= 0 || ny_sys <= 0 || nz_row <= 0) return PPF_ERR_INVALID_ARG;
    ppf_tdma_plan *p = new (std::nothrow) ppf_tdma_plan{nx_sys, ny_sys, nz_row, nullptr};
    if (!p) return PPF_ERR_ALLOC;
    // the cyclic solve's correction vector, as the reference plan's e_buff
    // for one rank (pascal_tdma_cuda.f90:108-201)
    if (hipMalloc(&p->e, sizeof(double) * (size_t)nx_sys * ny_sys * nz_row) != hipSuccess) {
        delete p;
        return PPF_ERR_ALLOC;
    }
    *plan = p;
    return PPF_OK;
}

int ppf_tdma_many_solve(ppf_tdma_plan *p, const double *a, const double *b, double *c, double *d, void *stream) {
    if (!p || !a || !b || !c || !d) return PPF_ERR_INVALID_ARG;
    const long nsys = (long)p->nx * p->ny;
    const int nt = 256;
    hipLaunchKernelGGL(k_tdma_many, dim3((unsigned)((nsys + nt - 1) / nt)), dim3(nt), 0,
                       static_cast<hipStream_t>(stream), a, b, c, d, nsys, p->nz);
    return hipGetLastError() == hipSuccess ? PPF_OK : PPF_ERR_HIP;
}

int ppf_tdma_many_solve_cycle(ppf_tdma_plan *p, const double *a, const double *b, double *c, double *d,
                              void *stream) {
    if (!p || !a || !b || !c || !d) return PPF_ERR_INVALID_ARG;
    if (p->nz < 2) return PPF_ERR_SIZE;
    const long nsys = (long)p->nx * p->ny;
    const int nt = 256;
    hipLaunchKernelGGL(k_tdma_cycl_many, dim3((unsigned)((nsys + nt - 1) / nt)), dim3(nt), 0,
                       static_cast<hipStream_t>(stream), a, b, c, d, p->e, nsys, p->nz);
    return hipGetLastError() == hipSuccess ? PPF_OK : PPF_ERR_HIP;
}

int ppf_tdma_plan_many_destroy(ppf_tdma_plan *p) {
    if (!p) return PPF_ERR_INVALID_ARG;
    if (p->e) (void)hipFree(p->e);
    delete p;
    return PPF_OK;
}

}  // extern "C"
