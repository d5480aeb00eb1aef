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

}  // namespace

extern "C" {

int ppf_tdma_plan_many_create(int nx_sys, int ny_sys, int nz_row, ppf_tdma_plan **plan) {
    if (!plan) return PPF_ERR_INVALID_ARG;
    *plan = nullptr;
    if (nx_sys <= 0 || ny_sys <= 0 || nz_row <= 0) return PPF_ERR_INVALID_ARG;
    ppf_tdma_plan *p = new (std::nothrow) ppf_tdma_plan{nx_sys, ny_sys, nz_row};
    if (!p) return PPF_ERR_ALLOC;
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

int ppf_tdma_plan_many_destroy(ppf_tdma_plan *p) {
    if (!p) return PPF_ERR_INVALID_ARG;
    delete p;
    return PPF_OK;
}

}  // extern "C"
