/*
 * pascal_poisson_fft.h -- C ABI of the MI355X (gfx950) Poisson solve path.
 *
 * Drop-in for the GPU solve path of PaScaL_POISSON_FFT (Fortran module
 * `fft_poisson`, 01_Fortran/src/fft_3d_poisson.f90) and the one-rank branch of
 * the batched Thomas library it calls (PaScaL_TDMA_cuda,
 * 01_Fortran/PaScaL_TDMA/src/pascal_tdma_cuda.f90).  The Fortran module in
 * pascal_poisson_fft_amd/fortran/fft_poisson.f90 binds these entry points with
 * ISO_C_BINDING under the reference's own procedure names.
 *
 * Conventions
 *   - every array is x-fastest (Fortran column-major), FP64, in device memory;
 *   - all work is ordered on the given HIP stream (NULL = default stream) and
 *     returns without synchronising, except where a function says otherwise;
 *   - functions return PPF_OK (0) or a negative PPF_ERR_* code, never abort,
 *     never throw; ppf_error_string() names a code;
 *   - a plan may be used by one host thread at a time.
 */
#ifndef PASCAL_POISSON_FFT_H
#define PASCAL_POISSON_FFT_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PPF_OK 0
#define PPF_ERR_INVALID_ARG (-1)   /* null pointer, negative size, bad enum */
#define PPF_ERR_SIZE (-2)          /* an extent the transforms do not support */
#define PPF_ERR_HIP (-3)           /* a HIP runtime call failed */
#define PPF_ERR_ALLOC (-4)         /* device allocation failed */
#define PPF_ERR_UNSUPPORTED (-5)   /* a configuration not implemented */
#define PPF_ERR_NOT_CONVERGED (-6) /* refinement hit max_sweeps above tol */
#define PPF_ERR_SCRATCH (-7)       /* beyond the static scratch of ppf_solve_noalloc */

#define PPF_PREC_FP64 0    /* FP64 transforms and Thomas */
#define PPF_PREC_FP32_IR 1 /* FP32 transforms/Thomas + FP64 residual refinement */

typedef struct ppf_plan ppf_plan;

/* Problem description.  Replaces the scalar arguments of
 * fft_poisson_plan_cuda_create (fft_3d_poisson.f90:70-108): the reference
 * passes n = nm+1 and derives nm = n-1 (:103-104); here the interior counts nm
 * are given directly.  Extents must be even (01_Fortran/examples/
 * global.f90:114-120): 8 <= nm <= 1024 in x and y with nm = 2^a times 1, 3,
 * 5, 7 or 15 (the lengths with compiled line transforms), 2 <= nm <= 4096 in
 * z. */
typedef struct ppf_desc {
    int n1m, n2m, n3m;    /* interior cells per axis */
    double L1, L2, L3;    /* domain lengths (x/y spacing L/nm is uniform) */
    int pbc1, pbc2, pbc3; /* nonzero: periodic; zero: Neumann walls */
    int precision;        /* PPF_PREC_FP64 or PPF_PREC_FP32_IR */
    double tol;           /* FP32_IR: target ||f-Lp||/||f|| (0: 1e-10; < 0:
                             run exactly max_sweeps corrections, no test) */
    int max_sweeps;       /* FP32_IR: cap on correction sweeps (<=0: 8) */
} ppf_desc;

typedef struct ppf_stats {
    int sweeps;          /* FP32_IR refinement sweeps taken (0 for FP64) */
    double rel_residual; /* final ||f-Lp||/||f|| when measured, else -1 */
} ppf_stats;

/* Plan + buffers + tables for one problem shape.  Replaces
 * fft_poisson_plan_cuda_create + cuda_Poisson_FFT_initial
 * (fft_3d_poisson.f90:70-121: BC type, buffers, cuFFT plans, wavenumbers,
 * PaScaL_TDMA plans).  All device memory the solve needs is allocated here. */
int ppf_plan_create(const ppf_desc *desc, ppf_plan **plan);

/* One solve.  Replaces cuda_Poisson_FFT_1D(PRHS_d, P_d, dmx1_d, dmx2_d,
 * dmx3_d, dx3_d) (fft_3d_poisson.f90:353-626).
 *   prhs  (n1m, n2m, n3m)              right-hand side, NOT modified (the
 *                                      reference uses it as scratch)
 *   p     (n1m+2, n2m+2, n3m+2)        solution with the reference's ghost
 *                                      layers: zero unweighted interior mean
 *                                      (:628-664), wall extrapolation
 *                                      (:798-852), periodic wrap (:666-796)
 *   dmx1, dmx2, dmx3, dx3 (nm+2 each)  grid metrics, element q = the
 *                                      reference's Fortran element q (0:nsub)
 *   stream  hipStream_t or NULL
 *   stats   may be NULL
 * FP64 plans run asynchronously.  FP32_IR plans: an FP32 solve, then
 * refinement sweeps x += solve32(f - L x) with the residual and the update in
 * FP64.  With tol >= 0 they synchronise once per sweep to test
 * ||f - L x||/||f|| <= tol (stats gets the sweeps and that residual;
 * PPF_ERR_NOT_CONVERGED after max_sweeps, p then holds the last iterate); with
 * tol < 0 they run exactly max_sweeps sweeps without synchronising (and
 * measure the residual only when stats is given). */
int ppf_solve(ppf_plan *plan, const double *prhs, double *p, const double *dmx1, const double *dmx2,
              const double *dmx3, const double *dx3, void *stream, ppf_stats *stats);

/* ppf_solve with HIP events recorded on `stream` around each pass; on return
 * (the stream is synchronised) stage_ms[0..4] hold the durations of the
 * x-forward, y-forward, z-Thomas, y-inverse and x-inverse passes (for an
 * FP32_IR plan: of one FP32 solve, no refinement).  The
 * counterpart of the reference's NVTX ranges Poisson-F_x, -F_y, -TDMA-z,
 * -B_y, -B_x (fft_3d_poisson.f90:377-624), for measurement. */
int ppf_solve_timed(ppf_plan *plan, const double *prhs, double *p, const double *dmx1, const double *dmx2,
                    const double *dmx3, const double *dx3, void *stream, float stage_ms[5]);

/* ppf_solve without a plan and without allocating: for callers that cannot
 * call hipMalloc (code under hipGraph capture, tight serving loops).  No
 * counterpart in the reference.  FP64 descriptors only (PPF_ERR_UNSUPPORTED
 * otherwise); same arguments, result and ghost layers as ppf_solve with a plan
 * of *desc.  The workspace, tables and scalars live in a static device array
 * of 64 MiB that each GPU holds once for the life of the process; a problem
 * whose workspace ((n1m or n1m+2 rounded up to 16) x (n2m+2) x (n3m+2)
 * doubles) and tables exceed it returns PPF_ERR_SCRATCH (e.g. 128 x 128 x 256
 * fits, 256 x 256 x 128 does not).  Only kernel launches on `stream`: the axis
 * tables are rebuilt on the device each call.  Calls on one GPU share the
 * array, so calls that overlap on one GPU (other streams, other threads)
 * race; serialise them. */
int ppf_solve_noalloc(const ppf_desc *desc, const double *prhs, double *p, const double *dmx1, const double *dmx2,
                      const double *dmx3, const double *dx3, void *stream);

/* Frees everything the plan owns.  Replaces cuda_Poisson_FFT_clean
 * (fft_3d_poisson.f90:123-130). */
int ppf_plan_destroy(ppf_plan *plan);

/* ppf_solve with exactly `sweeps` refinement corrections and no convergence
 * test (FP32_IR plans; FP64 plans need sweeps = 0).  Only kernel launches on
 * `stream`: safe inside hipStreamBeginCapture.  No counterpart in the
 * reference (its solve is FP64 only); used by the batched solve's graph. */
int ppf_solve_sweeps(ppf_plan *plan, const double *prhs, double *p, const double *dmx1, const double *dmx2,
                     const double *dmx3, const double *dx3, int sweeps, void *stream);

/* A solve that never synchronises the host, for callers that capture it in a
 * hipGraph or replay it on new right-hand sides.  FP32_IR plans: the FP32
 * solve, min_sweeps corrections, then an FP64 residual test on the device
 * before each of up to extra_sweeps further corrections (extra_sweeps < 0:
 * up to max_sweeps corrections in all); once the relative residual is <= tol
 * every later correction and test is a no-op, though still a kernel launch
 * (tol < 0: all of them run).  Ghost layers are re-derived at the end.
 * Only kernel launches on `stream` (no memset, no copy).  FP64 plans: a plain
 * solve.  No counterpart in the reference (its solve is FP64 only and
 * synchronous, fft_3d_poisson.f90:353-626); used by the batched solve. */
int ppf_solve_async(ppf_plan *plan, const double *prhs, double *p, const double *dmx1, const double *dmx2,
                    const double *dmx3, const double *dx3, int min_sweeps, int extra_sweeps, void *stream);

/* After a ppf_solve_async on `stream`: synchronises the stream and reports
 * the corrections it ran and the last relative residual its device-side test
 * measured.  Returns PPF_ERR_NOT_CONVERGED when that residual is above the
 * plan's tol (tol >= 0).  FP64 plans: sweeps 0, rel_residual -1. */
int ppf_solve_async_result(ppf_plan *plan, void *stream, ppf_stats *stats);

/* Autotuner: times, with full solves of the given data on this plan's
 * extents, the x- and y-pass variants (threads per line: one radix-R
 * butterfly per thread per stage, or half the threads with two) and the
 * z-pass kernel shapes (partitioned Thomas with K rows per thread and S
 * columns per workgroup, pipelined wavefront cyclic reduction); keeps the
 * fastest of each pass for later solves and returns the z shape's id in
 * *zvariant (may be NULL).  Results are cached per problem shape (extents,
 * boundary conditions, precision) for the life of the process: a second plan
 * of a tuned shape is set without timing.  p is overwritten.  Synchronises
 * the stream.  Replaces the reference's fixed thread shapes threads_tdma /
 * threads_fft (fft_3d_poisson.f90:70-89) and its fixed TDMA launch
 * (pascal_tdma_cuda.f90:168-199), which the caller had to choose by hand. */
int ppf_plan_tune(ppf_plan *plan, const double *prhs, double *p, const double *dmx1, const double *dmx2,
                  const double *dmx3, const double *dx3, void *stream, int *zvariant);

/* Pins the z-pass kernel shape (an id returned by ppf_plan_tune, 1..17; 0 =
 * the default chosen by n3m).  A shape that does not fit the plan's z extent
 * makes later solves return PPF_ERR_UNSUPPORTED. */
int ppf_plan_set_zvariant(ppf_plan *plan, int zvariant);

/* The plan's pass variants: variants[0] x passes, [1] y pass (0/1 as for
 * x), [2] z shape (as ppf_plan_set_zvariant).  The x forms are 0: one
 * butterfly per thread per stage, 1: half the threads with two, 2: as 0
 * without holding the next tile in registers, 3: as 2 with one workgroup per
 * tile instead of a persistent grid.  variants[0] is f when both x passes use
 * form f, and f + 16 i when the forward pass uses f and the inverse i (the
 * tuner picks the two apart). */
int ppf_plan_variants(const ppf_plan *plan, int variants[3]);
/* Pins all three (x: 0..3 for both x passes, or f + 16 i with f != i in
 * 0..3; y: 0 or 1). */
int ppf_plan_set_variants(ppf_plan *plan, int xvariant, int yvariant, int zvariant);

/* ---- batched solves of independent problems ----------------------------- */
/* Device buffers of one problem of a batch (as in ppf_solve). */
typedef struct ppf_batch_io {
    const double *prhs;
    double *p;
    const double *dmx1, *dmx2, *dmx3, *dx3;
} ppf_batch_io;

#define PPF_BATCH_NO_GRAPH 1 /* launch every call directly (no hipGraph) */
#define PPF_BATCH_NO_TUNE 2  /* skip the per-shape autotuner */

typedef struct ppf_batch ppf_batch;

/* Many independent problems of any extents and boundary conditions, bound to
 * their device buffers.  One plan per problem; problems with equal extents
 * and BCs share one autotuner result.  The reference solves one problem per
 * plan and per call (module singleton, fft_3d_poisson.f90:43-55); this has no
 * counterpart there.
 * The first ppf_batch_solve solves every problem directly (FP32_IR problems
 * with their convergence test), records the sweeps each needed, and captures
 * the whole batch -- every problem as its own branch, FP32_IR problems as
 * ppf_solve_async with those sweep counts as min_sweeps and one guarded extra
 * sweep -- into one hipGraph; later calls replay the graph. */
int ppf_batch_create(int nprob, const ppf_desc *descs, const ppf_batch_io *io, int flags, ppf_batch **batch);

/* Solves every problem of the batch on `stream`.  stats (nprob entries or
 * NULL) receives per-problem sweeps and the measured relative residual (which
 * synchronises): for FP32_IR problems of a replay, the sweeps and residual of
 * their device-side test, and PPF_ERR_NOT_CONVERGED is returned if any of
 * them stayed above its tol.  Without stats a replay returns without
 * synchronising. */
int ppf_batch_solve(ppf_batch *batch, void *stream, ppf_stats *stats);

int ppf_batch_destroy(ppf_batch *batch);

/* Relative FP64 residual ||f - L p||_2 / ||f||_2 of the 7-point operator the
 * solve inverts (x/y: uniform spacing with zero-flux walls or wrap, the
 * operators whose eigenvalues are fft_3d_poisson.f90:288-351; z: the stretched
 * rows of :1137-1149), on interior cells of p.  Synchronises the stream. */
int ppf_residual(ppf_plan *plan, const double *f, const double *p, const double *dmx3, const double *dx3,
                 void *stream, double *rel_residual);

/* Number of device bytes the plan holds (workspace + tables). */
size_t ppf_plan_workspace_bytes(const ppf_plan *plan);

const char *ppf_error_string(int code);

/* ---- batched Thomas, one rank ------------------------------------------ */
/* Replaces PaScaL_TDMA_plan_many_create_cuda / _destroy_cuda
 * (pascal_tdma_cuda.f90:108-220) and, for nprocs = 1,
 * PaScaL_TDMA_many_solve_cuda (:230-269) -> tdma_many_cuda
 * (tdmas_cuda.f90:35-102).  Arrays are (nx_sys, ny_sys, nz_row), x fastest;
 * c and d are overwritten, d holds the solution, with the reference's
 * operation order (bitwise equal to the reference Thomas on the same inputs).
 * Unlike the reference, nx_sys and ny_sys need not be multiples of a thread
 * block (the reference aborts, :168-190). */
typedef struct ppf_tdma_plan ppf_tdma_plan;
int ppf_tdma_plan_many_create(int nx_sys, int ny_sys, int nz_row, ppf_tdma_plan **plan);
int ppf_tdma_many_solve(ppf_tdma_plan *plan, const double *a, const double *b, double *c, double *d, void *stream);
int ppf_tdma_plan_many_destroy(ppf_tdma_plan *plan);
/* Cyclic systems (row 0 couples to x_{n-1} through a, row n-1 to x_0 through
 * c).  Replaces PaScaL_TDMA_many_solve_cycle_cuda for nprocs = 1
 * (pascal_tdma_cuda.f90:279-318) -> tdma_cycl_many_cuda (tdmas_cuda.f90:
 * 114-217), same operation order; needs nz_row >= 2. */
int ppf_tdma_many_solve_cycle(ppf_tdma_plan *plan, const double *a, const double *b, double *c, double *d,
                              void *stream);

/* ---- device memory shim for hosts without HIP bindings (Fortran) -------- */
/* Stand-ins for the CUDA Fortran `device` arrays, allocate/deallocate and
 * array assignment host<->device the reference relies on
 * (01_Fortran/examples/cuda_subdomain.f90:96-180). */
#define PPF_MEMCPY_HOST_TO_DEVICE 1
#define PPF_MEMCPY_DEVICE_TO_HOST 2
#define PPF_MEMCPY_DEVICE_TO_DEVICE 3
int ppf_malloc(void **ptr, size_t bytes);
int ppf_free(void *ptr);
int ppf_memcpy(void *dst, const void *src, size_t bytes, int kind);
int ppf_memset(void *ptr, int value, size_t bytes);
int ppf_sync(void *stream);
int ppf_device_count(int *count);
int ppf_set_device(int device);

/* ---- measurement helper (no counterpart in the reference) --------------- */
/* Streams `count` doubles (count even and > 0, 16-B aligned buffers) from src
 * to dst `reps` times with a streaming copy kernel (one nontemporal 16-byte
 * load and store per thread) on `stream` and returns the mean
 * duration of one copy in *ms (HIP events; synchronises).  The measured
 * bandwidth 16*count/ms is the yardstick the pass kernels are compared with,
 * and the known byte count calibrates rocprofv3's HBM byte counters. */
int ppf_copy_timed(double *dst, const double *src, size_t count, int reps, void *stream, float *ms);

/* Synthetic projection-step right-hand side (measurement workload, no
 * counterpart in the reference): f (n1m, n2m, n3m) = the discrete divergence
 * of a staggered velocity field -- u on x-faces (periodic in x), v on y-faces
 * and w on z-faces (both zero on the walls) -- each a sum of nmodes separable
 * modes from `tables` (device, 9 blocks of nmodes x stride doubles: u's
 * x-face / y-centre / z-centre factors, v's x-centre / y-face / z-centre, w's
 * x-centre / y-centre / z-face; stride >= max(n1m, n2m, n3m) + 1) plus
 * seeded uniform face noise of amplitude `noise`.  x and y spacings are
 * L1/n1m and L2/n2m; dx3 (device, n3m+2) gives the z cell widths.  Zero
 * wall fluxes make sum f dx3 = 0 (compatible by construction). */
int ppf_synth_divergence_rhs(double *f, int n1m, int n2m, int n3m, double L1, double L2, const double *dx3,
                             const double *tables, int nmodes, int stride, unsigned long long seed, double noise,
                             void *stream);

/* The reference example's right-hand side on the device: f(i,j,k) =
 * -cos(x1(i) pi) cos(x2(j) pi) cos(x3(k) pi) 3 pi^2 at the grid nodes, i.e.
 * cuda_pressure_RHS_kernel (01_Fortran/examples/cuda_pressure.f90:66-90).
 * x1, x2, x3 (device, nm+2 each) are the node arrays x(0:nm+1); f is
 * (n1m, n2m, n3m). */
int ppf_example_rhs(double *f, const double *x1, const double *x2, const double *x3, int n1m, int n2m, int n3m,
                    void *stream);

#ifdef __cplusplus
}
#endif

#endif /* PASCAL_POISSON_FFT_H */
