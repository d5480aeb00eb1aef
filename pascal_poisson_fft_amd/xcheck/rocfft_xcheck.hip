// rocfft_xcheck.hip -- a timing cross-check, not part of the solve path: the
// vendor FFT (hipFFT over rocFFT) on the x lines of a field, timed with HIP
// events, for comparison with the solver's own x passes.  Built into its own
// library (libppf_rocfft_xcheck.so); the solve library does not link it.
//
// The reference's x step is cuFFT's batched D2Z / Z2D along the contiguous
// axis (01_Fortran/src/fft_3d_poisson.f90:195-286 plans, :386-608 calls),
// plus separate DCT pre/post passes for Neumann axes; this times the
// transform alone.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

extern "C" {

// Batched double-precision real FFTs of nlines contiguous lines of n points:
// the mean time per launch of D2Z (n reals -> n/2+1 complex per line) and of
// Z2D (back), over reps launches each after two untimed ones.  Returns 0, or
// -1 on a bad argument, -3 on a HIP or hipFFT failure.
int ppfx_rocfft_lines(int n, int nlines, int reps, float *ms_fwd, float *ms_inv) {
    if (n < 2 || nlines < 1 || reps < 1 || !ms_fwd || !ms_inv) return -1;
    const size_t nin = (size_t)n * nlines, nout = (size_t)(n / 2 + 1) * nlines;
    double *in = nullptr;
    hipfftDoubleComplex *out = nullptr;
    hipfftHandle pf = 0, pi = 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = -3;
    int len[1] = {n};
    float ms = 0.f;
    if (hipMalloc(&in, nin * sizeof(double)) != hipSuccess) goto done;
    if (hipMalloc(&out, nout * sizeof(hipfftDoubleComplex)) != hipSuccess) goto done;
    if (hipMemset(in, 0, nin * sizeof(double)) != hipSuccess) goto done;
    if (hipfftPlanMany(&pf, 1, len, nullptr, 1, n, nullptr, 1, n / 2 + 1, HIPFFT_D2Z, nlines) != HIPFFT_SUCCESS) goto done;
    if (hipfftPlanMany(&pi, 1, len, nullptr, 1, n / 2 + 1, nullptr, 1, n, HIPFFT_Z2D, nlines) != HIPFFT_SUCCESS) goto done;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) goto done;
    for (int i = 0; i < 2; ++i)
        if (hipfftExecD2Z(pf, in, out) != HIPFFT_SUCCESS || hipfftExecZ2D(pi, out, in) != HIPFFT_SUCCESS) goto done;
    if (hipDeviceSynchronize() != hipSuccess) goto done;
    hipEventRecord(e0, 0);
    for (int i = 0; i < reps; ++i)
        if (hipfftExecD2Z(pf, in, out) != HIPFFT_SUCCESS) goto done;
    hipEventRecord(e1, 0);
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) goto done;
    *ms_fwd = ms / reps;
    hipEventRecord(e0, 0);
    for (int i = 0; i < reps; ++i)
        if (hipfftExecZ2D(pi, out, in) != HIPFFT_SUCCESS) goto done;
    hipEventRecord(e1, 0);
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) goto done;
    *ms_inv = ms / reps;
    rc = 0;
done:
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (pf) (void)hipfftDestroy(pf);
    if (pi) (void)hipfftDestroy(pi);
    if (in) (void)hipFree(in);
    if (out) (void)hipFree(out);
    return rc;
}

}  // extern "C"
