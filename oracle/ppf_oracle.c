/*
 * ppf_oracle.c -- FP64 CPU restatement of the PaScaL_POISSON_FFT GPU solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle and the CPU
 * baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it, and only as the checker; the product path never calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference checkout).  Conventions follow the reference:
 *   - arrays are x-fastest (Fortran column-major); a (n1,n2,n3) field is stored
 *     at ((k*n2)+j)*n1+i with 0-based i,j,k;
 *   - the solution P carries one ghost layer per side: (n1m+2)(n2m+2)(n3m+2),
 *     interior cell (i,j,k) (1-based, as in the reference) sits at 0-based
 *     index (i,j,k) of that array;
 *   - metric arrays dmx*, dx* are 0-based with length nm+2, element q holding
 *     the reference's Fortran element q (declared 0:nsub).
 *
 * The Makhoul reorder of fft_3d_poisson.f90:924-933 / :1022-1031 maps
 * (1-based) out(i+n/2) <- in(2*(n-(i+n/2)+1)) = in(n-2i+2), i.e. the textbook
 * v[N-1-m] = x[2m+1] (0-based); with the twiddle passes at :937-1011 this is
 * exactly the DCT-II / DCT-III(1/2N) pair of scipy.fft.dct/idct (type 2,
 * norm=None).  tests/test_oracle.py restates the reference's index
 * expressions verbatim and checks them against scipy.
 *
 * Parity pins (tests/test_oracle_*.py): grid closed form and invariants,
 * eigenvalues vs dense eigen-decomposition, DCT/FFT vs scipy/numpy, Thomas vs
 * the reference's own tdmas.cpp compiled from source (bitwise) and vs dense
 * solves, full solve vs a dense brute-force 7-point solve on tiny grids and vs
 * discrete eigen-mode manufactured solutions.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

#define IDX3(i, j, k, n1, n2) ((((size_t)(k)) * (size_t)(n2) + (size_t)(j)) * (size_t)(n1) + (size_t)(i))

static const double OR_PI = 3.14159265358979323846264338327950288;

int or_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void or_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------------- */
/* Grid and metrics: 01_Fortran/examples/mpi_subdomain.f90:397-472 (1 rank).  */
/* x, dx, dmx have nm+2 entries (Fortran 0:nm+1).                              */
/* ------------------------------------------------------------------------- */
void or_grid(int uniform, int nm, double L, double gamma, int pbc, double *x, double *dx, double *dmx) {
    const int nsub = nm + 1;
    /* calculate_x (:397-415): non_uniform_channel_grid (:417-429) or uniform_grid (:431-441) */
    for (int idx = 0; idx <= nsub; ++idx) {
        if (uniform)
            x[idx] = (double)(idx - 1) * L / (double)nm;
        else
            x[idx] = L * 0.5 * (1.0 + tanh(0.5 * gamma * (2.0 * (double)(idx - 1) / (double)nm - 1.0)) / tanh(gamma * 0.5));
    }
    if (!pbc) x[0] = x[1]; /* :413 */
    /* calculate_dx_dmx (:443-472) with mpi_subdomain_ghostcell_1d (:783-798) on one rank */
    for (int q = 0; q <= nsub; ++q) { dx[q] = 0.0; dmx[q] = 0.0; }
    for (int q = 1; q <= nm; ++q) dx[q] = x[q + 1] - x[q];
    if (pbc) { dx[0] = dx[nm]; dx[nsub] = dx[1]; }
    else { dx[0] = 0.0; dx[nsub] = 0.0; }
    for (int q = 1; q <= nm; ++q) dmx[q] = 0.5 * (dx[q - 1] + dx[q]);
    if (pbc) { dmx[0] = dmx[nm]; dmx[nsub] = dmx[1]; }
    else dmx[nsub] = 0.5 * (dx[nm] + dx[nsub]);
}

/* ------------------------------------------------------------------------- */
/* Modified wavenumbers: 01_Fortran/src/fft_3d_poisson.f90:288-351.            */
/* dxk2 has n1m entries (reference i=1..n1m -> [i-1]); dyk2 has n2m.           */
/* ------------------------------------------------------------------------- */
static double eig_neumann(int m, int n, double h) {
    double s = sin(0.5 * (double)m * OR_PI / (double)n);
    return 4.0 * s * s / (h * h);
}
static double eig_periodic(int m, int n, double h) {
    return 2.0 * (1.0 - cos(2.0 * (double)m * OR_PI / (double)n)) / (h * h);
}

void or_eigen(int pbc1, int pbc2, int n1m, int n2m, double L1, double L2, double *dxk2, double *dyk2) {
    const double h1 = L1 / (double)n1m, h2 = L2 / (double)n2m;
    for (int i = 1; i <= n1m; ++i)
        dxk2[i - 1] = pbc1 ? eig_periodic(i - 1, n1m, h1) : eig_neumann(i - 1, n1m, h1);
    for (int j = 1; j <= n2m; ++j) {
        int jm = j - 1;
        /* P-P: the C2C axis wraps the mode index (:339-345) */
        if (pbc1 && pbc2) jm = (j <= n2m / 2 + 1) ? j - 1 : n2m - j + 1;
        dyk2[j - 1] = pbc2 ? eig_periodic(jm, n2m, h2) : eig_neumann(jm, n2m, h2);
    }
}

/* ------------------------------------------------------------------------- */
/* Complex FFT (unnormalised, sign -1 forward / +1 backward), mixed radix       */
/* Stockham autosort with radices 4, 2, 3 and a generic odd radix.  Stands in  */
/* for the cuFFT/FFTW executes at fft_3d_poisson.f90:386-608.                  */
/* ------------------------------------------------------------------------- */
/* tw[m] = exp(-2 pi i m / n), m = 0..n-1 */
static void twiddles(int n, cplx *tw) {
    for (int m = 0; m < n; ++m) {
        double ang = -2.0 * OR_PI * (double)m / (double)n;
        tw[m] = cos(ang) + I * sin(ang);
    }
}

static void fft_rec(int n, int sign, cplx *a, cplx *tmp, const cplx *tw) {
    int ns = 1; /* current sub-transform length */
    cplx *src = a, *dst = tmp;
    int rem = n;
    while (rem > 1) {
        int R;
        if (rem % 4 == 0) R = 4;
        else if (rem % 2 == 0) R = 2;
        else if (rem % 3 == 0) R = 3;
        else {
            R = rem;
            for (int f = 5; f * f <= rem; f += 2) if (rem % f == 0) { R = f; break; }
        }
        const int S = n / R;
        const int step = n / (ns * R);
        for (int b = 0; b < S; ++b) {
            const int j = b % ns, kk = b / ns;
            cplx v[R], vv[R];
            for (int r = 0; r < R; ++r) {
                cplx w = tw[(j * r * step) % n];
                if (sign > 0) w = conj(w);
                v[r] = src[b + r * S] * w;
            }
            for (int q = 0; q < R; ++q) {
                cplx s = 0;
                for (int r = 0; r < R; ++r) {
                    cplx w = tw[((q * r) % R) * (n / R)];
                    if (sign > 0) w = conj(w);
                    s += v[r] * w;
                }
                vv[q] = s;
            }
            for (int q = 0; q < R; ++q) dst[kk * ns * R + j + q * ns] = vv[q];
        }
        cplx *t = src; src = dst; dst = t;
        ns *= R;
        rem /= R;
    }
    if (src != a) memcpy(a, src, sizeof(cplx) * (size_t)n);
}

void or_fft(int n, cplx *data, int sign) {
    cplx *tmp = (cplx *)malloc(sizeof(cplx) * 2 * (size_t)n);
    twiddles(n, tmp + n);
    fft_rec(n, sign, data, tmp, tmp + n);
    free(tmp);
}

/* R2C, unnormalised: out[0..n/2] (cufftExecD2Z semantics). */
static void rfft_line(int n, const double *in, cplx *out, cplx *work, cplx *tmp, const cplx *tw) {
    for (int q = 0; q < n; ++q) work[q] = in[q];
    fft_rec(n, -1, work, tmp, tw);
    for (int q = 0; q <= n / 2; ++q) out[q] = work[q];
}
/* C2R, unnormalised (cufftExecZ2D): Hermitian extension, imaginary parts of the
 * DC and Nyquist terms ignored. */
static void irfft_line(int n, const cplx *in, double *out, cplx *work, cplx *tmp, const cplx *tw) {
    work[0] = creal(in[0]);
    for (int q = 1; q < n / 2; ++q) { work[q] = in[q]; work[n - q] = conj(in[q]); }
    work[n / 2] = creal(in[n / 2]);
    fft_rec(n, +1, work, tmp, tw);
    for (int q = 0; q < n; ++q) out[q] = creal(work[q]);
}

void or_rfft(int n, const double *in, cplx *out) {
    cplx *w = (cplx *)malloc(sizeof(cplx) * 3 * (size_t)n);
    twiddles(n, w + 2 * n);
    rfft_line(n, in, out, w, w + n, w + 2 * n);
    free(w);
}
void or_irfft(int n, const cplx *in, double *out) {
    cplx *w = (cplx *)malloc(sizeof(cplx) * 3 * (size_t)n);
    twiddles(n, w + 2 * n);
    irfft_line(n, in, out, w, w + n, w + 2 * n);
    free(w);
}

/* DCT-II, scipy norm=None: X_k = 2 sum_n x_n cos(pi k (2n+1) / 2N).
 * Makhoul: reorder (fft_3d_poisson.f90:915-935), R2C, post
 * twiddle X_k = 2 Re(exp(-i pi k / 2N) V_k) (:937-969). */
static void dct2_line(int n, const double *in, double *out, double *v, cplx *spec, cplx *work, cplx *tmp, const cplx *tw) {
    for (int q = 0; q < n / 2; ++q) {
        v[q] = in[2 * q];
        v[n - 1 - q] = in[2 * q + 1];
    }
    rfft_line(n, v, spec, work, tmp, tw);
    for (int k = 0; k <= n / 2; ++k) {
        double arg = -OR_PI * (double)k / (2.0 * (double)n);
        out[k] = 2.0 * (cos(arg) * creal(spec[k]) - sin(arg) * cimag(spec[k]));
    }
    for (int k = n / 2 + 1; k < n; ++k) {
        double arg = -OR_PI * (double)k / (2.0 * (double)n);
        out[k] = 2.0 * (cos(arg) * creal(spec[n - k]) + sin(arg) * cimag(spec[n - k]));
    }
}

/* Inverse: DCT-III scaled by 1/(2N) (scipy idct type 2).  Pre-twiddle
 * V_k = exp(i pi k/2N)(X_k - i X_{N-k}), X_N = 0 (:971-1011), C2R, inverse
 * reorder (:1013-1033), scale 1/(2N) (:517, :579). */
static void idct2_line(int n, const double *in, double *out, double *v, cplx *spec, cplx *work, cplx *tmp, const cplx *tw) {
    for (int k = 0; k <= n / 2; ++k) {
        double arg = -OR_PI * (double)k / (2.0 * (double)n);
        double xk = in[k];
        double xnk = (k == 0) ? 0.0 : in[n - k];
        spec[k] = (cos(arg) * xk - sin(arg) * xnk) + I * (-sin(arg) * xk - cos(arg) * xnk);
    }
    irfft_line(n, spec, v, work, tmp, tw);
    const double s = 1.0 / (2.0 * (double)n);
    for (int q = 0; q < n / 2; ++q) {
        out[2 * q] = v[q] * s;
        out[2 * q + 1] = v[n - 1 - q] * s;
    }
}

void or_dct2(int n, const double *in, double *out) {
    double *v = (double *)malloc(sizeof(double) * (size_t)n);
    cplx *w = (cplx *)malloc(sizeof(cplx) * (4 * (size_t)n + 2));
    twiddles(n, w + 3 * n + 1);
    dct2_line(n, in, out, v, w, w + n + 1, w + 2 * n + 1, w + 3 * n + 1);
    free(v); free(w);
}
void or_idct2(int n, const double *in, double *out) {
    double *v = (double *)malloc(sizeof(double) * (size_t)n);
    cplx *w = (cplx *)malloc(sizeof(cplx) * (4 * (size_t)n + 2));
    twiddles(n, w + 3 * n + 1);
    idct2_line(n, in, out, v, w, w + n + 1, w + 2 * n + 1, w + 3 * n + 1);
    free(v); free(w);
}

/* ------------------------------------------------------------------------- */
/* Thomas algorithm for many systems, op order of                              */
/* 01_Fortran/PaScaL_TDMA/src/tdmas_cuda.f90:35-102 and tdmas.f90:114-147.      */
/* Layout: element (s, k) at k*nsys + s (system index fastest, row slowest).   */
/* c and d are overwritten (d holds the solution), as in the reference.         */
/* ------------------------------------------------------------------------- */
void or_tdma_many(const double *a, const double *b, double *c, double *d, long nsys, int nrow) {
#pragma omp parallel for schedule(static)
    for (long s = 0; s < nsys; ++s) {
        d[s] = d[s] / b[s];
        c[s] = c[s] / b[s];
        for (int k = 1; k < nrow; ++k) {
            const size_t q = (size_t)k * (size_t)nsys + (size_t)s, qm = q - (size_t)nsys;
            double r = 1.0 / (b[q] - a[q] * c[qm]);
            d[q] = r * (d[q] - a[q] * d[qm]);
            c[q] = r * c[q];
        }
        for (int k = nrow - 2; k >= 0; --k) {
            const size_t q = (size_t)k * (size_t)nsys + (size_t)s;
            d[q] = d[q] - c[q] * d[q + (size_t)nsys];
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Cyclic Thomas for many systems (periodic rows: a_0 couples to x_{n-1},       */
/* c_{n-1} to x_0), op order of tdma_cycl_many, 01_Fortran/PaScaL_TDMA/src/     */
/* tdmas.f90:160-213 (= tdmas_cuda.f90:114-217): rows 1..n-1 are solved for     */
/* d' and the correction vector e (e_1 = -a_1, e_{n-1} = -c_{n-1}), then x_0,   */
/* then x_k = d'_k + x_0 e_k.  Same layout as or_tdma_many; c, d overwritten.   */
/* Needs nrow >= 2.                                                            */
/* ------------------------------------------------------------------------- */
void or_tdma_cycl_many(const double *a, const double *b, double *c, double *d, long nsys, int nrow) {
    const size_t ns = (size_t)nsys;
#pragma omp parallel
    {
        double *e = (double *)malloc(sizeof(double) * (size_t)nrow);
#pragma omp for schedule(static)
        for (long s = 0; s < nsys; ++s) {
            for (int k = 0; k < nrow; ++k) e[k] = 0.0;
            e[1] = -a[ns + (size_t)s];
            e[nrow - 1] = -c[(size_t)(nrow - 1) * ns + (size_t)s];
            {
                const size_t q = ns + (size_t)s;
                d[q] = d[q] / b[q];
                e[1] = e[1] / b[q];
                c[q] = c[q] / b[q];
            }
            for (int k = 2; k < nrow; ++k) {
                const size_t q = (size_t)k * ns + (size_t)s, qm = q - ns;
                const double rr = 1.0 / (b[q] - a[q] * c[qm]);
                d[q] = rr * (d[q] - a[q] * d[qm]);
                e[k] = rr * (e[k] - a[q] * e[k - 1]);
                c[q] = rr * c[q];
            }
            for (int k = nrow - 2; k >= 1; --k) {
                const size_t q = (size_t)k * ns + (size_t)s;
                d[q] = d[q] - c[q] * d[q + ns];
                e[k] = e[k] - c[q] * e[k + 1];
            }
            const size_t q0 = (size_t)s, q1 = ns + (size_t)s, qn = (size_t)(nrow - 1) * ns + (size_t)s;
            d[q0] = (d[q0] - a[q0] * d[qn] - c[q0] * d[q1]) / (b[q0] + a[q0] * e[nrow - 1] + c[q0] * e[1]);
            for (int k = 1; k < nrow; ++k) {
                const size_t q = (size_t)k * ns + (size_t)s;
                d[q] = d[q] + d[q0] * e[k];
            }
        }
        free(e);
    }
}

/* ------------------------------------------------------------------------- */
/* Line-transform drivers over a 3-D field.                                    */
/* ------------------------------------------------------------------------- */
typedef struct { double *v; cplx *spec, *work, *tmp, *tw; double *lin, *lout; cplx *cin; } scratch_t;
static scratch_t scratch_new(int n) {
    scratch_t s;
    s.v = (double *)malloc(sizeof(double) * 3 * (size_t)n);
    s.lin = s.v + n; s.lout = s.v + 2 * n;
    s.spec = (cplx *)malloc(sizeof(cplx) * (5 * (size_t)n + 2));
    s.work = s.spec + n + 1; s.tmp = s.work + n; s.cin = s.tmp + n; s.tw = s.cin + n + 1;
    twiddles(n, s.tw);
    return s;
}
static void scratch_free(scratch_t s) { free(s.v); free(s.spec); }

/* Apply a real transform along axis `ax` (0=x,1=y) of a real (n1,n2,n3) array. */
static void real_transform(int kind /*0 dct2, 1 idct2*/, int ax, double *a, int n1, int n2, int n3) {
    const int n = ax == 0 ? n1 : n2;
    const long nl = (long)(ax == 0 ? n2 : n1) * n3;
#pragma omp parallel
    {
        scratch_t s = scratch_new(n);
#pragma omp for schedule(static)
        for (long l = 0; l < nl; ++l) {
            long o, st;
            if (ax == 0) { o = l * n1; st = 1; }
            else { long i = l % n1, k = l / n1; o = (long)k * n1 * n2 + i; st = n1; }
            for (int q = 0; q < n; ++q) s.lin[q] = a[o + q * st];
            if (kind == 0) dct2_line(n, s.lin, s.lout, s.v, s.spec, s.work, s.tmp, s.tw);
            else idct2_line(n, s.lin, s.lout, s.v, s.spec, s.work, s.tmp, s.tw);
            for (int q = 0; q < n; ++q) a[o + q * st] = s.lout[q];
        }
        scratch_free(s);
    }
}

/* R2C along axis ax of real (n1,n2,n3) into complex out; complex dims are
 * (n1/2+1, n2, n3) for ax=0 and (n1, n2/2+1, n3) for ax=1. */
static void r2c_axis(int ax, const double *a, cplx *out, int n1, int n2, int n3) {
    const int n = ax == 0 ? n1 : n2;
    const int h = n / 2 + 1;
    const long nl = (long)(ax == 0 ? n2 : n1) * n3;
#pragma omp parallel
    {
        scratch_t s = scratch_new(n);
#pragma omp for schedule(static)
        for (long l = 0; l < nl; ++l) {
            long oi, sti, oo, sto;
            if (ax == 0) { oi = l * n1; sti = 1; oo = l * h; sto = 1; }
            else {
                long i = l % n1, k = l / n1;
                oi = (long)k * n1 * n2 + i; sti = n1;
                oo = (long)k * n1 * h + i; sto = n1;
            }
            for (int q = 0; q < n; ++q) s.lin[q] = a[oi + q * sti];
            rfft_line(n, s.lin, s.cin, s.work, s.tmp, s.tw);
            for (int q = 0; q < h; ++q) out[oo + q * sto] = s.cin[q];
        }
        scratch_free(s);
    }
}

static void c2r_axis(int ax, const cplx *a, double *out, int n1, int n2, int n3, double scale) {
    const int n = ax == 0 ? n1 : n2;
    const int h = n / 2 + 1;
    const long nl = (long)(ax == 0 ? n2 : n1) * n3;
#pragma omp parallel
    {
        scratch_t s = scratch_new(n);
#pragma omp for schedule(static)
        for (long l = 0; l < nl; ++l) {
            long oi, sti, oo, sto;
            if (ax == 0) { oo = l * n1; sto = 1; oi = l * h; sti = 1; }
            else {
                long i = l % n1, k = l / n1;
                oo = (long)k * n1 * n2 + i; sto = n1;
                oi = (long)k * n1 * h + i; sti = n1;
            }
            for (int q = 0; q < h; ++q) s.cin[q] = a[oi + q * sti];
            irfft_line(n, s.cin, s.lout, s.work, s.tmp, s.tw);
            for (int q = 0; q < n; ++q) out[oo + q * sto] = s.lout[q] * scale;
        }
        scratch_free(s);
    }
}

/* C2C along y of complex (n1c, n2, n3). */
static void c2c_y(cplx *a, int n1c, int n2, int n3, int sign, double scale) {
    const long nl = (long)n1c * n3;
#pragma omp parallel
    {
        cplx *w = (cplx *)malloc(sizeof(cplx) * 3 * (size_t)n2);
        twiddles(n2, w + 2 * n2);
#pragma omp for schedule(static)
        for (long l = 0; l < nl; ++l) {
            long i = l % n1c, k = l / n1c;
            long o = (long)k * n1c * n2 + i;
            for (int q = 0; q < n2; ++q) w[q] = a[o + (long)q * n1c];
            fft_rec(n2, sign, w, w + n2, w + 2 * n2);
            for (int q = 0; q < n2; ++q) a[o + (long)q * n1c] = w[q] * scale;
        }
        free(w);
    }
}

/* ------------------------------------------------------------------------- */
/* z-direction operator rows: cuda_Poisson_TDMA_z_real / _complex              */
/* (fft_3d_poisson.f90:1121-1170, :1172-1344).                                 */
/* ------------------------------------------------------------------------- */
static void z_coeffs(int n3m, const double *dx3, const double *dmx3, double *am, double *ap) {
    for (int k = 1; k <= n3m; ++k) {
        double m = 1.0 / dx3[k] / dmx3[k];
        double p = 1.0 / dx3[k] / dmx3[k + 1];
        if (k == 1) m = 0.0;
        if (k == n3m) p = 0.0;
        am[k - 1] = m;
        ap[k - 1] = p;
    }
}

/* ------------------------------------------------------------------------- */
/* Mean removal, wall BC and periodic ghosts:                                  */
/* cuda_Poisson_average_elimination (fft_3d_poisson.f90:628-664),               */
/* cuda_neumann_BC / apply_bc_on_direction (:798-852),                         */
/* cuda_ghostcell_update on one rank (:666-796, periodic wrap :750-757).        */
/* ------------------------------------------------------------------------- */
static void finish_solution(const double *sol, double *P, int n1m, int n2m, int n3m, int pbc1, int pbc2, int pbc3,
                            const double *dmx1, const double *dmx2, const double *dmx3) {
    const int n1 = n1m + 2, n2 = n2m + 2, n3 = n3m + 2;
    double sum = 0.0;
    /* sequential sum (the reference's CUF reduction order is unspecified) */
    for (size_t q = 0; q < (size_t)n1m * n2m * n3m; ++q) sum += sol[q];
    const double aver = sum / (double)n1m / (double)n2m / (double)n3m;
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= n3m; ++k)
        for (int j = 1; j <= n2m; ++j)
            for (int i = 1; i <= n1m; ++i)
                P[IDX3(i, j, k, n1, n2)] = sol[IDX3(i - 1, j - 1, k - 1, n1m, n2m)] - aver;

    const int pbc[3] = {pbc1, pbc2, pbc3};
    const int nm[3] = {n1m, n2m, n3m};
    const double *dmx[3] = {dmx1, dmx2, dmx3};
    /* neumann BC, directions 1,2,3 in order, full (0:nsub) ranges of the other two */
    for (int dir = 0; dir < 3; ++dir) {
        if (pbc[dir]) continue;
        const int nsub = nm[dir] + 1, nmsub = nm[dir];
        const double *d = dmx[dir];
        const double s0 = d[1] + d[2];
        const double bca0 = s0 * s0 / (s0 * s0 - d[1] * d[1]);
        const double bcb0 = d[1] * d[1] / (s0 * s0 - d[1] * d[1]);
        const double s1 = d[nsub] + d[nmsub];
        const double bca1 = s1 * s1 / (s1 * s1 - d[nsub] * d[nsub]);
        const double bcb1 = d[nsub] * d[nsub] / (s1 * s1 - d[nsub] * d[nsub]);
        for (int c = 0; c < (dir == 2 ? n2 : n3); ++c)
            for (int b = 0; b < (dir == 0 ? n2 : n1); ++b) {
                size_t q0, q1, q2, qn, qnm, qn2;
                if (dir == 0) {
                    q0 = IDX3(0, b, c, n1, n2); q1 = IDX3(1, b, c, n1, n2); q2 = IDX3(2, b, c, n1, n2);
                    qn = IDX3(nsub, b, c, n1, n2); qnm = IDX3(nmsub, b, c, n1, n2); qn2 = IDX3(nsub - 2, b, c, n1, n2);
                } else if (dir == 1) {
                    q0 = IDX3(b, 0, c, n1, n2); q1 = IDX3(b, 1, c, n1, n2); q2 = IDX3(b, 2, c, n1, n2);
                    qn = IDX3(b, nsub, c, n1, n2); qnm = IDX3(b, nmsub, c, n1, n2); qn2 = IDX3(b, nsub - 2, c, n1, n2);
                } else {
                    q0 = IDX3(b, c, 0, n1, n2); q1 = IDX3(b, c, 1, n1, n2); q2 = IDX3(b, c, 2, n1, n2);
                    qn = IDX3(b, c, nsub, n1, n2); qnm = IDX3(b, c, nmsub, n1, n2); qn2 = IDX3(b, c, nsub - 2, n1, n2);
                }
                P[q0] = bca0 * P[q1] - bcb0 * P[q2];
                P[qn] = bca1 * P[qnm] - bcb1 * P[qn2];
            }
    }
    /* periodic wrap, directions 1,2,3 in order */
    for (int dir = 0; dir < 3; ++dir) {
        if (!pbc[dir]) continue;
        const int nsub = nm[dir] + 1;
        for (int c = 0; c < (dir == 2 ? n2 : n3); ++c)
            for (int b = 0; b < (dir == 0 ? n2 : n1); ++b) {
                size_t q0, q1, qn, qnm;
                if (dir == 0) { q0 = IDX3(0, b, c, n1, n2); q1 = IDX3(1, b, c, n1, n2); qn = IDX3(nsub, b, c, n1, n2); qnm = IDX3(nsub - 1, b, c, n1, n2); }
                else if (dir == 1) { q0 = IDX3(b, 0, c, n1, n2); q1 = IDX3(b, 1, c, n1, n2); qn = IDX3(b, nsub, c, n1, n2); qnm = IDX3(b, nsub - 1, c, n1, n2); }
                else { q0 = IDX3(b, c, 0, n1, n2); q1 = IDX3(b, c, 1, n1, n2); qn = IDX3(b, c, nsub, n1, n2); qnm = IDX3(b, c, nsub - 1, n1, n2); }
                const double lo = P[q1], hi = P[qnm];
                P[q0] = hi;
                P[qn] = lo;
            }
    }
}

/* ------------------------------------------------------------------------- */
/* The solve: cuda_Poisson_FFT_1D (fft_3d_poisson.f90:353-626), one rank.       */
/* rhs: (n1m,n2m,n3m) x-fastest, not modified.  P: (n1m+2,n2m+2,n3m+2).         */
/* Returns 0, or a negative code for unsupported input (odd n).                */
/* Periodic z (pbc3) follows the reference as written: the z rows are built   */
/* exactly as for walls (am = 0 at k = 1 and ap = 0 at k = n3m on the only     */
/* rank, :1143-1146, whatever pbc3 says) and solved by the cyclic Thomas        */
/* (cuda_ptdma_core, :898-913 -> tdma_cycl_many_cuda), so the corner            */
/* couplings are zero; the z ghost planes are wrapped instead of extrapolated. */
/* ------------------------------------------------------------------------- */
int or_solve(int n1m, int n2m, int n3m, double L1, double L2, double L3, int pbc1, int pbc2, int pbc3,
             const double *rhs, double *P, const double *dmx1, const double *dmx2, const double *dmx3, const double *dx3) {
    (void)L3;
    if (n1m < 2 || n2m < 2 || n3m < 2 || (n1m & 1) || (n2m & 1) || (n3m & 1)) return -2;
    const size_t N = (size_t)n1m * n2m * n3m;
    double *dxk2 = (double *)malloc(sizeof(double) * (size_t)n1m);
    double *dyk2 = (double *)malloc(sizeof(double) * (size_t)n2m);
    double *am = (double *)malloc(sizeof(double) * (size_t)n3m);
    double *ap = (double *)malloc(sizeof(double) * (size_t)n3m);
    or_eigen(pbc1, pbc2, n1m, n2m, L1, L2, dxk2, dyk2);
    z_coeffs(n3m, dx3, dmx3, am, ap);
    double *sol = (double *)malloc(sizeof(double) * N);

    if (!pbc1 && !pbc2) {
        /* N-N: DCT x (:380-391), DCT y (:424-444), real TDMA (:1121-1170), inverse y (:497-518), inverse x (:562-579) */
        memcpy(sol, rhs, sizeof(double) * N);
        real_transform(0, 0, sol, n1m, n2m, n3m);
        real_transform(0, 1, sol, n1m, n2m, n3m);
        const long nsys = (long)n1m * n2m;
        {
            double *a = (double *)malloc(sizeof(double) * N), *b = (double *)malloc(sizeof(double) * N), *c = (double *)malloc(sizeof(double) * N);
#pragma omp parallel for schedule(static)
            for (int k = 0; k < n3m; ++k) {
                const double ac = -am[k] - ap[k];
                for (int j = 0; j < n2m; ++j)
                    for (int i = 0; i < n1m; ++i) {
                        size_t q = IDX3(i, j, k, n1m, n2m);
                        a[q] = am[k]; b[q] = ac - dxk2[i] - dyk2[j]; c[q] = ap[k];
                    }
            }
            a[0] = 0.0; b[0] = 1.0; c[0] = 0.0; sol[0] = 0.0; /* pin (:1152-1157) */
            if (pbc3) or_tdma_cycl_many(a, b, c, sol, nsys, n3m);
            else or_tdma_many(a, b, c, sol, nsys, n3m);
            free(a); free(b); free(c);
        }
        real_transform(1, 1, sol, n1m, n2m, n3m);
        real_transform(1, 0, sol, n1m, n2m, n3m);
    } else {
        /* complex branches: TDMA on real and imaginary parts with shared coefficients (:1172-1344) */
        int c1, c2; /* complex array dims */
        cplx *S;
        if (pbc1) { c1 = n1m / 2 + 1; c2 = n2m; }
        else { c1 = n1m; c2 = n2m / 2 + 1; }
        const size_t NC = (size_t)c1 * c2 * n3m;
        S = (cplx *)malloc(sizeof(cplx) * NC);
        double *tmp = (double *)malloc(sizeof(double) * N);
        memcpy(tmp, rhs, sizeof(double) * N);
        if (pbc1 && pbc2) {
            /* P-P: R2C x (:410-417), C2C y forward (:470-480) */
            r2c_axis(0, tmp, S, n1m, n2m, n3m);
            c2c_y(S, c1, n2m, n3m, -1, 1.0);
        } else if (pbc1 && !pbc2) {
            /* P-N: DCT y on real data (:393-408), then R2C x (:460-468) */
            real_transform(0, 1, tmp, n1m, n2m, n3m);
            r2c_axis(0, tmp, S, n1m, n2m, n3m);
        } else {
            /* N-P: DCT x (:380-391), then R2C y (:446-458) */
            real_transform(0, 0, tmp, n1m, n2m, n3m);
            r2c_axis(1, tmp, S, n1m, n2m, n3m);
        }
        /* split, build coefficients, pin, two Thomas solves */
        const long nsys = (long)c1 * c2;
        double *dr = (double *)malloc(sizeof(double) * NC), *di = (double *)malloc(sizeof(double) * NC);
        for (size_t q = 0; q < NC; ++q) { dr[q] = creal(S[q]); di[q] = cimag(S[q]); }
        double *a = (double *)malloc(sizeof(double) * NC), *b = (double *)malloc(sizeof(double) * NC), *c = (double *)malloc(sizeof(double) * NC);
        for (int part = 0; part < 2; ++part) {
            double *d = part ? di : dr;
#pragma omp parallel for schedule(static)
            for (int k = 0; k < n3m; ++k) {
                const double ac = -am[k] - ap[k];
                for (int j = 0; j < c2; ++j)
                    for (int i = 0; i < c1; ++i) {
                        size_t q = IDX3(i, j, k, c1, c2);
                        a[q] = am[k]; b[q] = ac - dxk2[i] - dyk2[j]; c[q] = ap[k];
                    }
            }
            a[0] = 0.0; b[0] = 1.0; c[0] = 0.0; d[0] = 0.0;
            if (pbc3) or_tdma_cycl_many(a, b, c, d, nsys, n3m);
            else or_tdma_many(a, b, c, d, nsys, n3m);
        }
        for (size_t q = 0; q < NC; ++q) S[q] = dr[q] + I * di[q];
        free(a); free(b); free(c); free(dr); free(di);
        if (pbc1 && pbc2) {
            /* inverse y C2C with 1/n2m (:544-554), C2R x with 1/n1m (:605-614) */
            c2c_y(S, c1, n2m, n3m, +1, 1.0 / (double)n2m);
            c2r_axis(0, S, sol, n1m, n2m, n3m, 1.0 / (double)n1m);
        } else if (pbc1 && !pbc2) {
            /* C2R x with 1/n1m (:534-542), inverse DCT y with 1/(2 n2m) (:581-603) */
            c2r_axis(0, S, sol, n1m, n2m, n3m, 1.0 / (double)n1m);
            real_transform(1, 1, sol, n1m, n2m, n3m);
        } else {
            /* C2R y with 1/n2m (:520-532), inverse DCT x with 1/(2 n1m) (:562-579) */
            c2r_axis(1, S, sol, n1m, n2m, n3m, 1.0 / (double)n2m);
            real_transform(1, 0, sol, n1m, n2m, n3m);
        }
        free(S); free(tmp);
    }
    finish_solution(sol, P, n1m, n2m, n3m, pbc1, pbc2, pbc3, dmx1, dmx2, dmx3);
    free(sol); free(dxk2); free(dyk2); free(am); free(ap);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* The discrete operator the solve inverts (7-point, FP64):                     */
/* x/y: uniform spacing L/nm, zero-flux (mirror) walls for Neumann axes, wrap   */
/* for periodic axes -- the operators diagonalised by the DCT-II / DFT with the */
/* eigenvalues of fft_3d_poisson.f90:288-351;                                   */
/* z: the stretched rows of fft_3d_poisson.f90:1137-1149.                       */
/* Lp is evaluated on interior values p (n1m,n2m,n3m).                          */
/* ------------------------------------------------------------------------- */
void or_apply_L(int n1m, int n2m, int n3m, double L1, double L2, int pbc1, int pbc2,
                const double *p, double *out, const double *dmx3, const double *dx3) {
    const double h1 = L1 / (double)n1m, h2 = L2 / (double)n2m;
    const double i1 = 1.0 / (h1 * h1), i2 = 1.0 / (h2 * h2);
    double *am = (double *)malloc(sizeof(double) * (size_t)n3m), *ap = (double *)malloc(sizeof(double) * (size_t)n3m);
    z_coeffs(n3m, dx3, dmx3, am, ap);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n3m; ++k)
        for (int j = 0; j < n2m; ++j)
            for (int i = 0; i < n1m; ++i) {
                const double c = p[IDX3(i, j, k, n1m, n2m)];
                int im = i - 1, ip = i + 1, jm = j - 1, jp = j + 1;
                double xm, xp, ym, yp;
                if (pbc1) { im = (im + n1m) % n1m; ip = ip % n1m; xm = p[IDX3(im, j, k, n1m, n2m)]; xp = p[IDX3(ip, j, k, n1m, n2m)]; }
                else { xm = im < 0 ? c : p[IDX3(im, j, k, n1m, n2m)]; xp = ip >= n1m ? c : p[IDX3(ip, j, k, n1m, n2m)]; }
                if (pbc2) { jm = (jm + n2m) % n2m; jp = jp % n2m; ym = p[IDX3(i, jm, k, n1m, n2m)]; yp = p[IDX3(i, jp, k, n1m, n2m)]; }
                else { ym = jm < 0 ? c : p[IDX3(i, jm, k, n1m, n2m)]; yp = jp >= n2m ? c : p[IDX3(i, jp, k, n1m, n2m)]; }
                const double zm = k > 0 ? p[IDX3(i, j, k - 1, n1m, n2m)] : 0.0;
                const double zp = k < n3m - 1 ? p[IDX3(i, j, k + 1, n1m, n2m)] : 0.0;
                out[IDX3(i, j, k, n1m, n2m)] = (xp - 2.0 * c + xm) * i1 + (yp - 2.0 * c + ym) * i2 +
                                               am[k] * zm - (am[k] + ap[k]) * c + ap[k] * zp;
            }
    free(am); free(ap);
}

/* ||f - L p||_2 / ||f||_2 over interior cells; p given as the ghosted P array. */
double or_rel_residual(int n1m, int n2m, int n3m, double L1, double L2, int pbc1, int pbc2,
                       const double *f, const double *P, const double *dmx3, const double *dx3) {
    const size_t N = (size_t)n1m * n2m * n3m;
    double *p = (double *)malloc(sizeof(double) * N), *lp = (double *)malloc(sizeof(double) * N);
    for (int k = 0; k < n3m; ++k)
        for (int j = 0; j < n2m; ++j)
            for (int i = 0; i < n1m; ++i)
                p[IDX3(i, j, k, n1m, n2m)] = P[IDX3(i + 1, j + 1, k + 1, n1m + 2, n2m + 2)];
    or_apply_L(n1m, n2m, n3m, L1, L2, pbc1, pbc2, p, lp, dmx3, dx3);
    double rn = 0.0, fn = 0.0;
    for (size_t q = 0; q < N; ++q) { double r = f[q] - lp[q]; rn += r * r; fn += f[q] * f[q]; }
    free(p); free(lp);
    return fn > 0 ? sqrt(rn / fn) : sqrt(rn);
}
