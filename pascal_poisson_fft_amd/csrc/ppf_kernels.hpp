// ppf_kernels.hpp -- the five passes of the solve for gfx950.
//
// Workspace W (type T) holds one spectral field laid out x fastest:
//   W[(kp * Ly + jp) * Lx + i],  Lx = n1m (Neumann x) or n1m+2 (periodic x,
//   half spectrum as interleaved re/im), Ly = n2m+2, Lz = n3m+2.
// The forward passes use rows jp = 1..n2m and planes kp = 1..n3m; the z pass
// adds the two spectral ghost planes kp = 0, n3m+1; the y-inverse fills the
// y ghost rows jp = 0, n2m+1; the x-inverse writes every row of P (with the x
// ghosts), so P's ghost layers come out of the passes with no extra sweep.
//
//   pass 1  x-forward   prhs rows -> W            DCT-II (N) or R2C (P)
//   pass 2  y-forward   W columns -> W            DCT-II, C2C (P-P) or R2C (N-P)
//   pass 3  z-Thomas    W z-lines -> W            one tridiagonal solve per mode
//   pass 4  y-inverse   W columns -> W            + y ghosts
//   pass 5  x-inverse   W rows    -> P            + x ghosts, mean removal
#pragma once

#include "ppf_lines.hpp"

namespace ppf {

constexpr int kMaxNT = 512;

// ---------------------------------------------------------------------------
// pass 1: x-forward.  One LDS line per prhs row; LPB rows per workgroup.
// PER: periodic x (R2C, Lx = N+2) else DCT-II (Lx = N).
// ---------------------------------------------------------------------------
template <typename T, int N, bool PER>
struct XGeom {
    using G = LineGeom<N, PER>;
    static constexpr int TPL = G::TPL;
    static constexpr int LPB = (256 / TPL) > 0 ? (256 / TPL) : 1;
    static constexpr int NT = LPB * TPL;
};

template <typename T, int N, bool PER>
__global__ void __launch_bounds__(256)
k_xfwd(const double *__restrict__ in, T *__restrict__ W, int nlines, int n2m, int Ly,
       const cx<T> *__restrict__ tw, const cx<T> *__restrict__ wk, const cx<T> *__restrict__ qk) {
    using X = XGeom<T, N, PER>;
    using G = typename X::G;
    constexpr int M = G::M, TPL = X::TPL, LPB = X::LPB, NT = X::NT, S = G::STRIDE;
    constexpr int LX = PER ? N + 2 : N;
    __shared__ cx<T> lds[LPB * S];
    const int tid = threadIdx.x, l = tid / TPL, t = tid - l * TPL;
    const int line0 = blockIdx.x * LPB;
    for (int e = tid; e < LPB * N; e += NT) {
        const int li = e / N, n = e - li * N;
        const int line = line0 + li;
        T v = T(0);
        if (line < nlines) v = (T)in[(size_t)line * N + n];
        T *re = reinterpret_cast<T *>(lds + li * S);
        re[PER ? n : makhoul_pos<N>(n)] = v;
    }
    __syncthreads();
    cx<T> *my = lds + l * S;
    fft_forward<T, M, TPL>(my, t, tw);
    if (PER)
        r2c_post<T, N, TPL>(my, t, wk);
    else
        dct_post<T, N, TPL>(my, t, wk, qk);
    for (int e = tid; e < LPB * LX; e += NT) {
        const int li = e / LX, n = e - li * LX;
        const int line = line0 + li;
        if (line < nlines) {
            const int j = line % n2m, k = line / n2m;
            const T *re = reinterpret_cast<const T *>(lds + li * S);
            W[((size_t)(k + 1) * Ly + (j + 1)) * LX + n] = re[n];
        }
    }
}

// ---------------------------------------------------------------------------
// pass 5: x-inverse.  Every W row (Ly*Lz of them) -> one P row of N+2 with the
// x ghosts; subtracts the mean found by the z pass.
// ---------------------------------------------------------------------------
template <typename T, int N, bool PER>
__global__ void __launch_bounds__(256)
k_xinv(const T *__restrict__ W, double *__restrict__ P, int nlines,
       const cx<T> *__restrict__ tw, const cx<T> *__restrict__ wk, const cx<T> *__restrict__ qk,
       const double *__restrict__ dmx1, const double *__restrict__ mean) {
    using X = XGeom<T, N, PER>;
    using G = typename X::G;
    constexpr int M = G::M, TPL = X::TPL, LPB = X::LPB, NT = X::NT, S = G::STRIDE;
    constexpr int LX = PER ? N + 2 : N;
    constexpr int LP = N + 2;
    __shared__ cx<T> lds[LPB * S];
    const int tid = threadIdx.x, l = tid / TPL, t = tid - l * TPL;
    const int line0 = blockIdx.x * LPB;
    for (int e = tid; e < LPB * LX; e += NT) {
        const int li = e / LX, n = e - li * LX;
        const int line = line0 + li;
        T v = T(0);
        if (line < nlines) v = W[(size_t)line * LX + n];
        reinterpret_cast<T *>(lds + li * S)[n] = v;
    }
    __syncthreads();
    cx<T> *my = lds + l * S;
    if (PER)
        c2r_pre<T, N, TPL>(my, t, wk);
    else
        idct_pre<T, N, TPL>(my, t, wk, qk);
    fft_forward<T, M, TPL>(my, t, tw);
    const double scale = PER ? 1.0 / (double)N : 1.0 / (2.0 * (double)N);
    const double mu = *mean;
    double bca0 = 0, bcb0 = 0, bca1 = 0, bcb1 = 0;
    if (!PER) {
        const double d1 = dmx1[1], d2 = dmx1[2], s0 = d1 + d2;
        bca0 = s0 * s0 / (s0 * s0 - d1 * d1);
        bcb0 = d1 * d1 / (s0 * s0 - d1 * d1);
        const double e1 = dmx1[N + 1], e2 = dmx1[N], s1 = e1 + e2;
        bca1 = s1 * s1 / (s1 * s1 - e1 * e1);
        bcb1 = e1 * e1 / (s1 * s1 - e1 * e1);
    }
    for (int e = tid; e < LPB * LP; e += NT) {
        const int li = e / LP, np = e - li * LP;
        const int line = line0 + li;
        if (line >= nlines) continue;
        const T *re = reinterpret_cast<const T *>(lds + li * S);
        // physical sample n (0-based) of this line, mean removed
        auto sample = [&](int n) -> double {
            const int tpos = PER ? n : makhoul_pos<N>(n);
            const double v = (double)re[tpos];
            return ((tpos & 1) ? -v : v) * scale - mu;
        };
        double out;
        if (np == 0)
            out = PER ? sample(N - 1) : bca0 * sample(0) - bcb0 * sample(1);
        else if (np == N + 1)
            out = PER ? sample(0) : bca1 * sample(N - 1) - bcb1 * sample(N - 2);
        else
            out = sample(np - 1);
        P[(size_t)line * LP + np] = out;
    }
}

// ---------------------------------------------------------------------------
// passes 2 and 4: y-forward / y-inverse on a tile of TX scalar columns of one
// plane.  MODE 0: DCT on real columns; 1: C2C on complex columns (P-P);
// 2: R2C / C2R on real columns (N-P, spectral rows 0..N+1 as re/im pairs).
// ---------------------------------------------------------------------------
template <typename T, int N, int MODE>
struct YGeom {
    static constexpr bool CPLX = MODE == 1;
    static constexpr int M = CPLX ? N : N / 2;
    static constexpr int TPL = (M / max_radix(M)) > 0 ? (M / max_radix(M)) : 1;
    static constexpr int S = CPLX ? M + 1 : (MODE == 2 ? M + 1 : M);
    static constexpr int LINE_BYTES = S * (int)sizeof(cx<T>);
    // scalar columns per tile: 128-byte rows, LDS kept within 64 KiB,
    // threads within kMaxNT
    static constexpr int TX0 = 128 / (int)sizeof(T);
    static constexpr int tx_fit(int tx) {
        return (tx > 2 && ((CPLX ? tx / 2 : tx) * LINE_BYTES > 65536 || (CPLX ? tx / 2 : tx) * TPL > kMaxNT))
                   ? tx_fit(tx / 2)
                   : tx;
    }
    static constexpr int TX = tx_fit(TX0);
    static constexpr int NL = CPLX ? TX / 2 : TX;  // LDS lines per tile
    static constexpr int NT = NL * TPL;
};

template <typename T, int N, int MODE>
__global__ void __launch_bounds__(kMaxNT)
k_yfwd(T *__restrict__ W, int Lx, int Ly, const cx<T> *__restrict__ tw, const cx<T> *__restrict__ wk,
       const cx<T> *__restrict__ qk) {
    using Y = YGeom<T, N, MODE>;
    constexpr int M = Y::M, TPL = Y::TPL, TX = Y::TX, NT = Y::NT, S = Y::S;
    __shared__ cx<T> lds[Y::NL * S];
    const int tid = threadIdx.x, l = tid / TPL, t = tid - l * TPL;
    const int c0 = blockIdx.x * TX;
    const size_t plane = (size_t)(blockIdx.y + 1) * Ly;
    for (int e = tid; e < TX * N; e += NT) {
        const int r = e / TX, c = e - r * TX;
        const int col = c0 + c;
        T v = T(0);
        if (col < Lx) v = W[(plane + 1 + r) * Lx + col];
        if (MODE == 0)
            reinterpret_cast<T *>(lds + c * S)[makhoul_pos<N>(r)] = v;
        else if (MODE == 1)
            reinterpret_cast<T *>(lds + (c >> 1) * S)[2 * r + (c & 1)] = v;
        else
            reinterpret_cast<T *>(lds + c * S)[r] = v;
    }
    __syncthreads();
    cx<T> *my = lds + l * S;
    fft_forward<T, M, TPL>(my, t, tw);
    if (MODE == 0) dct_post<T, N, TPL>(my, t, wk, qk);
    if (MODE == 2) r2c_post<T, N, TPL>(my, t, wk);
    constexpr int NR = MODE == 2 ? N + 2 : N;  // spectral rows written
    constexpr int R0 = MODE == 2 ? 0 : 1;
    for (int e = tid; e < TX * NR; e += NT) {
        const int r = e / TX, c = e - r * TX;
        const int col = c0 + c;
        if (col >= Lx) continue;
        T v;
        if (MODE == 1)
            v = reinterpret_cast<const T *>(lds + (c >> 1) * S)[2 * r + (c & 1)];
        else
            v = reinterpret_cast<const T *>(lds + c * S)[r];
        W[(plane + R0 + r) * Lx + col] = v;
    }
}

template <typename T, int N, int MODE>
__global__ void __launch_bounds__(kMaxNT)
k_yinv(T *__restrict__ W, int Lx, int Ly, const cx<T> *__restrict__ tw, const cx<T> *__restrict__ wk,
       const cx<T> *__restrict__ qk, const double *__restrict__ dmx2) {
    using Y = YGeom<T, N, MODE>;
    constexpr int M = Y::M, TPL = Y::TPL, TX = Y::TX, NT = Y::NT, S = Y::S;
    __shared__ cx<T> lds[Y::NL * S];
    const int tid = threadIdx.x, l = tid / TPL, t = tid - l * TPL;
    const int c0 = blockIdx.x * TX;
    const size_t plane = (size_t)blockIdx.y * Ly;  // every plane incl. z ghosts
    constexpr int NR = MODE == 2 ? N + 2 : N;
    constexpr int R0 = MODE == 2 ? 0 : 1;
    for (int e = tid; e < TX * NR; e += NT) {
        const int r = e / TX, c = e - r * TX;
        const int col = c0 + c;
        T v = T(0);
        if (col < Lx) v = W[(plane + R0 + r) * Lx + col];
        if (MODE == 1)  // conjugate on load (inverse via the forward FFT)
            reinterpret_cast<T *>(lds + (c >> 1) * S)[2 * r + (c & 1)] = (c & 1) ? -v : v;
        else
            reinterpret_cast<T *>(lds + c * S)[r] = v;
    }
    __syncthreads();
    cx<T> *my = lds + l * S;
    if (MODE == 0) idct_pre<T, N, TPL>(my, t, wk, qk);
    if (MODE == 2) c2r_pre<T, N, TPL>(my, t, wk);
    fft_forward<T, M, TPL>(my, t, tw);
    const T scale = MODE == 0 ? T(1.0 / (2.0 * N)) : T(1.0 / (double)N);
    // y ghosts: Neumann for MODE 0, periodic wrap otherwise
    T bca0 = 0, bcb0 = 0, bca1 = 0, bcb1 = 0;
    if (MODE == 0) {
        const double d1 = dmx2[1], d2 = dmx2[2], s0 = d1 + d2;
        bca0 = (T)(s0 * s0 / (s0 * s0 - d1 * d1));
        bcb0 = (T)(d1 * d1 / (s0 * s0 - d1 * d1));
        const double e1 = dmx2[N + 1], e2 = dmx2[N], s1 = e1 + e2;
        bca1 = (T)(s1 * s1 / (s1 * s1 - e1 * e1));
        bcb1 = (T)(e1 * e1 / (s1 * s1 - e1 * e1));
    }
    for (int e = tid; e < TX * (N + 2); e += NT) {
        const int rp = e / TX, c = e - rp * TX;
        const int col = c0 + c;
        if (col >= Lx) continue;
        auto sample = [&](int n) -> T {  // physical row n (0-based)
            if (MODE == 1) {
                const T v = reinterpret_cast<const T *>(lds + (c >> 1) * S)[2 * n + (c & 1)];
                return ((c & 1) ? -v : v) * scale;
            }
            const int tpos = MODE == 0 ? makhoul_pos<N>(n) : n;
            const T v = reinterpret_cast<const T *>(lds + c * S)[tpos];
            return ((tpos & 1) ? -v : v) * scale;
        };
        T out;
        if (rp == 0)
            out = MODE == 0 ? bca0 * sample(0) - bcb0 * sample(1) : sample(N - 1);
        else if (rp == N + 1)
            out = MODE == 0 ? bca1 * sample(N - 1) - bcb1 * sample(N - 2) : sample(0);
        else
            out = sample(rp - 1);
        W[(plane + rp) * Lx + col] = out;
    }
}

// ---------------------------------------------------------------------------
// pass 3: z Thomas, one thread per spectral column (scalar column kx of row
// jp).  Coefficients of cuda_Poisson_TDMA_z_real/_complex
// (fft_3d_poisson.f90:1137-1161, :1216-1254): am_k = 1/dx3_k/dmx3_k (0 at the
// bottom), ap_k = 1/dx3_k/dmx3_{k+1} (0 at the top), b_k = (-am-ap) - lx - ly,
// with the (0,0) mode's first row pinned (a=c=0, b=1, d=0).  Thomas in the op
// order of tdma_many_cuda (tdmas_cuda.f90:35-102).  The column's two z ghost
// planes (wall extrapolation, :809-852) are written in spectral space, and
// the (0,0) real column stores the solution mean for pass 5 (:628-664).
// ---------------------------------------------------------------------------
struct ZArgs {
    int Lx, Ly, n3m, jlo, njr;
    const double *lamx, *lamy;        // per scalar column / per row
    const unsigned char *zx, *zy;     // zero-mode flags
    int mean_kx, mean_jp;             // real (0,0) column
    double mean_scale;                // 1/(fx fy n1m n2m n3m)
};

template <typename T>
__global__ void __launch_bounds__(256)
k_zthomas(T *__restrict__ W, T *__restrict__ cs, ZArgs z, const double *__restrict__ dx3,
          const double *__restrict__ dmx3, double *__restrict__ mean_out) {
    extern __shared__ double zsh[];
    double *am = zsh, *ap = zsh + z.n3m;
    for (int k = threadIdx.x; k < z.n3m; k += blockDim.x) {
        const int kk = k + 1;
        double m = 1.0 / dx3[kk] / dmx3[kk];
        double p = 1.0 / dx3[kk] / dmx3[kk + 1];
        if (kk == 1) m = 0.0;
        if (kk == z.n3m) p = 0.0;
        am[k] = m;
        ap[k] = p;
    }
    __syncthreads();
    // wall extrapolation of cuda_neumann_BC (fft_3d_poisson.f90:829-837)
    double bca0, bcb0, bca1, bcb1;
    {
        const double d1 = dmx3[1], d2 = dmx3[2], s0 = d1 + d2;
        bca0 = s0 * s0 / (s0 * s0 - d1 * d1);
        bcb0 = d1 * d1 / (s0 * s0 - d1 * d1);
        const double e1 = dmx3[z.n3m + 1], e2 = dmx3[z.n3m], s1 = e1 + e2;
        bca1 = s1 * s1 / (s1 * s1 - e1 * e1);
        bcb1 = e1 * e1 / (s1 * s1 - e1 * e1);
    }
    const long col = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= (long)z.Lx * z.njr) return;
    const int kx = (int)(col % z.Lx);
    const int jp = z.jlo + (int)(col / z.Lx);
    const size_t pstride = (size_t)z.Ly * z.Lx;
    T *w = W + (size_t)jp * z.Lx + kx;       // plane 0 of this column
    T *c = cs + (size_t)(col);               // scratch, stride Lx*njr
    const size_t cstride = (size_t)z.Lx * z.njr;
    const double lx = z.lamx[kx], ly = z.lamy[jp];
    const bool pin = z.zx[kx] && z.zy[jp];
    const int n = z.n3m;
    // forward elimination
    double b1 = (-am[0] - ap[0] - lx) - ly;
    double c1 = ap[0];
    double d1 = (double)w[pstride];
    if (pin) { b1 = 1.0; c1 = 0.0; d1 = 0.0; }
    double dprev = d1 / b1;
    double cprev = c1 / b1;
    w[pstride] = (T)dprev;
    c[0] = (T)cprev;
    for (int k = 1; k < n; ++k) {
        const double a = am[k], b = (-am[k] - ap[k] - lx) - ly, cc = ap[k];
        const double d = (double)w[(size_t)(k + 1) * pstride];
        const double r = 1.0 / (b - a * cprev);
        dprev = r * (d - a * dprev);
        cprev = r * cc;
        w[(size_t)(k + 1) * pstride] = (T)dprev;
        c[(size_t)k * cstride] = (T)cprev;
    }
    // back substitution
    double xn = dprev, xk = xn, xk1 = 0.0;
    double sum = xn;
    for (int k = n - 2; k >= 0; --k) {
        const double ck = (double)c[(size_t)k * cstride];
        const double dk = (double)w[(size_t)(k + 1) * pstride];
        xk1 = xk;
        xk = dk - ck * xk1;
        w[(size_t)(k + 1) * pstride] = (T)xk;
        sum += xk;
    }
    // now xk = x_1, xk1 = x_2; x_n and x_{n-1} are re-read (n >= 2)
    const double x1 = xk, x2 = xk1;
    const double xm1 = (double)w[(size_t)(n - 1) * pstride];
    w[0] = (T)(bca0 * x1 - bcb0 * x2);
    w[(size_t)(n + 1) * pstride] = (T)(bca1 * xn - bcb1 * xm1);
    if (kx == z.mean_kx && jp == z.mean_jp) *mean_out = sum * z.mean_scale;
}

// ---------------------------------------------------------------------------
// FP64 residual r = f - L p on interior cells of the ghosted P, with partial
// sums of r^2 and f^2 per workgroup (deterministic two-level reduction).
// ---------------------------------------------------------------------------
struct ResArgs {
    int n1m, n2m, n3m, pbc1, pbc2;
    double ih1, ih2;  // 1/h^2
};

__global__ void __launch_bounds__(256)
k_residual(const double *__restrict__ f, const double *__restrict__ P, ResArgs a, const double *__restrict__ dx3,
           const double *__restrict__ dmx3, double *__restrict__ rout, double *__restrict__ partial) {
    __shared__ double s_r[256], s_f[256];
    const int n1m = a.n1m, n2m = a.n2m, n3m = a.n3m;
    const size_t N = (size_t)n1m * n2m * n3m;
    const int L1 = n1m + 2, L2 = n2m + 2;
    double sr = 0.0, sf = 0.0;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < N; q += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(q % n1m);
        const size_t rest = q / n1m;
        const int j = (int)(rest % n2m), k = (int)(rest / n2m);
        const size_t pc = ((size_t)(k + 1) * L2 + (j + 1)) * L1 + (i + 1);
        const double c = P[pc];
        double xm, xp, ym, yp;
        if (a.pbc1) {
            xm = P[pc - 1 + (i == 0 ? n1m : 0)];
            xp = P[pc + 1 - (i == n1m - 1 ? n1m : 0)];
        } else {
            xm = i == 0 ? c : P[pc - 1];
            xp = i == n1m - 1 ? c : P[pc + 1];
        }
        if (a.pbc2) {
            ym = P[pc - L1 + (j == 0 ? (size_t)n2m * L1 : 0)];
            yp = P[pc + L1 - (j == n2m - 1 ? (size_t)n2m * L1 : 0)];
        } else {
            ym = j == 0 ? c : P[pc - L1];
            yp = j == n2m - 1 ? c : P[pc + L1];
        }
        const size_t pl = (size_t)L1 * L2;
        const int kk = k + 1;
        const double amk = kk == 1 ? 0.0 : 1.0 / dx3[kk] / dmx3[kk];
        const double apk = kk == n3m ? 0.0 : 1.0 / dx3[kk] / dmx3[kk + 1];
        const double zm = kk == 1 ? 0.0 : P[pc - pl];
        const double zp = kk == n3m ? 0.0 : P[pc + pl];
        const double lp = (xp - 2.0 * c + xm) * a.ih1 + (yp - 2.0 * c + ym) * a.ih2 + amk * zm - (amk + apk) * c + apk * zp;
        const double fv = f[q];
        const double r = fv - lp;
        if (rout) rout[q] = r;
        sr += r * r;
        sf += fv * fv;
    }
    s_r[threadIdx.x] = sr;
    s_f[threadIdx.x] = sf;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            s_r[threadIdx.x] += s_r[threadIdx.x + s];
            s_f[threadIdx.x] += s_f[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = s_r[0];
        partial[2 * blockIdx.x + 1] = s_f[0];
    }
}

__global__ void __launch_bounds__(256) k_reduce_pairs(const double *__restrict__ partial, int n, double *__restrict__ out) {
    __shared__ double s_r[256], s_f[256];
    double a = 0.0, b = 0.0;
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
        a += partial[2 * q];
        b += partial[2 * q + 1];
    }
    s_r[threadIdx.x] = a;
    s_f[threadIdx.x] = b;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            s_r[threadIdx.x] += s_r[threadIdx.x + s];
            s_f[threadIdx.x] += s_f[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = s_r[0];
        out[1] = s_f[0];
    }
}

}  // namespace ppf
