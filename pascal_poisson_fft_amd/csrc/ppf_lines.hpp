// ppf_lines.hpp -- real/complex line transforms built on the LDS FFT.
//
// A "line" of real length N lives in LDS as M = N/2 complex values (real view:
// 2M scalars, plus room for one more complex value so half spectra of N+2
// scalars fit).  The operations and their conventions:
//
//   DCT-II forward   x[N] -> X[N]   X_k = 2 sum_n x_n cos(pi k (2n+1) / 2N)
//   DCT inverse      X[N] -> x[N]   DCT-III / (2N)  (scipy idct type 2)
//   R2C              x[N] -> V[N/2+1] complex, unnormalised
//   C2R              V[N/2+1] -> x[N], unnormalised, Im V_0, Im V_{N/2} ignored
//   C2C              complex line of length N (the FFT itself)
//
// DCT via Makhoul: v_n = x_{2n}, v_{N-1-n} = x_{2n+1} (the reorder of
// 01_Fortran/src/fft_3d_poisson.f90:924-933), real FFT of v, post twiddle
// X_k = 2 Re(exp(-i pi k / 2N) V_k) (:937-969); inverse: pre twiddle
// V_k = exp(i pi k / 2N)(X_k - i X_{N-k}) (:971-1011), C2R, inverse reorder
// (:1013-1033).  The length-N real FFT is a length-M complex FFT of
// z_m = v_{2m} + i v_{2m+1} with the usual even/odd split, so each real line
// costs one half-length complex FFT.
//
// Table layout per axis (device, type cx<T>):
//   tw      per-stage FFT twiddles, M-1 entries (ppf_fft.hpp)
//   wk[k]   = exp(-2 pi i k / N), k <= M       (real FFT split)
//   qk[k]   = exp(-i pi k / (2N)), k <= M      (DCT twiddle)
#pragma once

#include "ppf_fft.hpp"

namespace ppf {

template <int N, bool HALF>
struct LineGeom {
    static constexpr int M = N / 2;
    static constexpr int TPL = (M / max_radix(M)) > 0 ? (M / max_radix(M)) : 1;
    // complex slots per LDS line: M for DCT lines, M+1 when a half spectrum
    // (N+2 scalars) has to fit
    static constexpr int STRIDE = HALF ? M + 1 : M;
};

template <int N>
struct CLineGeom {  // complex line of length N (C2C)
    static constexpr int M = N;
    static constexpr int TPL = (M / max_radix(M)) > 0 ? (M / max_radix(M)) : 1;
    static constexpr int STRIDE = M + 1;
};

// Makhoul position of input sample n within v (also the scalar index inside
// the packed z array).
template <int N>
__device__ __forceinline__ int makhoul_pos(int n) {
    return (n & 1) ? (N - 1 - (n >> 1)) : (n >> 1);
}

// ---- forward post-processing -------------------------------------------
// After fft_forward on z, produce (for each k in [0, M]) V_k = E_k + w^k O_k.
// Pair (k, M-k) is handled by one thread.  Result written back as V[0..M]
// (M+1 complex).  Barrier inside.
template <typename T, int N, int TPL>
__device__ __forceinline__ void r2c_post(LLine<T> line, int t, const cx<T> *__restrict__ wk, bool act = true) {
    constexpr int M = N / 2;
    constexpr int NP = M / 2 + 1;  // pairs k = 0..M/2
    constexpr int NB = (NP + TPL - 1) / TPL;
    cx<T> va[NB], vb[NB];
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const int k = t + q * TPL;
        if (act && k < NP) {
            va[q] = line[k];
            vb[q] = line[(M - k) % M];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const int k = t + q * TPL;
        if (act && k < NP) {
            const cx<T> zk = va[q], zc = cconj(vb[q]);       // Z_k, conj(Z_{M-k})
            const cx<T> E = cscale(cadd(zk, zc), T(0.5));
            const cx<T> O = cmi(cscale(csub(zk, zc), T(0.5)));  // (Z_k - conj Z_{M-k}) / 2i
            if (k == 0) {
                line[0] = {E.x + O.x, T(0)};
                line[M] = {E.x - O.x, T(0)};
            } else {
                const cx<T> Vk = cadd(E, cmul(wk[k], O));
                line[k] = Vk;
                if (k != M - k) {
                    // V_{M-k} = conj(E) + w^{M-k} conj(O)
                    const cx<T> Vm = cadd(cconj(E), cmul(wk[M - k], cconj(O)));
                    line[M - k] = Vm;
                }
            }
        }
    }
    __syncthreads();
}

// DCT-II post: from Z (FFT of packed Makhoul sequence) to X[0..N-1] written as
// the real view of the line.  Barrier inside.
template <typename T, int N, int TPL>
__device__ __forceinline__ void dct_post(LLine<T> line, int t, const cx<T> *__restrict__ wk, const cx<T> *__restrict__ qk,
                                         bool act = true) {
    constexpr int M = N / 2;
    constexpr int NP = M / 2 + 1;
    constexpr int NB = (NP + TPL - 1) / TPL;
    cx<T> va[NB], vb[NB];
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const int k = t + q * TPL;
        if (act && k < NP) {
            va[q] = line[k];
            vb[q] = line[(M - k) % M];
        }
    }
    __syncthreads();
    const LReal<T> re = real_view(line);
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const int k = t + q * TPL;
        if (act && k < NP) {
            const cx<T> zk = va[q], zc = cconj(vb[q]);
            const cx<T> E = cscale(cadd(zk, zc), T(0.5));
            const cx<T> O = cmi(cscale(csub(zk, zc), T(0.5)));
            if (k == 0) {
                re[0] = T(2) * (E.x + O.x);
                // X_M = 2 Re(exp(-i pi/4) V_M), V_M = E_0 - O_0 real
                re[M] = T(2) * qk[M].x * (E.x - O.x);
            } else {
                const cx<T> Vk = cadd(E, cmul(wk[k], O));
                const cx<T> Wk = cmul(qk[k], Vk);
                re[k] = T(2) * Wk.x;
                re[N - k] = -T(2) * Wk.y;
                if (k != M - k) {
                    const int k2 = M - k;
                    const cx<T> V2 = cadd(cconj(E), cmul(wk[k2], cconj(O)));
                    const cx<T> W2 = cmul(qk[k2], V2);
                    re[k2] = T(2) * W2.x;
                    re[N - k2] = -T(2) * W2.y;
                }
            }
        }
    }
    __syncthreads();
}

// ---- inverse pre-processing ----------------------------------------------
// From half spectrum V[0..M] (in the line) to conj(Y) where
// Y_k = (V_k + conj V_{M-k}) + i exp(2 pi i k/N) (V_k - conj V_{M-k}), k < M,
// so that a forward FFT followed by conjugation yields the unnormalised
// inverse z (z_m = v_{2m} + i v_{2m+1}).  Im V_0 and Im V_M are ignored.
template <typename T, int N, int TPL>
__device__ __forceinline__ void c2r_pre(LLine<T> line, int t, const cx<T> *__restrict__ wk, bool act = true) {
    constexpr int M = N / 2;
    constexpr int NP = M / 2 + 1;
    constexpr int NB = (NP + TPL - 1) / TPL;
    cx<T> va[NB], vb[NB];
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const int k = t + q * TPL;
        if (act && k < NP) {
            va[q] = line[k];
            vb[q] = line[M - k];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const int k = t + q * TPL;
        if (act && k < NP) {
            cx<T> Vk = va[q], Vm = vb[q];
            if (k == 0) { Vk.y = T(0); Vm.y = T(0); }
            // Y_k
            {
                const cx<T> s = cadd(Vk, cconj(Vm));
                const cx<T> d = csub(Vk, cconj(Vm));
                const cx<T> e = cconj(wk[k]);                 // exp(+2 pi i k / N)
                const cx<T> ied = cmul(cx<T>{-e.y, e.x}, d);  // i e d
                line[k] = cconj(cadd(s, ied));
            }
            if (k != 0 && k != M - k) {
                const int k2 = M - k;
                const cx<T> s = cadd(Vm, cconj(Vk));
                const cx<T> d = csub(Vm, cconj(Vk));
                const cx<T> e = cconj(wk[k2]);
                const cx<T> ied = cmul(cx<T>{-e.y, e.x}, d);
                line[k2] = cconj(cadd(s, ied));
            }
        }
    }
    __syncthreads();
}

// DCT inverse pre: from X[0..N-1] (real view of the line) to conj(Y) with
// V_k = exp(i pi k/2N)(X_k - i X_{N-k}), X_N = 0.  Barrier inside.
template <typename T, int N, int TPL>
__device__ __forceinline__ void idct_pre(LLine<T> line, int t, const cx<T> *__restrict__ wk, const cx<T> *__restrict__ qk,
                                         bool act = true) {
    constexpr int M = N / 2;
    constexpr int NP = M / 2 + 1;
    constexpr int NB = (NP + TPL - 1) / TPL;
    const LReal<T> re = real_view(line);
    T x0[NB], x1[NB], x2[NB], x3[NB];  // X_k, X_{N-k}, X_{M-k}, X_{M+k}
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const int k = t + q * TPL;
        if (act && k < NP) {
            x0[q] = re[k];
            x1[q] = (k == 0) ? T(0) : re[N - k];
            x2[q] = re[M - k];
            x3[q] = re[M + k];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const int k = t + q * TPL;
        if (act && k < NP) {
            const int k2 = M - k;
            // V_k = conj(q_k) (X_k - i X_{N-k});  V_{M-k} likewise with X_{M-k}, X_{M+k}
            const cx<T> Vk = cmul(cconj(qk[k]), cx<T>{x0[q], -x1[q]});
            const cx<T> Vm = cmul(cconj(qk[k2]), cx<T>{x2[q], -x3[q]});
            {
                const cx<T> s = cadd(Vk, cconj(Vm));
                const cx<T> d = csub(Vk, cconj(Vm));
                const cx<T> e = cconj(wk[k]);
                line[k] = cconj(cadd(s, cmul(cx<T>{-e.y, e.x}, d)));
            }
            if (k != 0 && k != k2) {
                const cx<T> s = cadd(Vm, cconj(Vk));
                const cx<T> d = csub(Vm, cconj(Vk));
                const cx<T> e = cconj(wk[k2]);
                line[k2] = cconj(cadd(s, cmul(cx<T>{-e.y, e.x}, d)));
            }
        }
    }
    __syncthreads();
}

}  // namespace ppf
