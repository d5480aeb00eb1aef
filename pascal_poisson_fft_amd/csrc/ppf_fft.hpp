// ppf_fft.hpp -- LDS-resident complex FFT for gfx950 (Stockham autosort).
//
// A workgroup transforms several independent lines at once; each line is M
// complex values in LDS and is served by TPL threads.  Stages use radix 8, 7,
// 5, 4, 3 or 2, chosen at compile time from M = 2^a 3^b 5^c 7^d.  Each stage reads its
// butterfly inputs into registers, barriers, and writes the outputs back in
// place (Stockham ordering), so one LDS buffer per line suffices and the
// result lands in natural order.
//
// Twiddles come from a per-stage table (stage_twiddles): the stage of span
// LS and radix R owns entries [LS-1, LS-1 + (R-1) LS) holding
// exp(-2 pi i j r / (LS R)) at (r-1) LS + j, so the lanes of a line (j
// consecutive) read consecutive entries.  The spans telescope, so the table
// has M-1 entries.  Forward transforms use sign -1; inverse transforms are
// obtained by the caller by conjugating on load and on store
// (IDFT(x) = conj(DFT(conj(x)))).
#pragma once

#include <hip/hip_runtime.h>

namespace ppf {

// complex value; PPF_CX_ALIGN aligns it to its size (16 B in FP64) so LDS
// and global accesses can be single 128-bit operations
#ifndef PPF_CX_ALIGN
#define PPF_CX_ALIGN 0
#endif
template <typename T>
struct alignas(PPF_CX_ALIGN ? 2 * sizeof(T) : alignof(T)) cx {
    T x, y;
};

template <typename T>
__host__ __device__ __forceinline__ cx<T> cadd(cx<T> a, cx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__host__ __device__ __forceinline__ cx<T> csub(cx<T> a, cx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__host__ __device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
// multiply by -i
template <typename T>
__host__ __device__ __forceinline__ cx<T> cmi(cx<T> a) { return {a.y, -a.x}; }
template <typename T>
__host__ __device__ __forceinline__ cx<T> cconj(cx<T> a) { return {a.x, -a.y}; }
template <typename T>
__host__ __device__ __forceinline__ cx<T> cscale(cx<T> a, T s) { return {a.x * s, a.y * s}; }

// ---- swizzled LDS lines ----------------------------------------------------
// A Stockham stage writes its butterfly outputs R complex values apart
// (R*16 bytes for FP64), which in a plain layout puts consecutive lanes on
// the same LDS banks.  Complex index i is stored at i ^ ((i >> 3) & 7): the
// low three bits are permuted by the next three, a bijection inside every
// 8-value block, so 8 consecutive lanes at a stride of 1, 2, 4 or 8 values
// land in 8 different 16-byte bank groups, at no extra LDS.
__host__ __device__ constexpr int lds_pad(int i) { return i ^ ((i >> 3) & 7); }
// slots for indices < m (an index past the last full block may move up to 7)
__host__ __device__ constexpr int lds_slots(int m) { return ((m + 7) / 8) * 8; }

// complex view of a swizzled line
template <typename T>
struct LLine {
    cx<T> *p;
    __device__ __forceinline__ cx<T> &operator[](int i) const { return p[lds_pad(i)]; }
};
// scalar view of the same line: scalar r is component r&1 of complex r>>1
template <typename T>
struct LReal {
    T *p;
    __device__ __forceinline__ T &operator[](int r) const { return p[2 * lds_pad(r >> 1) + (r & 1)]; }
};
template <typename T>
__device__ __forceinline__ LReal<T> real_view(LLine<T> l) {
    return LReal<T>{reinterpret_cast<T *>(l.p)};
}

// ---- compile-time radix plan -------------------------------------------
constexpr bool fft_size_ok(int m) {
    if (m < 1) return false;
    for (int p : {3, 5, 7})
        while (m % p == 0) m /= p;
    while (m % 2 == 0) m /= 2;
    return m == 1;
}
// radix of the next stage given the remaining length: the odd factors first,
// then 8s (16 as 4 x 4), 4s and a 2
constexpr int pick_radix(int rem) {
    return (rem % 7 == 0)   ? 7
           : (rem % 5 == 0) ? 5
           : (rem % 3 == 0) ? 3
                            : ((rem % 8 == 0 && rem != 16) ? 8 : ((rem % 4 == 0) ? 4 : 2));
}
constexpr int max_radix(int m) {
    int r = 1;
    while (m > 1) {
        int p = pick_radix(m);
        if (p > r) r = p;
        m /= p;
    }
    return r;
}

// ---- forward DFT butterflies ---------------------------------------------
template <typename T, int R>
struct Dft;

template <typename T>
struct Dft<T, 2> {
    __host__ __device__ __forceinline__ static void run(cx<T> *v) {
        cx<T> a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <typename T>
struct Dft<T, 3> {
    __host__ __device__ __forceinline__ static void run(cx<T> *v) {
        const T h = T(0.86602540378443864676372317075294);  // sqrt(3)/2
        cx<T> a = v[0], b = v[1], c = v[2];
        cx<T> s = cadd(b, c);
        cx<T> t = {a.x - T(0.5) * s.x, a.y - T(0.5) * s.y};
        cx<T> z = {h * (b.x - c.x), h * (b.y - c.y)};
        v[0] = cadd(a, s);
        v[1] = cadd(t, cmi(z));
        v[2] = csub(t, cmi(z));
    }
};

template <typename T>
struct Dft<T, 5> {
    // X_m = b_m -+ i c_m over the conjugate pairs (1,4), (2,3)
    __host__ __device__ __forceinline__ static void run(cx<T> *v) {
        const T c1 = T(0.30901699437494742410229341718281906), c2 = T(-0.80901699437494742410229341718281906);
        const T s1 = T(0.95105651629515357211643933337938214), s2 = T(0.58778525229247312916870595463907277);
        const cx<T> a0 = v[0];
        const cx<T> t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
        const cx<T> u1 = csub(v[1], v[4]), u2 = csub(v[2], v[3]);
        const cx<T> b1 = {a0.x + c1 * t1.x + c2 * t2.x, a0.y + c1 * t1.y + c2 * t2.y};
        const cx<T> b2 = {a0.x + c2 * t1.x + c1 * t2.x, a0.y + c2 * t1.y + c1 * t2.y};
        const cx<T> e1 = cmi(cx<T>{s1 * u1.x + s2 * u2.x, s1 * u1.y + s2 * u2.y});  // -i c_1
        const cx<T> e2 = cmi(cx<T>{s2 * u1.x - s1 * u2.x, s2 * u1.y - s1 * u2.y});  // -i c_2
        v[0] = cadd(a0, cadd(t1, t2));
        v[1] = cadd(b1, e1);
        v[4] = csub(b1, e1);
        v[2] = cadd(b2, e2);
        v[3] = csub(b2, e2);
    }
};

template <typename T>
struct Dft<T, 7> {
    // X_m = b_m -+ i c_m over the conjugate pairs (1,6), (2,5), (3,4)
    __host__ __device__ __forceinline__ static void run(cx<T> *v) {
        const T c1 = T(0.62348980185873353052500488400423981), c2 = T(-0.22252093395631440428890256449679476),
                c3 = T(-0.90096886790241912623610231950744505);
        const T s1 = T(0.78183148246802980870844452667405775), s2 = T(0.97492791218182360701813168299393122),
                s3 = T(0.43388373911755812047576833284835875);
        const cx<T> a0 = v[0];
        const cx<T> t1 = cadd(v[1], v[6]), t2 = cadd(v[2], v[5]), t3 = cadd(v[3], v[4]);
        const cx<T> u1 = csub(v[1], v[6]), u2 = csub(v[2], v[5]), u3 = csub(v[3], v[4]);
        auto lin = [](T a, cx<T> p, T b, cx<T> q, T c, cx<T> r) {
            return cx<T>{a * p.x + b * q.x + c * r.x, a * p.y + b * q.y + c * r.y};
        };
        const cx<T> b1 = cadd(a0, lin(c1, t1, c2, t2, c3, t3));
        const cx<T> b2 = cadd(a0, lin(c2, t1, c3, t2, c1, t3));
        const cx<T> b3 = cadd(a0, lin(c3, t1, c1, t2, c2, t3));
        const cx<T> e1 = cmi(lin(s1, u1, s2, u2, s3, u3));
        const cx<T> e2 = cmi(lin(s2, u1, -s3, u2, -s1, u3));
        const cx<T> e3 = cmi(lin(s3, u1, -s1, u2, s2, u3));
        v[0] = cadd(a0, cadd(t1, cadd(t2, t3)));
        v[1] = cadd(b1, e1);
        v[6] = csub(b1, e1);
        v[2] = cadd(b2, e2);
        v[5] = csub(b2, e2);
        v[3] = cadd(b3, e3);
        v[4] = csub(b3, e3);
    }
};

template <typename T>
struct Dft<T, 4> {
    __host__ __device__ __forceinline__ static void run(cx<T> *v) {
        cx<T> s0 = cadd(v[0], v[2]), d0 = csub(v[0], v[2]);
        cx<T> s1 = cadd(v[1], v[3]), d1 = cmi(csub(v[1], v[3]));
        v[0] = cadd(s0, s1);
        v[2] = csub(s0, s1);
        v[1] = cadd(d0, d1);
        v[3] = csub(d0, d1);
    }
};

template <typename T>
struct Dft<T, 8> {
    __host__ __device__ __forceinline__ static void run(cx<T> *v) {
        const T r = T(0.70710678118654752440084436210485);
        cx<T> e[4] = {v[0], v[2], v[4], v[6]};
        cx<T> o[4] = {v[1], v[3], v[5], v[7]};
        Dft<T, 4>::run(e);
        Dft<T, 4>::run(o);
        // o_k *= exp(-2 pi i k / 8)
        cx<T> o1 = {r * (o[1].x + o[1].y), r * (o[1].y - o[1].x)};
        cx<T> o2 = cmi(o[2]);
        cx<T> o3 = {r * (o[3].y - o[3].x), -r * (o[3].x + o[3].y)};
        v[0] = cadd(e[0], o[0]);
        v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], o1);
        v[5] = csub(e[1], o1);
        v[2] = cadd(e[2], o2);
        v[6] = csub(e[2], o2);
        v[3] = cadd(e[3], o3);
        v[7] = csub(e[3], o3);
    }
};

// ---- one Stockham stage ---------------------------------------------------
// line: this thread's LDS line (M complex, padded), t: thread within the line.
// act = false: a thread with no line in this workgroup; it only takes the
// barriers (every thread of the workgroup must).
template <typename T, int M, int TPL, int R, int LS>
__device__ __forceinline__ void fft_stage(LLine<T> line, int t, const cx<T> *__restrict__ tw, bool act = true) {
    constexpr int S = M / R;                  // butterflies per line
    constexpr int NB = (S + TPL - 1) / TPL;   // butterflies per thread
    cx<T> v[NB][R];
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const int b = t + q * TPL;
        if (act && ((S % TPL == 0) || b < S)) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[q][r] = line[b + r * S];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const int b = t + q * TPL;
        if (act && ((S % TPL == 0) || b < S)) {
            const int j = b % LS;
            const int kk = b / LS;
            if (LS > 1) {
                const cx<T> *w = tw + (LS - 1) + j;  // w[(r - 1) * LS] = exp(-2 pi i j r / (LS R))
                if constexpr (sizeof(T) == 8 && (R == 8 || R == 4)) {
                    // FP64: powers 3, 5, 6, 7 as products of loaded powers 1, 2, 4
                    // (a few ulp), trading LDS reads shared by the CU for VALU work
                    const cx<T> w1 = w[0], w2 = w[LS];
                    const cx<T> w3 = cmul(w1, w2);
                    v[q][1] = cmul(v[q][1], w1);
                    v[q][2] = cmul(v[q][2], w2);
                    v[q][3] = cmul(v[q][3], w3);
                    if constexpr (R == 8) {
                        const cx<T> w4 = w[3 * LS];
                        v[q][4] = cmul(v[q][4], w4);
                        v[q][5] = cmul(v[q][5], cmul(w1, w4));
                        v[q][6] = cmul(v[q][6], cmul(w2, w4));
                        v[q][7] = cmul(v[q][7], cmul(w3, w4));
                    }
                } else {
#pragma unroll
                    for (int r = 1; r < R; ++r) v[q][r] = cmul(v[q][r], w[(r - 1) * LS]);
                }
            }
            Dft<T, R>::run(v[q]);
#pragma unroll
            for (int r = 0; r < R; ++r) line[kk * LS * R + j + r * LS] = v[q][r];
        }
    }
    __syncthreads();
}

template <typename T, int M, int TPL, int LS>
struct FftStages {
    __device__ __forceinline__ static void run(LLine<T> line, int t, const cx<T> *__restrict__ tw, bool act) {
        constexpr int R = pick_radix(M / LS);
        fft_stage<T, M, TPL, R, LS>(line, t, tw, act);
        FftStages<T, M, TPL, LS * R>::run(line, t, tw, act);
    }
};
template <typename T, int M, int TPL>
struct FftStages<T, M, TPL, M> {
    __device__ __forceinline__ static void run(LLine<T>, int, const cx<T> *__restrict__, bool) {}
};

// Forward FFT (sign -1) of one LDS line, in place, natural order in and out.
// Every thread of the workgroup must call it (it contains barriers).
template <typename T, int M, int TPL>
__device__ __forceinline__ void fft_forward(LLine<T> line, int t, const cx<T> *__restrict__ tw, bool act = true) {
    static_assert(fft_size_ok(M), "FFT length must be 2^a 3^b 5^c 7^d");
    FftStages<T, M, TPL, 1>::run(line, t, tw, act);
}

}  // namespace ppf
