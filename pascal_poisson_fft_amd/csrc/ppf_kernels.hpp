// ppf_kernels.hpp -- the five passes of the solve for gfx950.
//
// Workspace W (type T) holds one spectral field as Lz planes of Ly rows of
// Lxs scalars (i < Lx = n1m (Neumann x) or n1m+2 (periodic x, half spectrum
// as interleaved re/im), Lxs = Lx rounded up to 128 bytes, Ly = n2m+2,
// Lz = n3m+2).  Inside a plane the rows are cut into 128-byte column tiles
// of WT = 128/sizeof(T) scalars, each tile a contiguous block of Ly rows:
//   W[kp * Ly * Lxs + (i / WT) * Ly * WT + jp * WT + i % WT]   (w_at)
// so the y passes' column strips and the x passes' 8-line stores are
// runs of consecutive bytes (a strip of the row-major layout is Ly segments
// of 128 bytes, Lxs apart).
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

// Row stride of the spectral workspace: the Lx scalars of a row rounded up to
// a multiple of 128 bytes, so every column tile of the y and z passes starts
// on a cache-line boundary (a periodic-x row of n1m+2 scalars otherwise
// straddles two lines in every tile segment).
__host__ __device__ constexpr int w_stride(int lx, int tbytes) { return (lx * tbytes + 127) / 128 * 128 / tbytes; }

// Column tiles of W: WT scalars (128 bytes) wide, Ly rows tall, contiguous.
template <typename T>
__host__ __device__ constexpr int w_tile() { return 128 / (int)sizeof(T); }
// offset of column c (row 0) inside a plane of Ly rows
template <typename T>
__host__ __device__ inline size_t w_col(int c, int Ly) {
    constexpr int WT = w_tile<T>();
    return (size_t)(c / WT) * (size_t)Ly * WT + (size_t)(c % WT);
}
// offset of (row j, column c) inside a plane; rows of one column are WT apart
template <typename T>
__host__ __device__ inline size_t w_at(int j, int c, int Ly) { return w_col<T>(c, Ly) + (size_t)j * w_tile<T>(); }

// Gate of the device-side refinement test: a pass launched with a gate runs
// only while gate[0] is nonzero (more sweeps needed); nullptr = always.  The
// test is uniform over the grid, so a gated kernel returns before any barrier.
__device__ __forceinline__ bool gated_off(const int *gate) { return gate && *gate == 0; }

// Streaming global accesses of the pass kernels: every element of a pass is
// read once and written once, so PPF_NT_LOAD / PPF_NT_STORE (default on)
// mark them nontemporal (no reuse expected in the caches).
#ifndef PPF_NT_LOAD
#define PPF_NT_LOAD 1
#endif
#ifndef PPF_NT_STORE
#define PPF_NT_STORE 1
#endif
template <typename T>
__device__ __forceinline__ T ld_stream(const T *p) {
    if constexpr (PPF_NT_LOAD) return __builtin_nontemporal_load(p);
    else return *p;
}
template <typename T>
__device__ __forceinline__ void st_stream(T *p, T v) {
    if constexpr (PPF_NT_STORE) __builtin_nontemporal_store(v, p);
    else *p = v;
}


// Scalars of padding between consecutive LDS lines (odd: lines start on
// different banks; 2: lines stay 16-byte aligned).
#ifndef PPF_LINE_PAD
#define PPF_LINE_PAD 2
#endif

// Rounds in which a y-pass workgroup transforms its tile's lines.
#ifndef PPF_YROUNDS
#define PPF_YROUNDS 1
#endif

// Bytes of each W row a y-pass tile covers (its scalar columns x sizeof(T)).
#ifndef PPF_YROW_BYTES
#define PPF_YROW_BYTES 128
#endif

// Register budgets of the line-transform kernels, as minimum waves per SIMD
// (occupancy is what hides the LDS and HBM latency of these kernels).
#ifndef PPF_XWPE
#define PPF_XWPE 3
#endif
#ifndef PPF_YWPE
#define PPF_YWPE 1
#endif

// LDS layout of the line-transform kernels: NL lines of LSD scalars each
// (LSD odd, so consecutive lines start on different banks and the
// column-to-line scatter of the y passes is conflict-free), followed by the
// axis tables tw[M] | wk[M+1] | qk[M+1], copied from global memory on entry
// together with the data so no FFT stage waits on a global round trip.  QK:
// the DCT's qk table is held (only DCT lines read it; the real-FFT and
// complex forms leave it out, which at 1024 points lets a third x-pass
// workgroup onto the CU).
template <int M, bool QK>
constexpr int lds_table_slots() { return M + (M + 1) + (QK ? M + 1 : 0); }
template <typename T, int M, int S, int NL, bool QK = true, int PAD = PPF_LINE_PAD>
struct LdsLines {
    static constexpr int LSD = 2 * lds_slots(S) + PAD;  // S complex values + pad scalars
    static constexpr int TABN = lds_table_slots<M, QK>();
    static constexpr int SIZE = NL * LSD + 2 * TABN;  // scalars of type T
    __device__ __forceinline__ static LLine<T> line(T *lds, int l) {
        return LLine<T>{reinterpret_cast<cx<T> *>(lds + l * LSD)};
    }
    __device__ __forceinline__ static LReal<T> real(T *lds, int l) { return LReal<T>{lds + l * LSD}; }
    __device__ __forceinline__ static cx<T> *tables(T *lds) { return reinterpret_cast<cx<T> *>(lds + NL * LSD); }
    __device__ __forceinline__ static void load_tables(T *lds, const cx<T> *__restrict__ tw, const cx<T> *__restrict__ wk,
                                                       const cx<T> *__restrict__ qk, int tid, int nt) {
        cx<T> *tab = tables(lds);
        for (int i = tid; i < TABN; i += nt) tab[i] = i < M ? tw[i] : (i <= 2 * M ? wk[i - M] : qk[i - 2 * M - 1]);
    }
};

// ---------------------------------------------------------------------------
// pass 1: x-forward.  One LDS line per prhs row; LPB rows per workgroup.
// PER: periodic x (R2C, Lx = N+2) else DCT-II (Lx = N).
// XV: the kernel form of the tuner's variant -- 0: one radix-R butterfly per
// thread per stage; 1: half the threads per line, two butterflies each
// (twice the lines per workgroup); 2: as 0 without the next tile's prefetch
// held in registers across the transform (fewer registers, more waves).
// Variant 3 is the XV = 2 kernel launched with one workgroup per tile.
// ---------------------------------------------------------------------------
#ifndef PPF_XLITE_WPE
#define PPF_XLITE_WPE 4
#endif
// Pad of an x-pass LDS line: the line stride is 128 bytes modulo 256, so in
// the column-tile order of the W accesses (8 lanes of 16 bytes per line,
// consecutive lines next) the two lines of one 16-lane LDS pass fall on
// disjoint banks.
template <typename T>
constexpr int x_line_pad(int s) {
    const int base = (2 * lds_slots(s) * (int)sizeof(T)) % 256;
    return ((128 - base + 256) % 256) / (int)sizeof(T);
}

template <typename T, int N, bool PER, int XV = 0>
struct XGeom {
    using G = LineGeom<N, PER>;
    static constexpr bool LITE = XV == 2;
    static constexpr int TPL = (G::TPL >> (XV == 1 ? 1 : 0)) > 0 ? (G::TPL >> (XV == 1 ? 1 : 0)) : 1;
    static constexpr int LPB = (256 / TPL) > 0 ? (256 / TPL) : 1;
    static constexpr int NT = LPB * TPL;
    // the bank pad where a thread keeps one line per tile (the longer lines:
    // from 128 points in FP64, 256 in FP32); on shorter lines the larger tile
    // cost more occupancy than the pad gained (96-problem batch 5.57-5.60
    // against 5.60-5.64 ms)
    static constexpr int PADX = x_line_pad<T>(G::STRIDE);
    static constexpr bool PAD_USE =
        NT % ((w_tile<T>() / 2) * LPB) == 0 &&
        LPB * (2 * lds_slots(G::STRIDE) + PADX) + 2 * lds_table_slots<G::M, !PER>() <= (160 * 1024) / (int)sizeof(T);
    using L = LdsLines<T, G::M, G::STRIDE, LPB, !PER, PAD_USE ? PADX : PPF_LINE_PAD>;
    // register target: PPF_XWPE waves per SIMD, or fewer where the LDS tile
    // (very short or 1024-point lines) admits fewer workgroups per CU
    static constexpr int WG_LDS = (160 * 1024) / (L::SIZE * (int)sizeof(T));
    static constexpr int WPE_LDS = (WG_LDS * (NT / 64)) / 4 > 0 ? (WG_LDS * (NT / 64)) / 4 : 1;
    static constexpr int WPE_REG = LITE ? PPF_XLITE_WPE : PPF_XWPE;
    static constexpr int WPE = WPE_REG < WPE_LDS ? WPE_REG : WPE_LDS;
};

// q = a / b for 0 <= a < 2^24 via a float reciprocal and one correction
__device__ __forceinline__ int fdiv_small(int a, int b, float rb) {
    int q = (int)((float)a * rb);
    q -= (q * b > a);
    q += ((q + 1) * b <= a);
    return q;
}

// z row k (0-based) of the stretched operator, fft_3d_poisson.f90:1143-1146:
// am couples to row k-1 (0 on the bottom row), ap to row k+1 (0 on the top).
__device__ __forceinline__ void z_row(const double *__restrict__ dx3, const double *__restrict__ dmx3, int n3m, int k,
                                      double &am, double &ap) {
    const int kk = k + 1;
    am = kk == 1 ? 0.0 : 1.0 / dx3[kk] / dmx3[kk];
    ap = kk == n3m ? 0.0 : 1.0 / dx3[kk] / dmx3[kk + 1];
}
// z wall extrapolation p0 = BC_a p1 - BC_b p2 (:829-837): bc[0..1] at the
// bottom wall, bc[2..3] at the top wall.
__device__ __forceinline__ void z_wall_bc(const double *__restrict__ dmx3, int n3m, double *bc) {
    const double d1 = dmx3[1], d2 = dmx3[2], s0 = d1 + d2;
    bc[0] = s0 * s0 / (s0 * s0 - d1 * d1);
    bc[1] = d1 * d1 / (s0 * s0 - d1 * d1);
    const double e1 = dmx3[n3m + 1], e2 = dmx3[n3m], s1 = e1 + e2;
    bc[2] = s1 * s1 / (s1 * s1 - e1 * e1);
    bc[3] = e1 * e1 / (s1 * s1 - e1 * e1);
}

// The per-solve z row table (zam[k], zap[k], then the 4 wall coefficients at
// zap[n3m..]); pass 1 fills it (workgroup 0) when zam is set.
struct ZRows {
    const double *dx3, *dmx3;
    double *zam, *zap;
    int n3m;
};
__device__ __forceinline__ void fill_z_rows(const ZRows &zr, int tid, int nt) {
    for (int k = tid; k < zr.n3m; k += nt) z_row(zr.dx3, zr.dmx3, zr.n3m, k, zr.zam[k], zr.zap[k]);
    if (tid == 0) z_wall_bc(zr.dmx3, zr.n3m, zr.zap + zr.n3m);
}

// Inputs of the FP64 residual r = f - L x on interior cells, x the ghosted
// FP64 iterate.  L is the operator the solve inverts: uniform x/y spacing with
// zero-flux walls (the DCT eigenvalues) or wrap, and the stretched z rows
// zam/zap of fft_3d_poisson.f90:1143-1146.
struct XRes {
    const double *P;          // (n1m+2)(n2m+2)(n3m+2) iterate
    const double *zam, *zap;  // z rows (zero at the walls)
    double *partial;          // per-workgroup sums of r^2, f^2 and x (3 per workgroup)
    double ih1, ih2;          // 1/h^2
    int pbc1, pbc2;
    float *rout;              // nullptr, or W32: r stored in FP32 where pass 1 reads line (j,k)
    int Ly, Lxs;              // W32 rows per plane and scalars per row (plane = Ly * Lxs)
};

// Sum of v over a workgroup of NT <= 256 threads (any count); thread 0 gets
// the total.  red holds NT doubles of LDS.  Fixed order: deterministic.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double *red) {
    const int tid = threadIdx.x;
    red[tid] = v;
    __syncthreads();
#pragma unroll
    for (int h = 128; h > 0; h >>= 1) {
        if (tid < h && tid + h < NT) red[tid] += red[tid + h];
        __syncthreads();
    }
    const double t = red[0];
    __syncthreads();
    return t;
}

// The line-transform kernels walk tiles tile = blockIdx.x, blockIdx.x +
// gridDim.x, ...  On a persistent grid (as many workgroups as fit) the forms
// with the prefetch are software-pipelined: the next tile's global loads are
// issued into registers right after the current tile has been scattered
// into LDS, so they are in flight while the current tile is transformed and
// stored (barriers wait on LDS only).  Every form issues its first tile's
// loads before staging the twiddle tables; on a grid of one workgroup per
// tile (x variant 3) that is the only overlap inside a workgroup, and the
// other workgroups on the CU supply the rest.
//
// RES: the refinement form -- the line is the residual r = f - L x that
// k_residual stored in FP32 in this line's own W row (in is unused), and the
// transform overwrites it there.
template <typename T, int N, bool PER, bool RES = false, int XV = 0>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(XGeom<T, N, PER, XV>::WPE)))
k_xfwd(const double *__restrict__ in, T *W, int nlines, int n2m, int Ly,
       const cx<T> *__restrict__ tw, const cx<T> *__restrict__ wk, const cx<T> *__restrict__ qk,
       ZRows zr, const int *__restrict__ gate) {
    if (gated_off(gate)) return;
    using X = XGeom<T, N, PER, XV>;
    using G = typename X::G;
    using L = typename X::L;
    constexpr int M = G::M, TPL = X::TPL, LPB = X::LPB, NT = X::NT;
    constexpr int LX = PER ? N + 2 : N;
    constexpr int LXS = w_stride(LX, (int)sizeof(T));  // scalars per W row (a plane is Ly * LXS)
    constexpr int H = N / 2, HX = LX / 2;          // sample pairs per line / per W row
    // W-side element order: PT pairs fill one 128-byte row of a column tile;
    // consecutive lanes take the PT pairs of one line, then the same pairs
    // of the next line, so a wave covers consecutive bytes of a tile.  The
    // prhs side (row-major lines) keeps line-major order.
    constexpr int PT = w_tile<T>() / 2;
    constexpr int HR = RES ? (H + PT - 1) / PT * PT : H;  // pairs per line in the load order
    constexpr int EPT = (LPB * HR + NT - 1) / NT;        // pair loads per thread per tile
    using D2 = double __attribute__((ext_vector_type(2)));
    using V2 = T __attribute__((ext_vector_type(2)));
    __shared__ __align__(16) T lds[L::SIZE];
    const int tid = threadIdx.x, l = tid / TPL, t = tid - l * TPL;
    const int ntiles = (nlines + LPB - 1) / LPB;
    const float rn2m = 1.0f / (float)n2m;
    constexpr int WT = w_tile<T>();
    constexpr bool ONE_ROW = NT % (PT * LPB) == 0;    // W side: thread -> one line per tile
    constexpr int HSTEP = NT / LPB;                   // its pair step (ONE_ROW)
    const int h0 = (tid / (PT * LPB)) * PT + tid % PT;  // its first pair (ONE_ROW)
    // element e of a tile's loads -> (line li of the tile, pair h)
    auto emap = [&](int e, int &li, int &h) {
        if constexpr (RES) {
            const int r = e / PT;
            li = r % LPB;
            h = (r / LPB) * PT + (e - r * PT);
        } else {
            li = e / H;
            h = e - li * H;
        }
    };
    D2 pre[EPT];
    // unconditional 16-byte loads of sample pairs from clamped addresses
    // (lines past the end are transformed but never stored); nothing
    // consumes them before the next tile, so they stay in flight
    // W base of line (j,k) of this tile: plane k+1, row j+1
    auto wline = [&](int line) -> T * {
        const int k = fdiv_small(line, n2m, rn2m);
        return W + (size_t)(k + 1) * Ly * LXS + (size_t)(line - k * n2m + 1) * WT;
    };
    auto fetch = [&](int tile) {
        if constexpr (RES && ONE_ROW) {
            int line = tile * LPB + (tid / PT) % LPB;
            line = line < nlines ? line : nlines - 1;
            const T *base = wline(line);
            const T *p0 = base + w_col<T>(2 * h0, Ly), *plast = base + w_col<T>(2 * (H - 1), Ly);
            const size_t qstep = (size_t)(HSTEP / PT) * Ly * WT;
#pragma unroll
            for (int q = 0; q < EPT; ++q) {
                const T *a = h0 + q * HSTEP < H ? p0 + q * qstep : plast;
                const V2 v = ld_stream(reinterpret_cast<const V2 *>(a));
                pre[q] = D2{(double)v.x, (double)v.y};
            }
        } else {
#pragma unroll
            for (int q = 0; q < EPT; ++q) {
                const int e0 = tid + q * NT;
                const int e = e0 < LPB * HR ? e0 : LPB * HR - 1;
                int li, h;
                emap(e, li, h);
                const int n = 2 * (h < H ? h : H - 1);
                const int line0 = tile * LPB + li;
                const int line = line0 < nlines ? line0 : nlines - 1;
                if constexpr (RES) {
                    const V2 v = ld_stream(reinterpret_cast<const V2 *>(wline(line) + w_col<T>(n, Ly)));
                    pre[q] = D2{(double)v.x, (double)v.y};
                } else {
                    pre[q] = ld_stream(reinterpret_cast<const D2 *>(&in[(size_t)line * N + n]));
                }
            }
        }
    };
    // the first tile's loads (every form) overlap the set-up below
    int tile = blockIdx.x;
    if (tile < ntiles) fetch(tile);
    // the z row table of this solve, for pass 3 (saves a launch)
    if (zr.zam && blockIdx.x == 0) fill_z_rows(zr, tid, NT);
    L::load_tables(lds, tw, wk, qk, tid, NT);
    const cx<T> *tab = L::tables(lds);
    const LLine<T> my = L::line(lds, l);
    for (; tile < ntiles; tile += gridDim.x) {
        __syncthreads();  // the previous tile's LDS lines have been stored
        if (X::LITE && tile != (int)blockIdx.x) fetch(tile);
#pragma unroll
        for (int q = 0; q < EPT; ++q) {
            const int e = tid + q * NT;
            int li, h;
            if constexpr (RES && ONE_ROW) {
                li = (tid / PT) % LPB;
                h = h0 + q * HSTEP;
            } else {
                li = 0;
                h = H;
                if (e < LPB * HR) emap(e, li, h);
            }
            if (h < H) {
                const D2 v = pre[q];
                if (PER) {  // samples 2h, 2h+1 are complex h of the packed line
                    L::line(lds, li)[h] = cx<T>{(T)v.x, (T)v.y};
                } else {    // Makhoul: x_2h -> v_h, x_2h+1 -> v_{N-1-h}
                    L::real(lds, li)[h] = (T)v.x;
                    L::real(lds, li)[N - 1 - h] = (T)v.y;
                }
            }
        }
        if (!X::LITE && tile + (int)gridDim.x < ntiles) fetch(tile + gridDim.x);
        __syncthreads();
        fft_forward<T, M, TPL>(my, t, tab);
        if (PER)
            r2c_post<T, N, TPL>(my, t, tab + M);
        else
            dct_post<T, N, TPL>(my, t, tab + M, tab + 2 * M + 1);
        // prhs line (j,k) goes to W plane k+1, row j+1, in the tile order
        // above.  The R2C row's Nyquist pair (h = N/2) is stored apart, so the
        // main loop covers N/2 pairs per line.
        constexpr int HM = PER ? HX - 1 : HX;
        constexpr int HMR = (HM + PT - 1) / PT * PT;
        auto store_pair = [&](int li, int h) {
            const int line = tile * LPB + li;
            if (line < nlines) {
                const cx<T> z = L::line(lds, li)[h];
                st_stream(reinterpret_cast<V2 *>(wline(line) + w_col<T>(2 * h, Ly)), V2{z.x, z.y});
            }
        };
        if constexpr (ONE_ROW) {
            const int li = (tid / PT) % LPB, line = tile * LPB + li;
            if (line < nlines) {
                T *a = wline(line) + w_col<T>(2 * h0, Ly);
                const size_t qstep = (size_t)(HSTEP / PT) * Ly * WT;
                for (int h = h0; h < HM; h += HSTEP, a += qstep) {
                    const cx<T> z = L::line(lds, li)[h];
                    st_stream(reinterpret_cast<V2 *>(a), V2{z.x, z.y});
                }
            }
        } else {
            for (int e = tid; e < LPB * HMR; e += NT) {
                const int r = e / PT, li = r % LPB, h = (r / LPB) * PT + (e - r * PT);
                if (h < HM) store_pair(li, h);
            }
        }
        if (PER && tid < LPB) store_pair(tid, HX - 1);
    }
}

// ---------------------------------------------------------------------------
// pass 5: x-inverse.  Every W row (Ly*Lz of them) -> one P row of N+2 with the
// x ghosts; subtracts the mean found by the z pass.
// ACC: the refinement form -- the row is added to P (x += e - mean(e) -
// mean(x), xsum[2] = sum of x's interior) instead of stored; P's rows are
// prefetched with the W rows.
// ---------------------------------------------------------------------------
// The ACC form carries the prefetched iterate: uncapped it takes one wave per
// SIMD (256 VGPRs + AGPRs).  From lines of PPF_XIR_CAP_N points on it is
// capped at two waves (0.65 -> 0.47 ms at 512^3), while on the 32..256-point
// lines of the 96-problem batch the spills cost more than the occupancy
// gains.
#ifndef PPF_XACC_WPE
#define PPF_XACC_WPE 2
#endif
#ifndef PPF_XIR_CAP_N
#define PPF_XIR_CAP_N 512
#endif
template <typename T, int N, bool PER, bool ACC = false, int XV = 0>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(ACC ? (N >= PPF_XIR_CAP_N ? PPF_XACC_WPE : 1) : XGeom<T, N, PER, XV>::WPE)))
k_xinv(const T *__restrict__ W, double *__restrict__ P, int nlines, int Ly,
       const cx<T> *__restrict__ tw, const cx<T> *__restrict__ wk, const cx<T> *__restrict__ qk,
       const double *__restrict__ dmx1, const double *__restrict__ mean, const double *__restrict__ xsum,
       double inv_cells, const int *__restrict__ gate) {
    if (gated_off(gate)) return;
    using X = XGeom<T, N, PER, XV>;
    using G = typename X::G;
    using L = typename X::L;
    constexpr int M = G::M, TPL = X::TPL, LPB = X::LPB, NT = X::NT;
    constexpr int LX = PER ? N + 2 : N;
    constexpr int LXS = w_stride(LX, (int)sizeof(T));  // scalars per W row (a plane is Ly * LXS)
    constexpr int LP = N + 2;
    constexpr int HX = LX / 2, HP = LP / 2;  // scalar pairs per W row / per P row
    // W rows are read in column-tile order (PT pairs of one row's tile, then
    // the same pairs of the next row), as k_xfwd stores them
    constexpr int PT = w_tile<T>() / 2;
    constexpr int HXR = (HX + PT - 1) / PT * PT;
    constexpr int EPT = (LPB * HXR + NT - 1) / NT;
    constexpr int EPTP = (LPB * HP + NT - 1) / NT;  // P pairs stored per thread per tile
    using V2 = T __attribute__((ext_vector_type(2)));
    using D2 = double __attribute__((ext_vector_type(2)));
    __shared__ __align__(16) T lds[L::SIZE];
    const int tid = threadIdx.x, l = tid / TPL, t = tid - l * TPL;
    const int ntiles = (nlines + LPB - 1) / LPB;
    const float rly = 1.0f / (float)Ly;
    auto emap = [&](int e, int &li, int &h) {
        const int r = e / PT;
        li = r % LPB;
        h = (r / LPB) * PT + (e - r * PT);
    };
    constexpr int WT = w_tile<T>();
    constexpr bool ONE_ROW = NT % (PT * LPB) == 0;    // thread -> one row per tile
    constexpr int HSTEP = NT / LPB;                   // its pair step (ONE_ROW)
    const int h0 = (tid / (PT * LPB)) * PT + tid % PT;  // its first pair (ONE_ROW)
    V2 pre[EPT];
    D2 pold[ACC ? EPTP : 1];
    // 16-byte (FP64) pair loads of whole rows, clamped to the last pair;
    // row r of W is plane r / Ly, row r % Ly.  When NT is a multiple of
    // PT * LPB a thread keeps one row for the whole tile (its row, plane
    // and base address are computed once) and steps its pair by NT / LPB.
    auto fetch = [&](int tile) {
        if constexpr (ONE_ROW) {
            int row = tile * LPB + (tid / PT) % LPB;
            row = row < nlines ? row : nlines - 1;
            const int pl = fdiv_small(row, Ly, rly);
            const T *base = W + (size_t)pl * Ly * LXS + (size_t)(row - pl * Ly) * WT;
            // pair h0 + q HSTEP is HSTEP/PT column tiles further each step
            const T *p0 = base + w_col<T>(2 * h0, Ly), *plast = base + w_col<T>(2 * (HX - 1), Ly);
            const size_t qstep = (size_t)(HSTEP / PT) * Ly * WT;
#pragma unroll
            for (int q = 0; q < EPT; ++q) {
                const T *a = h0 + q * HSTEP < HX ? p0 + q * qstep : plast;
                pre[q] = ld_stream(reinterpret_cast<const V2 *>(a));
            }
        } else {
#pragma unroll
            for (int q = 0; q < EPT; ++q) {
                int e = tid + q * NT;
                e = e < LPB * HXR ? e : LPB * HXR - 1;
                int li, h;
                emap(e, li, h);
                h = h < HX ? h : HX - 1;
                int row = tile * LPB + li;
                row = row < nlines ? row : nlines - 1;
                const int pl = fdiv_small(row, Ly, rly);
                pre[q] = ld_stream(
                    reinterpret_cast<const V2 *>(W + (size_t)pl * Ly * LXS + w_at<T>(row - pl * Ly, 2 * h, Ly)));
            }
        }
        if constexpr (ACC) {
#pragma unroll
            for (int q = 0; q < EPTP; ++q) {
                int e = tid + q * NT;
                e = e < LPB * HP ? e : LPB * HP - 1;
                const size_t idx = (size_t)tile * LPB * HP + e;
                const size_t last = (size_t)nlines * HP - 1;
                pold[q] = ld_stream(reinterpret_cast<const D2 *>(P) + (idx < last ? idx : last));
            }
        }
    };
    // the first tile's loads (every form) overlap the set-up below
    int tile = blockIdx.x;
    if (tile < ntiles) fetch(tile);
    L::load_tables(lds, tw, wk, qk, tid, NT);
    // ACC: the iterate's own mean (sum from the residual pass) goes too, so
    // the refined solution keeps the zero mean exactly (the residual is blind
    // to a constant, and FP32 storage of the (0,0) mode leaves one)
    const double mu = *mean + (ACC ? xsum[2] * inv_cells : 0.0);
    double bca0 = 0, bcb0 = 0, bca1 = 0, bcb1 = 0;
    if (!PER) {
        const double d1 = dmx1[1], d2 = dmx1[2], e1 = dmx1[N + 1], e2 = dmx1[N];
        const double s0 = d1 + d2;
        bca0 = s0 * s0 / (s0 * s0 - d1 * d1);
        bcb0 = d1 * d1 / (s0 * s0 - d1 * d1);
        const double s1 = e1 + e2;
        bca1 = s1 * s1 / (s1 * s1 - e1 * e1);
        bcb1 = e1 * e1 / (s1 * s1 - e1 * e1);
    }
    const double scale = PER ? 1.0 / (double)N : 1.0 / (2.0 * (double)N);
    const cx<T> *tab = L::tables(lds);
    const LLine<T> my = L::line(lds, l);
    for (; tile < ntiles; tile += gridDim.x) {
        __syncthreads();  // the previous tile's LDS lines have been stored
        if (X::LITE && tile != (int)blockIdx.x) fetch(tile);
#pragma unroll
        for (int q = 0; q < EPT; ++q) {
            const int e = tid + q * NT;
            int li = (tid / PT) % LPB, h = h0 + q * HSTEP;
            if constexpr (!ONE_ROW) {
                h = HX;
                if (e < LPB * HXR) emap(e, li, h);
            }
            if (h < HX) {
                L::line(lds, li)[h] = cx<T>{pre[q].x, pre[q].y};
            }
        }
        D2 pcur[ACC ? EPTP : 1];
        if constexpr (ACC) {
#pragma unroll
            for (int q = 0; q < EPTP; ++q) pcur[q] = pold[q];
        }
        if (!X::LITE && tile + (int)gridDim.x < ntiles) fetch(tile + gridDim.x);
        __syncthreads();
        if (PER)
            c2r_pre<T, N, TPL>(my, t, tab + M);
        else
            idct_pre<T, N, TPL>(my, t, tab + M, tab + 2 * M + 1);
        fft_forward<T, M, TPL>(my, t, tab);
        // P element e of the tile: physical sample of its line, mean removed,
        // with the wall extrapolation or the wrap in the two x ghosts
        auto value = [&](int li, int np) -> double {
            const LReal<T> re = L::real(lds, li);
            auto sample = [&](int n) -> double {
                const int tpos = PER ? n : makhoul_pos<N>(n);
                const double v = (double)re[tpos];
                return ((tpos & 1) ? -v : v) * scale - mu;
            };
            if (np == 0) return PER ? sample(N - 1) : bca0 * sample(0) - bcb0 * sample(1);
            if (np == N + 1) return PER ? sample(0) : bca1 * sample(N - 1) - bcb1 * sample(N - 2);
            return sample(np - 1);
        };
        if constexpr (ACC) {
#pragma unroll
            for (int q = 0; q < EPTP; ++q) {
                const int e = tid + q * NT;
                if (e >= LPB * HP) break;
                const int li = e / HP, h = e - li * HP;
                const int line = tile * LPB + li;
                if (line < nlines)
                    st_stream(reinterpret_cast<D2 *>(&P[(size_t)line * LP + 2 * h]),
                              D2{value(li, 2 * h), value(li, 2 * h + 1)} + pcur[q]);
            }
        } else {
            for (int e = tid; e < LPB * HP; e += NT) {
                const int li = e / HP, h = e - li * HP;
                const int line = tile * LPB + li;
                if (line < nlines)
                    st_stream(reinterpret_cast<D2 *>(&P[(size_t)line * LP + 2 * h]),
                              D2{value(li, 2 * h), value(li, 2 * h + 1)});
            }
        }
    }
}

// ---------------------------------------------------------------------------
// passes 2 and 4: y-forward / y-inverse on a tile of TX scalar columns of one
// plane.  MODE 0: DCT on real columns; 1: C2C on complex columns (P-P);
// 2: R2C / C2R on real columns (N-P, spectral rows 0..N+1 as re/im pairs).
// YV: the tuner's variant -- threads per line halved YV times (each thread
// then runs 2^YV butterflies per stage; fewer threads per tile, fewer rounds).
// NTCAP: the most threads a tile may use (more lines go to further rounds).
// ---------------------------------------------------------------------------
template <typename T, int N, int MODE, int YV = 0, int NTCAP = kMaxNT>
struct YGeom {
    static constexpr bool CPLX = MODE == 1;
    static constexpr int M = CPLX ? N : N / 2;
    // threads per line: one radix-R butterfly each per stage (YV = 0)
    static constexpr int TPL0 = (M / max_radix(M)) > 0 ? (M / max_radix(M)) : 1;
    static constexpr int TPL = (TPL0 >> YV) > 0 ? (TPL0 >> YV) : 1;
    static constexpr int S = CPLX ? M + 1 : (MODE == 2 ? M + 1 : M);
    // scalar columns per tile: PPF_YROW_BYTES-wide rows (halved only if the
    // LDS tile -- lines + tables -- would exceed the CU's 160 KiB); lines
    // beyond kMaxNT / TPL are transformed in further rounds
    static constexpr int lds_bytes(int tx) {
        return ((CPLX ? tx / 2 : tx) * (2 * lds_slots(S) + PPF_LINE_PAD) + 2 * lds_table_slots<M, MODE == 0>()) *
               (int)sizeof(T);
    }
    static constexpr int TX0 = PPF_YROW_BYTES / (int)sizeof(T);
    static constexpr int tx_fit(int tx) { return (tx > 2 && lds_bytes(tx) > 160 * 1024 - 1024) ? tx_fit(tx / 2) : tx; }
    static constexpr int TX = tx_fit(TX0);
    static constexpr int NL = CPLX ? TX / 2 : TX;  // LDS lines per tile
    static constexpr int rounds_for(int r) {
        return (NL / r) * TPL <= NTCAP || NL % (2 * r) != 0 ? r : rounds_for(2 * r);
    }
    static constexpr int ROUNDS = rounds_for(PPF_YROUNDS);
    static constexpr int NT = NL / ROUNDS * TPL;
    static_assert(NT <= NTCAP, "y-pass workgroup exceeds its thread cap");
    static_assert(lds_bytes(TX) <= 160 * 1024, "y-pass tile exceeds the CU's LDS");
    using L = LdsLines<T, M, S, NL, MODE == 0>;
    // Row-blocked mapping (real columns, DCT or R2C): thread (g, c2) owns the
    // RPT consecutive rows RPT*g .. RPT*g+RPT-1 of column pair c2.  In the
    // DCT's Makhoul order a thread's rows fill 2*(RPT/4) complex LDS slots
    // per column, in natural order RPT/2 slots; either way the slots are
    // known once per thread and LDS is read and written as whole complex
    // values (16 bytes in FP64).  The R2C spectrum's last slot (rows N, N+1)
    // belongs to the last group.  Complex (C2C) columns keep the
    // element-strided mapping: row-blocked, their inverse pass measured
    // 0.427 ms against 0.385 ms at 512^3.
    static constexpr int TP = TX / 2;
    static constexpr int G = NT / TP;
    static constexpr int RPT = G > 0 ? N / G : 0;
    static constexpr bool RB = NT % TP == 0 && G * RPT == N &&
                               ((MODE == 0 && RPT % 4 == 0) || (MODE == 2 && RPT % 2 == 0));
    // registers: uncapped (PPF_YWPE = 1).  The row-blocked 512-point FP64
    // kernels hold 178-186 VGPRs; a cap of 128 (two 512-thread workgroups per
    // CU) spills about 50 of them and measured 1.7x slower.
    static constexpr int WPE = PPF_YWPE;
};

template <typename T, int N, int MODE, int YV = 0>
__global__ void __launch_bounds__(kMaxNT) __attribute__((amdgpu_waves_per_eu(YGeom<T, N, MODE, YV>::WPE)))
k_yfwd(T *__restrict__ W, int Lx, int Lxs, int Ly, int nplanes, const cx<T> *__restrict__ tw, const cx<T> *__restrict__ wk,
       const cx<T> *__restrict__ qk, const int *__restrict__ gate) {
    if (gated_off(gate)) return;
    using Y = YGeom<T, N, MODE, YV>;
    using L = typename Y::L;
    constexpr int M = Y::M, TPL = Y::TPL, TX = Y::TX, NT = Y::NT;
    constexpr int TP = TX / 2;  // column pairs: global accesses are 2 adjacent scalars
    constexpr int EPT = (TP * N + NT - 1) / NT;
    using V2 = T __attribute__((ext_vector_type(2)));
    constexpr int WT = w_tile<T>();  // W row step inside a column tile
    __shared__ __align__(16) T lds[L::SIZE];
    const int tid = threadIdx.x, l = tid / TPL, t = tid - l * TPL;
    const int nct = (Lx + TX - 1) / TX;
    const int ntiles = nct * nplanes;
    constexpr int RPT = Y::RPT, NPRE = Y::RB ? RPT : EPT;
    const int bg = tid / TP, bc2 = tid - bg * TP;  // row-blocked: group, column pair
    V2 pre[NPRE];
    auto fetch = [&](int tile) {  // unconditional, clamped: see k_xfwd
        const int ct = tile % nct, pl = tile / nct;
        const size_t plane = (size_t)(pl + 1) * Ly;
        if constexpr (Y::RB) {
            int col = ct * TX + 2 * bc2;
            col = col < Lx ? col : Lx - 2;  // Lx is even
            const T *src = &W[plane * Lxs + w_at<T>(1 + (size_t)RPT * bg, col, Ly)];
#pragma unroll
            for (int q = 0; q < RPT; ++q) pre[q] = ld_stream(reinterpret_cast<const V2 *>(src + (size_t)q * WT));
        } else {
#pragma unroll
            for (int q = 0; q < EPT; ++q) {
                int e = tid + q * NT;
                e = e < TP * N ? e : TP * N - 1;
                const int r = e / TP, c2 = e - r * TP;
                int col = ct * TX + 2 * c2;
                col = col < Lx ? col : Lx - 2;  // Lx is even
                pre[q] = ld_stream(reinterpret_cast<const V2 *>(&W[plane * Lxs + w_at<T>(1 + r, col, Ly)]));
            }
        }
    };
    int tile = blockIdx.x;
    if (tile < ntiles) fetch(tile);
    L::load_tables(lds, tw, wk, qk, tid, NT);
    const cx<T> *tab = L::tables(lds);
    for (; tile < ntiles; tile += gridDim.x) {
    const int c0 = (tile % nct) * TX;
    const size_t plane = (size_t)(tile / nct + 1) * Ly;
    __syncthreads();  // the previous tile's LDS lines have been stored
    if constexpr (Y::RB && MODE == 0) {
        // Makhoul order: row RPT g + 4j (+2) is component 0 (1) of complex slot
        // (RPT/4) g + j; row RPT g + 4j + 3 (+1) is component 0 (1) of slot
        // N/2 - 1 - (RPT/4) g - j
        const LLine<T> c0l = L::line(lds, 2 * bc2), c1l = L::line(lds, 2 * bc2 + 1);
#pragma unroll
        for (int j = 0; j < RPT / 4; ++j) {
            const int ce = (RPT / 4) * bg + j, co = N / 2 - 1 - (RPT / 4) * bg - j;
            c0l[ce] = cx<T>{pre[4 * j].x, pre[4 * j + 2].x};
            c1l[ce] = cx<T>{pre[4 * j].y, pre[4 * j + 2].y};
            c0l[co] = cx<T>{pre[4 * j + 3].x, pre[4 * j + 1].x};
            c1l[co] = cx<T>{pre[4 * j + 3].y, pre[4 * j + 1].y};
        }
    } else if constexpr (Y::RB) {
        // natural order: rows RPT g + 2k, +1 are the components of slot (RPT/2) g + k
        const LLine<T> c0l = L::line(lds, 2 * bc2), c1l = L::line(lds, 2 * bc2 + 1);
#pragma unroll
        for (int k = 0; k < RPT / 2; ++k) {
            c0l[(RPT / 2) * bg + k] = cx<T>{pre[2 * k].x, pre[2 * k + 1].x};
            c1l[(RPT / 2) * bg + k] = cx<T>{pre[2 * k].y, pre[2 * k + 1].y};
        }
    } else {
#pragma unroll
        for (int q = 0; q < EPT; ++q) {
            const int e = tid + q * NT;
            if (e < TP * N) {
                const int r = e / TP, c2 = e - r * TP;
                const V2 v = pre[q];
                if (MODE == 0) {
                    L::real(lds, 2 * c2)[makhoul_pos<N>(r)] = v.x;
                    L::real(lds, 2 * c2 + 1)[makhoul_pos<N>(r)] = v.y;
                } else if (MODE == 1) {  // a pair is one complex column
                    L::line(lds, c2)[r] = cx<T>{v.x, v.y};
                } else {
                    L::real(lds, 2 * c2)[r] = v.x;
                    L::real(lds, 2 * c2 + 1)[r] = v.y;
                }
            }
        }
    }
    if (tile + (int)gridDim.x < ntiles) fetch(tile + gridDim.x);
    __syncthreads();
#pragma unroll 1
    for (int rd = 0; rd < Y::ROUNDS; ++rd) {
        const LLine<T> my = L::line(lds, l + rd * (Y::NL / Y::ROUNDS));
        fft_forward<T, M, TPL>(my, t, tab);
        if (MODE == 0) dct_post<T, N, TPL>(my, t, tab + M, tab + 2 * M + 1);
        if (MODE == 2) r2c_post<T, N, TPL>(my, t, tab + M);
    }
    constexpr int NR = MODE == 2 ? N + 2 : N;  // spectral rows written
    constexpr int R0 = MODE == 2 ? 0 : 1;
    if constexpr (Y::RB) {
        // natural order: rows RPT g + 2k, +1 are the components of slot (RPT/2) g + k
        const int col = c0 + 2 * bc2;
        if (col < Lx) {
            const LLine<T> c0l = L::line(lds, 2 * bc2), c1l = L::line(lds, 2 * bc2 + 1);
            T *dst = &W[plane * Lxs + w_at<T>(R0 + (size_t)RPT * bg, col, Ly)];
#pragma unroll
            for (int k = 0; k < RPT / 2; ++k) {
                const cx<T> z0 = c0l[(RPT / 2) * bg + k], z1 = c1l[(RPT / 2) * bg + k];
                st_stream(reinterpret_cast<V2 *>(dst + (size_t)(2 * k) * WT), V2{z0.x, z1.x});
                st_stream(reinterpret_cast<V2 *>(dst + (size_t)(2 * k + 1) * WT), V2{z0.y, z1.y});
            }
            if (MODE == 2 && bg == Y::G - 1) {  // R2C: spectral rows N, N+1 (slot N/2)
                const cx<T> z0 = c0l[N / 2], z1 = c1l[N / 2];
                st_stream(reinterpret_cast<V2 *>(dst + (size_t)RPT * WT), V2{z0.x, z1.x});
                st_stream(reinterpret_cast<V2 *>(dst + (size_t)(RPT + 1) * WT), V2{z0.y, z1.y});
            }
        }
    } else {
        for (int e = tid; e < TP * NR; e += NT) {
            const int r = e / TP, c2 = e - r * TP;
            const int col = c0 + 2 * c2;
            if (col >= Lx) continue;
            V2 v;
            if (MODE == 1) {
                const cx<T> z = L::line(lds, c2)[r];
                v.x = z.x;
                v.y = z.y;
            } else {
                v.x = L::real(lds, 2 * c2)[r];
                v.y = L::real(lds, 2 * c2 + 1)[r];
            }
            st_stream(reinterpret_cast<V2 *>(&W[plane * Lxs + w_at<T>(R0 + r, col, Ly)]), v);
        }
    }
    }
}

template <typename T, int N, int MODE, int YV = 0>
__global__ void __launch_bounds__(kMaxNT) __attribute__((amdgpu_waves_per_eu(YGeom<T, N, MODE, YV>::WPE)))
k_yinv(T *__restrict__ W, int Lx, int Lxs, int Ly, int nplanes, const cx<T> *__restrict__ tw, const cx<T> *__restrict__ wk,
       const cx<T> *__restrict__ qk, const double *__restrict__ dmx2, const int *__restrict__ gate) {
    if (gated_off(gate)) return;
    using Y = YGeom<T, N, MODE, YV>;
    using L = typename Y::L;
    constexpr int M = Y::M, TPL = Y::TPL, TX = Y::TX, NT = Y::NT;
    constexpr int NR = MODE == 2 ? N + 2 : N;
    constexpr int R0 = MODE == 2 ? 0 : 1;
    constexpr int TP = TX / 2;  // column pairs: global accesses are 2 adjacent scalars
    constexpr int EPT = (TP * NR + NT - 1) / NT;
    using V2 = T __attribute__((ext_vector_type(2)));
    constexpr int WT = w_tile<T>();  // W row step inside a column tile
    __shared__ __align__(16) T lds[L::SIZE];
    const int tid = threadIdx.x, l = tid / TPL, t = tid - l * TPL;
    const int nct = (Lx + TX - 1) / TX;
    const int ntiles = nct * nplanes;  // every plane incl. the z ghost planes
    constexpr int RPT = Y::RPT, NPRE = Y::RB ? RPT : EPT;
    const int bg = tid / TP, bc2 = tid - bg * TP;  // row-blocked: group, column pair
    V2 pre[NPRE];
    V2 pre_x[2] = {V2{0, 0}, V2{0, 0}};  // row-blocked R2C: rows N, N+1 (last group)
    auto fetch = [&](int tile) {  // unconditional, clamped: see k_xfwd
        const int ct = tile % nct, pl = tile / nct;
        const size_t plane = (size_t)pl * Ly;
        if constexpr (Y::RB) {
            int col = ct * TX + 2 * bc2;
            col = col < Lx ? col : Lx - 2;  // Lx is even
            const T *src = &W[plane * Lxs + w_at<T>(R0 + (size_t)RPT * bg, col, Ly)];
#pragma unroll
            for (int q = 0; q < RPT; ++q) pre[q] = ld_stream(reinterpret_cast<const V2 *>(src + (size_t)q * WT));
            if (MODE == 2 && bg == Y::G - 1) {  // R2C: spectral rows N, N+1
                pre_x[0] = ld_stream(reinterpret_cast<const V2 *>(src + (size_t)RPT * WT));
                pre_x[1] = ld_stream(reinterpret_cast<const V2 *>(src + (size_t)(RPT + 1) * WT));
            }
        } else {
#pragma unroll
            for (int q = 0; q < EPT; ++q) {
                int e = tid + q * NT;
                e = e < TP * NR ? e : TP * NR - 1;
                const int r = e / TP, c2 = e - r * TP;
                int col = ct * TX + 2 * c2;
                col = col < Lx ? col : Lx - 2;  // Lx is even
                pre[q] = ld_stream(reinterpret_cast<const V2 *>(&W[plane * Lxs + w_at<T>(R0 + r, col, Ly)]));
            }
        }
    };
    int tile = blockIdx.x;
    if (tile < ntiles) fetch(tile);
    L::load_tables(lds, tw, wk, qk, tid, NT);
    const T scale = MODE == 0 ? T(1.0 / (2.0 * N)) : T(1.0 / (double)N);
    // y ghosts: Neumann for MODE 0, periodic wrap otherwise
    T bca0 = 0, bcb0 = 0, bca1 = 0, bcb1 = 0;
    if (MODE == 0) {
        const double d1 = dmx2[1], d2 = dmx2[2], e1 = dmx2[N + 1], e2 = dmx2[N];
        const double s0 = d1 + d2;
        bca0 = (T)(s0 * s0 / (s0 * s0 - d1 * d1));
        bcb0 = (T)(d1 * d1 / (s0 * s0 - d1 * d1));
        const double s1 = e1 + e2;
        bca1 = (T)(s1 * s1 / (s1 * s1 - e1 * e1));
        bcb1 = (T)(e1 * e1 / (s1 * s1 - e1 * e1));
    }
    const cx<T> *tab = L::tables(lds);
    for (; tile < ntiles; tile += gridDim.x) {
    const int c0 = (tile % nct) * TX;
    const size_t plane = (size_t)(tile / nct) * Ly;
    __syncthreads();  // the previous tile's LDS lines have been stored
    if constexpr (Y::RB) {
        // natural order: rows RPT g + 2k, +1 are the components of slot (RPT/2) g + k
        const LLine<T> c0l = L::line(lds, 2 * bc2), c1l = L::line(lds, 2 * bc2 + 1);
#pragma unroll
        for (int k = 0; k < RPT / 2; ++k) {
            c0l[(RPT / 2) * bg + k] = cx<T>{pre[2 * k].x, pre[2 * k + 1].x};
            c1l[(RPT / 2) * bg + k] = cx<T>{pre[2 * k].y, pre[2 * k + 1].y};
        }
        if (MODE == 2 && bg == Y::G - 1) {
            c0l[N / 2] = cx<T>{pre_x[0].x, pre_x[1].x};
            c1l[N / 2] = cx<T>{pre_x[0].y, pre_x[1].y};
        }
    } else {
#pragma unroll
        for (int q = 0; q < EPT; ++q) {
            const int e = tid + q * NT;
            if (e < TP * NR) {
                const int r = e / TP, c2 = e - r * TP;
                const V2 v = pre[q];
                if (MODE == 1) {  // a complex column, conjugated (inverse via the forward FFT)
                    L::line(lds, c2)[r] = cx<T>{v.x, -v.y};
                } else {
                    L::real(lds, 2 * c2)[r] = v.x;
                    L::real(lds, 2 * c2 + 1)[r] = v.y;
                }
            }
        }
    }
    if (tile + (int)gridDim.x < ntiles) fetch(tile + gridDim.x);
    __syncthreads();
#pragma unroll 1
    for (int rd = 0; rd < Y::ROUNDS; ++rd) {
        const LLine<T> my = L::line(lds, l + rd * (Y::NL / Y::ROUNDS));
        if (MODE == 0) idct_pre<T, N, TPL>(my, t, tab + M, tab + 2 * M + 1);
        if (MODE == 2) c2r_pre<T, N, TPL>(my, t, tab + M);
        fft_forward<T, M, TPL>(my, t, tab);
    }
    if constexpr (Y::RB && MODE == 2) {
        // natural order; odd samples change sign (the C2R's conjugate trick);
        // periodic ghost rows 0 <- sample N-1 (last group), N+1 <- sample 0
        const int col = c0 + 2 * bc2;
        if (col < Lx) {
            const LLine<T> c0l = L::line(lds, 2 * bc2), c1l = L::line(lds, 2 * bc2 + 1);
            V2 v[RPT];
#pragma unroll
            for (int k = 0; k < RPT / 2; ++k) {
                const cx<T> a0 = c0l[(RPT / 2) * bg + k], a1 = c1l[(RPT / 2) * bg + k];
                v[2 * k] = V2{a0.x, a1.x} * scale;
                v[2 * k + 1] = V2{a0.y, a1.y} * -scale;
            }
            T *dst = &W[plane * Lxs + w_at<T>(1 + (size_t)RPT * bg, col, Ly)];
#pragma unroll
            for (int q = 0; q < RPT; ++q) st_stream(reinterpret_cast<V2 *>(dst + (size_t)q * WT), v[q]);
            if (bg == 0) st_stream(reinterpret_cast<V2 *>(&W[plane * Lxs + w_at<T>(N + 1, col, Ly)]), v[0]);
            if (bg == Y::G - 1) st_stream(reinterpret_cast<V2 *>(&W[plane * Lxs + w_at<T>(0, col, Ly)]), v[RPT - 1]);
        }
    } else if constexpr (Y::RB) {
        // Makhoul order (as in k_yfwd), the signs of the DCT-III's conjugate
        // trick, and the scale; physical row n goes to W row n + 1
        const int col = c0 + 2 * bc2;
        if (col < Lx) {
            const LLine<T> c0l = L::line(lds, 2 * bc2), c1l = L::line(lds, 2 * bc2 + 1);
            V2 v[RPT];
#pragma unroll
            for (int j = 0; j < RPT / 4; ++j) {
                const int ce = (RPT / 4) * bg + j, co = N / 2 - 1 - (RPT / 4) * bg - j;
                const cx<T> a0 = c0l[ce], a1 = c1l[ce], b0 = c0l[co], b1 = c1l[co];
                v[4 * j] = V2{a0.x, a1.x} * scale;
                v[4 * j + 2] = V2{a0.y, a1.y} * -scale;
                v[4 * j + 3] = V2{b0.x, b1.x} * scale;
                v[4 * j + 1] = V2{b0.y, b1.y} * -scale;
            }
            T *dst = &W[plane * Lxs + w_at<T>(1 + (size_t)RPT * bg, col, Ly)];
#pragma unroll
            for (int q = 0; q < RPT; ++q) st_stream(reinterpret_cast<V2 *>(dst + (size_t)q * WT), v[q]);
            // wall ghost rows (fft_3d_poisson.f90:809-852)
            if (bg == 0) st_stream(reinterpret_cast<V2 *>(dst - WT), bca0 * v[0] - bcb0 * v[1]);
            if (bg == Y::G - 1)
                st_stream(reinterpret_cast<V2 *>(dst + (size_t)RPT * WT), bca1 * v[RPT - 1] - bcb1 * v[RPT - 2]);
        }
    } else
    for (int e = tid; e < TP * (N + 2); e += NT) {
        const int rp = e / TP, c2 = e - rp * TP;
        const int col = c0 + 2 * c2;
        if (col >= Lx) continue;
        // physical row n (0-based) of the pair's two columns
        auto sample = [&](int n) -> V2 {
            V2 v;
            if (MODE == 1) {
                const cx<T> z = L::line(lds, c2)[n];
                v.x = z.x * scale;
                v.y = -z.y * scale;
            } else {
                const int tpos = MODE == 0 ? makhoul_pos<N>(n) : n;
                const T sg = (tpos & 1) ? T(-1) : T(1);
                v.x = sg * L::real(lds, 2 * c2)[tpos] * scale;
                v.y = sg * L::real(lds, 2 * c2 + 1)[tpos] * scale;
            }
            return v;
        };
        V2 out;
        if (rp == 0)
            out = MODE == 0 ? bca0 * sample(0) - bcb0 * sample(1) : sample(N - 1);
        else if (rp == N + 1)
            out = MODE == 0 ? bca1 * sample(N - 1) - bcb1 * sample(N - 2) : sample(0);
        else
            out = sample(rp - 1);
        st_stream(reinterpret_cast<V2 *>(&W[plane * Lxs + w_at<T>(rp, col, Ly)]), out);
    }
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
    int Lx, Lxs, Ly, n3m, jlo, njr;   // Lx columns per row, Lxs scalars per row (a plane is Ly * Lxs)
    const double *lamx, *lamy;        // per scalar column / per row
    const unsigned char *zx, *zy;     // zero-mode flags
    int mean_kx, mean_jp;             // real (0,0) column
    double mean_scale;                // 1/(fx fy n1m n2m n3m)
    int zper;                         // periodic z: ghost planes wrap instead of wall extrapolation
    const int *gate;                  // refinement gate (gated_off), nullptr = always run
};

template <typename T>
__global__ void __launch_bounds__(256)
k_zthomas(T *__restrict__ W, T *__restrict__ cs, ZArgs z, const double *__restrict__ dx3,
          const double *__restrict__ dmx3, double *__restrict__ mean_out) {
    if (gated_off(z.gate)) return;
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
    const size_t pstride = (size_t)z.Ly * z.Lxs;
    T *w = W + w_at<T>(jp, kx, z.Ly);      // plane 0 of this column
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
    if (z.zper) {
        w[0] = (T)xn;
        w[(size_t)(n + 1) * pstride] = (T)x1;
    } else {
        w[0] = (T)(bca0 * x1 - bcb0 * x2);
        w[(size_t)(n + 1) * pstride] = (T)(bca1 * xn - bcb1 * xm1);
    }
    if (kx == z.mean_kx && jp == z.mean_jp) *mean_out = sum * z.mean_scale;
}

// ---------------------------------------------------------------------------
// pass 3, register-resident partitioned Thomas.
//
// A workgroup owns S adjacent spectral columns of one row jp and splits each
// z line into P partitions of K rows; thread (s, p) keeps its K values in
// registers, so the line is read once and written once.  Per partition (rows
// l = 0..K-1, outside neighbours x_{-1}, x_K):
//   reverse sweep (upper diagonal eliminated towards row 0, x_{K-1} kept):
//       row 0 becomes  ah x_{-1} + x_0 + ch x_{K-1} = dh
//   forward sweep (lower diagonal eliminated, x_0 kept), rows l >= 1:
//       A_l x_0 + x_l + C_l x_{l+1} = D_l,  row K-1 couples to x_K
// These two rows of every partition form a 2P-row tridiagonal system in the
// partition end values (the reduced system of PaScaL_TDMA's modified Thomas,
// pascal_tdma_cuda.f90:332-473 and :230-269 for many ranks).  Substituting the
// first-row equations leaves P rows in the last values x_{K-1}, solved per
// column by cyclic reduction in LDS; x_0 follows from its first-row
// equation.  With x_0 and x_{K-1} known, a Thomas sweep over rows 1..K-2 that
// reuses the forward multipliers C_l recovers the interior.  Both reduction
// sweeps carry O(1) state, so a thread holds 2K values (its right-hand sides
// and C).  Rows past n3m are identity rows.  Coefficients, pin, ghost planes
// and mean as in k_zthomas.
// ---------------------------------------------------------------------------
#ifndef PPF_RCP_NEWTON
#define PPF_RCP_NEWTON 2
#endif
// FP32: the hardware reciprocal is correct to about 1 ulp
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double frcp(double x) {
    // reciprocal: hardware estimate refined by Newton steps (each doubles the
    // correct bits; two give about 1 ulp)
    double r = __builtin_amdgcn_rcp(x);
#pragma unroll
    for (int i = 0; i < PPF_RCP_NEWTON; ++i) {
        const double e = __builtin_fma(-x, r, 1.0);
        r = __builtin_fma(r, e, r);
    }
    return r;
}

// WPE: minimum waves per SIMD the register allocation must allow (occupancy
// is what lets one workgroup's loads overlap another's arithmetic here).
// CT: arithmetic type (FP64 for every plan: FP32 arithmetic on FP32 storage
// measured no faster and costs the refinement a sweep).
template <typename T, int K, int S, int NTMAX, int WPE, typename CT = double>
__global__ void __launch_bounds__(NTMAX) __attribute__((amdgpu_waves_per_eu(WPE)))
k_zpart(T *__restrict__ W, ZArgs z, const double *__restrict__ zam, const double *__restrict__ zap,
        double *__restrict__ mean_out, int P) {
    if (gated_off(z.gate)) return;
    extern __shared__ double zsh[];
    const int n3m = z.n3m;
    constexpr int KS = (K + S - 1) / S;           // table rows per thread
    const int NPAD = KS * (int)blockDim.x;        // >= NP: every thread writes KS entries
    double *am = zsh;               // [NPAD], 0 past n3m
    double *ap = am + NPAD;         // [NPAD], 0 past n3m
    const int PS = P * S;
    CT *ra = reinterpret_cast<CT *>(ap + NPAD);  // [3][P][S]: first-row equations, then reduction buffer
    CT *rc = ra + 3 * PS;           // [3][P][S]: reduction buffer
    CT *xe = rc + 3 * PS;           // [P][S]: partition end values x_{K-1}
    CT *ps = ra;                    // [P][S]: partial sums, once the reduction is done
    const int tid = threadIdx.x, s = tid % S, p = tid / S;
    const int blk = (int)blockIdx.x;
    // rows fastest inside a column tile: consecutive workgroups take
    // consecutive rows of one 128-byte column tile of W (adjacent bytes)
    const int ct = blk / z.njr;
    const int jp = z.jlo + (blk - ct * z.njr);
    const int kx = ct * S + s;
    const bool active = kx < z.Lx;
    const size_t pstride = (size_t)z.Ly * z.Lxs;
    // columns past Lx (the last tile's padding) load the last real column,
    // which lies in the same 128-byte rows, and are masked below
    T *w = W + w_at<T>(jp, active ? kx : z.Lx - 1, z.Ly);
    const CT lx = active ? (CT)z.lamx[kx] : CT(0);
    const CT ly = (CT)z.lamy[jp];
    const bool pin = active && z.zx[kx] && z.zy[jp];
    const int g0 = p * K;
    // loads are clamped in bounds and selected, so the sweep code is branch-free
    const int glast = n3m - 1;
    // Row coefficients (the per-solve table filled by pass 1) are loaded
    // first: loads complete in order, so waiting for them does not wait for
    // the line loads issued after them.  Their LDS writes are unconditional
    // (the arrays are padded to KS entries per thread), which keeps the
    // loads ahead of the line loads.
    double ca[KS], cp[KS];
#pragma unroll
    for (int j = 0; j < KS; ++j) {
        const int k = tid + j * (int)blockDim.x, kc = k < glast ? k : glast;
        ca[j] = zam[kc];
        cp[j] = zap[kc];
    }
    double bcd[4];  // z wall extrapolation, from the same table
#pragma unroll
    for (int q = 0; q < 4; ++q) bcd[q] = zap[n3m + q];
    // line loads in the order the two sweeps below consume them (rows 1 and
    // K-2 first, then inwards; rows 0 and K-1 last), so the sweeps start as
    // soon as the first rows arrive; nothing consumes them before that
    CT d[K];
#pragma unroll
    for (int q = 0; q < K; ++q) {
        const int l = q < K - 2 ? ((q & 1) ? K - 2 - q / 2 : 1 + q / 2) : (q == K - 2 ? 0 : K - 1);
        const int g = g0 + l;
        d[l] = (CT)ld_stream(&w[(size_t)((g < glast ? g : glast) + 1) * pstride]);
    }
#pragma unroll
    for (int j = 0; j < KS; ++j) {
        const int k = tid + j * (int)blockDim.x;
        am[k] = k < n3m ? ca[j] : 0.0;
        ap[k] = k < n3m ? cp[j] : 0.0;
    }
    CT bc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) bc[q] = (CT)bcd[q];
    __syncthreads();
    // a row's value is masked where it is first used: zero past n3m, in
    // inactive columns, and in the pinned first row of the (0,0) mode
    auto dmask = [&](int l) {
        const int g = g0 + l;
        d[l] = (active && g < n3m && !(pin && g == 0)) ? d[l] : CT(0);
    };
    // identity rows: padding past n3m, and the pinned first row of the (0,0) mode
    auto coef = [&](int l, CT &a, CT &b, CT &c) {
        const int g = g0 + l;
        const bool id = (g >= n3m) || (pin && g == 0);
        const CT am_ = (CT)am[g], ap_ = (CT)ap[g];
        a = am_;
        c = id ? CT(0) : ap_;
        b = id ? CT(1) : (-am_ - ap_ - lx) - ly;
    };
    // Reverse sweep (first-row equation) and forward sweep (last-row
    // equation) interleaved: two independent dependency chains per step.
    // The forward pivots r_l are kept for the final pass.
    CT rr[K];
    CT ah, ch, dh, Af, Df, Cf;
    {
        CT a, b, c;
        dmask(K - 2);
        dmask(1);
        coef(K - 2, a, b, c);
        double r = frcp(b);
        ah = r * a; ch = r * c; dh = r * d[K - 2];
        coef(1, a, b, c);
        r = frcp(b);
        rr[1] = r;
        Af = r * a; Cf = r * c; Df = r * d[1];
#pragma unroll
        for (int i = 1; i <= K - 2; ++i) {
            const int lb = K - 2 - i;  // reverse row
            const int lf = 1 + i;      // forward row
            dmask(lb);
            dmask(lf);
            CT a1, b1, c1, a2, b2, c2;
            coef(lb, a1, b1, c1);
            coef(lf, a2, b2, c2);
            const double r1 = frcp(b1 - c1 * ah);
            const double r2 = frcp(b2 - a2 * Cf);
            ch = -r1 * c1 * ch;
            dh = r1 * (d[lb] - c1 * dh);
            ah = r1 * a1;
            rr[lf] = r2;
            Af = -r2 * a2 * Af;
            Cf = r2 * c2;
            Df = r2 * (d[lf] - a2 * Df);
            // keep coefficient reads next to their use (register pressure)
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    // Reduced system in the partition end values e_p = x_{K-1} of partition p.
    // The first-row equation gives x_0 = dh - ah e_{p-1} - ch e_p; put into the
    // last-row equation Af x_0 + e_p + Cf x_0(p+1) = Df it leaves one
    // tridiagonal row per partition,
    //   (-Af ah_p) e_{p-1} + (1 - Af ch_p - Cf ah_{p+1}) e_p + (-Cf ch_{p+1}) e_{p+1}
    //     = Df - Af dh_p - Cf dh_{p+1},
    // kept with a unit diagonal as (el, eu, er).
    const int o = p * S + s;
    ra[o] = ah;
    ra[PS + o] = ch;
    ra[2 * PS + o] = dh;
    __syncthreads();
    CT el, eu, er;
    {
        CT ah1 = 0, ch1 = 0, dh1 = 0;
        if (p + 1 < P) {
            ah1 = ra[o + S];
            ch1 = ra[PS + o + S];
            dh1 = ra[2 * PS + o + S];
        }
        const CT inv = frcp(CT(1) - Af * ch - Cf * ah1);
        el = -Af * ah * inv;
        eu = -Cf * ch1 * inv;
        er = (Df - Af * dh - Cf * dh1) * inv;
    }
    // parallel cyclic reduction over the P rows of each column: step h couples
    // row p to p +- 2h; rows outside [0, P) are zero.  The rows alternate
    // between two LDS buffers, so one barrier per step separates every write
    // from the reads of the same buffer.
    {
        CT *buf = rc;
        for (int h = 1; h < P; h <<= 1) {
            buf[o] = el;
            buf[PS + o] = eu;
            buf[2 * PS + o] = er;
            __syncthreads();
            CT lm = 0, um = 0, rm = 0, lp = 0, up = 0, rp = 0;
            if (p >= h) {
                const int q = o - h * S;
                lm = buf[q]; um = buf[PS + q]; rm = buf[2 * PS + q];
            }
            if (p + h < P) {
                const int q = o + h * S;
                lp = buf[q]; up = buf[PS + q]; rp = buf[2 * PS + q];
            }
            const CT inv = frcp(CT(1) - el * um - eu * lp);
            er = (er - el * rm - eu * rp) * inv;
            el = -el * lm * inv;
            eu = -eu * up * inv;
            buf = (buf == rc) ? ra : rc;
        }
    }
    const CT xl = er;
    xe[o] = xl;
    __syncthreads();  // xe complete; the reduction buffers (and ps) are free
    const CT x0 = dh - ah * (p > 0 ? xe[o - S] : CT(0)) - ch * xl;
    // interior rows 1..K-2 with both ends known: forward elimination with the
    // stored pivots, back substitution with C_l = r_l c_l (results in d)
    CT *x = d;
    {
        CT e = x0;
#pragma unroll
        for (int l = 1; l <= K - 2; ++l) {
            CT a, b, c;
            coef(l, a, b, c);
            e = rr[l] * (d[l] - a * e);
            d[l] = e;
            __builtin_amdgcn_sched_barrier(0);
        }
        x[0] = x0;
        x[K - 1] = xl;
#pragma unroll
        for (int l = K - 2; l >= 1; --l) {
            CT a, b, c;
            coef(l, a, b, c);
            x[l] = d[l] - rr[l] * c * x[l + 1];
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    double part = 0.0;  // mean: FP64 (only the (0,0) tile's sum is used)
#pragma unroll
    for (int l = 0; l < K; ++l) {
        const int g = g0 + l;
        if (g < n3m) {
            part += x[l];
            if (active) st_stream(&w[(size_t)(g + 1) * pstride], (T)x[l]);
        }
    }
    // spectral z ghost planes: wall extrapolation (fft_3d_poisson.f90:809-852)
    // or, for periodic z, the wrapped end planes (:666-796)
    if (active) {
        const int gt = n3m - 1, pt = gt / K, lt = gt - pt * K;
        if (p == 0) {
            if (z.zper) w[(size_t)(n3m + 1) * pstride] = (T)x[0];
            else w[0] = (T)(bc[0] * x[0] - bc[1] * x[1]);
        }
        if (p == pt) {
            CT xt = 0, xb = 0;
#pragma unroll
            for (int l = 0; l < K; ++l) {
                if (l == lt) xt = x[l];
                if (l == lt - 1) xb = x[l];
            }
            if (lt == 0) xb = xe[(pt - 1) * S + s];
            if (z.zper) w[0] = (T)xt;
            else w[(size_t)(n3m + 1) * pstride] = (T)(bc[2] * xt - bc[3] * xb);
        }
    }
    // mean of the solution from the (0,0) real column (:628-664)
    const bool mean_tile = (jp == z.mean_jp) && ct * S <= z.mean_kx && z.mean_kx < ct * S + S;
    if (mean_tile) {
        ps[p * S + s] = part;
        __syncthreads();
        if (p == 0 && kx == z.mean_kx) {
            double sum = 0.0;
            for (int q = 0; q < P; ++q) sum += ps[q * S + s];
            *mean_out = sum * z.mean_scale;
        }
    }
}

// ---------------------------------------------------------------------------
// pass 3, pipelined form (an alternative shape; the default is k_zpart with
// K = 32, S = 16, which measured faster at 512^3 on MI355X).
//
// Same partitioned Thomas as k_zpart, with two changes that matter on
// gfx950:
//  - persistent workgroups walk column tiles and prefetch the next tile's K
//    values per thread into registers before solving the current tile, so
//    HBM traffic overlaps the arithmetic instead of every resident
//    workgroup loading, then computing, in lockstep;
//  - the 2P-row reduced system of each column is transposed once through LDS
//    so that one wavefront owns a whole column (lane i = reduced row i) and
//    cyclic reduction runs with cross-lane shuffles: two barriers per tile.
// ---------------------------------------------------------------------------
template <typename T, int K, int S, int WPE>
__global__ void __launch_bounds__(S * 32) __attribute__((amdgpu_waves_per_eu(WPE)))
k_zpipe(T *__restrict__ W, ZArgs z, const double *__restrict__ zam, const double *__restrict__ zap,
        double *__restrict__ mean_out, int P) {
    if (gated_off(z.gate)) return;
    constexpr int RS = 65;  // LDS column stride of the reduced system (2P <= 64, padded)
    extern __shared__ double zsh[];
    const int n3m = z.n3m;
    const int NP = P * K;
    const int R = 2 * P;
    double *am = zsh;
    double *ap = am + NP;
    double *red_l = ap + NP;       // [S][RS]
    double *red_u = red_l + S * RS;
    double *red_d = red_u + S * RS;
    double *ps = red_d + S * RS;   // [P][S] partial sums
    const int tid = threadIdx.x, s = tid % S, p = tid / S;
    // A column's reduced system goes to one wavefront, lane i holding row i,
    // so only waves with at least R lanes take columns: every full wave, and
    // the last (partial) wave only if it still has R lanes.  R = 2P <= S P, so
    // wave 0 always qualifies.
    const int lane = tid & 63, wv = tid >> 6;
    const int nwv = (int)(blockDim.x >> 6) + ((int)(blockDim.x & 63) >= 2 * P ? 1 : 0);
    for (int k = tid; k < NP; k += blockDim.x) {
        am[k] = k < n3m ? zam[k] : 0.0;
        ap[k] = k < n3m ? zap[k] : 0.0;
    }
    double bc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) bc[q] = zap[n3m + q];
    const int tiles_x = (z.Lx + S - 1) / S;
    const int ntiles = tiles_x * z.njr;
    const size_t pstride = (size_t)z.Ly * z.Lxs;
    const int g0 = p * K;
    const int glast = n3m - 1;
    // Prefetch loads go to clamped, always valid addresses and nothing consumes
    // them until the next tile starts: a select on a loaded value right after
    // the load would make every load a full round trip (s_waitcnt per row).
    T pre[K];
    double pre_lx, pre_ly;
    unsigned char pre_zx, pre_zy;
    auto fetch = [&](int tile) {
        const int jp = z.jlo + tile % z.njr;  // rows fastest inside a column tile
        int kx = (tile / z.njr) * S + s;
        kx = kx < z.Lx ? kx : z.Lx - 1;
        const T *w = W + w_at<T>(jp, kx, z.Ly);
#pragma unroll
        for (int l = 0; l < K; ++l) {
            const int g = g0 + l;
            pre[l] = w[(size_t)((g < glast ? g : glast) + 1) * pstride];
        }
        pre_lx = z.lamx[kx];
        pre_ly = z.lamy[jp];
        pre_zx = z.zx[kx];
        pre_zy = z.zy[jp];
    };
    int tile = blockIdx.x;
    if (tile < ntiles) fetch(tile);
    __syncthreads();
    for (; tile < ntiles; tile += gridDim.x) {
        const int jp = z.jlo + tile % z.njr;
        const int kx = (tile / z.njr) * S + s;
        const bool active = kx < z.Lx;
        T *w = W + w_at<T>(jp, active ? kx : 0, z.Ly);
        const double lx = active ? pre_lx : 0.0;
        const double ly = pre_ly;
        const bool pin = active && pre_zx && pre_zy;
        double d[K];
#pragma unroll
        for (int l = 0; l < K; ++l) d[l] = (active && g0 + l < n3m) ? (double)pre[l] : 0.0;
        if (pin && g0 == 0) d[0] = 0.0;
        if (tile + (int)gridDim.x < ntiles) fetch(tile + gridDim.x);
        auto coef = [&](int l, double &a, double &b, double &c) {
            const int g = g0 + l;
            const bool id = (g >= n3m) || (pin && g == 0);
            const double am_ = am[g], ap_ = ap[g];
            a = am_;
            c = id ? 0.0 : ap_;
            b = id ? 1.0 : (-am_ - ap_ - lx) - ly;
        };
        double rr[K];
        double ah, ch, dh, Af, Df, Cf;
        {
            double a, b, c;
            coef(K - 2, a, b, c);
            double r = frcp(b);
            ah = r * a; ch = r * c; dh = r * d[K - 2];
            coef(1, a, b, c);
            r = frcp(b);
            rr[1] = r;
            Af = r * a; Cf = r * c; Df = r * d[1];
#pragma unroll
            for (int i = 1; i <= K - 2; ++i) {
                const int lb = K - 2 - i, lf = 1 + i;
                double a1, b1, c1, a2, b2, c2;
                coef(lb, a1, b1, c1);
                coef(lf, a2, b2, c2);
                const double r1 = frcp(b1 - c1 * ah);
                const double r2 = frcp(b2 - a2 * Cf);
                ch = -r1 * c1 * ch;
                dh = r1 * (d[lb] - c1 * dh);
                ah = r1 * a1;
                rr[lf] = r2;
                Af = -r2 * a2 * Af;
                Cf = r2 * c2;
                Df = r2 * (d[lf] - a2 * Df);
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        red_l[s * RS + 2 * p] = ah;
        red_u[s * RS + 2 * p] = ch;
        red_d[s * RS + 2 * p] = dh;
        red_l[s * RS + 2 * p + 1] = Af;
        red_u[s * RS + 2 * p + 1] = Cf;
        red_d[s * RS + 2 * p + 1] = Df;
        __syncthreads();
        // cyclic reduction of each column's reduced system inside one wavefront
        // (waves past nwv take none; a shuffle source past R is masked below)
        for (int col = wv; wv < nwv && col < S; col += nwv) {
            const bool in = lane < R;
            double l_ = in ? red_l[col * RS + lane] : 0.0;
            double u_ = in ? red_u[col * RS + lane] : 0.0;
            double d_ = in ? red_d[col * RS + lane] : 0.0;
            double b_ = 1.0;
            for (int h = 1; h < R; h <<= 1) {
                const int im = lane - h, ip = lane + h;
                const double lm = __shfl(l_, im < 0 ? 0 : im), bm = __shfl(b_, im < 0 ? 0 : im);
                const double um = __shfl(u_, im < 0 ? 0 : im), dm = __shfl(d_, im < 0 ? 0 : im);
                const double lp = __shfl(l_, ip > 63 ? 63 : ip), bp = __shfl(b_, ip > 63 ? 63 : ip);
                const double up = __shfl(u_, ip > 63 ? 63 : ip), dp = __shfl(d_, ip > 63 ? 63 : ip);
                const bool hm = im >= 0, hp = ip < R;
                const double al = hm ? -l_ * frcp(bm) : 0.0;
                const double ga = hp ? -u_ * frcp(bp) : 0.0;
                const double nb = b_ + al * (hm ? um : 0.0) + ga * (hp ? lp : 0.0);
                const double nl = al * (hm ? lm : 0.0);
                const double nu = ga * (hp ? up : 0.0);
                const double nd = d_ + al * (hm ? dm : 0.0) + ga * (hp ? dp : 0.0);
                b_ = nb; l_ = nl; u_ = nu; d_ = nd;
            }
            if (in) red_d[col * RS + lane] = d_ * frcp(b_);
        }
        __syncthreads();
        const double x0 = red_d[s * RS + 2 * p];
        const double xl = red_d[s * RS + 2 * p + 1];
        double *x = d;
        {
            double e = x0;
#pragma unroll
            for (int l = 1; l <= K - 2; ++l) {
                double a, b, c;
                coef(l, a, b, c);
                e = rr[l] * (d[l] - a * e);
                d[l] = e;
                __builtin_amdgcn_sched_barrier(0);
            }
            x[0] = x0;
            x[K - 1] = xl;
#pragma unroll
            for (int l = K - 2; l >= 1; --l) {
                double a, b, c;
                coef(l, a, b, c);
                x[l] = d[l] - rr[l] * c * x[l + 1];
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        double part = 0.0;
#pragma unroll
        for (int l = 0; l < K; ++l) {
            const int g = g0 + l;
            if (g < n3m) {
                part += x[l];
                if (active) st_stream(&w[(size_t)(g + 1) * pstride], (T)x[l]);
            }
        }
        if (active) {
            if (p == 0) w[0] = (T)(bc[0] * x[0] - bc[1] * x[1]);
            const int gt = n3m - 1, pt = gt / K, lt = gt - pt * K;
            if (p == pt) {
                double xt = 0.0, xb = 0.0;
#pragma unroll
                for (int l = 0; l < K; ++l) {
                    if (l == lt) xt = x[l];
                    if (l == lt - 1) xb = x[l];
                }
                if (lt == 0) xb = red_d[s * RS + 2 * (pt - 1) + 1];
                w[(size_t)(n3m + 1) * pstride] = (T)(bc[2] * xt - bc[3] * xb);
            }
        }
        const int kx0 = (tile / z.njr) * S;
        if (jp == z.mean_jp && kx0 <= z.mean_kx && z.mean_kx < kx0 + S) {
            ps[p * S + s] = part;
            __syncthreads();
            if (p == 0 && kx == z.mean_kx) {
                double sum = 0.0;
                for (int q = 0; q < P; ++q) sum += ps[q * S + s];
                *mean_out = sum * z.mean_scale;
            }
        }
        __syncthreads();  // red_* is rewritten by the next tile
    }
}

// ---------------------------------------------------------------------------
// FP64 residual r = f - L p on interior cells of the ghosted P (the operator
// of XRes), with per-workgroup sums of r^2, f^2 and p (deterministic
// two-level reduction with k_reduce3); with xr.rout set, r is also stored in
// FP32 in the W row pass 1 transforms next (the refinement sweep's input).
// A tile is 256 consecutive x cells of TY consecutive rows j (TY divides n2m)
// over KZ consecutive planes.  Each thread walks its TY columns up the planes
// keeping p(k-1), p(k), p(k+1) of all TY rows in registers:
//  * y neighbours inside the block come from those registers; only the two
//    halo rows (j0-1 and j0+TY, or the wall row itself, or the periodic
//    wrap) are loaded;
//  * x neighbours come from the adjacent lanes of the wavefront (DPP wave
//    shifts); only a wave's edge lanes and the domain's edge cells load them;
//  * the tiles of one XCD (workgroup b runs on XCD b mod 8) are consecutive,
//    so a halo row is usually in that XCD's L2 already.
// P is fetched about 1.1 times at 512^3 (one row per thread, with x and y
// neighbours loaded, fetched it about 2.7 times: 35 against 16 B per cell,
// 0.70 against 0.49 ms).  Workgroups stride over the tiles.
// ---------------------------------------------------------------------------
// value of lane l-1 (l+1) of the wavefront: DPP wave_shr:1 (wave_shl:1) per dword
__device__ __forceinline__ double dpp_lane_minus1(double v) {
    const int2 w = __builtin_bit_cast(int2, v);
    const int2 r{__builtin_amdgcn_update_dpp(0, w.x, 0x138, 0xf, 0xf, false),
                 __builtin_amdgcn_update_dpp(0, w.y, 0x138, 0xf, 0xf, false)};
    return __builtin_bit_cast(double, r);
}
__device__ __forceinline__ double dpp_lane_plus1(double v) {
    const int2 w = __builtin_bit_cast(int2, v);
    const int2 r{__builtin_amdgcn_update_dpp(0, w.x, 0x130, 0xf, 0xf, false),
                 __builtin_amdgcn_update_dpp(0, w.y, 0x130, 0xf, 0xf, false)};
    return __builtin_bit_cast(double, r);
}

template <int KZ, int TY, bool STORE>
__global__ void __launch_bounds__(256)
k_residual(const double *__restrict__ f, XRes xr, int n1m, int n2m, int n3m, const int *__restrict__ gate) {
    if (gated_off(gate)) return;
    __shared__ double red[256];
    const int nxc = (n1m + 255) / 256, nyb = n2m / TY, nkb = (n3m + KZ - 1) / KZ;
    const long ntiles = (long)nxc * nyb * nkb;
    const long L1 = (long)n1m + 2, L2 = (long)n2m + 2, pl = L1 * L2;
    const double *__restrict__ P = xr.P;
    double sr = 0.0, sf = 0.0, sx = 0.0;
    // XCD-grouped tile order (a permutation of the workgroups when 8 | grid)
    const long vb = (gridDim.x % 8 == 0) ? (long)(blockIdx.x % 8) * (gridDim.x / 8) + blockIdx.x / 8 : blockIdx.x;
    for (long tile = vb; tile < ntiles; tile += gridDim.x) {
        const int xc = (int)(tile % nxc);
        const long rest = tile / nxc;
        const int jb = (int)(rest % nyb), kb = (int)(rest / nyb);
        const int i = xc * 256 + (int)threadIdx.x;
        if (i >= n1m) continue;
        // x neighbour offsets: zero-flux walls reuse the centre, periodic x wraps
        long dxm = -1, dxp = 1;
        if (xr.pbc1) {
            if (i == 0) dxm = n1m - 1;
            if (i == n1m - 1) dxp = -(long)(n1m - 1);
        } else {
            if (i == 0) dxm = 0;
            if (i == n1m - 1) dxp = 0;
        }
        const int lane = (int)threadIdx.x & 63;
        const bool xedge_m = lane == 0 || i == 0, xedge_p = lane == 63 || i == n1m - 1;
        // halo rows of the block: the row below j0 and the row above j0+TY-1
        // (at a wall the edge row itself, for periodic y the wrapped row)
        const int j0 = jb * TY;
        const int jlo = j0 > 0 ? j0 - 1 : (xr.pbc2 ? n2m - 1 : 0);
        const int jhi = j0 + TY < n2m ? j0 + TY : (xr.pbc2 ? 0 : n2m - 1);
        const long dlo = (long)(jlo - j0) * L1, dhi = (long)(jhi - j0) * L1;
        const int k0 = kb * KZ;
        const long pc0 = ((long)(k0 + 1) * L2 + (j0 + 1)) * L1 + (i + 1);
        double zm[TY], c[TY];  // planes k-1 and k of the TY rows
#pragma unroll
        for (int t = 0; t < TY; ++t) {
            zm[t] = P[pc0 + t * L1 - pl];  // ghost plane below the first (finite, weight 0 at the wall)
            c[t] = P[pc0 + t * L1];
        }
        // planes past n3m load the last plane again (clamped addresses, no
        // branch); their terms are masked out of the sums
#pragma unroll 2
        for (int kk = 0; kk < KZ; ++kk) {
            const int k0k = k0 + kk;
            const bool in = k0k < n3m;
            const int k = in ? k0k : n3m - 1;
            const long pc = pc0 + (long)(k - k0) * pl;
            double zp[TY];
#pragma unroll
            for (int t = 0; t < TY; ++t) zp[t] = P[pc + t * L1 + pl];
            const double hlo = P[pc + dlo], hhi = P[pc + dhi];
            const double am = xr.zam[k], ap = xr.zap[k];
            // STORE: r of row j0+t goes to W32 plane k+1, row j0+t+1 (rows of a
            // column tile are one tile width apart)
            float *rrow = nullptr;
            if constexpr (STORE)
                rrow = xr.rout + (size_t)(k + 1) * (size_t)xr.Ly * (size_t)xr.Lxs + w_at<float>(j0 + 1, i, xr.Ly);
#pragma unroll
            for (int t = 0; t < TY; ++t) {
                const long q = pc + t * L1;
                double xm = dpp_lane_minus1(c[t]), xp = dpp_lane_plus1(c[t]);
                if (xedge_m) xm = P[q + dxm];
                if (xedge_p) xp = P[q + dxp];
                const double ym = t > 0 ? c[t - 1] : hlo, yp = t < TY - 1 ? c[t + 1] : hhi;
                const double fv = ld_stream(&f[((size_t)k * n2m + (j0 + t)) * n1m + i]);
                const double lp = (xp - 2.0 * c[t] + xm) * xr.ih1 + (yp - 2.0 * c[t] + ym) * xr.ih2 + am * zm[t] -
                                  (am + ap) * c[t] + ap * zp[t];
                const double r = fv - lp;
                if (in) {
                    sr += r * r;
                    sf += fv * fv;
                    sx += c[t];
                    if constexpr (STORE) st_stream(rrow + (size_t)t * w_tile<float>(), (float)r);
                }
            }
#pragma unroll
            for (int t = 0; t < TY; ++t) {
                zm[t] = c[t];
                c[t] = zp[t];
            }
        }
    }
    const double tr = block_sum<256>(sr, red);
    const double tf = block_sum<256>(sf, red);
    const double tx = block_sum<256>(sx, red);
    if (threadIdx.x == 0) {
        xr.partial[3 * blockIdx.x] = tr;
        xr.partial[3 * blockIdx.x + 1] = tf;
        xr.partial[3 * blockIdx.x + 2] = tx;
    }
}

// Planes per residual tile, and the tile count for TY rows per tile.
constexpr int kResKZ = 32;
inline long residual_tiles(int n1m, int n2m, int n3m, int ty) {
    return (long)((n1m + 255) / 256) * (n2m / ty) * ((n3m + kResKZ - 1) / kResKZ);
}
// Rows per residual tile: the largest of 8, 4, 2 dividing n2m that still
// leaves min_tiles tiles (small problems need the parallelism more than the
// halo saving), else 1.
inline int residual_rows(int n1m, int n2m, int n3m, long min_tiles) {
    for (int ty = 8; ty > 1; ty >>= 1)
        if (n2m % ty == 0 && residual_tiles(n1m, n2m, n3m, ty) >= min_tiles) return ty;
    return 1;
}
// The launch (grid g <= tiles, one partial triple per workgroup).
template <bool STORE>
inline void launch_k_residual_s(int ty, int g, hipStream_t s, const double *f, const XRes &xr, int n1m, int n2m,
                                int n3m, const int *gate) {
    switch (ty) {
        case 8: hipLaunchKernelGGL((k_residual<kResKZ, 8, STORE>), dim3(g), dim3(256), 0, s, f, xr, n1m, n2m, n3m, gate); break;
        case 4: hipLaunchKernelGGL((k_residual<kResKZ, 4, STORE>), dim3(g), dim3(256), 0, s, f, xr, n1m, n2m, n3m, gate); break;
        case 2: hipLaunchKernelGGL((k_residual<kResKZ, 2, STORE>), dim3(g), dim3(256), 0, s, f, xr, n1m, n2m, n3m, gate); break;
        default: hipLaunchKernelGGL((k_residual<kResKZ, 1, STORE>), dim3(g), dim3(256), 0, s, f, xr, n1m, n2m, n3m, gate); break;
    }
}
// r is stored (FP32, into xr.rout) when xr.rout is set
inline void launch_k_residual(int ty, int g, hipStream_t s, const double *f, const XRes &xr, int n1m, int n2m, int n3m,
                              const int *gate) {
    if (xr.rout) launch_k_residual_s<true>(ty, g, s, f, xr, n1m, n2m, n3m, gate);
    else launch_k_residual_s<false>(ty, g, s, f, xr, n1m, n2m, n3m, gate);
}

}  // namespace ppf
