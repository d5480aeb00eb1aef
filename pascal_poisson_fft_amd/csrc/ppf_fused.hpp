// ppf_fused.hpp -- one persistent launch that solves a whole batch of FP64
// problems: a device work queue of (stage, problem, tile) items.
//
// The five passes of every problem are cut into tiles (the tiles of the
// persistent pass kernels: LPB x-lines, one y column strip of one plane, S z
// columns of one y row).  Items are queued stage-major -- every problem's
// x-forward tiles, then every problem's y-forward tiles, ... -- and a
// workgroup takes the next item with one atomic.  Before running a tile of
// stage s of problem q it waits until all tiles of stage s-1 of q are done
// (per-problem stage counters); those items were queued earlier, so running
// workgroups already hold them and the wait always ends.  Small problems that
// cannot fill 256 CUs alone run side by side, a problem's next stage starts
// as soon as its own previous stage is done, and the batch is one kernel
// launch (plus a counter reset) in a graph.
//
// The tile functions are the pass kernels' tiles for one workgroup of
// kFusedNT threads: threads beyond a tile's own thread count take the
// barriers and nothing else (act = false).  LDS is one dynamic buffer, carved
// per item.  Cross-workgroup visibility: every wave of the producing
// workgroup waits for its stores, then one thread releases them at agent
// scope (write-back of the XCD's L2) with the counting atomic; the consumer
// polls the counter with coherent loads and, once satisfied, acquires once
// (invalidating its CU's and XCD's caches).  An acquire on every poll would
// invalidate the whole L2 each time; with hundreds of workgroups waiting at a
// stage boundary that starves the very tiles they wait for.
//
// Status: switched off (kFusedBatch in ppf_batch.hip).  The launch did not
// complete on the GPU (24 problems of the 96-problem workload), and the
// cause has not been shown: the waits are bounded, the barriers are reached
// by every thread and the ticket counter only grows.  Batches run one plan
// per problem as graph branches instead.
#pragma once

#include "ppf_kernels.hpp"

namespace ppf {

constexpr int kFusedNT = 256;                 // threads per workgroup
constexpr int kFusedK = 32, kFusedS = 16;     // z tiles: K rows per thread, S columns
constexpr int kFusedMaxP = kFusedNT / kFusedS;  // z partitions: n3m <= kFusedK * kFusedMaxP
// A dependency wait ends after kFusedWaitTicks of the 100 MHz device clock
// (2 s) with the timeout flag set; once the flag is set every wait ends and
// the queue drains, so no launch outlives a few seconds whatever goes wrong.
constexpr long long kFusedWaitTicks = 200000000LL;
// line lengths of the fused kernel (x and y)
#define PPF_FUSED_SIZES(X) X(32) X(48) X(64) X(96) X(128) X(192) X(256)
// The tile functions are calls, not inlined: inlined into one kernel they
// spill about 3 KB per lane; as calls the kernel holds 256 VGPRs (two waves
// per SIMD) and saves the callees' registers, about 760 bytes per lane.
#ifndef PPF_FUSED_INLINE
#define PPF_FUSED_INLINE __noinline__
#endif
#ifndef PPF_FUSED_WPE
#define PPF_FUSED_WPE 2
#endif

struct FusedProb {
    const double *in;                          // prhs (n1m, n2m, n3m)
    double *P;                                 // (n1m+2)(n2m+2)(n3m+2)
    const double *dmx1, *dmx2, *dmx3, *dx3;
    double *W;                                 // workspace (Lxs, Ly, Lz), row-major: private to
                                               // this kernel's own stages, not the tiled layout
                                               // of the per-problem passes
    const cx<double> *xtw, *xwk, *xqk, *ytw, *ywk, *yqk;
    const double *lamx, *lamy;
    const unsigned char *zx, *zy;
    double *zcoef;                             // zam[n3m], zap[n3m] + 4 wall coefficients
    double *mean;
    int n1m, n2m, n3m, Lx, Lxs, Ly, Lz;
    int per_x, ymode, zper;
    int zjlo, znjr, mean_kx, mean_jp;
    double mean_scale;
    int nzp;                                   // z partitions
    int ntiles[5];                             // tiles of each stage
};

struct FusedArgs {
    const FusedProb *probs;
    int nprob;
    const int *gstart;  // first item of group g = stage * nprob + problem (5 nprob + 1 entries)
    int nitems;
    int *ctr;           // [0] queue head, [1] timeout flag, [2 + 5 q + s] tiles of stage s of q done
};

// ---- pass 1: x-forward, one tile of LPB lines -------------------------------
template <int N, bool PER>
__device__ PPF_FUSED_INLINE void fused_xfwd(const FusedProb &q, int tile, double *lds) {
    using X = XGeom<double, N, PER>;
    using G = typename X::G;
    using L = typename X::L;
    constexpr int M = G::M, TPL = X::TPL, LPB = X::LPB, NT = X::NT;
    constexpr int LX = PER ? N + 2 : N;
    constexpr int LXS = w_stride(LX, (int)sizeof(double));
    constexpr int H = N / 2, HX = LX / 2;
    constexpr int EPT = (LPB * H + NT - 1) / NT;
    using D2 = double __attribute__((ext_vector_type(2)));
    const int tid = threadIdx.x;
    const bool act = tid < NT;
    const int l = act ? tid / TPL : 0, t = act ? tid - l * TPL : 0;
    const int n2m = q.n2m, nlines = n2m * q.n3m;
    const float rn2m = 1.0f / (float)n2m;
    L::load_tables(lds, q.xtw, q.xwk, q.xqk, tid, kFusedNT);
    if (tile == 0) fill_z_rows(ZRows{q.dx3, q.dmx3, q.zcoef, q.zcoef + q.n3m, q.n3m}, tid, kFusedNT);
    D2 pre[EPT];
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
        const int e = tid + k * NT;
        if (act && e < LPB * H) {
            const int li = e / H, n = 2 * (e - li * H);
            const int line0 = tile * LPB + li, line = line0 < nlines ? line0 : nlines - 1;
            pre[k] = ld_stream(reinterpret_cast<const D2 *>(&q.in[(size_t)line * N + n]));
        }
    }
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
        const int e = tid + k * NT;
        if (act && e < LPB * H) {
            const int li = e / H, h = e - li * H;
            if (PER) {
                L::line(lds, li)[h] = cx<double>{pre[k].x, pre[k].y};
            } else {  // Makhoul: x_2h -> v_h, x_2h+1 -> v_{N-1-h}
                L::real(lds, li)[h] = pre[k].x;
                L::real(lds, li)[N - 1 - h] = pre[k].y;
            }
        }
    }
    __syncthreads();
    const cx<double> *tab = L::tables(lds);
    const LLine<double> my = L::line(lds, l);
    fft_forward<double, M, TPL>(my, t, tab, act);
    if (PER) r2c_post<double, N, TPL>(my, t, tab + M, act);
    else dct_post<double, N, TPL>(my, t, tab + M, tab + 2 * M + 1, act);
    constexpr int HM = PER ? HX - 1 : HX;
    auto store_pair = [&](int li, int h) {
        const int line = tile * LPB + li;
        if (line < nlines) {
            const int kk = fdiv_small(line, n2m, rn2m);
            const cx<double> z = L::line(lds, li)[h];
            st_stream(reinterpret_cast<D2 *>(&q.W[((size_t)line + (size_t)(q.Ly + 1) + 2 * (size_t)kk) * LXS + 2 * h]),
                      D2{z.x, z.y});
        }
    };
    if (act) {
        for (int e = tid; e < LPB * HM; e += NT) {
            const int li = e / HM;
            store_pair(li, e - li * HM);
        }
        if (PER && tid < LPB) store_pair(tid, HX - 1);
    }
}

// ---- pass 5: x-inverse, one tile of LPB W rows -> P rows ---------------------
template <int N, bool PER>
__device__ PPF_FUSED_INLINE void fused_xinv(const FusedProb &q, int tile, double *lds) {
    using X = XGeom<double, N, PER>;
    using G = typename X::G;
    using L = typename X::L;
    constexpr int M = G::M, TPL = X::TPL, LPB = X::LPB, NT = X::NT;
    constexpr int LX = PER ? N + 2 : N;
    constexpr int LXS = w_stride(LX, (int)sizeof(double));
    constexpr int LP = N + 2;
    constexpr int HX = LX / 2, HP = LP / 2;
    constexpr int EPT = (LPB * HX + NT - 1) / NT;
    using D2 = double __attribute__((ext_vector_type(2)));
    const int tid = threadIdx.x;
    const bool act = tid < NT;
    const int l = act ? tid / TPL : 0, t = act ? tid - l * TPL : 0;
    const int nlines = q.Ly * q.Lz;
    L::load_tables(lds, q.xtw, q.xwk, q.xqk, tid, kFusedNT);
    D2 pre[EPT];
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
        const int e = tid + k * NT;
        if (act && e < LPB * HX) {
            const int li = e / HX, h = e - li * HX;
            int row = tile * LPB + li;
            row = row < nlines ? row : nlines - 1;
            pre[k] = ld_stream(reinterpret_cast<const D2 *>(q.W + (size_t)row * LXS) + h);
        }
    }
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
        const int e = tid + k * NT;
        if (act && e < LPB * HX) {
            const int li = e / HX, h = e - li * HX;
            L::line(lds, li)[h] = cx<double>{pre[k].x, pre[k].y};
        }
    }
    const double mu = *q.mean;
    double bca0 = 0, bcb0 = 0, bca1 = 0, bcb1 = 0;
    if (!PER) {  // wall extrapolation (fft_3d_poisson.f90:809-852)
        const double d1 = q.dmx1[1], d2 = q.dmx1[2], e1 = q.dmx1[N + 1], e2 = q.dmx1[N];
        const double s0 = d1 + d2, s1 = e1 + e2;
        bca0 = s0 * s0 / (s0 * s0 - d1 * d1);
        bcb0 = d1 * d1 / (s0 * s0 - d1 * d1);
        bca1 = s1 * s1 / (s1 * s1 - e1 * e1);
        bcb1 = e1 * e1 / (s1 * s1 - e1 * e1);
    }
    const double scale = PER ? 1.0 / (double)N : 1.0 / (2.0 * (double)N);
    __syncthreads();
    const cx<double> *tab = L::tables(lds);
    const LLine<double> my = L::line(lds, l);
    if (PER) c2r_pre<double, N, TPL>(my, t, tab + M, act);
    else idct_pre<double, N, TPL>(my, t, tab + M, tab + 2 * M + 1, act);
    fft_forward<double, M, TPL>(my, t, tab, act);
    auto value = [&](int li, int np) -> double {
        const LReal<double> re = L::real(lds, li);
        auto sample = [&](int n) -> double {
            const int tpos = PER ? n : makhoul_pos<N>(n);
            const double v = re[tpos];
            return ((tpos & 1) ? -v : v) * scale - mu;
        };
        if (np == 0) return PER ? sample(N - 1) : bca0 * sample(0) - bcb0 * sample(1);
        if (np == N + 1) return PER ? sample(0) : bca1 * sample(N - 1) - bcb1 * sample(N - 2);
        return sample(np - 1);
    };
    if (act) {
        for (int e = tid; e < LPB * HP; e += NT) {
            const int li = e / HP, h = e - li * HP;
            const int line = tile * LPB + li;
            if (line < nlines)
                st_stream(reinterpret_cast<D2 *>(&q.P[(size_t)line * LP + 2 * h]),
                          D2{value(li, 2 * h), value(li, 2 * h + 1)});
        }
    }
}

// ---- passes 2 and 4: one y column strip of one plane -------------------------
template <int N, int MODE>
__device__ PPF_FUSED_INLINE void fused_yfwd(const FusedProb &q, int tile, double *lds) {
    using Y = YGeom<double, N, MODE, 0, kFusedNT>;
    using L = typename Y::L;
    constexpr int M = Y::M, TPL = Y::TPL, TX = Y::TX, NT = Y::NT;
    constexpr int TP = TX / 2;
    constexpr int EPT = (TP * N + NT - 1) / NT;
    constexpr int RPT = Y::RPT, NPRE = Y::RB ? RPT : EPT;
    using V2 = double __attribute__((ext_vector_type(2)));
    const int tid = threadIdx.x;
    const bool act = tid < NT;
    const int l = act ? tid / TPL : 0, t = act ? tid - l * TPL : 0;
    const int Lx = q.Lx, Lxs = q.Lxs;
    const int nct = (Lx + TX - 1) / TX;
    const int ct = tile % nct;
    const int c0 = ct * TX;
    const size_t plane = (size_t)(tile / nct + 1) * q.Ly;
    double *W = q.W;
    const int bg = act ? tid / TP : 0, bc2 = act ? tid - bg * TP : 0;
    L::load_tables(lds, q.ytw, q.ywk, q.yqk, tid, kFusedNT);
    V2 pre[NPRE];
    if (act) {
        if constexpr (Y::RB) {
            int col = c0 + 2 * bc2;
            col = col < Lx ? col : Lx - 2;
            const double *src = &W[(plane + 1 + (size_t)RPT * bg) * Lxs + col];
#pragma unroll
            for (int k = 0; k < RPT; ++k) pre[k] = ld_stream(reinterpret_cast<const V2 *>(src + (size_t)k * Lxs));
        } else {
#pragma unroll
            for (int k = 0; k < EPT; ++k) {
                int e = tid + k * NT;
                e = e < TP * N ? e : TP * N - 1;
                const int r = e / TP, c2 = e - r * TP;
                int col = c0 + 2 * c2;
                col = col < Lx ? col : Lx - 2;
                pre[k] = ld_stream(reinterpret_cast<const V2 *>(&W[(plane + 1 + r) * Lxs + col]));
            }
        }
        if constexpr (Y::RB && MODE == 0) {
            const LLine<double> c0l = L::line(lds, 2 * bc2), c1l = L::line(lds, 2 * bc2 + 1);
#pragma unroll
            for (int j = 0; j < RPT / 4; ++j) {
                const int ce = (RPT / 4) * bg + j, co = N / 2 - 1 - (RPT / 4) * bg - j;
                c0l[ce] = cx<double>{pre[4 * j].x, pre[4 * j + 2].x};
                c1l[ce] = cx<double>{pre[4 * j].y, pre[4 * j + 2].y};
                c0l[co] = cx<double>{pre[4 * j + 3].x, pre[4 * j + 1].x};
                c1l[co] = cx<double>{pre[4 * j + 3].y, pre[4 * j + 1].y};
            }
        } else if constexpr (Y::RB) {
            const LLine<double> c0l = L::line(lds, 2 * bc2), c1l = L::line(lds, 2 * bc2 + 1);
#pragma unroll
            for (int k = 0; k < RPT / 2; ++k) {
                c0l[(RPT / 2) * bg + k] = cx<double>{pre[2 * k].x, pre[2 * k + 1].x};
                c1l[(RPT / 2) * bg + k] = cx<double>{pre[2 * k].y, pre[2 * k + 1].y};
            }
        } else {
#pragma unroll
            for (int k = 0; k < EPT; ++k) {
                const int e = tid + k * NT;
                if (e < TP * N) {
                    const int r = e / TP, c2 = e - r * TP;
                    const V2 v = pre[k];
                    if (MODE == 0) {
                        L::real(lds, 2 * c2)[makhoul_pos<N>(r)] = v.x;
                        L::real(lds, 2 * c2 + 1)[makhoul_pos<N>(r)] = v.y;
                    } else if (MODE == 1) {
                        L::line(lds, c2)[r] = cx<double>{v.x, v.y};
                    } else {
                        L::real(lds, 2 * c2)[r] = v.x;
                        L::real(lds, 2 * c2 + 1)[r] = v.y;
                    }
                }
            }
        }
    }
    __syncthreads();
    const cx<double> *tab = L::tables(lds);
#pragma unroll 1
    for (int rd = 0; rd < Y::ROUNDS; ++rd) {
        const LLine<double> my = L::line(lds, l + rd * (Y::NL / Y::ROUNDS));
        fft_forward<double, M, TPL>(my, t, tab, act);
        if (MODE == 0) dct_post<double, N, TPL>(my, t, tab + M, tab + 2 * M + 1, act);
        if (MODE == 2) r2c_post<double, N, TPL>(my, t, tab + M, act);
    }
    constexpr int NR = MODE == 2 ? N + 2 : N;
    constexpr int R0 = MODE == 2 ? 0 : 1;
    if (!act) return;
    if constexpr (Y::RB) {
        const int col = c0 + 2 * bc2;
        if (col < Lx) {
            const LLine<double> c0l = L::line(lds, 2 * bc2), c1l = L::line(lds, 2 * bc2 + 1);
            double *dst = &W[(plane + R0 + (size_t)RPT * bg) * Lxs + col];
#pragma unroll
            for (int k = 0; k < RPT / 2; ++k) {
                const cx<double> z0 = c0l[(RPT / 2) * bg + k], z1 = c1l[(RPT / 2) * bg + k];
                st_stream(reinterpret_cast<V2 *>(dst + (size_t)(2 * k) * Lxs), V2{z0.x, z1.x});
                st_stream(reinterpret_cast<V2 *>(dst + (size_t)(2 * k + 1) * Lxs), V2{z0.y, z1.y});
            }
            if (MODE == 2 && bg == Y::G - 1) {
                const cx<double> z0 = c0l[N / 2], z1 = c1l[N / 2];
                st_stream(reinterpret_cast<V2 *>(dst + (size_t)RPT * Lxs), V2{z0.x, z1.x});
                st_stream(reinterpret_cast<V2 *>(dst + (size_t)(RPT + 1) * Lxs), V2{z0.y, z1.y});
            }
        }
    } else {
        for (int e = tid; e < TP * NR; e += NT) {
            const int r = e / TP, c2 = e - r * TP;
            const int col = c0 + 2 * c2;
            if (col >= Lx) continue;
            V2 v;
            if (MODE == 1) {
                const cx<double> z = L::line(lds, c2)[r];
                v.x = z.x;
                v.y = z.y;
            } else {
                v.x = L::real(lds, 2 * c2)[r];
                v.y = L::real(lds, 2 * c2 + 1)[r];
            }
            st_stream(reinterpret_cast<V2 *>(&W[(plane + R0 + r) * Lxs + col]), v);
        }
    }
}

template <int N, int MODE>
__device__ PPF_FUSED_INLINE void fused_yinv(const FusedProb &q, int tile, double *lds) {
    using Y = YGeom<double, N, MODE, 0, kFusedNT>;
    using L = typename Y::L;
    constexpr int M = Y::M, TPL = Y::TPL, TX = Y::TX, NT = Y::NT;
    constexpr int NR = MODE == 2 ? N + 2 : N;
    constexpr int R0 = MODE == 2 ? 0 : 1;
    constexpr int TP = TX / 2;
    constexpr int EPT = (TP * NR + NT - 1) / NT;
    constexpr int RPT = Y::RPT, NPRE = Y::RB ? RPT : EPT;
    using V2 = double __attribute__((ext_vector_type(2)));
    const int tid = threadIdx.x;
    const bool act = tid < NT;
    const int l = act ? tid / TPL : 0, t = act ? tid - l * TPL : 0;
    const int Lx = q.Lx, Lxs = q.Lxs;
    const int nct = (Lx + TX - 1) / TX;
    const int c0 = (tile % nct) * TX;
    const size_t plane = (size_t)(tile / nct) * q.Ly;
    double *W = q.W;
    const int bg = act ? tid / TP : 0, bc2 = act ? tid - bg * TP : 0;
    L::load_tables(lds, q.ytw, q.ywk, q.yqk, tid, kFusedNT);
    const double scale = MODE == 0 ? 1.0 / (2.0 * N) : 1.0 / (double)N;
    double bca0 = 0, bcb0 = 0, bca1 = 0, bcb1 = 0;
    if (MODE == 0) {  // y wall ghosts (fft_3d_poisson.f90:809-852)
        const double d1 = q.dmx2[1], d2 = q.dmx2[2], e1 = q.dmx2[N + 1], e2 = q.dmx2[N];
        const double s0 = d1 + d2, s1 = e1 + e2;
        bca0 = s0 * s0 / (s0 * s0 - d1 * d1);
        bcb0 = d1 * d1 / (s0 * s0 - d1 * d1);
        bca1 = s1 * s1 / (s1 * s1 - e1 * e1);
        bcb1 = e1 * e1 / (s1 * s1 - e1 * e1);
    }
    V2 pre[NPRE];
    V2 pre_x[2] = {V2{0, 0}, V2{0, 0}};
    if (act) {
        if constexpr (Y::RB) {
            int col = c0 + 2 * bc2;
            col = col < Lx ? col : Lx - 2;
            const double *src = &W[(plane + R0 + (size_t)RPT * bg) * Lxs + col];
#pragma unroll
            for (int k = 0; k < RPT; ++k) pre[k] = ld_stream(reinterpret_cast<const V2 *>(src + (size_t)k * Lxs));
            if (MODE == 2 && bg == Y::G - 1) {
                pre_x[0] = ld_stream(reinterpret_cast<const V2 *>(src + (size_t)RPT * Lxs));
                pre_x[1] = ld_stream(reinterpret_cast<const V2 *>(src + (size_t)(RPT + 1) * Lxs));
            }
            const LLine<double> c0l = L::line(lds, 2 * bc2), c1l = L::line(lds, 2 * bc2 + 1);
#pragma unroll
            for (int k = 0; k < RPT / 2; ++k) {
                c0l[(RPT / 2) * bg + k] = cx<double>{pre[2 * k].x, pre[2 * k + 1].x};
                c1l[(RPT / 2) * bg + k] = cx<double>{pre[2 * k].y, pre[2 * k + 1].y};
            }
            if (MODE == 2 && bg == Y::G - 1) {
                c0l[N / 2] = cx<double>{pre_x[0].x, pre_x[1].x};
                c1l[N / 2] = cx<double>{pre_x[0].y, pre_x[1].y};
            }
        } else {
#pragma unroll
            for (int k = 0; k < EPT; ++k) {
                int e = tid + k * NT;
                e = e < TP * NR ? e : TP * NR - 1;
                const int r = e / TP, c2 = e - r * TP;
                int col = c0 + 2 * c2;
                col = col < Lx ? col : Lx - 2;
                pre[k] = ld_stream(reinterpret_cast<const V2 *>(&W[(plane + R0 + r) * Lxs + col]));
            }
#pragma unroll
            for (int k = 0; k < EPT; ++k) {
                const int e = tid + k * NT;
                if (e < TP * NR) {
                    const int r = e / TP, c2 = e - r * TP;
                    const V2 v = pre[k];
                    if (MODE == 1) {
                        L::line(lds, c2)[r] = cx<double>{v.x, -v.y};
                    } else {
                        L::real(lds, 2 * c2)[r] = v.x;
                        L::real(lds, 2 * c2 + 1)[r] = v.y;
                    }
                }
            }
        }
    }
    __syncthreads();
    const cx<double> *tab = L::tables(lds);
#pragma unroll 1
    for (int rd = 0; rd < Y::ROUNDS; ++rd) {
        const LLine<double> my = L::line(lds, l + rd * (Y::NL / Y::ROUNDS));
        if (MODE == 0) idct_pre<double, N, TPL>(my, t, tab + M, tab + 2 * M + 1, act);
        if (MODE == 2) c2r_pre<double, N, TPL>(my, t, tab + M, act);
        fft_forward<double, M, TPL>(my, t, tab, act);
    }
    if (!act) return;
    if constexpr (Y::RB && MODE == 2) {
        const int col = c0 + 2 * bc2;
        if (col < Lx) {
            const LLine<double> c0l = L::line(lds, 2 * bc2), c1l = L::line(lds, 2 * bc2 + 1);
            V2 v[RPT];
#pragma unroll
            for (int k = 0; k < RPT / 2; ++k) {
                const cx<double> a0 = c0l[(RPT / 2) * bg + k], a1 = c1l[(RPT / 2) * bg + k];
                v[2 * k] = V2{a0.x, a1.x} * scale;
                v[2 * k + 1] = V2{a0.y, a1.y} * -scale;
            }
            double *dst = &W[(plane + 1 + (size_t)RPT * bg) * Lxs + col];
#pragma unroll
            for (int k = 0; k < RPT; ++k) st_stream(reinterpret_cast<V2 *>(dst + (size_t)k * Lxs), v[k]);
            if (bg == 0) st_stream(reinterpret_cast<V2 *>(&W[(plane + N + 1) * Lxs + col]), v[0]);
            if (bg == Y::G - 1) st_stream(reinterpret_cast<V2 *>(&W[plane * Lxs + col]), v[RPT - 1]);
        }
    } else if constexpr (Y::RB) {
        const int col = c0 + 2 * bc2;
        if (col < Lx) {
            const LLine<double> c0l = L::line(lds, 2 * bc2), c1l = L::line(lds, 2 * bc2 + 1);
            V2 v[RPT];
#pragma unroll
            for (int j = 0; j < RPT / 4; ++j) {
                const int ce = (RPT / 4) * bg + j, co = N / 2 - 1 - (RPT / 4) * bg - j;
                const cx<double> a0 = c0l[ce], a1 = c1l[ce], b0 = c0l[co], b1 = c1l[co];
                v[4 * j] = V2{a0.x, a1.x} * scale;
                v[4 * j + 2] = V2{a0.y, a1.y} * -scale;
                v[4 * j + 3] = V2{b0.x, b1.x} * scale;
                v[4 * j + 1] = V2{b0.y, b1.y} * -scale;
            }
            double *dst = &W[(plane + 1 + (size_t)RPT * bg) * Lxs + col];
#pragma unroll
            for (int k = 0; k < RPT; ++k) st_stream(reinterpret_cast<V2 *>(dst + (size_t)k * Lxs), v[k]);
            if (bg == 0) st_stream(reinterpret_cast<V2 *>(dst - Lxs), bca0 * v[0] - bcb0 * v[1]);
            if (bg == Y::G - 1)
                st_stream(reinterpret_cast<V2 *>(dst + (size_t)RPT * Lxs), bca1 * v[RPT - 1] - bcb1 * v[RPT - 2]);
        }
    } else {
        for (int e = tid; e < TP * (N + 2); e += NT) {
            const int rp = e / TP, c2 = e - rp * TP;
            const int col = c0 + 2 * c2;
            if (col >= Lx) continue;
            auto sample = [&](int n) -> V2 {
                V2 v;
                if (MODE == 1) {
                    const cx<double> z = L::line(lds, c2)[n];
                    v.x = z.x * scale;
                    v.y = -z.y * scale;
                } else {
                    const int tpos = MODE == 0 ? makhoul_pos<N>(n) : n;
                    const double sg = (tpos & 1) ? -1.0 : 1.0;
                    v.x = sg * L::real(lds, 2 * c2)[tpos] * scale;
                    v.y = sg * L::real(lds, 2 * c2 + 1)[tpos] * scale;
                }
                return v;
            };
            V2 out;
            if (rp == 0) out = MODE == 0 ? bca0 * sample(0) - bcb0 * sample(1) : sample(N - 1);
            else if (rp == N + 1) out = MODE == 0 ? bca1 * sample(N - 1) - bcb1 * sample(N - 2) : sample(0);
            else out = sample(rp - 1);
            st_stream(reinterpret_cast<V2 *>(&W[(plane + rp) * Lxs + col]), out);
        }
    }
}

// ---- pass 3: z Thomas, S columns of one spectral row -------------------------
// The partitioned register Thomas of k_zpart (K = 32 rows per thread, S = 16
// columns), on the S * P threads of the tile.
__device__ PPF_FUSED_INLINE void fused_z(const FusedProb &q, int tile, double *lds) {
    constexpr int K = kFusedK, S = kFusedS;
    constexpr int KS = (K + S - 1) / S;
    const int n3m = q.n3m, P = q.nzp, PS = P * S;
    const int NPAD = KS * kFusedNT;
    double *am = lds, *ap = am + NPAD;
    double *ra = ap + NPAD, *rc = ra + 3 * PS, *xe = rc + 3 * PS, *ps = ra;
    const int tid = threadIdx.x;
    const bool act = tid < PS;
    const int s = tid % S, p = act ? tid / S : 0;
    const int tiles = (q.Lx + S - 1) / S;
    const int jp = q.zjlo + tile / tiles;
    const int kx = (tile % tiles) * S + s;
    const bool active = act && kx < q.Lx;
    const size_t pstride = (size_t)q.Ly * q.Lxs;
    double *w = q.W + (size_t)jp * q.Lxs + (active ? kx : 0);
    const double lx = active ? q.lamx[kx] : 0.0, ly = q.lamy[jp];
    const bool pin = active && q.zx[kx] && q.zy[jp];
    const int g0 = p * K, glast = n3m - 1;
    const double *zam = q.zcoef, *zap = q.zcoef + n3m;
#pragma unroll
    for (int j = 0; j < KS; ++j) {
        const int k = tid + j * kFusedNT, kc = k < glast ? k : glast;
        am[k] = k < n3m ? zam[kc] : 0.0;
        ap[k] = k < n3m ? zap[kc] : 0.0;
    }
    double bc[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) bc[k] = zap[n3m + k];
    double d[K];
#pragma unroll
    for (int l = 0; l < K; ++l) {
        const int g = g0 + l;
        d[l] = act ? ld_stream(&w[(size_t)((g < glast ? g : glast) + 1) * pstride]) : 0.0;
        d[l] = (active && g < n3m && !(pin && g == 0)) ? d[l] : 0.0;
    }
    __syncthreads();
    auto coef = [&](int l, double &a, double &b, double &c) {
        const int g = g0 + l;
        const bool id = (g >= n3m) || (pin && g == 0);
        const double am_ = am[g], ap_ = ap[g];
        a = am_;
        c = id ? 0.0 : ap_;
        b = id ? 1.0 : (-am_ - ap_ - lx) - ly;
    };
    double rr[K];
    double ah = 0, ch = 0, dh = 0, Af = 0, Df = 0, Cf = 0;
    if (act) {
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
        }
    }
    const int o = p * S + s;
    if (act) {
        ra[o] = ah;
        ra[PS + o] = ch;
        ra[2 * PS + o] = dh;
    }
    __syncthreads();
    double el = 0, eu = 0, er = 0;
    if (act) {
        double ah1 = 0, ch1 = 0, dh1 = 0;
        if (p + 1 < P) {
            ah1 = ra[o + S];
            ch1 = ra[PS + o + S];
            dh1 = ra[2 * PS + o + S];
        }
        const double inv = frcp(1.0 - Af * ch - Cf * ah1);
        el = -Af * ah * inv;
        eu = -Cf * ch1 * inv;
        er = (Df - Af * dh - Cf * dh1) * inv;
    }
    {
        double *buf = rc;
        for (int h = 1; h < P; h <<= 1) {
            if (act) {
                buf[o] = el;
                buf[PS + o] = eu;
                buf[2 * PS + o] = er;
            }
            __syncthreads();
            if (act) {
                double lm = 0, um = 0, rm = 0, lp = 0, up = 0, rp = 0;
                if (p >= h) {
                    const int qq = o - h * S;
                    lm = buf[qq]; um = buf[PS + qq]; rm = buf[2 * PS + qq];
                }
                if (p + h < P) {
                    const int qq = o + h * S;
                    lp = buf[qq]; up = buf[PS + qq]; rp = buf[2 * PS + qq];
                }
                const double inv = frcp(1.0 - el * um - eu * lp);
                er = (er - el * rm - eu * rp) * inv;
                el = -el * lm * inv;
                eu = -eu * up * inv;
            }
            buf = (buf == rc) ? ra : rc;
        }
    }
    const double xl = er;
    if (act) xe[o] = xl;
    __syncthreads();
    double part = 0.0;
    if (act) {
        const double x0 = dh - ah * (p > 0 ? xe[o - S] : 0.0) - ch * xl;
        double *x = d;
        double e = x0;
#pragma unroll
        for (int l = 1; l <= K - 2; ++l) {
            double a, b, c;
            coef(l, a, b, c);
            e = rr[l] * (d[l] - a * e);
            d[l] = e;
        }
        x[0] = x0;
        x[K - 1] = xl;
#pragma unroll
        for (int l = K - 2; l >= 1; --l) {
            double a, b, c;
            coef(l, a, b, c);
            x[l] = d[l] - rr[l] * c * x[l + 1];
        }
#pragma unroll
        for (int l = 0; l < K; ++l) {
            const int g = g0 + l;
            if (g < n3m) {
                part += x[l];
                if (active) st_stream(&w[(size_t)(g + 1) * pstride], x[l]);
            }
        }
        if (active) {
            const int gt = n3m - 1, pt = gt / K, lt = gt - pt * K;
            if (p == 0) {
                if (q.zper) w[(size_t)(n3m + 1) * pstride] = x[0];
                else w[0] = bc[0] * x[0] - bc[1] * x[1];
            }
            if (p == pt) {
                double xt = 0, xb = 0;
#pragma unroll
                for (int l = 0; l < K; ++l) {
                    if (l == lt) xt = x[l];
                    if (l == lt - 1) xb = x[l];
                }
                if (lt == 0) xb = xe[(pt - 1) * S + s];
                if (q.zper) w[0] = xt;
                else w[(size_t)(n3m + 1) * pstride] = bc[2] * xt - bc[3] * xb;
            }
        }
    }
    // mean of the solution from the (0,0) real column (:628-664)
    const bool mean_tile = (jp == q.mean_jp) && (tile % tiles) * S <= q.mean_kx && q.mean_kx < (tile % tiles) * S + S;
    if (mean_tile) {
        if (act) ps[p * S + s] = part;
        __syncthreads();
        if (act && p == 0 && kx == q.mean_kx) {
            double sum = 0.0;
            for (int k = 0; k < P; ++k) sum += ps[k * S + s];
            *q.mean = sum * q.mean_scale;
        }
    }
}

// ---- the queue -----------------------------------------------------------------
__device__ __forceinline__ void fused_item(const FusedProb &q, int stage, int tile, double *lds) {
    switch (stage) {
        case 0:
            switch (q.n1m) {
#define PPF_FX(n) case n: if (q.per_x) fused_xfwd<n, true>(q, tile, lds); else fused_xfwd<n, false>(q, tile, lds); break;
                PPF_FUSED_SIZES(PPF_FX)
#undef PPF_FX
            }
            break;
        case 1:
        case 3:
            switch (q.n2m) {
#define PPF_FY(n)                                                                        \
    case n:                                                                              \
        if (stage == 1) {                                                                \
            if (q.ymode == 0) fused_yfwd<n, 0>(q, tile, lds);                            \
            else if (q.ymode == 1) fused_yfwd<n, 1>(q, tile, lds);                       \
            else fused_yfwd<n, 2>(q, tile, lds);                                         \
        } else {                                                                         \
            if (q.ymode == 0) fused_yinv<n, 0>(q, tile, lds);                            \
            else if (q.ymode == 1) fused_yinv<n, 1>(q, tile, lds);                       \
            else fused_yinv<n, 2>(q, tile, lds);                                         \
        }                                                                                \
        break;
                PPF_FUSED_SIZES(PPF_FY)
#undef PPF_FY
            }
            break;
        case 2: fused_z(q, tile, lds); break;
        default:
            switch (q.n1m) {
#define PPF_FX(n) case n: if (q.per_x) fused_xinv<n, true>(q, tile, lds); else fused_xinv<n, false>(q, tile, lds); break;
                PPF_FUSED_SIZES(PPF_FX)
#undef PPF_FX
            }
            break;
    }
}

__global__ void __launch_bounds__(kFusedNT) __attribute__((amdgpu_waves_per_eu(PPF_FUSED_WPE))) k_fused_batch(FusedArgs a) {
    extern __shared__ double fsh[];
    __shared__ int sh[2];
    const int tid = threadIdx.x;
    for (;;) {
        if (tid == 0) {
            int it = atomicAdd(&a.ctr[0], 1);
            // after a timeout the remaining items are skipped (the queue drains)
            if (__hip_atomic_load(&a.ctr[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) it = a.nitems;
            int g = 0;
            if (it < a.nitems) {
                int lo = 0, hi = 5 * a.nprob - 1;  // last group starting at or before it
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (a.gstart[mid] <= it) lo = mid;
                    else hi = mid - 1;
                }
                g = lo;
                const int stage = g / a.nprob, prob = g - stage * a.nprob;
                if (stage > 0) {
                    const int need = a.probs[prob].ntiles[stage - 1];
                    int *done = &a.ctr[2 + 5 * prob + stage - 1];
                    const long long t0 = wall_clock64();
                    while (__hip_atomic_load(done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
                        __builtin_amdgcn_s_sleep(8);
                        // a bounded wait: flag it (which ends every other wait), never hang
                        if (wall_clock64() - t0 > kFusedWaitTicks ||
                            __hip_atomic_load(&a.ctr[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                            atomicOr(&a.ctr[1], 1);
                            break;
                        }
                    }
                    // once: the previous stage's stores, from any XCD
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                }
            }
            sh[0] = it;
            sh[1] = g;
        }
        __syncthreads();
        const int it = sh[0], g = sh[1];
        if (it >= a.nitems) break;
        const int stage = g / a.nprob, prob = g - stage * a.nprob;
        fused_item(a.probs[prob], stage, it - a.gstart[g], fsh);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");  // this wave's stores are out
        __syncthreads();
        // this tile's stores, to every XCD (write-back completed), then the count
        if (tid == 0) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            __hip_atomic_fetch_add(&a.ctr[2 + 5 * prob + stage], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// counters back to zero before a launch (a kernel: graphs replay it)
__global__ void k_fused_reset(int *ctr, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) ctr[i] = 0;
}

}  // namespace ppf
