// Full-GPU kernels for one large branch node: batched complex128 GEMMs on the FP64 matrix
// cores, used by the node-level entry points (node.cpp) that apply the effective
// Hamiltonian, build mean fields and run the Krylov propagator of a single node tensor.
//
// zgemm_kernel: C[b] (M x N) = alpha op(A[b]) (M x K) B[b] (K x N) (+ C), b = b1 * nb0 + b0.
//   * one output tile per workgroup (256 threads, 4 wave64 in a 2 x 2 grid, each wave a
//     WM x WN block of 16 x 16 MFMA tiles: 32 x 32 up to 128 x 64 and 64 x 128, the shape
//     chosen per GEMM by the node path's autotuner), K in chunks of 16 staged through LDS
//     with re and im planes kept apart so every MFMA operand is one LDS read;
//   * the next chunk's global loads are issued before the current chunk's MFMAs;
//   * K may be blocked: k = kblk * kb + kin, with a separate stride per block for A and B, so
//     a sum of products over terms (each term a contiguous K block) is one GEMM;
//   * split K: gridDim.z > 1 writes partial products to a workspace that zreduce_kernel sums
//     in a fixed order (results do not depend on scheduling).
// V_MFMA_F64_16X16X4F64 operand layout (one double per lane): A (16 x 4): lane l holds
// A[l % 16][l / 16]; B (4 x 16): lane l holds B[l / 16][l % 16]; D (16 x 16): lane l holds
// D[l / 16 + 4 j][l % 16] in accumulator element j.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "node_gemm.hpp"

namespace ttnd {

namespace {

constexpr int GT = 256;     // threads
constexpr int BK = 16;

typedef double d4 __attribute__((ext_vector_type(4)));

struct Acc {
    d4 re, im;
};

__device__ __forceinline__ void mfma_c(Acc& acc, double ar, double ai, double br, double bi) {
    acc.re = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, br, acc.re, 0, 0, 0);
    acc.re = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai, bi, acc.re, 0, 0, 0);
    acc.im = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, bi, acc.im, 0, 0, 0);
    acc.im = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, br, acc.im, 0, 0, 0);
}

// LDS row length (doubles) of a panel of width W: rows kq and kq + 1 fall 32 banks apart
template <int W>
struct Pad {
    static constexpr int v = W + 16;
};

template <int BM, int BN>
struct Panels {
    double ar[BK][Pad<BM>::v], ai[BK][Pad<BM>::v];  // A(row, k) at [k][row]
    double br[BK][Pad<BN>::v], bi[BK][Pad<BN>::v];  // B(k, col) at [k][col]
};

// Global -> register staging of one K chunk: BM / 16 A and BN / 16 B elements per thread.
template <int BM, int BN>
struct Stage {
    double2 a[BM / 16], b[BN / 16];
};

__device__ __forceinline__ int64_t a_off(const GemmDesc& g, int row, int k) {
    const int kb = k / g.kb, ki = k - kb * g.kb;
    return static_cast<int64_t>(row) * g.a_r + static_cast<int64_t>(ki) * g.a_k + static_cast<int64_t>(kb) * g.a_kblk;
}
__device__ __forceinline__ int64_t b_off(const GemmDesc& g, int k, int col) {
    const int kb = k / g.kb, ki = k - kb * g.kb;
    return static_cast<int64_t>(ki) * g.b_k + static_cast<int64_t>(col) * g.b_c + static_cast<int64_t>(kb) * g.b_kblk;
}

// element e (0 .. BM/16 - 1) of this thread in the A panel: (row, k) within the tile; with
// k fastest in memory 16 consecutive threads take one row's 16 k, otherwise BM consecutive
// threads take one k's rows
template <int BM>
__device__ __forceinline__ void a_slot(bool kfast, int t, int e, int& row, int& k) {
    if (kfast) {
        k = t & 15;
        row = (t >> 4) + 16 * e;
    } else {
        row = t % BM;
        k = t / BM + (GT / BM) * e;
    }
}
template <int BN>
__device__ __forceinline__ void b_slot(bool cfast, int t, int e, int& k, int& col) {
    if (cfast) {
        col = t % BN;
        k = t / BN + (GT / BN) * e;
    } else {
        k = t & 15;
        col = (t >> 4) + 16 * e;
    }
}

template <int BM, int BN>
__device__ __forceinline__ void load_stage(const GemmDesc& g, const double2* A, const double2* B, int row0, int col0,
                                           int k0, int kend, bool kfast, bool cfast, Stage<BM, BN>& s) {
    const int t = threadIdx.x;
    const double2 z = make_double2(0.0, 0.0);
#pragma unroll
    for (int e = 0; e < BM / 16; ++e) {
        int row, k;
        a_slot<BM>(kfast, t, e, row, k);
        const int gr = row0 + row, gk = k0 + k;
        s.a[e] = (gr < g.M && gk < kend) ? A[a_off(g, gr, gk)] : z;
    }
#pragma unroll
    for (int e = 0; e < BN / 16; ++e) {
        int kk, col;
        b_slot<BN>(cfast, t, e, kk, col);
        const int gk2 = k0 + kk, gc = col0 + col;
        s.b[e] = (gc < g.N && gk2 < kend) ? B[b_off(g, gk2, gc)] : z;
    }
}

template <int BM, int BN>
__device__ __forceinline__ void store_stage(Panels<BM, BN>& p, bool kfast, bool cfast, bool conja,
                                            const Stage<BM, BN>& s) {
    const int t = threadIdx.x;
#pragma unroll
    for (int e = 0; e < BM / 16; ++e) {
        int row, k;
        a_slot<BM>(kfast, t, e, row, k);
        p.ar[k][row] = s.a[e].x;
        p.ai[k][row] = conja ? -s.a[e].y : s.a[e].y;
    }
#pragma unroll
    for (int e = 0; e < BN / 16; ++e) {
        int kk, col;
        b_slot<BN>(cfast, t, e, kk, col);
        p.br[kk][col] = s.b[e].x;
        p.bi[kk][col] = s.b[e].y;
    }
}

}  // namespace

// One (2 WM 16) x (2 WN 16) output tile per workgroup: four waves in a 2 x 2 grid, each a
// WM x WN block of 16 x 16 MFMA tiles.  Every output element sums its K range in the same
// order for every WM, WN, so all tilings give the same bits.
template <int WM, int WN>
__global__ __launch_bounds__(GT) void zgemm_kernel(GemmDesc g) {
    constexpr int BM = 32 * WM, BN = 32 * WN;
    __shared__ Panels<BM, BN> P;
    const int tiles_n = (g.N + BN - 1) / BN;
    const int tile = blockIdx.x;
    const int row0 = (tile / tiles_n) * BM, col0 = (tile % tiles_n) * BN;
    const int b = blockIdx.y;
    const int b1 = b / g.nb0, b0 = b - b1 * g.nb0;
    const double2* A = g.A + b0 * g.a_b0 + b1 * g.a_b1;
    const double2* B = g.B + b0 * g.b_b0 + b1 * g.b_b1;
    // split K: this workgroup's k range
    const int nsplit = gridDim.z, sp = blockIdx.z;
    const int kchunks = (g.K + BK - 1) / BK;
    const int c0 = static_cast<int>(static_cast<int64_t>(kchunks) * sp / nsplit);
    const int c1 = static_cast<int>(static_cast<int64_t>(kchunks) * (sp + 1) / nsplit);
    const bool kfast = g.a_k == 1;
    const bool cfast = g.b_c == 1;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int wr = wv >> 1, wc = wv & 1;
    const int li = lane & 15, kq = lane >> 4;

    Acc acc[WM][WN];
#pragma unroll
    for (int i = 0; i < WM; ++i)
#pragma unroll
        for (int j = 0; j < WN; ++j) {
            acc[i][j].re = d4{0.0, 0.0, 0.0, 0.0};
            acc[i][j].im = d4{0.0, 0.0, 0.0, 0.0};
        }
    Stage<BM, BN> st;
    if (c0 < c1) load_stage<BM, BN>(g, A, B, row0, col0, c0 * BK, g.K, kfast, cfast, st);
    for (int c = c0; c < c1; ++c) {
        __syncthreads();
        store_stage<BM, BN>(P, kfast, cfast, g.conja != 0, st);
        __syncthreads();
        if (c + 1 < c1) load_stage<BM, BN>(g, A, B, row0, col0, (c + 1) * BK, g.K, kfast, cfast, st);
#pragma unroll
        for (int ks = 0; ks < BK; ks += 4) {
            double ar[WM], ai[WM], br[WN], bi[WN];
#pragma unroll
            for (int i = 0; i < WM; ++i) {
                ar[i] = P.ar[ks + kq][wr * 16 * WM + i * 16 + li];
                ai[i] = P.ai[ks + kq][wr * 16 * WM + i * 16 + li];
            }
#pragma unroll
            for (int j = 0; j < WN; ++j) {
                br[j] = P.br[ks + kq][wc * 16 * WN + j * 16 + li];
                bi[j] = P.bi[ks + kq][wc * 16 * WN + j * 16 + li];
            }
#pragma unroll
            for (int i = 0; i < WM; ++i)
#pragma unroll
                for (int j = 0; j < WN; ++j) mfma_c(acc[i][j], ar[i], ai[i], br[j], bi[j]);
        }
    }
    // epilogue: workspace (split K) or C
    double2* C = g.C + b0 * g.c_b0 + b1 * g.c_b1;
    if (nsplit > 1) C = g.W + sp * g.w_split + static_cast<int64_t>(b) * g.w_batch;
    const int64_t cr = nsplit > 1 ? g.N : g.c_r;
    const int64_t cc = nsplit > 1 ? 1 : g.c_c;
    const bool direct = nsplit == 1;
#pragma unroll
    for (int i = 0; i < WM; ++i)
#pragma unroll
        for (int j = 0; j < WN; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int row = row0 + wr * 16 * WM + i * 16 + kq + 4 * q;
                const int col = col0 + wc * 16 * WN + j * 16 + li;
                if (row < g.M && col < g.N) {
                    double2 v = make_double2(acc[i][j].re[q], acc[i][j].im[q]);
                    double2* p = C + row * cr + col * cc;
                    if (direct) {
                        v = make_double2(g.alpha.x * v.x - g.alpha.y * v.y, g.alpha.x * v.y + g.alpha.y * v.x);
                        if (g.accumulate) {
                            const double2 o = *p;
                            v.x += o.x;
                            v.y += o.y;
                        }
                    }
                    *p = v;
                }
            }
}

// C[b](row, col) (=|+=) alpha * sum_s W[s][b](row, col), s in order
__global__ __launch_bounds__(GT) void zreduce_kernel(GemmDesc g, int nsplit) {
    const int64_t per = static_cast<int64_t>(g.M) * g.N;
    const int64_t total = per * g.nb0 * g.nb1;
    for (int64_t e = blockIdx.x * static_cast<int64_t>(GT) + threadIdx.x; e < total;
         e += static_cast<int64_t>(gridDim.x) * GT) {
        const int64_t b = e / per, rc = e - b * per;
        const int row = static_cast<int>(rc / g.N), col = static_cast<int>(rc - static_cast<int64_t>(row) * g.N);
        double sx = 0.0, sy = 0.0;
        for (int s = 0; s < nsplit; ++s) {
            const double2 w = g.W[s * g.w_split + b * g.w_batch + rc];
            sx += w.x;
            sy += w.y;
        }
        const int b1 = static_cast<int>(b / g.nb0), b0 = static_cast<int>(b - static_cast<int64_t>(b1) * g.nb0);
        double2* p = g.C + b0 * g.c_b0 + b1 * g.c_b1 + row * g.c_r + col * g.c_c;
        double2 v = make_double2(g.alpha.x * sx - g.alpha.y * sy, g.alpha.x * sy + g.alpha.y * sx);
        if (g.accumulate) {
            const double2 o = *p;
            v.x += o.x;
            v.y += o.y;
        }
        *p = v;
    }
}

// ---------------------------------------------------------------- Krylov vector kernels
// partial[blk][i] = sum over this block's entries of conj(V_i) w, i < m (two doubles each)
template <int MB>
__global__ __launch_bounds__(GT) void zdots_kernel(const double2* V, int64_t ld, int m, const double2* w, int64_t n,
                                                   double* partial) {
    __shared__ double red[GT / 64][2 * MB];
    double acc[2 * MB];
#pragma unroll
    for (int i = 0; i < 2 * MB; ++i) acc[i] = 0.0;
    for (int64_t q = blockIdx.x * static_cast<int64_t>(GT) + threadIdx.x; q < n;
         q += static_cast<int64_t>(gridDim.x) * GT) {
        // the m basis loads are issued before any is used (m <= MB, uniform)
        const double2 x = w[q];
        double2 v[MB];
#pragma unroll
        for (int i = 0; i < MB; ++i) v[i] = i < m ? V[i * ld + q] : make_double2(0.0, 0.0);
#pragma unroll
        for (int i = 0; i < MB; ++i) {
            acc[2 * i] += v[i].x * x.x + v[i].y * x.y;
            acc[2 * i + 1] += v[i].x * x.y - v[i].y * x.x;
        }
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < 2 * MB; ++i) {
        if (i < 2 * m) {
            double s = acc[i];
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
            if (lane == 0) red[wv][i] = s;
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 2 * m; i += GT) {
        double s = 0.0;
        for (int w2 = 0; w2 < GT / 64; ++w2) s += red[w2][i];
        partial[static_cast<int64_t>(i) * gridDim.x + blockIdx.x] = s;  // entry-major
    }
}

// h[i] = sum over blocks of partial[i][blk]: one workgroup per entry, each thread a strided
// share in block order, then a fixed-shape tree (the same sums every run)
__global__ __launch_bounds__(GT) void zdots_finish_kernel(const double* partial, int nblk, double* h) {
    __shared__ double red[GT];
    const int i = blockIdx.x;
    double s = 0.0;
    for (int b = threadIdx.x; b < nblk; b += GT) s += partial[static_cast<int64_t>(i) * nblk + b];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = GT / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) h[i] = red[0];
}

// w -= sum_i h[i] V_i ; then partial norms of w per block
template <int MB>
__global__ __launch_bounds__(GT) void zorth_kernel(const double2* V, int64_t ld, int m, const double* h, double2* w,
                                                   int64_t n, double* partial) {
    __shared__ double red[GT / 64];
    double hr[MB], hi[MB];
#pragma unroll
    for (int i = 0; i < MB; ++i) {
        hr[i] = i < m ? h[2 * i] : 0.0;
        hi[i] = i < m ? h[2 * i + 1] : 0.0;
    }
    double nn = 0.0;
    for (int64_t q = blockIdx.x * static_cast<int64_t>(GT) + threadIdx.x; q < n;
         q += static_cast<int64_t>(gridDim.x) * GT) {
        double2 x = w[q];
        double2 v[MB];
#pragma unroll
        for (int i = 0; i < MB; ++i) v[i] = i < m ? V[i * ld + q] : make_double2(0.0, 0.0);
#pragma unroll
        for (int i = 0; i < MB; ++i) {
            if (i < m) {
                x.x -= hr[i] * v[i].x - hi[i] * v[i].y;
                x.y -= hr[i] * v[i].y + hi[i] * v[i].x;
            }
        }
        w[q] = x;
        nn += x.x * x.x + x.y * x.y;
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) nn += __shfl_xor(nn, o, 64);
    if (lane == 0) red[wv] = nn;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w2 = 0; w2 < GT / 64; ++w2) s += red[w2];
        partial[blockIdx.x] = s;
    }
}

// y = a x (scale) or y = sum_i c[i] V_i scaled by beta (combine)
__global__ __launch_bounds__(GT) void zscale_kernel(double2* y, const double2* x, double ar, double ai, int64_t n) {
    for (int64_t q = blockIdx.x * static_cast<int64_t>(GT) + threadIdx.x; q < n;
         q += static_cast<int64_t>(gridDim.x) * GT) {
        const double2 v = x[q];
        y[q] = make_double2(ar * v.x - ai * v.y, ar * v.y + ai * v.x);
    }
}
__global__ __launch_bounds__(GT) void zcombine_kernel(double2* y, const double2* V, int64_t ld, int m, const double* c,
                                                      double beta, int64_t n) {
    for (int64_t q = blockIdx.x * static_cast<int64_t>(GT) + threadIdx.x; q < n;
         q += static_cast<int64_t>(gridDim.x) * GT) {
        double sx = 0.0, sy = 0.0;
        for (int i = 0; i < m; ++i) {
            const double2 v = V[i * ld + q];
            const double cr = c[2 * i], ci = c[2 * i + 1];
            sx += cr * v.x - ci * v.y;
            sy += cr * v.y + ci * v.x;
        }
        y[q] = make_double2(beta * sx, beta * sy);
    }
}
// partial norms of x per block
__global__ __launch_bounds__(GT) void znorm_kernel(const double2* x, int64_t n, double* partial) {
    __shared__ double red[GT / 64];
    double nn = 0.0;
    for (int64_t q = blockIdx.x * static_cast<int64_t>(GT) + threadIdx.x; q < n;
         q += static_cast<int64_t>(gridDim.x) * GT) {
        const double2 v = x[q];
        nn += v.x * v.x + v.y * v.y;
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) nn += __shfl_xor(nn, o, 64);
    if (lane == 0) red[wv] = nn;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w2 = 0; w2 < GT / 64; ++w2) s += red[w2];
        partial[blockIdx.x] = s;
    }
}

// Stages 2 and 3 of the H-action fused, for d0 <= 64 and r <= 64:
//   res[(i0, i1)][c] (=|+=) sum_t sum_k (sum_j0 op0_t[i0][j0] Y_t[j0][i1][k]) mft_t[k][c]
// One workgroup per (i1, group of terms).  Per term, Y_t[:, i1, :] is staged in LDS, the mode-0
// product Z = op0_t Y_t[:, i1, :] is formed in registers and written back over it (k-major, after
// a barrier: nothing reads Y any more), and the mean-field product Z mft_t accumulates in
// registers over the group's terms.  Each wave owns up to four 16 x 16 output tiles (the same
// ones in both products) and advances all of them per k-step, so its MFMAs form eight
// independent accumulation chains; one 64 x 72 plane pair (74 KB) leaves room for two
// workgroups per CU.  With several groups the partial sums go to W[g] (D x r each) for
// zreduce_kernel, which adds them in group order.
constexpr int FL = 72;  // LDS row length (doubles)

__global__ __launch_bounds__(GT, 2) void zheff_fused_kernel(FusedDesc f) {
    __shared__ double pr[64][FL], pi[64][FL];  // Y_t[j0][i1][k] at [j0][k], then Z[i0][k] at [k][i0]
    const int i1 = blockIdx.x, g = blockIdx.y, ng = gridDim.y;
    const int t0 = static_cast<int>(static_cast<int64_t>(f.nt) * g / ng);
    const int t1 = static_cast<int>(static_cast<int64_t>(f.nt) * (g + 1) / ng);
    const int d0 = f.d0, r = f.r;
    const int64_t size = static_cast<int64_t>(d0) * f.d1 * r;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, li = lane & 15, kq = lane >> 4;
    const int mt = (d0 + 15) >> 4, nt = (r + 15) >> 4, tiles = mt * nt;
    // this wave's tiles wv, wv + 4, ...  The mode-0 product's tiles (zm row block, zn column
    // block) share the wave's row block when mt divides 4, so one op0 load per k-step serves all
    // of them; the mean-field product's tiles (om, on) share the column block of mft likewise.
    const int cnt = tiles > wv ? min(4, (tiles - wv + 3) / 4) : 0;
    const bool zshare = (4 % mt) == 0, oshare = (4 % nt) == 0;
    int zm[4], zn[4], om[4], on[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int it = wv + 4 * q;
        zm[q] = it % mt;
        zn[q] = it / mt;
        om[q] = it / nt;
        on[q] = it % nt;
    }
    Acc acc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        acc[q].re = d4{0.0, 0.0, 0.0, 0.0};
        acc[q].im = d4{0.0, 0.0, 0.0, 0.0};
    }
    for (int t = t0; t < t1; ++t) {
        const double2* Y = f.Y + t * size + static_cast<int64_t>(i1) * r;
        for (int e = threadIdx.x; e < 64 * 64; e += GT) {
            const int j0 = e >> 6, k = e & 63;
            const double2 y = (j0 < d0 && k < r) ? Y[static_cast<int64_t>(j0) * f.d1 * r + k] : make_double2(0.0, 0.0);
            pr[j0][k] = y.x;
            pi[j0][k] = y.y;
        }
        __syncthreads();
        // Z tiles = op0_t (global) x Y (LDS)
        const double2* O = f.op0 + static_cast<int64_t>(t) * d0 * d0;
        Acc z[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            z[q].re = d4{0.0, 0.0, 0.0, 0.0};
            z[q].im = d4{0.0, 0.0, 0.0, 0.0};
        }
        // the operator loads of two k-steps are issued before their MFMAs
        for (int k0 = 0; k0 < d0; k0 += 8) {
            double2 a[2][4];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int k = k0 + 4 * u + kq;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    if (q > 0 && zshare) {
                        a[u][q] = a[u][0];
                    } else {
                        const int row = 16 * zm[q] + li;
                        a[u][q] = (q < cnt && row < d0 && k < d0) ? O[static_cast<int64_t>(row) * d0 + k]
                                                                   : make_double2(0.0, 0.0);
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int k = k0 + 4 * u + kq;
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (q < cnt) mfma_c(z[q], a[u][q].x, a[u][q].y, pr[k][16 * zn[q] + li], pi[k][16 * zn[q] + li]);
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (q < cnt)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    pr[16 * zn[q] + li][16 * zm[q] + kq + 4 * j] = z[q].re[j];
                    pi[16 * zn[q] + li][16 * zm[q] + kq + 4 * j] = z[q].im[j];
                }
        __syncthreads();
        // acc += Z (LDS) x mft_t (global)
        const double2* M = f.mft + static_cast<int64_t>(t) * r * r;
        for (int k0 = 0; k0 < r; k0 += 8) {
            double2 b[2][4];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int k = k0 + 4 * u + kq;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    if (q > 0 && oshare) {
                        b[u][q] = b[u][0];
                    } else {
                        const int col = 16 * on[q] + li;
                        b[u][q] = (q < cnt && col < r && k < r) ? M[static_cast<int64_t>(k) * r + col]
                                                                 : make_double2(0.0, 0.0);
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int k = k0 + 4 * u + kq;
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (q < cnt) mfma_c(acc[q], pr[k][16 * om[q] + li], pi[k][16 * om[q] + li], b[u][q].x, b[u][q].y);
            }
        }
        __syncthreads();
    }
    // out[(i0 d1 + i1) r + c]
    double2* out = ng > 1 ? f.W + static_cast<int64_t>(g) * f.w_split : f.res;
    const bool direct = ng == 1;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (q >= cnt) break;
        const int col = 16 * on[q] + li;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int i0 = 16 * om[q] + kq + 4 * j;
            if (i0 < d0 && col < r) {
                double2* p = out + (static_cast<int64_t>(i0) * f.d1 + i1) * r + col;
                double2 v = make_double2(acc[q].re[j], acc[q].im[j]);
                if (direct && f.accumulate) {
                    const double2 o = *p;
                    v.x += o.x;
                    v.y += o.y;
                }
                *p = v;
            }
        }
    }
}

// The mean-field update's last two stages fused (MfDesc).  One workgroup per (chunk of slabs,
// term).  Per slab, Z_s = Y_s MF_t^T is formed with MFMAs whose operands come straight from L2 (Y_s
// rows and MF_t rows), written to LDS transposed (c-major) once every wave has finished the previous
// slab, and conj(A_s) Z_s^T accumulates in registers over the chunk's slabs.  Waves own up to four
// 16 x 16 tiles of each product and advance them together per k-step.
__global__ __launch_bounds__(GT, 2) void zmf_fused_kernel(MfDesc f) {
    __shared__ double zr[64][FL], zi[64][FL];  // Z_s^T at [c][j]
    const int g = blockIdx.x, ng = gridDim.x, t = blockIdx.y;
    const int s0 = static_cast<int>(static_cast<int64_t>(f.ns) * g / ng);
    const int s1 = static_cast<int>(static_cast<int64_t>(f.ns) * (g + 1) / ng);
    const int dc = f.dc, r = f.r;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, li = lane & 15, kq = lane >> 4;
    const int md = (dc + 15) >> 4, mr = (r + 15) >> 4;
    const int tz = md * mr, to = md * md;  // tiles of Z_s (dc x r) and of out (dc x dc)
    const int cz = tz > wv ? min(4, (tz - wv + 3) / 4) : 0;
    const int co = to > wv ? min(4, (to - wv + 3) / 4) : 0;
    // both products take their L2 operand by rows (Y_s rows, A_s rows): the tiles of a wave share
    // its row block when md divides 4, so one load per k-step serves all of them
    const bool share = (4 % md) == 0;
    int zm[4], zn[4], om[4], on[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int it = wv + 4 * q;
        zm[q] = it % md;
        zn[q] = it / md;
        om[q] = it % md;
        on[q] = it / md;
    }
    Acc acc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        acc[q].re = d4{0.0, 0.0, 0.0, 0.0};
        acc[q].im = d4{0.0, 0.0, 0.0, 0.0};
    }
    const double2* M = f.mf + static_cast<int64_t>(t) * r * r;
    for (int s = s0; s < s1; ++s) {
        // Z_s(j, c) = sum_p Y_s(j, p) MF_t(c, p)
        const double2* Ys = f.Y + static_cast<int64_t>(t) * f.size + s * f.ss;
        Acc z[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            z[q].re = d4{0.0, 0.0, 0.0, 0.0};
            z[q].im = d4{0.0, 0.0, 0.0, 0.0};
        }
        // the operands of two k-steps are loaded before their MFMAs
        for (int k0 = 0; k0 < r; k0 += 8) {
            double2 a[2][4], b[2][4];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int k = k0 + 4 * u + kq;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int row = 16 * zm[q] + li, col = 16 * zn[q] + li;
                    if (q > 0 && share)
                        a[u][q] = a[u][0];
                    else
                        a[u][q] = (q < cz && row < dc && k < r) ? Ys[row * f.si + k] : make_double2(0.0, 0.0);
                    b[u][q] = (q < cz && col < r && k < r) ? M[static_cast<int64_t>(col) * r + k] : make_double2(0.0, 0.0);
                }
            }
#pragma unroll
            for (int u = 0; u < 2; ++u)
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (q < cz) mfma_c(z[q], a[u][q].x, a[u][q].y, b[u][q].x, b[u][q].y);
        }
        __syncthreads();  // the previous slab's readers of the LDS planes are done
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (q < cz)
#pragma unroll
                for (int jj = 0; jj < 4; ++jj) {
                    zr[16 * zn[q] + li][16 * zm[q] + kq + 4 * jj] = z[q].re[jj];
                    zi[16 * zn[q] + li][16 * zm[q] + kq + 4 * jj] = z[q].im[jj];
                }
        __syncthreads();
        // out(i, j) += sum_c conj(A_s(i, c)) Z_s(j, c)
        const double2* As = f.A + s * f.ss;
        for (int k0 = 0; k0 < r; k0 += 8) {
            double2 a[2][4];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int k = k0 + 4 * u + kq;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int row = 16 * om[q] + li;
                    if (q > 0 && share)
                        a[u][q] = a[u][0];
                    else
                        a[u][q] = (q < co && row < dc && k < r) ? As[row * f.si + k] : make_double2(0.0, 0.0);
                }
            }
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int k = k0 + 4 * u + kq;
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (q < co) mfma_c(acc[q], a[u][q].x, -a[u][q].y, zr[k][16 * on[q] + li], zi[k][16 * on[q] + li]);
            }
        }
    }
    double2* out = ng > 1 ? f.W + static_cast<int64_t>(g) * f.w_split : f.out;
    out += static_cast<int64_t>(t) * dc * dc;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (q >= co) break;
        const int col = 16 * on[q] + li;
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
            const int row = 16 * om[q] + kq + 4 * jj;
            if (row < dc && col < dc) out[row * dc + col] = make_double2(acc[q].re[jj], acc[q].im[jj]);
        }
    }
}

// out[t](p, j) = coeff[t] * mf[t](j, p): the mean-field factors of the term-blocked sum
__global__ __launch_bounds__(GT) void zmft_kernel(const double2* coeff, const double2* mf, int64_t nt, int64_t r,
                                                  double2* out) {
    const int64_t rr = r * r, total = nt * rr;
    for (int64_t e = blockIdx.x * static_cast<int64_t>(GT) + threadIdx.x; e < total;
         e += static_cast<int64_t>(gridDim.x) * GT) {
        const int64_t t = e / rr, pj = e - t * rr, p = pj / r, j = pj - p * r;
        const double2 c = coeff[t], m = mf[t * rr + j * r + p];
        out[e] = make_double2(c.x * m.x - c.y * m.y, c.x * m.y + c.y * m.x);
    }
}

}  // namespace ttnd

// ---------------------------------------------------------------- launchers
extern "C" hipError_t ttnd_zmft(const double2* coeff, const double2* mf, int64_t nt, int64_t r, double2* out,
                                hipStream_t s) {
    using namespace ttnd;
    const int64_t total = nt * r * r;
    const int blocks = static_cast<int>(total / GT + 1 < 1024 ? total / GT + 1 : 1024);
    hipLaunchKernelGGL(zmft_kernel, dim3(blocks), dim3(GT), 0, s, coeff, mf, nt, r, out);
    return hipGetLastError();
}
extern "C" hipError_t ttnd_zmf_fused(const ttnd::MfDesc* f, int nchunks, hipStream_t s) {
    using namespace ttnd;
    if (f->dc > 64 || f->r > 64 || f->dc < 1 || f->r < 1 || f->nt < 1 || f->nt > 65535 || nchunks < 1 ||
        nchunks > f->ns)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(zmf_fused_kernel, dim3(nchunks, f->nt), dim3(GT), 0, s, *f);
    if (nchunks > 1) {
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        GemmDesc g{};
        g.M = g.N = f->dc;
        g.nb0 = f->nt;
        g.nb1 = 1;
        g.W = f->W;
        g.w_split = f->w_split;
        g.w_batch = static_cast<int64_t>(f->dc) * f->dc;
        g.C = f->out;
        g.c_b0 = static_cast<int64_t>(f->dc) * f->dc;
        g.c_r = f->dc;
        g.c_c = 1;
        g.alpha = make_double2(1.0, 0.0);
        const int64_t total = static_cast<int64_t>(f->dc) * f->dc * f->nt;
        const int blocks = static_cast<int>(total / GT + 1 < 4096 ? total / GT + 1 : 4096);
        hipLaunchKernelGGL(zreduce_kernel, dim3(blocks), dim3(GT), 0, s, g, nchunks);
    }
    return hipGetLastError();
}
extern "C" hipError_t ttnd_zheff_fused(const ttnd::FusedDesc* f, int ngroups, hipStream_t s) {
    using namespace ttnd;
    if (f->d0 > 64 || f->r > 64 || f->d0 < 1 || f->r < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(zheff_fused_kernel, dim3(f->d1, ngroups), dim3(GT), 0, s, *f);
    if (ngroups > 1) {
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        GemmDesc g{};
        g.M = static_cast<int>(f->d0 * f->d1);
        g.N = static_cast<int>(f->r);
        g.nb0 = g.nb1 = 1;
        g.W = f->W;
        g.w_split = f->w_split;
        g.w_batch = 0;
        g.C = f->res;
        g.c_r = f->r;
        g.c_c = 1;
        g.accumulate = f->accumulate;
        g.alpha = make_double2(1.0, 0.0);
        const int64_t total = static_cast<int64_t>(g.M) * g.N;
        const int blocks = static_cast<int>(total / GT + 1 < 4096 ? total / GT + 1 : 4096);
        hipLaunchKernelGGL(zreduce_kernel, dim3(blocks), dim3(GT), 0, s, g, ngroups);
    }
    return hipGetLastError();
}

namespace {
template <int WM, int WN>
void launch_tiles(const ttnd::GemmDesc* g, int nsplit, hipStream_t s) {
    constexpr int BM = 32 * WM, BN = 32 * WN;
    const int tiles = ((g->M + BM - 1) / BM) * ((g->N + BN - 1) / BN);
    const int batch = g->nb0 * g->nb1;
    hipLaunchKernelGGL((ttnd::zgemm_kernel<WM, WN>), dim3(tiles, batch, nsplit), dim3(ttnd::GT), 0, s, *g);
}
}  // namespace

// tile: one of ttnd::GEMM_TILES (output tile of the workgroup, see node_gemm.hpp)
extern "C" hipError_t ttnd_zgemm(const ttnd::GemmDesc* g, int nsplit, int tile, hipStream_t s) {
    using namespace ttnd;
    switch (tile) {
        case TILE_32x32: launch_tiles<1, 1>(g, nsplit, s); break;
        case TILE_64x32: launch_tiles<2, 1>(g, nsplit, s); break;
        case TILE_32x64: launch_tiles<1, 2>(g, nsplit, s); break;
        case TILE_64x64: launch_tiles<2, 2>(g, nsplit, s); break;
        case TILE_128x64: launch_tiles<4, 2>(g, nsplit, s); break;
        case TILE_64x128: launch_tiles<2, 4>(g, nsplit, s); break;
        default: return hipErrorInvalidValue;
    }
    const int batch = g->nb0 * g->nb1;
    if (nsplit > 1) {
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        const int64_t total = static_cast<int64_t>(g->M) * g->N * batch;
        const int blocks = static_cast<int>(total / GT + 1 < 4096 ? total / GT + 1 : 4096);
        hipLaunchKernelGGL(zreduce_kernel, dim3(blocks), dim3(GT), 0, s, *g, nsplit);
    }
    return hipGetLastError();
}

extern "C" hipError_t ttnd_zdots(const double2* V, int64_t ld, int m, const double2* w, int64_t n, double* partial,
                                 int nblk, double* h, hipStream_t s) {
    using namespace ttnd;
    if (m < 1 || m > MAXV) return hipErrorInvalidValue;
    if (m <= 8)
        hipLaunchKernelGGL(zdots_kernel<8>, dim3(nblk), dim3(GT), 0, s, V, ld, m, w, n, partial);
    else if (m <= 16)
        hipLaunchKernelGGL(zdots_kernel<16>, dim3(nblk), dim3(GT), 0, s, V, ld, m, w, n, partial);
    else
        hipLaunchKernelGGL(zdots_kernel<MAXV>, dim3(nblk), dim3(GT), 0, s, V, ld, m, w, n, partial);
    hipLaunchKernelGGL(zdots_finish_kernel, dim3(2 * m), dim3(GT), 0, s, partial, nblk, h);
    return hipGetLastError();
}
extern "C" hipError_t ttnd_zorth(const double2* V, int64_t ld, int m, const double* h, double2* w, int64_t n,
                                 double* partial, int nblk, hipStream_t s) {
    using namespace ttnd;
    if (m < 1 || m > MAXV) return hipErrorInvalidValue;
    if (m <= 8)
        hipLaunchKernelGGL(zorth_kernel<8>, dim3(nblk), dim3(GT), 0, s, V, ld, m, h, w, n, partial);
    else if (m <= 16)
        hipLaunchKernelGGL(zorth_kernel<16>, dim3(nblk), dim3(GT), 0, s, V, ld, m, h, w, n, partial);
    else
        hipLaunchKernelGGL(zorth_kernel<MAXV>, dim3(nblk), dim3(GT), 0, s, V, ld, m, h, w, n, partial);
    return hipGetLastError();
}
extern "C" hipError_t ttnd_zscale(double2* y, const double2* x, double ar, double ai, int64_t n, int nblk,
                                  hipStream_t s) {
    using namespace ttnd;
    hipLaunchKernelGGL(zscale_kernel, dim3(nblk), dim3(GT), 0, s, y, x, ar, ai, n);
    return hipGetLastError();
}
extern "C" hipError_t ttnd_zcombine(double2* y, const double2* V, int64_t ld, int m, const double* c, double beta,
                                    int64_t n, int nblk, hipStream_t s) {
    using namespace ttnd;
    hipLaunchKernelGGL(zcombine_kernel, dim3(nblk), dim3(GT), 0, s, y, V, ld, m, c, beta, n);
    return hipGetLastError();
}
extern "C" hipError_t ttnd_znorm(const double2* x, int64_t n, double* partial, int nblk, hipStream_t s) {
    using namespace ttnd;
    hipLaunchKernelGGL(znorm_kernel, dim3(nblk), dim3(GT), 0, s, x, n, partial);
    return hipGetLastError();
}
