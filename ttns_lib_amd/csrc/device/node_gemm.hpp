// Descriptor of a batched, optionally K-blocked and K-split complex128 GEMM on the FP64
// matrix cores (node_gemm.hip); plain struct shared by host and device code.
//   C[b](i, j) (=|+=) alpha * sum_k op(A)[b](i, k) B[b](k, j),   b = b1 * nb0 + b0
//   A[b](i, k) at A + b0*a_b0 + b1*a_b1 + i*a_r + (k % kb)*a_k + (k / kb)*a_kblk
//   B[b](k, j) at B + b0*b_b0 + b1*b_b1 + (k % kb)*b_k + (k / kb)*b_kblk + j*b_c
//   C[b](i, j) at C + b0*c_b0 + b1*c_b1 + i*c_r + j*c_c
// op(A) = conj(A) when conja.  With a K split of S > 1 the partial products go to the
// workspace W (split s, batch b at W + s*w_split + b*w_batch, row-major M x N) and are
// summed in split order.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ttnd {

constexpr int MAXV = 32;  // Krylov vectors handled by the node-level Gram-Schmidt kernels

// output tiles of zgemm_kernel (rows x columns per workgroup)
enum GemmTile { TILE_32x32, TILE_64x32, TILE_32x64, TILE_64x64, TILE_128x64, TILE_64x128, GEMM_TILES };

struct GemmDesc {
    int M, N, K;
    int nb0, nb1;
    int kb;
    int conja, accumulate;
    const double2* A;
    int64_t a_b0, a_b1, a_r, a_k, a_kblk;
    const double2* B;
    int64_t b_b0, b_b1, b_k, b_c, b_kblk;
    double2* C;
    int64_t c_b0, c_b1, c_r, c_c;
    double2* W;
    int64_t w_split, w_batch;
    double2 alpha;
};

// The fused stages 2 and 3 of a node's H-action (zheff_fused_kernel), d0, r <= 64:
//   res[(i0, i1)][c] (=|+=) sum_t sum_k (sum_j0 op0[t][i0][j0] Y[t][j0][i1][k]) mft[t][k][c]
// Y[t] row-major d0 x d1 x r, op0[t] d0 x d0, mft[t] r x r; groups > 1 write partial sums to
// W (w_split entries apart, D x r each) and sum them in group order.
struct FusedDesc {
    int d0, d1, r, nt;
    int accumulate, pad;
    const double2* Y;
    const double2* op0;
    const double2* mft;
    double2* res;
    double2* W;
    int64_t w_split;
};

// The fused last two stages of a child's mean-field update (zmf_fused_kernel), dc, r <= 64:
//   out[t](i, j) = sum_s sum_c conj(A[s, i, c]) sum_p Y[t][s, j, p] MF[t](c, p)
// over the slabs s of the other child's index; element (s, i, c) of A and of each Y[t] at
// s * ss + i * si + c.  chunks > 1 write partial sums to W (w_split entries apart, nt dc x dc
// each) and sum them in chunk order.
struct MfDesc {
    int dc, ns, r, nt;
    int64_t ss, si, size;
    const double2* Y;
    const double2* A;
    const double2* mf;
    double2* out;
    double2* W;
    int64_t w_split;
};

}  // namespace ttnd
