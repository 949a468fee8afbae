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

}  // namespace ttnd
