// Workgroup-level complex128 primitives for the persistent TDVP sweep.
//
// Every function is executed by all NT threads of one workgroup on data in HBM/L2
// (node tensors, work buffers) or LDS (small Krylov matrices).  Functions that write
// results other threads read end with __syncthreads(); callers never need their own
// barrier between two primitives.
#pragma once

#include <hip/hip_runtime.h>

#include "desc.hpp"
#include "zgemm.hpp"

namespace ttnd {

constexpr int NT = 256;          // threads per workgroup (4 wave64)
constexpr int NWAVE = NT / 64;

// ---------------------------------------------------------------- complex helpers
__device__ __forceinline__ double2 cz() { return make_double2(0.0, 0.0); }
__device__ __forceinline__ double2 cmk(double a, double b) { return make_double2(a, b); }
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ __forceinline__ double cabs2(double2 a) { return a.x * a.x + a.y * a.y; }
__device__ __forceinline__ double cabsv(double2 a) { return hypot(a.x, a.y); }
// a / b with Smith's range reduction
__device__ __forceinline__ double2 cdiv(double2 a, double2 b) {
    if (fabs(b.x) >= fabs(b.y)) {
        const double r = b.y / b.x, den = b.x + b.y * r;
        return make_double2((a.x + a.y * r) / den, (a.y - a.x * r) / den);
    }
    const double r = b.x / b.y, den = b.x * r + b.y;
    return make_double2((a.x * r + a.y) / den, (a.y * r - a.x) / den);
}

// ---------------------------------------------------------------- reductions
struct Shared {
    double red[NWAVE * 2 * (MAXK + 1)];
    double2 bc[4];      // broadcast slots
    double2 qtau[64];   // Householder tau (up to 64 columns)
    double2 qw[64];     // Householder trailing dot products
    // Krylov small matrices
    double2 Hm[(MAXK + 1) * MAXK];
    double2 S[MAXK * MAXK], F[MAXK * MAXK];
    double2 E[8][MAXK * MAXK];
    double2 lu_f[MAXK];
    int ipiv;
    int fail;           // set (uniformly) when a numerical failure is detected
    double hb[2][2 * MAXK];  // Gram-Schmidt coefficients of the two CGS passes
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// sum of K doubles over the workgroup; every thread receives the totals in out[]
template <int K>
__device__ void block_sum(Shared& sh, double (&v)[K], double (&out)[K]) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const double s = wave_sum(v[k]);
        if (lane == 0) sh.red[wv * K + k] = s;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < NWAVE; ++w) s += sh.red[w * K + k];
        out[k] = s;
    }
    __syncthreads();
}

// sum of K doubles over the workgroup, totals left in the LDS array dst[K]
template <int K>
__device__ void block_sum_lds(Shared& sh, double (&v)[K], double* dst) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const double s = wave_sum(v[k]);
        if (lane == 0) sh.red[wv * K + k] = s;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < K; k += NT) {
        double s = 0.0;
        for (int w = 0; w < NWAVE; ++w) s += sh.red[w * K + k];
        dst[k] = s;
    }
    __syncthreads();
}

// maximum over the workgroup; every thread receives it
__device__ double block_max(Shared& sh, double v) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    if (lane == 0) sh.red[wv] = v;
    __syncthreads();
    double m = sh.red[0];
#pragma unroll
    for (int w = 1; w < NWAVE; ++w) m = fmax(m, sh.red[w]);
    __syncthreads();
    return m;
}

__device__ double block_norm2(Shared& sh, const double2* x, int64_t n) {
    double p[1] = {0.0}, o[1];
    for (int64_t b = threadIdx.x; b < n; b += 4 * NT) {
        double2 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = (b + u * NT < n) ? x[b + u * NT] : cz();
#pragma unroll
        for (int u = 0; u < 4; ++u) p[0] += cabs2(v[u]);
    }
    block_sum<1>(sh, p, o);
    return o[0];
}

// ---------------------------------------------------------------- elementwise
// One workgroup has a single wave per SIMD, so latency is hidden by issuing the loads
// of UNR consecutive strides before using any of them.
constexpr int UNR = 4;

__device__ void wg_zero(double2* y, int64_t n) {
    for (int64_t i = threadIdx.x; i < n; i += NT) y[i] = cz();
    __syncthreads();
}
__device__ void wg_copy(double2* __restrict__ y, const double2* __restrict__ x, int64_t n) {
    for (int64_t b = threadIdx.x; b < n; b += UNR * NT) {
        double2 v[UNR];
#pragma unroll
        for (int u = 0; u < UNR; ++u) v[u] = (b + u * NT < n) ? x[b + u * NT] : cz();
#pragma unroll
        for (int u = 0; u < UNR; ++u)
            if (b + u * NT < n) y[b + u * NT] = v[u];
    }
    __syncthreads();
}
// y += a * x
__device__ void wg_axpy(double2* __restrict__ y, double2 a, const double2* __restrict__ x, int64_t n) {
    for (int64_t b = threadIdx.x; b < n; b += UNR * NT) {
        double2 v[UNR], o[UNR];
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            const bool ok = b + u * NT < n;
            v[u] = ok ? x[b + u * NT] : cz();
            o[u] = ok ? y[b + u * NT] : cz();
        }
#pragma unroll
        for (int u = 0; u < UNR; ++u)
            if (b + u * NT < n) y[b + u * NT] = cadd(o[u], cmul(a, v[u]));
    }
    __syncthreads();
}
// y = a * x
__device__ void wg_scale_to(double2* __restrict__ y, double2 a, const double2* __restrict__ x, int64_t n) {
    for (int64_t b = threadIdx.x; b < n; b += UNR * NT) {
        double2 v[UNR];
#pragma unroll
        for (int u = 0; u < UNR; ++u) v[u] = (b + u * NT < n) ? x[b + u * NT] : cz();
#pragma unroll
        for (int u = 0; u < UNR; ++u)
            if (b + u * NT < n) y[b + u * NT] = cmul(a, v[u]);
    }
    __syncthreads();
}

// ---------------------------------------------------------------- contractions (FP64 MFMA)
// dst(l,i,r) (=|+=) alpha * sum_j op(i,j) src(l,j,r) on the (L, d, R) view.
// Reference: contract(op, 1, A3, 1) in kronecker_product_operator_helper.hpp:292-303.
// One 16x16 tile of (i, r) per wave work item; op rows come from L1.
__device__ void wg_mode_contract(const double2* op, int64_t L, int d, int64_t R, const double2* src, double2* dst,
                                 double2 alpha, bool accumulate) {
    const int wv = threadIdx.x >> 6;
    const int mt = (d + 15) >> 4;
    const int64_t nt = (R + 15) >> 4;
    const int64_t items = L * mt * nt;
    for (int64_t it = wv; it < items; it += NWAVE) {
        const int64_t n = it % nt;
        const int m = static_cast<int>((it / nt) % mt);
        const int64_t l = it / (nt * mt);
        ZAcc acc;
        zacc_zero(acc);
        const int mv = min(16, d - 16 * m);
        const int nv = static_cast<int>(R - 16 * n < 16 ? R - 16 * n : 16);
        const double2* s = src + l * d * R + 16 * n;
        ztile<false>(acc, op + static_cast<int64_t>(16 * m) * d, d, 1, mv, s, R, 1, nv, d);
        ztile_store(acc, dst + l * d * R + static_cast<int64_t>(16 * m) * R + 16 * n, R, 1, mv, nv, alpha,
                    accumulate);
    }
    __syncthreads();
}

// dst (=|+=) alpha * X M^T, X: rows x r, M: r x r  (v * trans(mf))
__device__ void wg_mul_transposed(const double2* X, int64_t rows, int r, const double2* M, double2* dst,
                                  double2 alpha, bool accumulate) {
    const int wv = threadIdx.x >> 6;
    const int64_t mt = (rows + 15) >> 4;
    const int nt = (r + 15) >> 4;
    const int64_t items = mt * nt;
    for (int64_t it = wv; it < items; it += NWAVE) {
        const int n = static_cast<int>(it % nt);
        const int64_t m = it / nt;
        ZAcc acc;
        zacc_zero(acc);
        const int mv = static_cast<int>(rows - 16 * m < 16 ? rows - 16 * m : 16);
        const int nv = min(16, r - 16 * n);
        ztile<false>(acc, X + 16 * m * r, r, 1, mv, M + static_cast<int64_t>(16 * n) * r, 1, r, nv, r);
        ztile_store(acc, dst + 16 * m * r + 16 * n, r, 1, mv, nv, alpha, accumulate);
    }
    __syncthreads();
}

// dst = X M, X: rows x r, M: r x r
__device__ void wg_mul(const double2* X, int64_t rows, int r, const double2* M, double2* dst) {
    const int wv = threadIdx.x >> 6;
    const int64_t mt = (rows + 15) >> 4;
    const int nt = (r + 15) >> 4;
    const int64_t items = mt * nt;
    for (int64_t it = wv; it < items; it += NWAVE) {
        const int n = static_cast<int>(it % nt);
        const int64_t m = it / nt;
        ZAcc acc;
        zacc_zero(acc);
        const int mv = static_cast<int>(rows - 16 * m < 16 ? rows - 16 * m : 16);
        const int nv = min(16, r - 16 * n);
        ztile<false>(acc, X + 16 * m * r, r, 1, mv, M + 16 * n, r, 1, nv, r);
        ztile_store(acc, dst + 16 * m * r + 16 * n, r, 1, mv, nv, make_double2(1.0, 0.0), false);
    }
    __syncthreads();
}

// out(i,j) (=|+=) alpha * sum_{l,r} conj(A(l,i,r)) B(l,j,r) on the (L, d, R) view: a
// long-K product.  With several output tiles the waves take whole tiles; with a single
// tile (d <= 16) they split K and the partial tiles are summed through LDS.
__device__ void wg_outer_contract(Shared& sh, const double2* A, const double2* B, int64_t L, int d, int64_t R,
                                  double2* out, double2 alpha, bool accumulate) {
    const int wv = threadIdx.x >> 6;
    const int mt = (d + 15) >> 4;
    const int tiles = mt * mt;
    const bool split = tiles < NWAVE;
    for (int t0 = 0; t0 < tiles; t0 += split ? 1 : NWAVE) {
        const int t = split ? t0 : t0 + wv;
        ZAcc acc;
        zacc_zero(acc);
        const int m = t / mt, n = t % mt;
        const int mv = t < tiles ? min(16, d - 16 * m) : 0;
        const int nv = t < tiles ? min(16, d - 16 * n) : 0;
        if (t < tiles) {
            if (R == 1) {
                // element (l, i) at l*d + i: K runs over l
                const int64_t k0 = split ? (L * wv) / NWAVE : 0, k1 = split ? (L * (wv + 1)) / NWAVE : L;
                ztile<true>(acc, A + k0 * d + 16 * m, 1, d, mv, B + k0 * d + 16 * n, d, 1, nv, k1 - k0);
            } else {
                for (int64_t l = split ? wv : 0; l < L; l += split ? NWAVE : 1) {
                    const double2* a = A + l * d * R + static_cast<int64_t>(16 * m) * R;
                    const double2* b = B + l * d * R + static_cast<int64_t>(16 * n) * R;
                    if (split && L < NWAVE) {
                        // too few slices: split each slice's K range instead (handled below)
                        break;
                    }
                    ztile<true>(acc, a, R, 1, mv, b, 1, R, nv, R);
                }
                if (split && L < NWAVE) {
                    for (int64_t l = 0; l < L; ++l) {
                        const int64_t k0 = (R * wv) / NWAVE, k1 = (R * (wv + 1)) / NWAVE;
                        const double2* a = A + l * d * R + static_cast<int64_t>(16 * m) * R + k0;
                        const double2* b = B + l * d * R + static_cast<int64_t>(16 * n) * R + k0;
                        ztile<true>(acc, a, R, 1, mv, b, 1, R, nv, k1 - k0);
                    }
                }
            }
        }
        if (!split) {
            if (t < tiles)
                ztile_store(acc, out + static_cast<int64_t>(16 * m) * d + 16 * n, d, 1, mv, nv, alpha, accumulate);
        } else {
            // sum the four partial tiles in LDS (reuses the Krylov scratch E[0..7])
            double* red = reinterpret_cast<double*>(&sh.E[0][0]);
            const int lane = threadIdx.x & 63;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                red[(wv * 64 + lane) * 8 + j] = acc.re[j];
                red[(wv * 64 + lane) * 8 + 4 + j] = acc.im[j];
            }
            __syncthreads();
            if (wv == 0) {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    double re = 0.0, im = 0.0;
                    for (int w = 0; w < NWAVE; ++w) {
                        re += red[(w * 64 + lane) * 8 + j];
                        im += red[(w * 64 + lane) * 8 + 4 + j];
                    }
                    acc.re[j] = re;
                    acc.im[j] = im;
                }
                ztile_store(acc, out + static_cast<int64_t>(16 * m) * d + 16 * n, d, 1, mv, nv, alpha, accumulate);
            }
            __syncthreads();
        }
    }
    __syncthreads();
}

// out = A^H B for A, B: rows x r
__device__ void wg_adjoint_mul(Shared& sh, const double2* A, const double2* B, int64_t rows, int r, double2* out,
                               double2 alpha, bool accumulate) {
    wg_outer_contract(sh, A, B, rows, r, 1, out, alpha, accumulate);
}

// ---------------------------------------------------------------- primitive leaf operators
// y = op x, x/y: dim x ncols row-major; kind: 0 dense, 1 diag, 2 csr
// y (=|+=) alpha * op x
__device__ void wg_prim_apply(const DK& k, const double2* base, const int64_t* ints, const double2* x, double2* y,
                              int ncols, double2 alpha = make_double2(1.0, 0.0), bool accumulate = false) {
    const int d = k.dim;
    const double2* mat = base + k.mat;
    const int64_t total = static_cast<int64_t>(d) * ncols;
    for (int64_t idx = threadIdx.x; idx < total; idx += NT) {
        const int i = static_cast<int>(idx / ncols), c = static_cast<int>(idx % ncols);
        double2 acc = cz();
        if (k.kind == 0) {
            for (int j = 0; j < d; ++j) acc = cadd(acc, cmul(mat[i * d + j], x[static_cast<int64_t>(j) * ncols + c]));
        } else if (k.kind == 1) {
            acc = cmul(mat[i], x[idx]);
        } else {
            const int64_t* rowptr = ints + k.idx;
            const int64_t* colind = rowptr + d + 1;
            for (int64_t q = rowptr[i]; q < rowptr[i + 1]; ++q)
                acc = cadd(acc, cmul(mat[q], x[colind[q] * ncols + c]));
        }
        acc = cmul(alpha, acc);
        y[idx] = accumulate ? cadd(y[idx], acc) : acc;
    }
    __syncthreads();
}

// ---------------------------------------------------------------- safe norms (LAPACK dznrm2/dlapy3 roles)
// The plain sum of squares is exact to rounding whenever it lies in [kNrmTiny, kNrmHuge];
// otherwise the column is rescaled by its largest component first.
constexpr double kNrmTiny = 1e-280, kNrmHuge = 1e280;
// safmin = (smallest normal) / (unit roundoff) = 2^-969, as in zlarfg
constexpr double kSafmin = 2.2250738585072014e-308 / 1.1102230246251565e-16;

// ||A(i+1:m, i)||_2 of a row-major m x n matrix
__device__ double wg_col_nrm2(Shared& sh, const double2* A, int i, int m, int n) {
    double p[1] = {0.0}, o[1];
    for (int rr = i + 1 + threadIdx.x; rr < m; rr += NT) p[0] += cabs2(A[static_cast<int64_t>(rr) * n + i]);
    block_sum<1>(sh, p, o);
    if (o[0] >= kNrmTiny && o[0] <= kNrmHuge) return sqrt(o[0]);
    double mx = 0.0;
    for (int rr = i + 1 + threadIdx.x; rr < m; rr += NT) {
        const double2 v = A[static_cast<int64_t>(rr) * n + i];
        mx = fmax(mx, fmax(fabs(v.x), fabs(v.y)));
    }
    const double scale = block_max(sh, mx);
    if (scale == 0.0) return 0.0;
    p[0] = 0.0;
    for (int rr = i + 1 + threadIdx.x; rr < m; rr += NT) {
        const double2 v = A[static_cast<int64_t>(rr) * n + i];
        p[0] += cabs2(make_double2(v.x / scale, v.y / scale));
    }
    block_sum<1>(sh, p, o);
    return scale * sqrt(o[0]);
}

// sqrt(a^2 + b^2 + c^2) scaled by the largest magnitude
__device__ __forceinline__ double lapy3(double a, double b, double c) {
    const double w = fmax(fabs(a), fmax(fabs(b), fabs(c)));
    if (w == 0.0) return 0.0;
    const double x = a / w, y = b / w, z = c / w;
    return w * sqrt(x * x + y * y + z * z);
}

// ---------------------------------------------------------------- Householder QR
// A (m x n, row-major, destroyed: holds the reflectors) = Q (m x n) R (n x n).
// LAPACK zgeqr2/zlarfg/zung2r conventions; for m < n, Q gets zero columns and R zero
// rows beyond m (decomposition_engine.hpp:124-158).
__device__ void wg_householder_qr(Shared& sh, double2* A, int m, int n, double2* Q, double2* Rout) {
    const int k = m < n ? m : n;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int i = 0; i < k; ++i) {
        // reflector of column i (zlarfg): beta from a scaled norm; tiny columns are
        // rescaled by 1/safmin until |beta| >= safmin
        double xnorm = wg_col_nrm2(sh, A, i, m, n);
        const double2 alpha = A[static_cast<int64_t>(i) * n + i];
        double ar = alpha.x, ai = alpha.y;
        double2 tau = cz(), scal = cz();
        double beta = ar;
        const bool refl = !(xnorm == 0.0 && ai == 0.0);
        int knt = 0;
        if (refl) {
            beta = -copysign(lapy3(ar, ai, xnorm), ar);
            if (fabs(beta) < kSafmin) {
                do {
                    ++knt;
                    for (int rr = i + 1 + threadIdx.x; rr < m; rr += NT) {
                        double2& v = A[static_cast<int64_t>(rr) * n + i];
                        v = cscale(v, 1.0 / kSafmin);
                    }
                    beta *= 1.0 / kSafmin;
                    ar *= 1.0 / kSafmin;
                    ai *= 1.0 / kSafmin;
                } while (fabs(beta) < kSafmin && knt < 20);
                __syncthreads();
                xnorm = wg_col_nrm2(sh, A, i, m, n);
                beta = -copysign(lapy3(ar, ai, xnorm), ar);
            }
            tau = make_double2((beta - ar) / beta, -ai / beta);
            scal = cdiv(make_double2(1.0, 0.0), make_double2(ar - beta, ai));
        }
        if (threadIdx.x == 0) sh.qtau[i] = tau;
        if (tau.x != 0.0 || tau.y != 0.0) {
            for (int rr = i + 1 + threadIdx.x; rr < m; rr += NT) {
                double2& v = A[static_cast<int64_t>(rr) * n + i];
                v = cmul(v, scal);
            }
            __syncthreads();
            // w_c = v^H A(i:m, c) with v0 = 1, for c > i; one wave per column
            for (int c = i + 1 + wv; c < n; c += NWAVE) {
                double2 acc = cz();
                for (int rr = i + 1 + lane; rr < m; rr += 64)
                    acc = cadd(acc, cmulc(A[static_cast<int64_t>(rr) * n + i], A[static_cast<int64_t>(rr) * n + c]));
                acc.x = wave_sum(acc.x);
                acc.y = wave_sum(acc.y);
                if (lane == 0) sh.qw[c] = cadd(acc, A[static_cast<int64_t>(i) * n + c]);
            }
            __syncthreads();
            const double2 tc = make_double2(tau.x, -tau.y);
            const int ncol = n - i - 1;
            const int64_t tot = static_cast<int64_t>(m - i) * ncol;
            for (int64_t e = threadIdx.x; e < tot; e += NT) {
                const int rr = i + static_cast<int>(e / ncol);
                const int c = i + 1 + static_cast<int>(e % ncol);
                const double2 v = rr == i ? make_double2(1.0, 0.0) : A[static_cast<int64_t>(rr) * n + i];
                double2& a = A[static_cast<int64_t>(rr) * n + c];
                a = csub(a, cmul(tc, cmul(v, sh.qw[c])));
            }
        }
        __syncthreads();
        if (refl) {
            for (int j = 0; j < knt; ++j) beta *= kSafmin;
            if (threadIdx.x == 0) A[static_cast<int64_t>(i) * n + i] = make_double2(beta, 0.0);
        }
        __syncthreads();
    }
    // R
    for (int64_t e = threadIdx.x; e < static_cast<int64_t>(n) * n; e += NT) {
        const int i = static_cast<int>(e / n), j = static_cast<int>(e % n);
        Rout[e] = (i < k && j >= i) ? A[static_cast<int64_t>(i) * n + j] : cz();
    }
    // Q = H_0 ... H_{k-1} [I_k; 0]
    for (int64_t e = threadIdx.x; e < static_cast<int64_t>(m) * n; e += NT) {
        const int i = static_cast<int>(e / n), j = static_cast<int>(e % n);
        Q[e] = (i == j && j < k) ? make_double2(1.0, 0.0) : cz();
    }
    __syncthreads();
    for (int i = k - 1; i >= 0; --i) {
        const double2 tau = sh.qtau[i];
        if (tau.x == 0.0 && tau.y == 0.0) continue;
        for (int c = wv; c < k; c += NWAVE) {
            double2 acc = cz();
            for (int rr = i + 1 + lane; rr < m; rr += 64)
                acc = cadd(acc, cmulc(A[static_cast<int64_t>(rr) * n + i], Q[static_cast<int64_t>(rr) * n + c]));
            acc.x = wave_sum(acc.x);
            acc.y = wave_sum(acc.y);
            if (lane == 0) sh.qw[c] = cmul(tau, cadd(acc, Q[static_cast<int64_t>(i) * n + c]));
        }
        __syncthreads();
        const int64_t tot = static_cast<int64_t>(m - i) * k;
        for (int64_t e = threadIdx.x; e < tot; e += NT) {
            const int rr = i + static_cast<int>(e / k);
            const int c = static_cast<int>(e % k);
            const double2 v = rr == i ? make_double2(1.0, 0.0) : A[static_cast<int64_t>(rr) * n + i];
            double2& q = Q[static_cast<int64_t>(rr) * n + c];
            q = csub(q, cmul(v, sh.qw[c]));
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- small dense expm (LDS)
// C = A B for n x n matrices in LDS, n*n <= NT
__device__ void sm_matmul(const double2* A, const double2* B, double2* C, int n) {
    const int t = threadIdx.x;
    double2 acc = cz();
    if (t < n * n) {
        const int i = t / n, j = t % n;
        for (int q = 0; q < n; ++q) acc = cadd(acc, cmul(A[i * n + q], B[q * n + j]));
    }
    __syncthreads();
    if (t < n * n) C[t] = acc;
    __syncthreads();
}

// E = exp(S), Pade(13) scaling and squaring (Higham 2005); S, E in LDS; uses sh.E[*]
__device__ void sm_expm(Shared& sh, const double2* S, double2* Eout, int n) {
    const double b[14] = {64764752532480000.0, 32382376266240000.0, 7771770303897600.0, 1187353796428800.0,
                          129060195264000.0,   10559470521600.0,   670442572800.0,     33522128640.0,
                          1323241920.0,        40840800.0,         960960.0,           16380.0,
                          182.0,               1.0};
    const double theta13 = 5.371920351148152;
    const int t = threadIdx.x, nn = n * n;
    double2 *A = sh.E[0], *A2 = sh.E[1], *A4 = sh.E[2], *A6 = sh.E[3], *T1 = sh.E[4], *T2 = sh.E[5], *U = sh.E[6],
            *V = sh.E[7];
    // 1-norm and scaling (every thread computes the same value from LDS)
    double norm1 = 0.0;
    for (int j = 0; j < n; ++j) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += cabsv(S[i * n + j]);
        norm1 = fmax(norm1, s);
    }
    int sc = 0;
    if (norm1 > theta13) sc = static_cast<int>(ceil(log2(norm1 / theta13)));
    const double scale = ldexp(1.0, -sc);
    if (t < nn) A[t] = cscale(S[t], scale);
    __syncthreads();
    sm_matmul(A, A, A2, n);
    sm_matmul(A2, A2, A4, n);
    sm_matmul(A4, A2, A6, n);
    if (t < nn) T1[t] = cadd(cadd(cscale(A6[t], b[13]), cscale(A4[t], b[11])), cscale(A2[t], b[9]));
    __syncthreads();
    sm_matmul(A6, T1, T2, n);
    if (t < nn) {
        double2 v = cadd(T2[t], cadd(cadd(cscale(A6[t], b[7]), cscale(A4[t], b[5])), cscale(A2[t], b[3])));
        if (t / n == t % n) v.x += b[1];
        T2[t] = v;
    }
    __syncthreads();
    sm_matmul(A, T2, U, n);
    if (t < nn) T1[t] = cadd(cadd(cscale(A6[t], b[12]), cscale(A4[t], b[10])), cscale(A2[t], b[8]));
    __syncthreads();
    sm_matmul(A6, T1, V, n);
    if (t < nn) {
        double2 v = cadd(V[t], cadd(cadd(cscale(A6[t], b[6]), cscale(A4[t], b[4])), cscale(A2[t], b[2])));
        if (t / n == t % n) v.x += b[0];
        V[t] = v;
    }
    __syncthreads();
    // P = V - U (into A), X = V + U (into T1)
    double2 *P = A, *X = T1;
    if (t < nn) {
        const double2 u = U[t], v = V[t];
        P[t] = csub(v, u);
        X[t] = cadd(v, u);
    }
    __syncthreads();
    for (int kk = 0; kk < n; ++kk) {
        if (t == 0) {
            int piv = kk;
            double best = cabsv(P[kk * n + kk]);
            for (int i = kk + 1; i < n; ++i) {
                const double a = cabsv(P[i * n + kk]);
                if (a > best) {
                    best = a;
                    piv = i;
                }
            }
            sh.ipiv = piv;
        }
        __syncthreads();
        const int piv = sh.ipiv;
        if (piv != kk && t < n) {
            double2 a = P[kk * n + t];
            P[kk * n + t] = P[piv * n + t];
            P[piv * n + t] = a;
            a = X[kk * n + t];
            X[kk * n + t] = X[piv * n + t];
            X[piv * n + t] = a;
        }
        __syncthreads();
        if (t > kk && t < n) sh.lu_f[t] = cdiv(P[t * n + kk], P[kk * n + kk]);
        __syncthreads();
        if (t < nn) {
            const int i = t / n, j = t % n;
            if (i > kk) {
                const double2 f = sh.lu_f[i];
                if (j >= kk) P[i * n + j] = csub(P[i * n + j], cmul(f, P[kk * n + j]));
                X[i * n + j] = csub(X[i * n + j], cmul(f, X[kk * n + j]));
            }
        }
        __syncthreads();
    }
    for (int kk = n - 1; kk >= 0; --kk) {
        if (t < n) {
            double2 s = X[kk * n + t];
            for (int i = kk + 1; i < n; ++i) s = csub(s, cmul(P[kk * n + i], X[i * n + t]));
            X[kk * n + t] = cdiv(s, P[kk * n + kk]);
        }
        __syncthreads();
    }
    double2* cur = X;
    double2* nxt = T2;
    for (int s = 0; s < sc; ++s) {
        sm_matmul(cur, cur, nxt, n);
        double2* tmp = cur;
        cur = nxt;
        nxt = tmp;
    }
    if (t < nn) Eout[t] = cur[t];
    __syncthreads();
}


// ---------------------------------------------------------------- small dense expm on one wave
// The same Pade(13) scaling-and-squaring algorithm as sm_expm, executed by a single
// wavefront: the 64 lanes run in lockstep, so LDS hand-offs need only a wavefront-scope
// fence instead of a workgroup barrier.  Called by wave 0 only.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ void wv_matmul(const double2* A, const double2* B, double2* C, int n, int lane) {
    double2 acc[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int e = lane + 64 * u;
        double2 s = cz();
        if (e < n * n) {
            const int i = e / n, j = e % n;
            for (int q = 0; q < n; ++q) s = cadd(s, cmul(A[i * n + q], B[q * n + j]));
        }
        acc[u] = s;
    }
    wave_sync();
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int e = lane + 64 * u;
        if (e < n * n) C[e] = acc[u];
    }
    wave_sync();
}

// max over the wave with the lowest lane winning ties
__device__ __forceinline__ void wave_argmax(double& v, int& idx) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double ov = __shfl_xor(v, o, 64);
        const int oi = __shfl_xor(idx, o, 64);
        if (ov > v || (ov == v && oi < idx)) {
            v = ov;
            idx = oi;
        }
    }
}

__device__ void wave_expm(Shared& sh, const double2* S, double2* Eout, int n) {
    const double b[14] = {64764752532480000.0, 32382376266240000.0, 7771770303897600.0, 1187353796428800.0,
                          129060195264000.0,   10559470521600.0,   670442572800.0,     33522128640.0,
                          1323241920.0,        40840800.0,         960960.0,           16380.0,
                          182.0,               1.0};
    const double theta13 = 5.371920351148152;
    const int lane = threadIdx.x & 63, nn = n * n;
    double2 *A = sh.E[0], *A2 = sh.E[1], *A4 = sh.E[2], *A6 = sh.E[3], *T1 = sh.E[4], *T2 = sh.E[5], *U = sh.E[6],
            *V = sh.E[7];
    // 1-norm: lane j sums column j, then a wave max
    double colsum = 0.0;
    if (lane < n)
        for (int i = 0; i < n; ++i) colsum += cabsv(S[i * n + lane]);
    int dummy = lane;
    wave_argmax(colsum, dummy);
    const double norm1 = colsum;
    int sc = 0;
    if (norm1 > theta13) sc = static_cast<int>(ceil(log2(norm1 / theta13)));
    const double scale = ldexp(1.0, -sc);
    for (int e = lane; e < nn; e += 64) A[e] = cscale(S[e], scale);
    wave_sync();
    wv_matmul(A, A, A2, n, lane);
    wv_matmul(A2, A2, A4, n, lane);
    wv_matmul(A4, A2, A6, n, lane);
    for (int e = lane; e < nn; e += 64)
        T1[e] = cadd(cadd(cscale(A6[e], b[13]), cscale(A4[e], b[11])), cscale(A2[e], b[9]));
    wave_sync();
    wv_matmul(A6, T1, T2, n, lane);
    for (int e = lane; e < nn; e += 64) {
        double2 v = cadd(T2[e], cadd(cadd(cscale(A6[e], b[7]), cscale(A4[e], b[5])), cscale(A2[e], b[3])));
        if (e / n == e % n) v.x += b[1];
        T2[e] = v;
    }
    wave_sync();
    wv_matmul(A, T2, U, n, lane);
    for (int e = lane; e < nn; e += 64)
        T1[e] = cadd(cadd(cscale(A6[e], b[12]), cscale(A4[e], b[10])), cscale(A2[e], b[8]));
    wave_sync();
    wv_matmul(A6, T1, V, n, lane);
    for (int e = lane; e < nn; e += 64) {
        double2 v = cadd(V[e], cadd(cadd(cscale(A6[e], b[6]), cscale(A4[e], b[4])), cscale(A2[e], b[2])));
        if (e / n == e % n) v.x += b[0];
        V[e] = v;
    }
    wave_sync();
    double2 *P = A, *X = T1;
    for (int e = lane; e < nn; e += 64) {
        const double2 u = U[e], v = V[e];
        P[e] = csub(v, u);
        X[e] = cadd(v, u);
    }
    wave_sync();
    for (int kk = 0; kk < n; ++kk) {
        // partial pivoting: lanes kk..n-1 take |P(i, kk)|, first maximum wins
        double best = (lane >= kk && lane < n) ? cabsv(P[lane * n + kk]) : -1.0;
        int piv = lane;
        wave_argmax(best, piv);
        if (piv != kk && lane < n) {
            double2 a = P[kk * n + lane];
            P[kk * n + lane] = P[piv * n + lane];
            P[piv * n + lane] = a;
            a = X[kk * n + lane];
            X[kk * n + lane] = X[piv * n + lane];
            X[piv * n + lane] = a;
        }
        wave_sync();
        if (lane > kk && lane < n) sh.lu_f[lane] = cdiv(P[lane * n + kk], P[kk * n + kk]);
        wave_sync();
        for (int e = lane; e < nn; e += 64) {
            const int i = e / n, j = e % n;
            if (i > kk) {
                const double2 f = sh.lu_f[i];
                if (j >= kk) P[i * n + j] = csub(P[i * n + j], cmul(f, P[kk * n + j]));
                X[i * n + j] = csub(X[i * n + j], cmul(f, X[kk * n + j]));
            }
        }
        wave_sync();
    }
    for (int kk = n - 1; kk >= 0; --kk) {
        if (lane < n) {
            double2 s = X[kk * n + lane];
            for (int i = kk + 1; i < n; ++i) s = csub(s, cmul(P[kk * n + i], X[i * n + lane]));
            X[kk * n + lane] = cdiv(s, P[kk * n + kk]);
        }
        wave_sync();
    }
    double2* cur = X;
    double2* nxt = T2;
    for (int s = 0; s < sc; ++s) {
        wv_matmul(cur, cur, nxt, n, lane);
        double2* tmp = cur;
        cur = nxt;
        nxt = tmp;
    }
    for (int e = lane; e < nn; e += 64) Eout[e] = cur[e];
    wave_sync();
}

}  // namespace ttnd
