// ORACLE — test infrastructure only (see oracle.hpp).  Dense complex128 kernels.
#include "oracle.hpp"

#include <cmath>

namespace orc {

using ttnh::ERR_NUMERIC;
using ttnh::fail;

void mode_contract(const cplx* op, i64 L, i64 d, i64 R, const cplx* A, cplx* res) {
    for (i64 l = 0; l < L; ++l) {
        const cplx* Al = A + l * d * R;
        cplx* rl = res + l * d * R;
        for (i64 i = 0; i < d; ++i) {
            cplx* ri = rl + i * R;
            for (i64 r = 0; r < R; ++r) ri[r] = cplx(0.0, 0.0);
            for (i64 j = 0; j < d; ++j) {
                const cplx o = op[i * d + j];
                const cplx* Aj = Al + j * R;
                for (i64 r = 0; r < R; ++r) ri[r] += o * Aj[r];
            }
        }
    }
}

void mul_transposed(const cplx* X, i64 rows, i64 r, const cplx* M, cplx* res) {
    for (i64 x = 0; x < rows; ++x)
        for (i64 i = 0; i < r; ++i) {
            cplx s(0.0, 0.0);
            for (i64 j = 0; j < r; ++j) s += X[x * r + j] * M[i * r + j];
            res[x * r + i] = s;
        }
}

void adjoint_mul(const cplx* A, const cplx* B, i64 rows, i64 r, cplx* res) {
    for (i64 i = 0; i < r * r; ++i) res[i] = cplx(0.0, 0.0);
    for (i64 x = 0; x < rows; ++x)
        for (i64 i = 0; i < r; ++i) {
            const cplx a = std::conj(A[x * r + i]);
            for (i64 j = 0; j < r; ++j) res[i * r + j] += a * B[x * r + j];
        }
}

void outer_contract(const cplx* A, const cplx* B, i64 L, i64 d, i64 R, cplx* res) {
    for (i64 i = 0; i < d * d; ++i) res[i] = cplx(0.0, 0.0);
    for (i64 l = 0; l < L; ++l)
        for (i64 i = 0; i < d; ++i)
            for (i64 j = 0; j < d; ++j) {
                cplx s(0.0, 0.0);
                const cplx* a = A + (l * d + i) * R;
                const cplx* b = B + (l * d + j) * R;
                for (i64 r = 0; r < R; ++r) s += std::conj(a[r]) * b[r];
                res[i * d + j] += s;
            }
}

// ||x||_2 of a complex vector without harmful underflow or overflow (the role of LAPACK's
// dznrm2): the plain sum of squares is exact to rounding whenever it lies in
// [kNrmTiny, kNrmHuge] -- squares that underflow then contribute below 1e-24 of it --
// otherwise the vector is rescaled by its largest component first.
constexpr double kNrmTiny = 1e-280, kNrmHuge = 1e280;

static double nrm2(i64 n, const cplx* x, i64 incx) {
    double s = 0.0;
    for (i64 i = 0; i < n; ++i) s += std::norm(x[i * incx]);
    if (s >= kNrmTiny && s <= kNrmHuge) return std::sqrt(s);
    double scale = 0.0;
    for (i64 i = 0; i < n; ++i)
        scale = std::max(scale, std::max(std::abs(x[i * incx].real()), std::abs(x[i * incx].imag())));
    if (scale == 0.0) return 0.0;
    double t = 0.0;
    for (i64 i = 0; i < n; ++i) t += std::norm(x[i * incx] / scale);
    return scale * std::sqrt(t);
}

// LAPACK dlapy3: sqrt(a^2 + b^2 + c^2) with scaling by the largest magnitude
static double lapy3(double a, double b, double c) {
    const double w = std::max(std::abs(a), std::max(std::abs(b), std::abs(c)));
    if (w == 0.0) return 0.0;
    const double x = a / w, y = b / w, z = c / w;
    return w * std::sqrt(x * x + y * y + z * z);
}

// LAPACK zlarfg: find H = I - tau v v^H (v0 = 1) with H^H [alpha; x] = [beta; 0], beta real.
// When |beta| falls below safmin = (smallest normal) / (unit roundoff) = 2^-969, x and
// alpha are rescaled by 1/safmin (at most 20 times) before the reflector is formed and
// beta is scaled back afterwards; this keeps 1 / (alpha - beta) finite for the nearly
// rank-deficient matrices a product-like tree state produces.
constexpr double kSafmin = 2.2250738585072014e-308 / 1.1102230246251565e-16;

static cplx larfg(i64 n, cplx& alpha, cplx* x, i64 incx) {
    double xnorm = nrm2(n - 1, x, incx);
    double ar = alpha.real(), ai = alpha.imag();
    if (xnorm == 0.0 && ai == 0.0) return cplx(0.0, 0.0);
    double beta = -std::copysign(lapy3(ar, ai, xnorm), ar);
    int knt = 0;
    if (std::abs(beta) < kSafmin) {
        const double rsafmn = 1.0 / kSafmin;
        do {
            ++knt;
            for (i64 i = 0; i < n - 1; ++i) x[i * incx] *= rsafmn;
            beta *= rsafmn;
            ar *= rsafmn;
            ai *= rsafmn;
        } while (std::abs(beta) < kSafmin && knt < 20);
        xnorm = nrm2(n - 1, x, incx);
        beta = -std::copysign(lapy3(ar, ai, xnorm), ar);
    }
    const cplx tau((beta - ar) / beta, -ai / beta);
    const cplx scal = 1.0 / (cplx(ar, ai) - beta);
    for (i64 i = 0; i < n - 1; ++i) x[i * incx] *= scal;
    for (int j = 0; j < knt; ++j) beta *= kSafmin;
    alpha = cplx(beta, 0.0);
    return tau;
}

void householder_qr(const cplx* Ain, i64 m, i64 n, cplx* Q, cplx* R) {
    std::vector<cplx> A(Ain, Ain + m * n);
    const i64 k = std::min(m, n);
    std::vector<cplx> tau(k);
    std::vector<cplx> w(n);
    for (i64 i = 0; i < k; ++i) {
        cplx alpha = A[i * n + i];
        tau[i] = larfg(m - i, alpha, &A[(i + 1) * n + i], n);
        A[i * n + i] = cplx(1.0, 0.0);
        // apply H(i)^H = I - conj(tau) v v^H to A(i:m, i+1:n)
        const cplx tc = std::conj(tau[i]);
        if (tc != cplx(0.0, 0.0)) {
            for (i64 c = i + 1; c < n; ++c) {
                cplx s(0.0, 0.0);
                for (i64 rr = i; rr < m; ++rr) s += std::conj(A[rr * n + i]) * A[rr * n + c];
                w[c] = s;
            }
            for (i64 rr = i; rr < m; ++rr) {
                const cplx v = A[rr * n + i];
                for (i64 c = i + 1; c < n; ++c) A[rr * n + c] -= tc * v * w[c];
            }
        }
        A[i * n + i] = alpha;
    }
    // R (n x n): upper trapezoid of A, zero rows beyond k
    for (i64 i = 0; i < n; ++i)
        for (i64 j = 0; j < n; ++j) R[i * n + j] = (i < k && j >= i) ? A[i * n + j] : cplx(0.0, 0.0);
    // Q (m x n): H_0 ... H_{k-1} [I_k; 0], columns >= k zero
    for (i64 i = 0; i < m; ++i)
        for (i64 j = 0; j < n; ++j) Q[i * n + j] = (i == j && j < k) ? cplx(1.0, 0.0) : cplx(0.0, 0.0);
    for (i64 i = k - 1; i >= 0; --i) {
        if (tau[i] == cplx(0.0, 0.0)) continue;
        // v = [1; A(i+1:m, i)]
        for (i64 c = 0; c < k; ++c) {
            cplx s = Q[i * n + c];
            for (i64 rr = i + 1; rr < m; ++rr) s += std::conj(A[rr * n + i]) * Q[rr * n + c];
            const cplx ts = tau[i] * s;
            Q[i * n + c] -= ts;
            for (i64 rr = i + 1; rr < m; ++rr) Q[rr * n + c] -= A[rr * n + i] * ts;
        }
    }
}

// ---- small matrix exponential: Pade(13) with scaling and squaring (Higham 2005) ----
static void matmul(const cplx* A, const cplx* B, int n, cplx* C) {
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            cplx s(0.0, 0.0);
            for (int k = 0; k < n; ++k) s += A[i * n + k] * B[k * n + j];
            C[i * n + j] = s;
        }
}

void expm(const cplx* M, int n, cplx* E) {
    static const double b[14] = {64764752532480000.0, 32382376266240000.0, 7771770303897600.0,
                                 1187353796428800.0,  129060195264000.0,   10559470521600.0,
                                 670442572800.0,      33522128640.0,       1323241920.0,
                                 40840800.0,          960960.0,            16380.0,
                                 182.0,               1.0};
    const double theta13 = 5.371920351148152;
    double norm1 = 0.0;
    for (int j = 0; j < n; ++j) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += std::abs(M[i * n + j]);
        norm1 = std::max(norm1, s);
    }
    if (!std::isfinite(norm1)) fail(ERR_NUMERIC, "expm: non-finite input");
    int sc = 0;
    if (norm1 > theta13) sc = static_cast<int>(std::ceil(std::log2(norm1 / theta13)));
    const double scale = std::ldexp(1.0, -sc);
    const int nn = n * n;
    std::vector<cplx> A(nn), A2(nn), A4(nn), A6(nn), T1(nn), T2(nn), U(nn), V(nn);
    for (int i = 0; i < nn; ++i) A[i] = M[i] * scale;
    matmul(A.data(), A.data(), n, A2.data());
    matmul(A2.data(), A2.data(), n, A4.data());
    matmul(A4.data(), A2.data(), n, A6.data());
    for (int i = 0; i < nn; ++i) T1[i] = b[13] * A6[i] + b[11] * A4[i] + b[9] * A2[i];
    matmul(A6.data(), T1.data(), n, T2.data());
    for (int i = 0; i < nn; ++i) T2[i] += b[7] * A6[i] + b[5] * A4[i] + b[3] * A2[i];
    for (int i = 0; i < n; ++i) T2[i * n + i] += b[1];
    matmul(A.data(), T2.data(), n, U.data());
    for (int i = 0; i < nn; ++i) T1[i] = b[12] * A6[i] + b[10] * A4[i] + b[8] * A2[i];
    matmul(A6.data(), T1.data(), n, V.data());
    for (int i = 0; i < nn; ++i) V[i] += b[6] * A6[i] + b[4] * A4[i] + b[2] * A2[i];
    for (int i = 0; i < n; ++i) V[i * n + i] += b[0];
    // solve (V - U) X = (V + U) by Gaussian elimination with partial pivoting
    std::vector<cplx> P(nn), X(nn);
    for (int i = 0; i < nn; ++i) {
        P[i] = V[i] - U[i];
        X[i] = V[i] + U[i];
    }
    for (int k = 0; k < n; ++k) {
        int piv = k;
        double best = std::abs(P[k * n + k]);
        for (int i = k + 1; i < n; ++i)
            if (std::abs(P[i * n + k]) > best) {
                best = std::abs(P[i * n + k]);
                piv = i;
            }
        if (best == 0.0) fail(ERR_NUMERIC, "expm: singular Pade denominator");
        if (piv != k)
            for (int j = 0; j < n; ++j) {
                std::swap(P[k * n + j], P[piv * n + j]);
                std::swap(X[k * n + j], X[piv * n + j]);
            }
        const cplx inv = 1.0 / P[k * n + k];
        for (int i = k + 1; i < n; ++i) {
            const cplx f = P[i * n + k] * inv;
            if (f == cplx(0.0, 0.0)) continue;
            for (int j = k; j < n; ++j) P[i * n + j] -= f * P[k * n + j];
            for (int j = 0; j < n; ++j) X[i * n + j] -= f * X[k * n + j];
        }
    }
    for (int k = n - 1; k >= 0; --k) {
        const cplx inv = 1.0 / P[k * n + k];
        for (int j = 0; j < n; ++j) {
            cplx s = X[k * n + j];
            for (int i = k + 1; i < n; ++i) s -= P[k * n + i] * X[i * n + j];
            X[k * n + j] = s * inv;
        }
    }
    for (int s = 0; s < sc; ++s) {
        matmul(X.data(), X.data(), n, T1.data());
        X.swap(T1);
    }
    for (int i = 0; i < nn; ++i) E[i] = X[i];
}

// Sub-step factor from the error model err(s) ~ s^m: 0.9 (tol / err)^(1/m), clamped to
// [lo, hi]; err == 0 (invariant Krylov space) gives hi.
double step_factor(double tol, double err, int m, double lo, double hi) {
    if (!(err > 0.0)) return hi;
    const double f = 0.9 * std::pow(tol / err, 1.0 / m);
    return std::min(hi, std::max(lo, f));
}

i64 expv(const std::function<void(const cplx*, cplx*)>& H, cplx* x, i64 n, cplx c, double tau, int kdim,
         double tol) {
    if (tau == 0.0 || n == 0) return 0;
    i64 napply = 0;
    std::vector<cplx> V(static_cast<size_t>(kdim + 1) * n), w(n);
    std::vector<cplx> Hm(static_cast<size_t>(kdim + 1) * kdim), S, F, h(kdim);
    // s_next: fraction of tau predicted for the next sub-step (starts with the whole step)
    double s_done = 0.0, s_next = 1.0;
    while (s_done < 1.0) {
        double s_step = std::min(s_next, 1.0 - s_done);
        double beta = 0.0;
        for (i64 i = 0; i < n; ++i) beta += std::norm(x[i]);
        beta = std::sqrt(beta);
        if (!std::isfinite(beta)) fail(ERR_NUMERIC, "expv: non-finite vector");
        if (beta == 0.0) return napply;
        for (i64 i = 0; i < n; ++i) V[i] = x[i] / beta;
        std::fill(Hm.begin(), Hm.end(), cplx(0.0, 0.0));
        int m = 0;
        double hn = 0.0, err = 0.0;
        bool found = false;
        auto small_exp = [&](int mm, double s) {
            S.assign(static_cast<size_t>(mm) * mm, cplx(0.0, 0.0));
            F.assign(static_cast<size_t>(mm) * mm, cplx(0.0, 0.0));
            const cplx f = c * (tau * s);
            for (int i = 0; i < mm; ++i)
                for (int j = 0; j < mm; ++j) S[i * mm + j] = f * Hm[i * kdim + j];
            expm(S.data(), mm, F.data());
            return hn * std::abs(F[(mm - 1) * mm]);
        };
        for (int j = 0; j < kdim; ++j) {
            H(&V[static_cast<size_t>(j) * n], w.data());
            ++napply;
            for (int pass = 0; pass < 2; ++pass) {
                for (int i = 0; i <= j; ++i) {
                    cplx s(0.0, 0.0);
                    const cplx* vi = &V[static_cast<size_t>(i) * n];
                    for (i64 q = 0; q < n; ++q) s += std::conj(vi[q]) * w[q];
                    h[i] = s;
                }
                for (int i = 0; i <= j; ++i) {
                    const cplx* vi = &V[static_cast<size_t>(i) * n];
                    for (i64 q = 0; q < n; ++q) w[q] -= h[i] * vi[q];
                    Hm[i * kdim + j] += h[i];
                }
            }
            double nn = 0.0;
            for (i64 q = 0; q < n; ++q) nn += std::norm(w[q]);
            hn = std::sqrt(nn);
            if (!std::isfinite(hn)) fail(ERR_NUMERIC, "expv: non-finite Krylov vector");
            if (j + 1 < kdim) Hm[(j + 1) * kdim + j] = hn;
            m = j + 1;
            // Leading Taylor term of the Saad estimate: (exp(S))_{m-1,0} starts with
            // prod_i S_{i+1,i} / (m-1)! for Hessenberg S.  The exponential is evaluated only
            // when this cheap estimate is within 1e3 of the tolerance (skipping a test can
            // only add Krylov vectors, never accept earlier).
            double lead = hn;
            for (int i = 0; i < j; ++i) lead *= std::abs(c) * std::abs(tau * s_step) * std::abs(Hm[(i + 1) * kdim + i]) / (i + 1);
            if (lead > 1e3 * tol && j + 1 < kdim) {
                for (i64 q = 0; q < n; ++q) V[static_cast<size_t>(j + 1) * n + q] = w[q] / hn;
                continue;
            }
            err = small_exp(m, s_step);
            if (err <= tol) {
                found = true;
                break;
            }
            if (j + 1 < kdim)
                for (i64 q = 0; q < n; ++q) V[static_cast<size_t>(j + 1) * n + q] = w[q] / hn;
        }
        // The full space does not reach the tolerance at s_step: shrink the sub-step with the
        // error model err(s) ~ s^m of the Saad estimate (Expokit's step-size rule, Sidje 1998,
        // safety factor 0.9), then fall back to halving.
        for (int k = 0; !found && k < 8; ++k) {
            s_step *= step_factor(tol, err, m, 0.0625, 0.9);
            err = small_exp(m, s_step);
            found = err <= tol;
        }
        for (int k = 0; !found && k < 60; ++k) {
            s_step *= 0.5;
            err = small_exp(m, s_step);
            found = err <= tol;
        }
        if (!found) fail(ERR_NUMERIC, "expv: Krylov step did not reach the tolerance");
        for (i64 q = 0; q < n; ++q) {
            cplx s(0.0, 0.0);
            for (int i = 0; i < m; ++i) s += F[i * m] * V[static_cast<size_t>(i) * n + q];
            x[q] = beta * s;
        }
        s_done += s_step;
        // the same model predicts the next sub-step, growing by at most a factor 2
        s_next = s_step * step_factor(tol, err, m, 1.0, 2.0);
    }
    return napply;
}

}  // namespace orc
