// exp(S) for a small dense complex matrix (n <= 32), host side: Pade(13) scaling and
// squaring (N. J. Higham, "The scaling and squaring method for the matrix exponential
// revisited", SIAM J. Matrix Anal. Appl. 26 (2005) 1179-1193), the same method the sweep
// kernel runs on one wavefront.  Used by the node-level Krylov driver (node.cpp), whose
// Hessenberg matrices live on the host between H applications.
#pragma once

#include <cmath>
#include <complex>
#include <vector>

namespace ttnh {

inline void small_expm(const std::complex<double>* S, int n, std::complex<double>* E) {
    using C = std::complex<double>;
    static const double b[14] = {64764752532480000.0, 32382376266240000.0, 7771770303897600.0,
                                 1187353796428800.0,  129060195264000.0,   10559470521600.0,
                                 670442572800.0,      33522128640.0,       1323241920.0,
                                 40840800.0,          960960.0,            16380.0,
                                 182.0,               1.0};
    const double theta13 = 5.371920351148152;
    const int nn = n * n;
    auto mul = [n](const C* X, const C* Y, C* Z) {
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                C s(0.0, 0.0);
                for (int q = 0; q < n; ++q) s += X[i * n + q] * Y[q * n + j];
                Z[i * n + j] = s;
            }
    };
    double norm1 = 0.0;
    for (int j = 0; j < n; ++j) {
        double c = 0.0;
        for (int i = 0; i < n; ++i) c += std::abs(S[i * n + j]);
        norm1 = std::max(norm1, c);
    }
    int sc = 0;
    if (norm1 > theta13) sc = static_cast<int>(std::ceil(std::log2(norm1 / theta13)));
    const double scale = std::ldexp(1.0, -sc);
    std::vector<C> A(nn), A2(nn), A4(nn), A6(nn), T1(nn), T2(nn), U(nn), V(nn);
    for (int e = 0; e < nn; ++e) A[e] = S[e] * scale;
    mul(A.data(), A.data(), A2.data());
    mul(A2.data(), A2.data(), A4.data());
    mul(A4.data(), A2.data(), A6.data());
    for (int e = 0; e < nn; ++e) T1[e] = b[13] * A6[e] + b[11] * A4[e] + b[9] * A2[e];
    mul(A6.data(), T1.data(), T2.data());
    for (int e = 0; e < nn; ++e) {
        T2[e] += b[7] * A6[e] + b[5] * A4[e] + b[3] * A2[e];
        if (e / n == e % n) T2[e] += b[1];
    }
    mul(A.data(), T2.data(), U.data());
    for (int e = 0; e < nn; ++e) T1[e] = b[12] * A6[e] + b[10] * A4[e] + b[8] * A2[e];
    mul(A6.data(), T1.data(), V.data());
    for (int e = 0; e < nn; ++e) {
        V[e] += b[6] * A6[e] + b[4] * A4[e] + b[2] * A2[e];
        if (e / n == e % n) V[e] += b[0];
    }
    // solve (V - U) X = (V + U) by LU with partial pivoting (first maximum wins)
    std::vector<C> P(nn), X(nn);
    for (int e = 0; e < nn; ++e) {
        P[e] = V[e] - U[e];
        X[e] = V[e] + U[e];
    }
    for (int k = 0; k < n; ++k) {
        int piv = k;
        double best = std::abs(P[k * n + k]);
        for (int i = k + 1; i < n; ++i)
            if (std::abs(P[i * n + k]) > best) {
                best = std::abs(P[i * n + k]);
                piv = i;
            }
        if (piv != k)
            for (int j = 0; j < n; ++j) {
                std::swap(P[k * n + j], P[piv * n + j]);
                std::swap(X[k * n + j], X[piv * n + j]);
            }
        for (int i = k + 1; i < n; ++i) {
            const C f = P[i * n + k] / P[k * n + k];
            for (int j = k; j < n; ++j) P[i * n + j] -= f * P[k * n + j];
            for (int j = 0; j < n; ++j) X[i * n + j] -= f * X[k * n + j];
        }
    }
    for (int k = n - 1; k >= 0; --k)
        for (int j = 0; j < n; ++j) {
            C s = X[k * n + j];
            for (int i = k + 1; i < n; ++i) s -= P[k * n + i] * X[i * n + j];
            X[k * n + j] = s / P[k * n + k];
        }
    for (int s = 0; s < sc; ++s) {
        mul(X.data(), X.data(), T1.data());
        X.swap(T1);
    }
    for (int e = 0; e < nn; ++e) E[e] = X[e];
}

}  // namespace ttnh
