// Subspace expansion inside the persistent sweep: the bond growth of the reference's
// expansion integrator (subspace_expansion_projector_splitting_engine.hpp:207-510,
// two_site_energy_variations.hpp:49-562), run by a team's leader workgroup between the gauge
// factorisation of a bond and the environment update that follows it.
//
// Node storage is laid out at capacity, so a bond grows in place: the child's matrix gains
// zero (downward) or new (upward) columns, the parent's slices of that mode gain new
// (downward) or zero (upward) slices, the gauge factor is zero-padded, and the child's live
// bond dimension (DInst::hr, mirrored in every member's LDS table) grows.  The new functions
// come from the dominant eigenpairs of the projected two-site operator M^H M (downward) or
// conj(M M^H) (upward), M = sum_t op1_t op2_t with the factors of
// construct_two_site_energy_terms_lower / _upper, found by restarted Lanczos from a random
// trial vector.  The random draws are the normal stream of std::mt19937 +
// std::normal_distribution, generated on the host and consumed here in order (two draws per
// complex entry, the first one the imaginary part).  Same algorithm as the oracle's
// expansion.cpp, step for step.  Part of the sweep kernel's translation unit: included by
// sweep.hip inside namespace ttnd, after the primitives it uses.
#pragma once

// Scratch of one expansion in the small propagator's LDS (idle while an expansion runs)
struct ExpLds {
    double2 s[64];         // a short vector of the two-site operator (one value per bond index)
    double T[MAXK * MAXK];  // the Lanczos tridiagonal matrix, then its eigenvalues on the diagonal
    double Z[MAXK * MAXK];  // its eigenvectors (columns)
    double al[MAXK], be[MAXK], th[MAXK];
    int mm, pad[3];
    double2 R[64 * 64];    // a gauge factor for its singular values
};
static_assert(sizeof(ExpLds) <= sizeof(SmallLds), "expansion scratch exceeds the small propagator's LDS");
__device__ __forceinline__ ExpLds& exp_lds() { return *reinterpret_cast<ExpLds*>(&g_small); }

__device__ __forceinline__ int ldc_int(const int* p) {
    return __hip_atomic_load((const __attribute__((address_space(1))) int*)p, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int64_t ldc_i64(const int64_t* p) {
    return __hip_atomic_load((const __attribute__((address_space(1))) int64_t*)p, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
}

// (L, d, R) of the rank-3 view of `mode` of a node; mode == nmodes is the bond index
__device__ __forceinline__ void rank3(const Shape& sh, int mode, int64_t& L, int& d, int64_t& R) {
    if (mode == sh.nmodes) {
        L = sh.dimen;
        d = sh.hrank;
        R = 1;
        return;
    }
    mode_view(sh, mode, L, R);
    d = sh.dims[mode];
}

// ---- workgroup products with a short operand (one workgroup; results visible afterwards)
// out[j] = sum_{l,k} op(A(l,j,k)) v(l,k), j < d <= 64, into LDS
__device__ void wg_contract_long(const double2* A, int64_t L, int d, int64_t R, const double2* v, bool conjA,
                                 double2* out) {
    const int64_t m = L * R;
    for (int j0 = 0; j0 < d; j0 += 8) {
        double p[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) p[u] = 0.0;
        for (int64_t q = threadIdx.x; q < m; q += NT) {
            const int64_t l = q / R, k = q - l * R;
            const double2 vv = ldc(v + q);
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (j0 + u < d) {
                    double2 a = ldc(A + (l * d + j0 + u) * R + k);
                    if (conjA) a.y = -a.y;
                    const double2 pr = cmul(a, vv);
                    p[2 * u] += pr.x;
                    p[2 * u + 1] += pr.y;
                }
        }
        double o[16];
        block_sum<16>(p, o);
        if (threadIdx.x == 0)
            for (int u = 0; u < 8 && j0 + u < d; ++u) out[j0 + u] = cmk(o[2 * u], o[2 * u + 1]);
    }
    __syncthreads();
}
// out(l,k) (=|+=) sum_j op(A(l,j,k)) s(j), s in LDS
__device__ void wg_apply_short(const double2* A, int64_t L, int d, int64_t R, const double2* s, bool conjA,
                               double2* out, bool accumulate) {
    const int64_t m = L * R;
    for (int64_t q = threadIdx.x; q < m; q += NT) {
        const int64_t l = q / R, k = q - l * R;
        double2 acc = cz();
        for (int j = 0; j < d; ++j) {
            double2 a = ldc(A + (l * d + j) * R + k);
            if (conjA) a.y = -a.y;
            acc = cadd(acc, cmul(a, s[j]));
        }
        out[q] = accumulate ? cadd(ldc(out + q), acc) : acc;
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
}

__device__ double wg_norm2(const double2* x, int64_t n) {
    double p[1] = {0.0}, o[1];
    for (int64_t q = threadIdx.x; q < n; q += NT) p[0] += cabs2(ldc(x + q));
    block_sum<1>(p, o);
    return o[0];
}
__device__ void wg_scale(double2* x, double s, int64_t n) {
    for (int64_t q = threadIdx.x; q < n; q += NT) x[q] = cscale(ldc(x + q), s);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
}

// ---- eigenpairs of the Lanczos tridiagonal matrix: cyclic Jacobi rotations on thread 0,
// eigenvalues descending (stable order), eigenvectors as the columns of Z
__device__ void tridiag_eigen_serial(ExpLds& E, int n) {
    double* T = E.T;
    double* Z = E.Z;
    for (int i = 0; i < n * n; ++i) {
        T[i] = 0.0;
        Z[i] = 0.0;
    }
    for (int i = 0; i < n; ++i) {
        T[i * n + i] = E.al[i];
        Z[i * n + i] = 1.0;
        if (i + 1 < n) T[i * n + i + 1] = T[(i + 1) * n + i] = E.be[i];
    }
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                tot += T[i * n + j] * T[i * n + j];
                if (i != j) off += T[i * n + j] * T[i * n + j];
            }
        if (off <= 1e-30 * tot || off == 0.0) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = T[p * n + q];
                if (apq == 0.0) continue;
                const double theta = (T[q * n + q] - T[p * n + p]) / (2.0 * apq);
                const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {
                    if (k == p || k == q) continue;
                    const double kp = T[k * n + p], kq = T[k * n + q];
                    T[k * n + p] = T[p * n + k] = c * kp - s * kq;
                    T[k * n + q] = T[q * n + k] = s * kp + c * kq;
                }
                T[p * n + p] -= t * apq;
                T[q * n + q] += t * apq;
                T[p * n + q] = T[q * n + p] = 0.0;
                for (int k = 0; k < n; ++k) {
                    const double zp = Z[k * n + p], zq = Z[k * n + q];
                    Z[k * n + p] = c * zp - s * zq;
                    Z[k * n + q] = s * zp + c * zq;
                }
            }
    }
    // stable insertion sort of the eigenvalues (descending), permuting Z's columns (T and be
    // serve as scratch once the diagonal is read)
    double* ev = E.be;
    int order[MAXK];
    for (int i = 0; i < n; ++i) {
        ev[i] = T[i * n + i];
        order[i] = i;
    }
    for (int i = 1; i < n; ++i) {
        const int key = order[i];
        int j = i - 1;
        while (j >= 0 && ev[order[j]] < ev[key]) {
            order[j + 1] = order[j];
            --j;
        }
        order[j + 1] = key;
    }
    for (int k = 0; k < n; ++k) {
        E.th[k] = ev[order[k]];
        for (int i = 0; i < n; ++i) T[i * n + k] = Z[i * n + order[k]];
    }
    for (int i = 0; i < n * n; ++i) Z[i] = T[i];
}

// Restarted Lanczos (full re-orthogonalisation by CGS2) for the dominant eigenpairs of the
// Hermitian map F from the start vector x (overwritten by restart vectors).  Leaves the
// Lanczos basis in Q (stride ld), the Ritz values in E.th[0 .. E.mm) and their coefficient
// vectors in the columns of E.Z.  Returns false after a non-finite value.
template <class FF>
__device__ bool wg_eigs(Ctx& c, FF&& F, double2* x, int64_t m, int kd, double tol, double2* Q, int64_t ld,
                        double2* w) {
    ExpLds& E = exp_lds();
    const int neig = c.I.neig;
    for (int cycle = 0;; ++cycle) {
        const double xn = sqrt(wg_norm2(x, m));
        if (!isfinite(xn) || !(xn > 0.0)) return false;
        for (int64_t q = threadIdx.x; q < m; q += NT) Q[q] = cscale(ldc(x + q), 1.0 / xn);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        double beta_last = 0.0, anorm = 0.0;
        bool invariant = false;
        int mm = 0;
        for (int j = 0; j < kd; ++j) {
            F(Q + j * ld, w);
            double h[2 * MAXK];
#pragma unroll
            for (int i = 0; i < 2 * MAXK; ++i) h[i] = 0.0;
            for (int pass = 0; pass < 2; ++pass) {
                double p[2 * MAXK];
#pragma unroll
                for (int i = 0; i < 2 * MAXK; ++i) p[i] = 0.0;
                for (int64_t q = threadIdx.x; q < m; q += NT) {
                    const double2 wq = ldc(w + q);
#pragma unroll
                    for (int i = 0; i < MAXK; ++i)
                        if (i <= j) {
                            const double2 pr = cmulc(ldc(Q + i * ld + q), wq);
                            p[2 * i] += pr.x;
                            p[2 * i + 1] += pr.y;
                        }
                }
                block_sum_lds<2 * MAXK>(p, g_sh.hb[pass], 2 * (j + 1));
                const double* hp = g_sh.hb[pass];
                for (int64_t q = threadIdx.x; q < m; q += NT) {
                    double2 wq = ldc(w + q);
                    for (int i = 0; i <= j; ++i) wq = csub(wq, cmul(cmk(hp[2 * i], hp[2 * i + 1]), ldc(Q + i * ld + q)));
                    w[q] = wq;
                }
                __builtin_amdgcn_s_waitcnt(0);
                __syncthreads();
#pragma unroll
                for (int i = 0; i < 2 * MAXK; ++i)
                    if (i < 2 * (j + 1)) h[i] += hp[i];
                __syncthreads();
            }
            const double bj = sqrt(wg_norm2(w, m));
            if (!isfinite(bj)) return false;
            const double aj = h[2 * j];
            anorm = fmax(anorm, fmax(fabs(aj), bj));
            if (threadIdx.x == 0) E.al[j] = aj;
            mm = j + 1;
            if (j + 1 == kd) {
                beta_last = bj;
                break;
            }
            if (bj <= 1e-12 * anorm) {
                invariant = true;
                break;
            }
            if (threadIdx.x == 0) E.be[j] = bj;
            for (int64_t q = threadIdx.x; q < m; q += NT) Q[(j + 1) * ld + q] = cscale(ldc(w + q), 1.0 / bj);
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            tridiag_eigen_serial(E, mm);
            E.mm = mm;
        }
        __syncthreads();
        const int k = min(neig, mm);
        bool converged = invariant;
        if (!converged) {
            converged = true;
            const double scale = fmax(fabs(E.th[0]), 1e-300);
            for (int i = 0; i < k; ++i)
                if (beta_last * fabs(E.Z[(mm - 1) * mm + i]) > tol * scale) converged = false;
        }
        if (converged || cycle == 99) return true;
        // restart from the sum of the leading Ritz vectors
        for (int64_t q = threadIdx.x; q < m; q += NT) {
            double2 acc = cz();
            for (int i = 0; i < k; ++i) {
                double2 y = cz();
                for (int j = 0; j < mm; ++j) y = cadd(y, cscale(ldc(Q + j * ld + q), E.Z[j * mm + i]));
                acc = cadd(acc, y);
            }
            x[q] = acc;
        }
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
    }
}

// entry q of Ritz vector i (after wg_eigs)
__device__ __forceinline__ double2 ritz(const double2* Q, int64_t ld, int i, int64_t q) {
    const ExpLds& E = exp_lds();
    double2 y = cz();
    for (int j = 0; j < E.mm; ++j) y = cadd(y, cscale(ldc(Q + j * ld + q), E.Z[j * E.mm + i]));
    return y;
}

// Singular values of the n x n gauge factor R, as populations relative to the largest:
// the number below the unoccupied threshold (get_nunoccupied, :577-585).  One-sided Jacobi
// (Hestenes) on wave 0, one row per lane.
__device__ int wg_n_unoccupied(const Ctx& c, const double2* R, int n) {
    ExpLds& E = exp_lds();
    __shared__ int cnt;
    for (int e = threadIdx.x; e < n * n; e += NT) E.R[e] = ldc(R + e);
    __syncthreads();
    if (threadIdx.x < 64) {
        const int i = threadIdx.x;
        for (int sweep = 0; sweep < 80; ++sweep) {
            bool rotated = false;
            for (int p = 0; p < n; ++p)
                for (int q = p + 1; q < n; ++q) {
                    const double2 ap = i < n ? E.R[i * n + p] : cz(), aq = i < n ? E.R[i * n + q] : cz();
                    const double alpha = wave_sum(cabs2(ap)), beta = wave_sum(cabs2(aq));
                    const double2 g = cmulc(ap, aq);
                    const double2 gamma = cmk(wave_sum(g.x), wave_sum(g.y));
                    const double gm = cabsv(gamma);
                    if (gm == 0.0 || gm <= 1e-15 * sqrt(alpha * beta)) continue;
                    rotated = true;
                    const double2 e = cscale(gamma, 1.0 / gm);
                    const double zeta = (beta - alpha) / (2.0 * gm);
                    const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                    const double cs = 1.0 / sqrt(1.0 + t * t), sn = cs * t;
                    if (i < n) {
                        const double2 aqr = cmulc(e, aq);  // conj(e) aq
                        E.R[i * n + p] = csub(cscale(ap, cs), cscale(aqr, sn));
                        E.R[i * n + q] = cmul(cadd(cscale(ap, sn), cscale(aqr, cs)), e);
                    }
                }
            if (!rotated) break;
        }
        double smax = 0.0;
        for (int j = 0; j < n; ++j) smax = fmax(smax, sqrt(wave_sum(i < n ? cabs2(E.R[i * n + j]) : 0.0)));
        int k = 0;
        for (int j = 0; j < n; ++j) {
            const double s = sqrt(wave_sum(i < n ? cabs2(E.R[i * n + j]) : 0.0));
            if (s / smax < c.I.unocc_thr) ++k;
        }
        if (i == 0) cnt = k;
    }
    __syncthreads();
    return cnt;
}

// generate_orthogonal_trial_vector (two_site_energy_variations.hpp:341-372) for node n's
// rank-3 view of `mode` (L, d, R): x gets L R entries from the normal stream, normalised, then
// each slice's unconjugated projection is removed and x renormalised; a draw without an
// orthogonal component is replaced by the next one.
__device__ bool wg_trial_vector(Ctx& c, int n, int mode, double2* x) {
    const Shape sh = c.shape(n);
    int64_t L, R;
    int d;
    rank3(sh, mode, L, d, R);
    const int64_t m = L * R;
    const double2* A = c.base + c.node(n).A;
    __shared__ int64_t pos_s;
    for (int attempt = 0; attempt < 1000; ++attempt) {
        if (threadIdx.x == 0) pos_s = ldc_i64(c.I.exp + EXP_RNG_POS);
        __syncthreads();
        const int64_t pos = pos_s;
        __syncthreads();
        if (pos + 2 * m > c.I.rng_n) {
            c.fail(FAIL_RNG_EXHAUSTED);
            return false;
        }
        for (int64_t q = threadIdx.x; q < m; q += NT) x[q] = cmk(c.I.rng[pos + 2 * q + 1], c.I.rng[pos + 2 * q]);
        if (threadIdx.x == 0) c.I.exp[EXP_RNG_POS] = pos + 2 * m;
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        wg_scale(x, 1.0 / sqrt(wg_norm2(x, m)), m);
        bool ok = true;
        for (int k = 0; k < d && ok; ++k) {
            double p[2] = {0.0, 0.0}, o[2];
            for (int64_t q = threadIdx.x; q < m; q += NT) {
                const int64_t l = q / R, j = q - l * R;
                const double2 pr = cmul(ldc(A + (l * d + k) * R + j), ldc(x + q));
                p[0] += pr.x;
                p[1] += pr.y;
            }
            block_sum<2>(p, o);
            const double2 uix = cmk(o[0], o[1]);
            for (int64_t q = threadIdx.x; q < m; q += NT) {
                const int64_t l = q / R, j = q - l * R;
                x[q] = csub(ldc(x + q), cmul(uix, ldc(A + (l * d + k) * R + j)));
            }
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            const double nrm = sqrt(wg_norm2(x, m));
            if (nrm > 1e-14)
                wg_scale(x, 1.0 / nrm, m);
            else
                ok = false;
        }
        if (ok) return true;
    }
    c.fail(FAIL_EXPANSION);
    return false;
}

// construct_two_site_energy_terms_lower (:78-142): out = (1 - A A^H) (sum_i a_i K_i A) r on
// the child cn (D1 x r1), r the r1 x r1 gauge factor; uses W1..W5
__device__ void ts_lower(Ctx& c, int cn, const DTerm& tm, const double2* r, double2* out) {
    const DNode& nd = c.node(cn);
    const Shape sh = c.shape(cn);
    const double2* A = c.base + nd.A;
    double2 *t1 = c.W(1), *t2 = c.W(2), *acc = c.W(3), *HA = c.W(4), *M = c.W(5);
    for (int i = 0; i < tm.nsub; ++i) {
        const DSub& s = c.I.subs[tm.sub0 + i];
        const double2 a = tm.nsub > 1 ? cmk(s.are, s.aim) : cmk(tm.a0re, tm.a0im);
        subterm_into(c, nd, sh, s, A, acc, a, i > 0, t1, t2);
    }
    if (tm.nsub == 0) t_zero(*c.T, acc, sh.size);
    t_mul(*c.T, acc, sh.dimen, sh.hrank, r, HA);
    t_adjoint_mul(*c.T, A, HA, sh.dimen, sh.hrank, M, cmk(1.0, 0.0), false);
    t_mul(*c.T, A, sh.dimen, sh.hrank, M, acc);
    t_copy(*c.T, out, HA, sh.size);
    t_axpy(*c.T, out, cmk(-1.0, 0.0), acc, sh.size);
}

// construct_two_site_energy_terms_upper (:144-222): out = sum_it a_it (siblings' spf) A_p
// mf_p^T on the parent, less its projection on the parent's slices of cn's mode; uses W1..W5
__device__ void ts_upper(Ctx& c, int cn, const DTerm& tm, double2* out) {
    const DNode& nd = c.node(cn);
    const DNode& pn = c.node(nd.parent);
    const Shape ps = c.shape(nd.parent);
    int64_t L, R;
    mode_view(ps, nd.child_id, L, R);
    const int d = ps.dims[nd.child_id];
    const double2* Ap = c.base + pn.A;
    double2 *t1 = c.W(1), *t2 = c.W(2), *G = c.W(4), *M = c.W(5);
    for (int it = 0; it < tm.nmf; ++it) {
        const DMf& mf = c.I.mfs[tm.mf0 + it];
        const double2 a = tm.nmf > 1 ? cmk(mf.are, mf.aim) : cmk(tm.a0re, tm.a0im);
        const double2* HA = Ap;
        if (mf.nk > 0) {
            kpo(c, ps, mf.k0, mf.nk, Ap, t1, t2);
            HA = t1;
        }
        if (mf.pmf >= 0)
            t_mul_transposed(*c.T, HA, ps.dimen, ps.hrank, c.base + mf.pmf, out, a, it > 0);
        else if (it > 0)
            t_axpy(*c.T, out, a, HA, ps.size);
        else
            t_scale_to(*c.T, out, a, HA, ps.size);
    }
    if (tm.nmf == 0) t_zero(*c.T, out, ps.size);
    // G(i, j) = sum conj(Ap(l,i,k)) out(l,j,k);  out(l,j,k) -= sum_j' G(j', j) Ap(l,j',k)
    t_outer_contract(*c.T, Ap, out, L, d, R, G, cmk(1.0, 0.0), false);
    for (int e = threadIdx.x; e < d * d; e += NT) {
        const int i = e / d, j = e - i * d;
        M[j * d + i] = ldc(G + e);
    }
    c.T->sync();
    t_mode_contract(*c.T, M, L, d, R, Ap, out, cmk(-1.0, 0.0), true);
}

// the two-site terms of node cn: indices (within its terms) of those with neither identity
// spf nor identity mean field, at most cap of them
__device__ int two_site_terms(const Ctx& c, int cn, int* idx, int cap) {
    const DNode& nd = c.node(cn);
    int k = 0;
    for (int ti = 0; ti < nd.nterms; ++ti) {
        const DTerm& tm = c.I.terms[nd.term0 + ti];
        if (!tm.id_spf && !tm.id_mf && k < cap) idx[k++] = ti;
    }
    return k;
}

// one record of the expansion log (thread 0 of the leader; dropped when the log is full): the
// spawning values sqrt(S) dt / 2 of the Ritz values S, the numbers the threshold is applied to
__device__ void log_attempt(const Ctx& c, int cn, int dir, int nadd, double dt) {
    if (c.I.exp_log == nullptr || !c.writer()) return;
    const int64_t k = c.I.exp[EXP_LOG_N];
    if (k >= c.I.exp_log_cap) return;
    const ExpLds& E = exp_lds();
    double* r = c.I.exp_log + k * EXP_LOG_REC;
    r[0] = cn;
    r[1] = dir;
    r[2] = E.mm;
    r[3] = nadd;
    for (int i = 0; i < MAXK; ++i) r[4 + i] = i < E.mm && E.th[i] > 0.0 ? sqrt(E.th[i]) * (dt / 2.0) : 0.0;
    c.I.exp[EXP_LOG_N] = k + 1;
}

// spawning test (:280-290): Ritz values S with sqrt(S) dt / 2 above the threshold
__device__ int count_spawned(const Ctx& c, double dt) {
    const ExpLds& E = exp_lds();
    int nadd = 0;
    for (int i = 0; i < E.mm; ++i) {
        const double sv = E.th[i] > 0.0 ? sqrt(E.th[i]) * (dt / 2.0) : 0.0;
        if (!isnan(sv) && sv > c.I.spawn) ++nadd;
    }
    return nadd;
}

// New layout of a (L, d, R) tensor with nadd slices appended to its middle index, from the
// copy in src; slice d + j from slice(j, l * R + k) (or zero when slice is null)
template <class SL>
__device__ void pad_middle(double2* dst, const double2* src, int64_t L, int d, int64_t R, int nadd, SL&& slice) {
    const int dn = d + nadd;
    const int64_t total = L * dn * R;
    for (int64_t e = threadIdx.x; e < total; e += NT) {
        const int64_t l = e / (dn * R), rem = e - l * dn * R;
        const int j = static_cast<int>(rem / R);
        const int64_t k = rem - static_cast<int64_t>(j) * R;
        dst[e] = j < d ? ldc(src + (l * d + j) * R + k) : slice(j - d, l * R + k);
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
}
// r (n x n) zero-padded to (n + nadd) squared, via the copy in tmp
__device__ void pad_square(double2* r, double2* tmp, int n, int nadd) {
    for (int e = threadIdx.x; e < n * n; e += NT) tmp[e] = ldc(r + e);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    const int m = n + nadd;
    for (int e = threadIdx.x; e < m * m; e += NT) {
        const int i = e / m, j = e - i * m;
        r[e] = (i < n && j < n) ? ldc(tmp + i * n + j) : cz();
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
}

// The child's live bond dimension becomes hnew (thread 0 of the leader writes it)
__device__ __forceinline__ void publish_hrank(const Ctx& c, int cn, int hnew) {
    if (c.writer()) c.I.hr[cn] = hnew;
}

// subspace_expansion_down (:207-358) for child `mode` of node p, run on one workgroup (team of
// one); returns the number of functions added.  The gauge factor of that mode is p's R (d x d).
__device__ __attribute__((noinline)) int expand_down_body(Ctx& c, int p, int mode, double dt) {
    const DNode& pn = c.node(p);
    const int cn = pn.children[mode];
    const DNode& nd = c.node(cn);
    const Shape ps = c.shape(p), cs = c.shape(cn);
    const int d = ps.dims[mode];
    if (d >= nd.cap_hrank || cs.hrank >= nd.cap_hrank) return 0;
    int max_add = nd.cap_hrank - d;
    int64_t L, R;
    mode_view(ps, mode, L, R);
    const int64_t maxdim = L * R;
    if (d >= maxdim) return 0;
    const bool twosite = cs.hrank < cs.dimen && cs.hrank != 1;
    double2* Rg = c.base + pn.R;
    int n_unocc = 0;
    if (c.I.only_unocc || !twosite) n_unocc = wg_n_unoccupied(c, Rg, d);
    if (c.I.only_unocc && n_unocc >= c.I.min_unocc) return 0;
    max_add = static_cast<int>(min<int64_t>(max_add, maxdim - d));
    int required = n_unocc < c.I.min_unocc ? c.I.min_unocc - n_unocc : 0;
    required = min(required, max_add);
    double2* ev = c.base + c.I.e_vec;
    const int64_t ld = c.I.maxsize;
    double2 *x = ev, *t1 = ev + ld;
    double2* A1 = c.base + nd.A;
    double2* A2 = c.base + pn.A;
    double2* tmp = c.W(5);
    if (!twosite) {
        // random one-site directions, one at a time (:319-350)
        int nadd = 0;
        while (nadd < required) {
            if (!wg_trial_vector(c, p, mode, x)) return nadd;
            const Shape ps2 = c.shape(p), cs2 = c.shape(cn);
            int64_t L2, R2;
            mode_view(ps2, mode, L2, R2);
            const int d2 = ps2.dims[mode];
            t_copy(*c.T, tmp, A1, cs2.size);
            pad_middle(A1, tmp, cs2.dimen, cs2.hrank, 1, 1, [&](int, int64_t) { return cz(); });
            pad_square(Rg, tmp, d2, 1);
            t_copy(*c.T, tmp, A2, ps2.size);
            pad_middle(A2, tmp, L2, d2, R2, 1, [&](int, int64_t q) { return ldc(x + q); });
            if (threadIdx.x == 0) g_hr[cn] = static_cast<unsigned char>(cs2.hrank + 1);
            __syncthreads();
            ++nadd;
        }
        if (c.writer()) c.I.exp[EXP_ONESITE] += 1;
        publish_hrank(c, cn, c.hr(cn));
        return nadd;
    }
    int idx[NC_TERMS];
    const int nt2 = two_site_terms(c, cn, idx, c.I.nt2);
    const int64_t D1 = cs.dimen;
    const int r1 = cs.hrank;
    for (int t = 0; t < nt2; ++t) {
        const DTerm& tm = c.I.terms[nd.term0 + idx[t]];
        ts_lower(c, cn, tm, Rg, c.base + c.I.e_op1 + t * ld);
        ts_upper(c, cn, tm, c.base + c.I.e_op2 + t * ld);
    }
    if (!wg_trial_vector(c, p, mode, x)) return 0;
    ExpLds& E = exp_lds();
    // M^H M v with M = sum_t op1_t (op2_t as d x LR)
    auto F = [&](const double2* v, double2* out) {
        if (nt2 == 0) {
            for (int64_t q = threadIdx.x; q < maxdim; q += NT) out[q] = cz();
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            return;
        }
        for (int t = 0; t < nt2; ++t) {
            wg_contract_long(c.base + c.I.e_op2 + t * ld, L, d, R, v, false, E.s);
            wg_apply_short(c.base + c.I.e_op1 + t * ld, D1, r1, 1, E.s, false, t1, t > 0);
        }
        for (int t = 0; t < nt2; ++t) {
            wg_contract_long(c.base + c.I.e_op1 + t * ld, D1, r1, 1, t1, true, E.s);
            wg_apply_short(c.base + c.I.e_op2 + t * ld, L, d, R, E.s, true, out, t > 0);
        }
    };
    const int kd = static_cast<int>(min<int64_t>(c.I.kdim, maxdim));
    double2* Q = c.base + c.I.krylov;
    if (!wg_eigs(c, F, x, maxdim, kd, c.I.eig_tol, Q, ld, c.W(0))) {
        c.fail(FAIL_EXPANSION);
        return 0;
    }
    const int nadd = min(count_spawned(c, dt), max_add);
    log_attempt(c, cn, 0, nadd, dt);
    if (c.writer()) c.I.exp[EXP_TWOSITE] += 1;
    if (nadd == 0) return 0;
    // child: zero columns; gauge factor: zero padding; parent: new slices conj(Ritz vectors)
    t_copy(*c.T, tmp, A1, cs.size);
    pad_middle(A1, tmp, D1, r1, 1, nadd, [&](int, int64_t) { return cz(); });
    pad_square(Rg, tmp, d, nadd);
    t_copy(*c.T, tmp, A2, ps.size);
    pad_middle(A2, tmp, L, d, R, nadd, [&](int j, int64_t q) {
        double2 y = ritz(Q, ld, j, q);
        y.y = -y.y;
        return y;
    });
    publish_hrank(c, cn, r1 + nadd);
    return nadd;
}

// subspace_expansion_up (:361-510) for node cn and its parent after cn's leaf-to-root
// factorisation (cn's R, r1 x r1), on one workgroup; returns the number of functions added
__device__ __attribute__((noinline)) int expand_up_body(Ctx& c, int cn, double dt) {
    const DNode& nd = c.node(cn);
    const int p = nd.parent;
    const DNode& pn = c.node(p);
    const Shape cs = c.shape(cn), ps = c.shape(p);
    if (cs.hrank >= nd.cap_hrank) return 0;
    int max_add = nd.cap_hrank - cs.hrank;
    const int64_t D1 = cs.dimen;
    const int r1 = cs.hrank;
    if (r1 >= D1) return 0;
    int64_t L, R;
    mode_view(ps, nd.child_id, L, R);
    const int d = ps.dims[nd.child_id];
    const bool twosite = d < L * R && d != 1;
    double2* Rg = c.base + nd.R;
    int n_unocc = 0;
    if (c.I.only_unocc || !twosite) n_unocc = wg_n_unoccupied(c, Rg, r1);
    if (c.I.only_unocc && n_unocc >= c.I.min_unocc) return 0;
    max_add = static_cast<int>(min<int64_t>(max_add, D1 - r1));
    int required = n_unocc < c.I.min_unocc ? c.I.min_unocc - n_unocc : 0;
    required = min(required, max_add);
    double2* ev = c.base + c.I.e_vec;
    const int64_t ld = c.I.maxsize;
    double2 *x = ev, *t1 = ev + ld;
    double2* A1 = c.base + nd.A;
    double2* A2 = c.base + pn.A;
    double2* tmp = c.W(5);
    if (!twosite) {
        int nadd = 0;
        while (nadd < required) {
            if (!wg_trial_vector(c, cn, nd.nmodes, x)) return nadd;
            const Shape ps2 = c.shape(p), cs2 = c.shape(cn);
            int64_t L2, R2;
            mode_view(ps2, nd.child_id, L2, R2);
            const int d2 = ps2.dims[nd.child_id];
            t_copy(*c.T, tmp, A2, ps2.size);
            pad_middle(A2, tmp, L2, d2, R2, 1, [&](int, int64_t) { return cz(); });
            pad_square(Rg, tmp, cs2.hrank, 1);
            t_copy(*c.T, tmp, A1, cs2.size);
            pad_middle(A1, tmp, D1, cs2.hrank, 1, 1, [&](int, int64_t q) { return ldc(x + q); });
            if (threadIdx.x == 0) g_hr[cn] = static_cast<unsigned char>(cs2.hrank + 1);
            __syncthreads();
            ++nadd;
        }
        if (c.writer()) c.I.exp[EXP_ONESITE] += 1;
        publish_hrank(c, cn, c.hr(cn));
        return nadd;
    }
    int idx[NC_TERMS];
    const int nt2 = two_site_terms(c, cn, idx, c.I.nt2);
    for (int t = 0; t < nt2; ++t) {
        const DTerm& tm = c.I.terms[nd.term0 + idx[t]];
        ts_upper(c, cn, tm, c.base + c.I.e_op2 + t * ld);
        ts_lower(c, cn, tm, Rg, c.base + c.I.e_op1 + t * ld);
    }
    if (!wg_trial_vector(c, cn, nd.nmodes, x)) return 0;
    ExpLds& E = exp_lds();
    // conj(M) M^T v
    auto F = [&](const double2* v, double2* out) {
        if (nt2 == 0) {
            for (int64_t q = threadIdx.x; q < D1; q += NT) out[q] = cz();
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            return;
        }
        for (int t = 0; t < nt2; ++t) {
            wg_contract_long(c.base + c.I.e_op1 + t * ld, D1, r1, 1, v, false, E.s);
            wg_apply_short(c.base + c.I.e_op2 + t * ld, L, d, R, E.s, false, t1, t > 0);
        }
        for (int t = 0; t < nt2; ++t) {
            wg_contract_long(c.base + c.I.e_op2 + t * ld, L, d, R, t1, true, E.s);
            wg_apply_short(c.base + c.I.e_op1 + t * ld, D1, r1, 1, E.s, true, out, t > 0);
        }
    };
    const int kd = static_cast<int>(min<int64_t>(c.I.kdim, D1));
    double2* Q = c.base + c.I.krylov;
    if (!wg_eigs(c, F, x, D1, kd, 1e-4, Q, ld, c.W(0))) {
        c.fail(FAIL_EXPANSION);
        return 0;
    }
    const int nadd = min(count_spawned(c, dt), max_add);
    log_attempt(c, cn, 1, nadd, dt);
    if (nadd == 0) return 0;
    if (c.writer()) c.I.exp[EXP_TWOSITE] += 1;
    // parent: zero slices; gauge factor: zero padding; child: new columns (Ritz vectors)
    t_copy(*c.T, tmp, A2, ps.size);
    pad_middle(A2, tmp, L, d, R, nadd, [&](int, int64_t) { return cz(); });
    pad_square(Rg, tmp, r1, nadd);
    t_copy(*c.T, tmp, A1, cs.size);
    pad_middle(A1, tmp, D1, r1, 1, nadd, [&](int j, int64_t q) { return ritz(Q, ld, j, q / 1); });
    publish_hrank(c, cn, r1 + nadd);
    return nadd;
}

// Runs an expansion on the team's leader and brings every member's live bond dimension of
// the child up to date; returns whether the bond grew (the same on every member).
template <class F>
__device__ bool expand_on_leader(Ctx& c, int cn, F&& body) {
    const int before = c.hr(cn);
    c.where = cn;
    c.bond = 1;
    on_team(c, 0, [&](Ctx& cc) { (void)body(cc); });
    if (threadIdx.x == 0) g_hr[cn] = static_cast<unsigned char>(ldc_int(c.I.hr + cn));
    __syncthreads();
    return c.hr(cn) != before;
}
