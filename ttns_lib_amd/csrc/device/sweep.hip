// Persistent TDVP sweep kernel for gfx950.
//
// A team of workgroups (one XCD) owns one TTNS instance and runs whole integrator calls
// on the device: prepare_evolution, the first-step ladder and any number of regular
// steps, each step a forward and a backward projector-splitting sweep (reference
// tdvp_core/tdvp_algorithm_engine.hpp:207-501) with the Krylov propagators, environment
// updates and QR gauge moves in between.  No host round trip happens inside a call.  A
// batch is one launch in which every instance has its own team.  Operations on large
// node tensors are spread over the whole team; small ones run on the team's leader while
// the other members wait at the next team barrier.
#include <hip/hip_runtime.h>

#include "desc.hpp"
#include "wg_ops.hpp"

namespace ttnd {

// LDS of the small-vector propagator (below); between its propagations the same 4608 entries
// stage the operands of one-workgroup bond products
constexpr int SMALL_E = 2;
constexpr int SMALL_N = SMALL_E * NT;
constexpr int SMALL_OPS = 2048;  // LDS operator arena (double2)

struct SmallLds {
    double2 v[SMALL_N], w[SMALL_N], t1[SMALL_N], t2[SMALL_N], t3[SMALL_N];
    double2 ops[SMALL_OPS];
};
__shared__ SmallLds g_small;
constexpr int LDS_STAGE = (5 * SMALL_N + SMALL_OPS) / 2;  // entries per staged operand

// Live bond dimension of every node of the workgroup's instance, loaded from DInst::hr when the
// kernel starts and updated by every member after an expansion grows a bond.
__shared__ unsigned char g_hr[MAX_NODES];

struct Ctx {
    const DInst& I;
    Team* T;
    double2* base;
    int64_t team_min;
    SmallLds* small = nullptr;  // LDS of the small-vector propagator
    int small_max = 0;          // propagations of at most this many entries use it
    // where: node whose propagation is running, bond: 1 for its bond matrix
    int where = 0, bond = 0;
    __device__ Ctx(const DInst& inst, Team* t, int64_t tmin)
        : I(inst), T(t), base(inst.base), team_min(tmin) {}
    __device__ double2* W(int k) const { return base + I.work[k]; }
    __device__ const DNode& node(int n) const { return I.nodes[n]; }
    __device__ int hr(int n) const { return g_hr[n]; }
    // live shape of node n (bond dimensions from the LDS table)
    __device__ Shape shape(int n) const {
        const DNode& nd = I.nodes[n];
        Shape s;
        s.hrank = g_hr[n];
        s.nmodes = nd.nmodes;
        if (nd.leaf >= 0) {
            s.dims[0] = nd.leafdim;
            for (int q = 1; q < MAXC; ++q) s.dims[q] = 1;
            s.dimen = nd.leafdim;
        } else {
            s.dimen = 1;
            for (int q = 0; q < MAXC; ++q) {
                s.dims[q] = q < nd.nmodes ? static_cast<int>(g_hr[nd.children[q]]) : 1;
                s.dimen *= s.dims[q];
            }
        }
        s.size = s.dimen * s.hrank;
        return s;
    }
    // failures are detected on team-uniform values, so every thread of every member takes
    // the same branch; the LDS flag is read only after a barrier
    __device__ bool failed() const { return g_sh.fail != 0; }
    __device__ void fail(int reason) const { T->fail(reason | (bond << 4) | (where << 5)); }
    // thread 0 of the team leader writes statistics and results
    __device__ bool writer() const { return T->rank == 0 && threadIdx.x == 0; }
    // operations on at least team_min elements are spread over the team
    __device__ bool spread(int64_t elems) const { return T->size > 1 && elems >= team_min; }
};

// Runs f on the team's leader alone (as a team of one) unless the operation is large
// enough to spread; the other members wait at the barrier that follows.  A failure the
// leader records reaches the others through the team word.
template <class F>
__device__ void on_team(Ctx& c, int64_t elems, F&& f) {
    if (c.T->size == 1 || c.spread(elems)) {
        f(c);
        return;
    }
    if (c.T->rank == 0) {
        Team s = c.T->single();
        Ctx cs = c;
        cs.T = &s;
        f(cs);
        if (threadIdx.x == 0 && g_sh.fail != 0)
            __hip_atomic_store(&c.T->w->fail, g_sh.fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    c.T->sync();
}

__device__ __forceinline__ uint64_t prof_now(const Ctx& c) { return c.I.prof ? wall_clock64() : 0; }
__device__ __forceinline__ void prof_add(const Ctx& c, int slot, uint64_t t0) {
    if (c.I.prof && c.writer()) c.I.prof[slot] += static_cast<int64_t>(wall_clock64() - t0);
}
#define PROFILE(slot, stmt)                  \
    {                                        \
        const uint64_t _pt0 = prof_now(c);   \
        stmt;                                \
        prof_add(c, slot, _pt0);             \
    }

// (L, R) of the rank-3 view of mode `mode` of a node
__device__ __forceinline__ void mode_view(const Shape& sh, int mode, int64_t& L, int64_t& R) {
    L = 1;
    R = sh.hrank;
    for (int c = 0; c < mode; ++c) L *= sh.dims[c];
    for (int c = mode + 1; c < sh.nmodes; ++c) R *= sh.dims[c];
}

// kronecker_product_operator::apply (kronecker_product_operator_helper.hpp:261-315):
// chain of mode contractions with spf matrices; the last contraction lands in out.
__device__ void kpo(Ctx& c, const Shape& sh, int k0, int nk, const double2* v, double2* out, double2* tmp) {
    if (nk == 0) {
        t_copy(*c.T, out, v, sh.size);
        return;
    }
    const double2* src = v;
    for (int q = 0; q < nk; ++q) {
        double2* dst = ((nk - 1 - q) % 2 == 0) ? out : tmp;
        const DK& k = c.I.ks[k0 + q];
        int64_t L, R;
        mode_view(sh, k.nu, L, R);
        t_mode_contract(*c.T, c.base + k.mat, L, sh.dims[k.nu], R, src, dst, cmk(1.0, 0.0), false);
        src = dst;
    }
}

// coefficient_evolution_{leaf,branch} (projector_splitting_evolution_functions.hpp:65-258)
// out (=|+=) alpha * K v for one spf sub-term: a leaf's primitive operator or a chain of
// mode contractions (intermediates in t1/t2, the last one lands in out with the epilogue)
__device__ void subterm_into(Ctx& c, const DNode& nd, const Shape& sh, const DSub& s, const double2* v, double2* out,
                             double2 alpha, bool accumulate, double2* t1, double2* t2) {
    if (s.nk == 0) {
        if (accumulate)
            t_axpy(*c.T, out, alpha, v, sh.size);
        else
            t_scale_to(*c.T, out, alpha, v, sh.size);
        return;
    }
    if (nd.leaf >= 0) {
        t_prim_apply(*c.T, c.I.ks[s.k0], c.base, c.I.ints, v, out, sh.hrank, alpha, accumulate);
        return;
    }
    const double2* src = v;
    for (int q = 0; q < s.nk; ++q) {
        const bool last = q == s.nk - 1;
        double2* dst = last ? out : (q % 2 == 0 ? t1 : t2);
        const DK& k = c.I.ks[s.k0 + q];
        int64_t L, R;
        mode_view(sh, k.nu, L, R);
        t_mode_contract(*c.T, c.base + k.mat, L, sh.dims[k.nu], R, src, dst, last ? alpha : cmk(1.0, 0.0),
                        last && accumulate);
        src = dst;
    }
}

// res = sum_t [sum_i c_i K_{t,i} v] mf_t^T with the identity shortcuts of the reference;
// accumulation and coefficients are folded into the GEMM epilogues.  Uses W1..W3.
__device__ void heff_node(Ctx& c, int n, const double2* v, double2* res) {
    const DNode& nd = c.node(n);
    const Shape sh = c.shape(n);
    const int r = sh.hrank;
    const int64_t D = sh.dimen;
    double2 *t1 = c.W(1), *t2 = c.W(2), *t3 = c.W(3);
    bool first = true;  // res not written yet
    for (int ti = 0; ti < nd.nterms; ++ti) {
        const DTerm& tm = c.I.terms[nd.term0 + ti];
        if (tm.id_spf) {
            const double2 coeff = cmk(tm.ire, tm.iim);
            if (tm.id_mf) {
                if (first)
                    t_scale_to(*c.T, res, coeff, v, sh.size);
                else
                    t_axpy(*c.T, res, coeff, v, sh.size);
            } else {
                t_mul_transposed(*c.T, v, D, r, c.base + tm.mf, res, coeff, !first);
            }
            first = false;
            continue;
        }
        for (int i = 0; i < tm.nsub; ++i) {
            const DSub& s = c.I.subs[tm.sub0 + i];
            const double2 ci = (i == 0) ? (tm.nsub > 1 ? cmk(s.are, s.aim) : cmk(tm.cre, tm.cim)) : cmk(s.are, s.aim);
            if (tm.id_mf) {
                subterm_into(c, nd, sh, s, v, res, ci, !first, t1, t2);
                first = false;
            } else {
                subterm_into(c, nd, sh, s, v, t3, ci, i > 0, t1, t2);
            }
        }
        if (!tm.id_mf) {
            t_mul_transposed(*c.T, t3, D, r, c.base + tm.mf, res, cmk(1.0, 0.0), !first);
            first = false;
        }
    }
    if (first) t_zero(*c.T, res, sh.size);
}

// dst (=|+=) alpha * P Q on one workgroup, P and Q r x r matrices copied into LDS first, with
// Q transposed when TRANS (Q^T: the mean-field product x mf^T).  The tiles and their k order
// are those of t_mode_contract (L = 1) and t_mul_transposed, so the sums are the same.  The copy
// issues the 16-byte coherent loads of UNR strides of both matrices before storing any of them.
template <bool TRANS>
__device__ void wg_square_product_lds(const double2* P, const double2* Q, int r, double2* dst, double2 alpha,
                                      bool accumulate) {
    double2* a = reinterpret_cast<double2*>(&g_small);  // the whole 4608-entry area
    double2* b = a + LDS_STAGE;
    const int rr = r * r;
    const Coh cp(P), cq(Q);
    for (int e0 = threadIdx.x; e0 < rr; e0 += UNR * NT) {
        double2 pv[UNR], qv[UNR];
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            const int e = e0 + u * NT;
            pv[u] = e < rr ? cp[e] : cz();
            qv[u] = e < rr ? cq[e] : cz();
        }
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            const int e = e0 + u * NT;
            if (e < rr) {
                a[e] = pv[u];
                b[e] = qv[u];
            }
        }
    }
    __syncthreads();
    const int wv = threadIdx.x >> 6;
    const int mt = (r + 15) >> 4;
    for (int it = wv; it < mt * mt; it += NWAVE) {
        const int n = it % mt, m = it / mt;
        const int mv = min(16, r - 16 * m), nv = min(16, r - 16 * n);
        ZAcc acc;
        zacc_zero(acc);
        if (TRANS)
            ztile<false>(acc, a + 16 * m * r, r, 1, mv, b + 16 * n * r, 1, r, nv, r, LoadLocal());
        else
            ztile<false>(acc, a + 16 * m * r, r, 1, mv, b + 16 * n, r, 1, nv, r, LoadLocal());
        ztile_store(acc, dst + 16 * m * r + 16 * n, r, 1, mv, nv, alpha, accumulate);
    }
    // the stores must have landed before other waves read dst with L1-bypassing loads
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
}

// full_hamiltonian_action (projector_splitting_evolution_functions.hpp:7-62); uses W1, W2
__device__ void heff_bond(Ctx& c, int n, const double2* v, double2* res) {
    const DNode& nd = c.node(n);
    const int r = c.hr(n);
    const int64_t sz = static_cast<int64_t>(r) * r;
    double2* t1 = c.W(1);
    t_zero(*c.T, res, sz);
    if (c.T->size == 1 && sz <= LDS_STAGE && c.small != nullptr) {
        // one workgroup: every product with both operands in LDS
        for (int ti = 0; ti < nd.nterms; ++ti) {
            const DTerm& tm = c.I.terms[nd.term0 + ti];
            const double2* x = v;
            if (!tm.id_spf) {
                wg_square_product_lds<false>(c.base + tm.spf, v, r, t1, cmk(1.0, 0.0), false);
                x = t1;
            }
            const double2 coeff = cmk(tm.cre, tm.cim);
            if (!tm.id_mf)
                wg_square_product_lds<true>(x, c.base + tm.mf, r, res, coeff, true);
            else
                t_axpy(*c.T, res, coeff, x, sz);
        }
        return;
    }
    for (int ti = 0; ti < nd.nterms; ++ti) {
        const DTerm& tm = c.I.terms[nd.term0 + ti];
        const double2* x = v;
        if (!tm.id_spf) {
            t_mode_contract(*c.T, c.base + tm.spf, 1, r, r, v, t1, cmk(1.0, 0.0), false);
            x = t1;
        }
        const double2 coeff = cmk(tm.cre, tm.cim);
        if (!tm.id_mf)
            t_mul_transposed(*c.T, x, r, r, c.base + tm.mf, res, coeff, true);
        else
            t_axpy(*c.T, res, coeff, x, sz);
    }
}

// single_particle_operator_engine::evaluate (single_particle_operator.hpp:26-103); uses W1, W2
__device__ void update_spf_body(Ctx& c, int n) {
    const DNode& nd = c.node(n);
    const Shape sh = c.shape(n);
    const int r = sh.hrank;
    const int64_t D = sh.dimen;
    const double2* a = c.base + nd.A;
    double2 *HA = c.W(1), *tmp = c.W(2);
    for (int ti = 0; ti < nd.nterms; ++ti) {
        const DTerm& tm = c.I.terms[nd.term0 + ti];
        if (tm.id_spf) continue;
        double2* out = c.base + tm.spf;
        for (int i = 0; i < tm.nsub; ++i) {
            const DSub& s = c.I.subs[tm.sub0 + i];
            if (nd.leaf >= 0 && s.nk > 0)
                t_prim_apply(*c.T, c.I.ks[s.k0], c.base, c.I.ints, a, HA, r);
            else
                kpo(c, sh, s.k0, s.nk, a, HA, tmp);
            if (tm.nsub == 1)
                t_adjoint_mul(*c.T, a, HA, D, r, out, cmk(1.0, 0.0), false);
            else
                t_adjoint_mul(*c.T, a, HA, D, r, out, cmk(s.are, s.aim), i > 0);
        }
    }
}
__device__ void update_spf(Ctx& c, int n) {
    on_team(c, c.shape(n).size, [&](Ctx& cc) { update_spf_body(cc, n); });
}

// mean_field_operator_engine::evaluate (mean_field_operator.hpp:37-135) for child n;
// uses W1, W2, W3
__device__ void update_mf_body(Ctx& c, int n) {
    const DNode& nd = c.node(n);
    const DNode& pn = c.node(nd.parent);
    const Shape ps = c.shape(nd.parent);
    const int mode = nd.child_id;
    int64_t L, R;
    mode_view(ps, mode, L, R);
    const int d = ps.dims[mode];
    const int64_t pD = ps.dimen;
    const int pr = ps.hrank;
    const double2* Ap = c.base + pn.A;
    double2 *t1 = c.W(1), *t2 = c.W(2), *t3 = c.W(3);
    for (int ti = 0; ti < nd.nterms; ++ti) {
        const DTerm& tm = c.I.terms[nd.term0 + ti];
        if (tm.id_mf) continue;
        double2* out = c.base + tm.mf;
        for (int it = 0; it < tm.nmf; ++it) {
            const DMf& m = c.I.mfs[tm.mf0 + it];
            const double2* HA = Ap;
            if (m.nk > 0) {
                kpo(c, ps, m.k0, m.nk, Ap, t1, t2);
                HA = t1;
            }
            if (m.pmf >= 0) {
                t_mul_transposed(*c.T, HA, pD, pr, c.base + m.pmf, t3, cmk(1.0, 0.0), false);
                HA = t3;
            }
            if (tm.nmf == 1)
                t_outer_contract(*c.T, Ap, HA, L, d, R, out, cmk(1.0, 0.0), false);
            else
                t_outer_contract(*c.T, Ap, HA, L, d, R, out, cmk(m.are, m.aim), it > 0);
        }
    }
}
__device__ void update_mf(Ctx& c, int n) {
    const DNode& nd = c.node(n);
    if (nd.parent < 0) return;
    on_team(c, c.shape(nd.parent).size, [&](Ctx& cc) { update_mf_body(cc, n); });
}

// ---------------------------------------------------------------- gauge moves
// root_to_leaf_decomposition_engine::evaluate + apply_to_node (root_to_leaf_decomposition.hpp:112-204)
__device__ void r2l_body(Ctx& c, int n, int mode) {
    Team& T = *c.T;
    const DNode& nd = c.node(n);
    const Shape sh = c.shape(n);
    const int64_t sz = sh.size;
    int64_t id1 = 1;
    for (int q = 0; q <= mode; ++q) id1 *= sh.dims[q];
    const int64_t id2 = sz / id1;
    const int d2 = sh.dims[mode];
    const int d1 = static_cast<int>(sz / d2);
    double2* A = c.base + nd.A;
    double2 *U = c.W(5), *Q = c.W(1), *Rm = c.W(2);
    for (int64_t e = T.tid(); e < sz; e += T.nthr()) {
        const int64_t i = e / id2, j = e % id2;
        U[j * id1 + i] = ldc(A + e);
    }
    T.sync();
    t_householder_qr(T, U, d1, d2, Q, Rm);
    for (int64_t e = T.tid(); e < sz; e += T.nthr()) {
        const int64_t i = e / id2, j = e % id2;
        A[e] = ldc(Q + j * id1 + i);
    }
    double2* Rn = c.base + nd.R;
    for (int64_t e = T.tid(); e < static_cast<int64_t>(d2) * d2; e += T.nthr()) {
        const int i = static_cast<int>(e / d2), j = static_cast<int>(e % d2);
        Rn[e] = ldc(Rm + static_cast<int64_t>(j) * d2 + i);
    }
    T.sync();
}
__device__ void r2l(Ctx& c, int n, int mode) {
    on_team(c, c.shape(n).size, [&](Ctx& cc) { r2l_body(cc, n, mode); });
}

// leaf_to_root_decomposition_engine::evaluate + apply_to_node (leaf_to_root_decomposition.hpp:56-101)
__device__ void l2r(Ctx& c, int n) {
    on_team(c, c.shape(n).size, [&](Ctx& cc) {
        const DNode& nd = cc.node(n);
        const Shape sh = cc.shape(n);
        double2* A = cc.base + nd.A;
        double2* U = cc.W(5);
        t_copy(*cc.T, U, A, sh.size);
        t_householder_qr(*cc.T, U, sh.dimen, sh.hrank, A, cc.base + nd.R);
    });
}

// root_to_leaf_decomposition_engine::apply_from_parent (:208-233): A <- A * R_parent
__device__ void apply_from_parent(Ctx& c, int n) {
    on_team(c, c.shape(n).size, [&](Ctx& cc) {
        const DNode& nd = cc.node(n);
        const Shape sh = cc.shape(n);
        double2* A = cc.base + nd.A;
        double2* Tm = cc.W(5);
        t_mul(*cc.T, A, sh.dimen, sh.hrank, cc.base + cc.node(nd.parent).R, Tm);
        t_copy(*cc.T, A, Tm, sh.size);
    });
}

// leaf_to_root_decomposition_engine::apply_to_parent (leaf_to_root_decomposition.hpp:105-138)
__device__ void apply_to_parent(Ctx& c, int n) {
    on_team(c, c.shape(c.node(n).parent).size, [&](Ctx& cc) {
        const DNode& nd = cc.node(n);
        const DNode& pn = cc.node(nd.parent);
        const Shape ps = cc.shape(nd.parent);
        int64_t L, R;
        mode_view(ps, nd.child_id, L, R);
        double2* Tm = cc.W(5);
        double2* Ap = cc.base + pn.A;
        t_mode_contract(*cc.T, cc.base + nd.R, L, ps.dims[nd.child_id], R, Ap, Tm, cmk(1.0, 0.0), false);
        t_copy(*cc.T, Ap, Tm, ps.size);
    });
}

// ---------------------------------------------------------------- Krylov propagator
// Sub-step factor from the error model err(s) ~ s^m: 0.9 (tol / err)^(1/m), clamped to
// [lo, hi]; err == 0 (invariant Krylov space) gives hi.
__device__ __forceinline__ double step_factor(double tol, double err, int m, double lo, double hi) {
    if (!(err > 0.0)) return hi;
    const double f = 0.9 * pow(tol / err, 1.0 / m);
    return fmin(hi, fmax(lo, f));
}

// x <- exp(coeff * tau * H) x: Arnoldi with CGS2, Saad error estimate, Expokit sub-step
// sizes, Pade(13) small exponential (same algorithm as the oracle's expv).  Returns the number
// of H applications.  H(v, w) may use W1..W4; the Krylov vector w is W0.  Every member
// owns the same vector entries in all three Gram-Schmidt sweeps, so only the reductions
// and the new basis vector need team barriers.
template <class HF>
__device__ int64_t expv(Ctx& c, HF&& H, double2* x, int64_t n, double2 coeff, double tau, int64_t& sumj) {
    if (tau == 0.0 || n == 0) return 0;
    Team& T = *c.T;
    const int kdim = c.I.kdim;
    const double tol = c.I.tol;
    double2* V = c.base + c.I.krylov;
    const int64_t ld = c.I.maxsize;
    double2* w = c.W(0);
    const int64_t q00 = T.tid(), st = T.nthr();
    int64_t napply = 0;
    // s_next: fraction of tau predicted for the next sub-step (starts with the whole step)
    double s_done = 0.0, s_next = 1.0;
    int guard = 0;
    while (s_done < 1.0 && guard++ < 4096) {
        double s_step = fmin(s_next, 1.0 - s_done);
        const double beta = sqrt(team_norm2(T, x, n));
        if (!isfinite(beta)) {
            c.fail(FAIL_INPUT_NONFINITE);
            return napply;
        }
        if (beta == 0.0) return napply;
        t_scale_to(T, V, cmk(1.0 / beta, 0.0), x, n);
        for (int e = threadIdx.x; e < (MAXK + 1) * MAXK; e += NT) g_sh.Hm[e] = cz();
        __syncthreads();
        int m = 0;
        double hn = 0.0, err = 0.0;
        bool found = false;
        auto small_exp = [&](int mm, double s) -> double {
            const double2 f = cscale(coeff, tau * s);
            for (int e = threadIdx.x; e < mm * mm; e += NT) {
                const int i = e / mm, j = e % mm;
                g_sh.S[e] = cmul(f, g_sh.Hm[i * MAXK + j]);
            }
            __syncthreads();
            block_expm(g_sh.S, g_sh.F, mm);
            return hn * cabsv(g_sh.F[(mm - 1) * mm]);
        };
        for (int j = 0; j < kdim; ++j) {
            H(V + j * ld, w);
            ++napply;
            sumj += j + 1;
            const uint64_t ot0 = prof_now(c);
            // classical Gram-Schmidt applied twice (CGS2): the second pass's inner
            // products accumulate while the first correction is written, the norm while
            // the second one is written (three sweeps over the basis instead of five)
            double p[2 * MAXK];
            const double* h1 = g_sh.hb[0];
            const double* h2 = g_sh.hb[1];
#pragma unroll
            for (int i = 0; i < 2 * MAXK; ++i) p[i] = 0.0;
            // two strides per iteration keep 2(j+2) loads in flight per thread
            const Coh cV(V), cw(w);
            for (int64_t q0 = q00; q0 < n; q0 += 2 * st) {
                const int64_t q1 = q0 + st;
                const bool ok1 = q1 < n;
                const double2 w0 = cw[q0], w1 = ok1 ? cw[q1] : cz();
#pragma unroll
                for (int i = 0; i < MAXK; ++i) {
                    if (i <= j) {
                        const double2 a0 = cV[i * ld + q0], a1 = ok1 ? cV[i * ld + q1] : cz();
                        const double2 pr = cadd(cmulc(a0, w0), cmulc(a1, w1));
                        p[2 * i] += pr.x;
                        p[2 * i + 1] += pr.y;
                    }
                }
            }
            team_sum_lds<2 * MAXK>(T, p, g_sh.hb[0], 2 * (j + 1));
#pragma unroll
            for (int i = 0; i < 2 * MAXK; ++i) p[i] = 0.0;
            for (int64_t q = q00; q < n; q += st) {
                double2 vi[MAXK];
                double2 wq = cw[q];
#pragma unroll
                for (int i = 0; i < MAXK; ++i)
                    if (i <= j) vi[i] = cV[i * ld + q];
#pragma unroll
                for (int i = 0; i < MAXK; ++i)
                    if (i <= j) wq = csub(wq, cmul(cmk(h1[2 * i], h1[2 * i + 1]), vi[i]));
                w[q] = wq;
#pragma unroll
                for (int i = 0; i < MAXK; ++i) {
                    if (i <= j) {
                        const double2 pr = cmulc(vi[i], wq);
                        p[2 * i] += pr.x;
                        p[2 * i + 1] += pr.y;
                    }
                }
            }
            team_sum_lds<2 * MAXK>(T, p, g_sh.hb[1], 2 * (j + 1));
            double nn[1] = {0.0}, nno[1];
            for (int64_t q0 = q00; q0 < n; q0 += 2 * st) {
                const int64_t q1 = q0 + st;
                const bool ok1 = q1 < n;
                double2 w0 = cw[q0], w1 = ok1 ? cw[q1] : cz();
#pragma unroll
                for (int i = 0; i < MAXK; ++i) {
                    if (i <= j) {
                        const double2 h = cmk(h2[2 * i], h2[2 * i + 1]);
                        const double2 a0 = cV[i * ld + q0], a1 = ok1 ? cV[i * ld + q1] : cz();
                        w0 = csub(w0, cmul(h, a0));
                        w1 = csub(w1, cmul(h, a1));
                    }
                }
                w[q0] = w0;
                if (ok1) w[q1] = w1;
                nn[0] += cabs2(w0) + cabs2(w1);
            }
            team_sum<1>(T, nn, nno);
            hn = sqrt(nno[0]);
            // column j of the Hessenberg matrix and its subdiagonal entry (different entries),
            // published by one barrier
            if (static_cast<int>(threadIdx.x) <= j) {
                const int i = threadIdx.x;
                g_sh.Hm[i * MAXK + j] =
                    cadd(cadd(g_sh.Hm[i * MAXK + j], cmk(h1[2 * i], h1[2 * i + 1])), cmk(h2[2 * i], h2[2 * i + 1]));
            }
            if (j + 1 < kdim && threadIdx.x == 0) g_sh.Hm[(j + 1) * MAXK + j] = cmk(hn, 0.0);
            __syncthreads();
            prof_add(c, PROF_ORTH, ot0);
            if (!isfinite(hn)) {
                c.fail(FAIL_KRYLOV_NONFINITE);
                return napply;
            }
            m = j + 1;
            // leading Taylor term of the Saad estimate (prod of subdiagonal entries / j!);
            // the exponential is evaluated only when it is within 1e3 of the tolerance
            double lead = hn;
            for (int i = 0; i < j; ++i)
                lead *= cabsv(coeff) * fabs(tau * s_step) * cabsv(g_sh.Hm[(i + 1) * MAXK + i]) / (i + 1);
            if (lead > 1e3 * tol && j + 1 < kdim) {
                t_scale_to(T, V + (j + 1) * ld, cmk(1.0 / hn, 0.0), w, n);
                continue;
            }
            PROFILE(PROF_EXPM, err = small_exp(m, s_step));
            if (err <= tol) {
                found = true;
                break;
            }
            if (j + 1 < kdim) t_scale_to(T, V + (j + 1) * ld, cmk(1.0 / hn, 0.0), w, n);
        }
        // the full space misses the tolerance at s_step: Expokit's step-size rule on the
        // error model err(s) ~ s^m, then halving (the oracle's expv, step for step)
        for (int k = 0; !found && k < 8; ++k) {
            s_step *= step_factor(tol, err, m, 0.0625, 0.9);
            PROFILE(PROF_EXPM, err = small_exp(m, s_step));
            found = err <= tol;
        }
        for (int k = 0; !found && k < 60; ++k) {
            s_step *= 0.5;
            PROFILE(PROF_EXPM, err = small_exp(m, s_step));
            found = err <= tol;
        }
        if (!found) {
            c.fail(FAIL_KRYLOV_STEP);
            return napply;
        }
        for (int64_t q0 = q00; q0 < n; q0 += 2 * st) {
            const int64_t q1 = q0 + st;
            const bool ok1 = q1 < n;
            double2 s0 = cz(), s1 = cz();
            for (int i = 0; i < m; ++i) {
                const double2 f = g_sh.F[i * m];
                s0 = cadd(s0, cmul(f, ldc(V + i * ld + q0)));
                if (ok1) s1 = cadd(s1, cmul(f, ldc(V + i * ld + q1)));
            }
            x[q0] = cscale(s0, beta);
            if (ok1) x[q1] = cscale(s1, beta);
        }
        T.sync();
        s_done += s_step;
        s_next = s_step * step_factor(tol, err, m, 1.0, 2.0);
    }
    return napply;
}

// ---------------------------------------------------------------- small propagations
// A propagation of at most SMALL_N entries (leaf nodes, bond matrices of small bonds, every
// node of a small tree) runs on one workgroup with the Krylov basis in registers (SMALL_E
// entries of every basis vector per thread) and the vector, the intermediates and every
// operator matrix of its H-action in LDS: a Krylov iteration then costs workgroup barriers
// instead of dependent round trips to L2.  Same algorithm as expv (Arnoldi, CGS2, Saad
// estimate, Expokit sub-steps, Pade(13)); the H-action mirrors heff_node / heff_bond term by
// term, with each output entry summed over k in order.

__device__ __forceinline__ int prim_count(const DK& k) { return k.kind == 0 ? k.dim * k.dim : (k.kind == 1 ? k.dim : k.nnz); }

// Every operator matrix the H-action reads, in evaluation order: f(arena offset of the
// matrix, element count).  bond: the bond H-action with node n's operators.
template <class F>
__device__ void small_walk(const Ctx& c, int n, bool bond, F&& f) {
    const DNode& nd = c.node(n);
    const int rr = c.hr(n) * c.hr(n);
    for (int ti = 0; ti < nd.nterms; ++ti) {
        const DTerm& tm = c.I.terms[nd.term0 + ti];
        if (bond) {
            if (!tm.id_spf) f(tm.spf, rr);
            if (!tm.id_mf) f(tm.mf, rr);
            continue;
        }
        if (tm.id_spf) {
            if (!tm.id_mf) f(tm.mf, rr);
            continue;
        }
        for (int i = 0; i < tm.nsub; ++i) {
            const DSub& sb = c.I.subs[tm.sub0 + i];
            if (sb.nk == 0) continue;
            if (nd.leaf >= 0) {
                const DK& k = c.I.ks[sb.k0];
                f(k.mat, prim_count(k));
            } else {
                for (int q = 0; q < sb.nk; ++q) {
                    const DK& k = c.I.ks[sb.k0 + q];
                    const int dk = c.hr(nd.children[k.nu]);
                    f(k.mat, dk * dk);
                }
            }
        }
        if (!tm.id_mf) f(tm.mf, rr);
    }
}

// The descriptor records of the node whose small propagation is running (node, terms, their
// sub-terms and operator chains, and the CSR index arrays of its leaf operators), copied into
// LDS once per propagation so that the H-action reads them with LDS loads instead of dependent
// global round trips in every Krylov iteration.  NC_INTS holds the CSR arrays of a leaf of up
// to ~340 states with two couplings per row (a 100-state mode needs 101 + 198).
constexpr int NC_TERMS = 8, NC_SUBS = 24, NC_KS = 32, NC_INTS = 1024;
struct NodeCache {
    DNode nd;
    Shape sh;  // its live shape
    DTerm terms[NC_TERMS];
    DSub subs[NC_SUBS];
    DK ks[NC_KS];
    int64_t ints[NC_INTS];  // CSR rowptr / colind of the cached leaf operators (DK::idx remapped)
    int64_t copy_src[NC_KS], copy_dst[NC_KS], copy_cnt[NC_KS];  // CSR arrays to copy into ints
    int sub_base, k_base, ncopy;
};
__shared__ NodeCache g_nc;

// ranges of node n's records: [s0, s1) sub-terms, [k0, k1) chain operators, CSR ints needed
__device__ void node_record_ranges(const Ctx& c, int n, int& s0, int& s1, int& k0, int& k1, int& nints) {
    const DNode& nd = c.node(n);
    s0 = 1 << 30, s1 = 0, k0 = 1 << 30, k1 = 0, nints = 0;
    for (int ti = 0; ti < nd.nterms; ++ti) {
        const DTerm& tm = c.I.terms[nd.term0 + ti];
        if (tm.nsub == 0) continue;
        s0 = min(s0, tm.sub0);
        s1 = max(s1, tm.sub0 + tm.nsub);
    }
    for (int s = s0; s < s1; ++s) {
        const DSub& sb = c.I.subs[s];
        if (sb.nk == 0) continue;
        k0 = min(k0, sb.k0);
        k1 = max(k1, sb.k0 + sb.nk);
    }
    for (int k = k0; k < k1; ++k) {
        const DK& dk = c.I.ks[k];
        if (nd.leaf >= 0 && dk.kind == 2) nints += dk.dim + 1 + dk.nnz;
    }
    if (s1 <= s0) s0 = s1 = 0;
    if (k1 <= k0) k0 = k1 = 0;
}

__device__ bool node_cacheable(const Ctx& c, int n) {
    int s0, s1, k0, k1, nints;
    node_record_ranges(c, n, s0, s1, k0, k1, nints);
    return c.node(n).nterms <= NC_TERMS && s1 - s0 <= NC_SUBS && k1 - k0 <= NC_KS && nints <= NC_INTS;
}

// fill g_nc for node n: thread 0 copies the records and lists the CSR arrays, then the
// whole workgroup copies those; the caller checked node_cacheable
__device__ void cache_node(const Ctx& c, int n) {
    if (threadIdx.x == 0) {
        int s0, s1, k0, k1, nints;
        node_record_ranges(c, n, s0, s1, k0, k1, nints);
        const DNode& nd = c.node(n);
        g_nc.nd = nd;
        g_nc.sh = c.shape(n);
        for (int ti = 0; ti < nd.nterms; ++ti) g_nc.terms[ti] = c.I.terms[nd.term0 + ti];
        for (int s = s0; s < s1; ++s) g_nc.subs[s - s0] = c.I.subs[s];
        int64_t at = 0;
        int nc = 0;
        for (int k = k0; k < k1; ++k) {
            DK dk = c.I.ks[k];
            if (nd.leaf >= 0 && dk.kind == 2) {
                const int64_t cnt = dk.dim + 1 + dk.nnz;
                g_nc.copy_src[nc] = dk.idx;
                g_nc.copy_dst[nc] = at;
                g_nc.copy_cnt[nc] = cnt;
                ++nc;
                dk.idx = at;
                at += cnt;
            }
            g_nc.ks[k - k0] = dk;
        }
        g_nc.ncopy = nc;
        g_nc.sub_base = s0;
        g_nc.k_base = k0;
    }
    __syncthreads();
    for (int q = 0; q < g_nc.ncopy; ++q) {
        const int64_t* src = c.I.ints + g_nc.copy_src[q];
        int64_t* dst = g_nc.ints + g_nc.copy_dst[q];
        const int64_t cnt = g_nc.copy_cnt[q];
        for (int64_t e = static_cast<int64_t>(threadIdx.x); e < cnt; e += NT) dst[e] = src[e];
    }
    __syncthreads();
}

__device__ bool small_eligible(const Ctx& c, int n, bool bond, int64_t size) {
    if (c.small == nullptr || size > c.small_max || size > SMALL_N) return false;
    if (!node_cacheable(c, n)) return false;
    int total = 0;
    small_walk(c, n, bond, [&](int64_t, int cnt) { total += cnt; });
    return total <= SMALL_OPS;
}

// ---- LDS elementary operations (all threads of the workgroup; each ends with a barrier)
__device__ __forceinline__ double2 s_out(double2 acc, double2 alpha, double2 old, bool accumulate) {
    const double2 v = cmul(alpha, acc);
    return accumulate ? cadd(old, v) : v;
}
__device__ __forceinline__ void s_axpby(double2* y, double2 a, const double2* x, int n, bool accumulate) {
    for (int e = threadIdx.x; e < n; e += NT) y[e] = accumulate ? cadd(y[e], cmul(a, x[e])) : cmul(a, x[e]);
    __syncthreads();
}
__device__ __forceinline__ void s_zero(double2* y, int n) {
    for (int e = threadIdx.x; e < n; e += NT) y[e] = cz();
    __syncthreads();
}
// dst(l,i,r) (=|+=) alpha sum_j op(i,j) src(l,j,r)
// sum_k a(k) b(k) for k = 0 .. K-1 in that order, the operands of four k loaded ahead
template <class FA, class FB>
__device__ __forceinline__ double2 s_dot(FA a, FB b, int K) {
    double2 acc = cz();
    int k = 0;
    for (; k + 4 <= K; k += 4) {
        double2 x[4], y[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            x[u] = a(k + u);
            y[u] = b(k + u);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) acc = cadd(acc, cmul(x[u], y[u]));
    }
    for (; k < K; ++k) acc = cadd(acc, cmul(a(k), b(k)));
    return acc;
}
__device__ __forceinline__ void s_mode_contract(const double2* op, int L, int d, int R, const double2* src, double2* dst, double2 alpha,
                                bool accumulate) {
    const int total = L * d * R;
    for (int e = threadIdx.x; e < total; e += NT) {
        const int l = e / (d * R), rem = e - l * d * R, i = rem / R, r = rem - i * R;
        const double2 acc = s_dot([&](int j) { return op[i * d + j]; },
                                  [&](int j) { return src[(l * d + j) * R + r]; }, d);
        dst[e] = s_out(acc, alpha, dst[e], accumulate);
    }
    __syncthreads();
}
// dst (=|+=) alpha X M^T, X: rows x r
__device__ __forceinline__ void s_mul_t(const double2* X, int rows, int r, const double2* M, double2* dst, double2 alpha,
                        bool accumulate) {
    const int total = rows * r;
    for (int e = threadIdx.x; e < total; e += NT) {
        const int x = e / r, i = e - x * r;
        const double2 acc = s_dot([&](int k) { return X[x * r + k]; }, [&](int k) { return M[i * r + k]; }, r);
        dst[e] = s_out(acc, alpha, dst[e], accumulate);
    }
    __syncthreads();
}
// y (=|+=) alpha op x for a primitive leaf operator (op values in LDS, CSR indices in the table)
__device__ __forceinline__ void s_prim(const DK& k, const double2* op, const int64_t* ints, const double2* x, double2* y, int ncols,
                       double2 alpha, bool accumulate) {
    const int d = k.dim, total = d * ncols;
    for (int e = threadIdx.x; e < total; e += NT) {
        const int i = e / ncols, cc = e - i * ncols;
        double2 acc = cz();
        if (k.kind == 0) {
            acc = s_dot([&](int j) { return op[i * d + j]; }, [&](int j) { return x[j * ncols + cc]; }, d);
        } else if (k.kind == 1) {
            acc = cmul(op[i], x[e]);
        } else {
            const int64_t* rowptr = ints + k.idx;
            const int64_t* colind = rowptr + d + 1;
            for (int64_t q = rowptr[i]; q < rowptr[i + 1]; ++q) acc = cadd(acc, cmul(op[q], x[colind[q] * ncols + cc]));
        }
        y[e] = s_out(acc, alpha, y[e], accumulate);
    }
    __syncthreads();
}

// heff_node on the LDS vector g_small.v into g_small.w
__device__ void small_heff_node() {
    SmallLds& L = g_small;
    const double2* v = L.v;
    double2* res = L.w;
    const DNode& nd = g_nc.nd;
    const Shape& sh = g_nc.sh;
    const int r = sh.hrank, D = sh.dimen, size = sh.size, rr = r * r;
    int cur = 0;
    bool first = true;
    for (int ti = 0; ti < nd.nterms; ++ti) {
        const DTerm& tm = g_nc.terms[ti];
        if (tm.id_spf) {
            const double2 coeff = cmk(tm.ire, tm.iim);
            if (tm.id_mf) {
                s_axpby(res, coeff, v, size, !first);
            } else {
                s_mul_t(v, D, r, L.ops + cur, res, coeff, !first);
                cur += rr;
            }
            first = false;
            continue;
        }
        for (int i = 0; i < tm.nsub; ++i) {
            const DSub& sb = g_nc.subs[tm.sub0 + i - g_nc.sub_base];
            const double2 ci = (i == 0) ? (tm.nsub > 1 ? cmk(sb.are, sb.aim) : cmk(tm.cre, tm.cim)) : cmk(sb.are, sb.aim);
            double2* dst = tm.id_mf ? res : L.t3;
            const bool acc = tm.id_mf ? !first : i > 0;
            if (sb.nk == 0) {
                s_axpby(dst, ci, v, size, acc);
            } else if (nd.leaf >= 0) {
                const DK& k = g_nc.ks[sb.k0 - g_nc.k_base];
                s_prim(k, L.ops + cur, g_nc.ints, v, dst, r, ci, acc);
                cur += prim_count(k);
            } else {
                const double2* src = v;
                for (int q = 0; q < sb.nk; ++q) {
                    const bool last = q == sb.nk - 1;
                    double2* out = last ? dst : (q % 2 == 0 ? L.t1 : L.t2);
                    const DK& k = g_nc.ks[sb.k0 + q - g_nc.k_base];
                    int64_t LL, RR;
                    mode_view(sh, k.nu, LL, RR);
                    const int dk = sh.dims[k.nu];
                    s_mode_contract(L.ops + cur, static_cast<int>(LL), dk, static_cast<int>(RR), src, out,
                                    last ? ci : cmk(1.0, 0.0), last && acc);
                    cur += dk * dk;
                    src = out;
                }
            }
            if (tm.id_mf) first = false;
        }
        if (!tm.id_mf) {
            s_mul_t(L.t3, D, r, L.ops + cur, res, cmk(1.0, 0.0), !first);
            cur += rr;
            first = false;
        }
    }
    if (first) s_zero(res, size);
}

// heff_bond on LDS data
__device__ void small_heff_bond() {
    SmallLds& L = g_small;
    const double2* v = L.v;
    double2* res = L.w;
    const DNode& nd = g_nc.nd;
    const int r = g_nc.sh.hrank, sz = r * r;
    int cur = 0;
    s_zero(res, sz);
    for (int ti = 0; ti < nd.nterms; ++ti) {
        const DTerm& tm = g_nc.terms[ti];
        const double2* x = v;
        if (!tm.id_spf) {
            s_mode_contract(L.ops + cur, 1, r, r, v, L.t1, cmk(1.0, 0.0), false);
            cur += sz;
            x = L.t1;
        }
        const double2 coeff = cmk(tm.cre, tm.cim);
        if (!tm.id_mf) {
            s_mul_t(x, r, r, L.ops + cur, res, coeff, true);
            cur += sz;
        } else {
            s_axpby(res, coeff, x, sz, true);
        }
    }
}

// x <- exp(coeff tau H) x for a vector of n <= SMALL_N entries on one workgroup (team of one)
template <int KM>
__device__ int64_t expv_small(Ctx& c, int n_op, bool bond, double2* x, int n, double2 coeff, double tau, int64_t& sumj) {
    if (tau == 0.0 || n == 0) return 0;
    SmallLds& L = g_small;
    const int kdim = c.I.kdim;
    const double tol = c.I.tol;
    cache_node(c, n_op);
    // stage the operators (written by other team members: coherent loads)
    int cur = 0;
    small_walk(c, n_op, bond, [&](int64_t off, int cnt) {
        for (int e = threadIdx.x; e < cnt; e += NT) L.ops[cur + e] = ldc(c.base + off + e);
        cur += cnt;
    });
    double2 xe[SMALL_E];
#pragma unroll
    for (int e = 0; e < SMALL_E; ++e) {
        const int q = threadIdx.x + e * NT;
        xe[e] = q < n ? ldc(x + q) : cz();
    }
    __syncthreads();
    double2 vb[KM + 1][SMALL_E];
    int64_t napply = 0;
    double s_done = 0.0, s_next = 1.0;
    int guard = 0;
    bool ok = true;
    while (s_done < 1.0 && guard++ < 4096) {
        double s_step = fmin(s_next, 1.0 - s_done);
        double b2[1] = {0.0}, bo[1];
#pragma unroll
        for (int e = 0; e < SMALL_E; ++e) b2[0] += cabs2(xe[e]);
        block_sum<1>(b2, bo);
        const double beta = sqrt(bo[0]);
        if (!isfinite(beta)) {
            c.fail(FAIL_INPUT_NONFINITE);
            ok = false;
            break;
        }
        if (beta == 0.0) break;
#pragma unroll
        for (int e = 0; e < SMALL_E; ++e) vb[0][e] = cscale(xe[e], 1.0 / beta);
        for (int e = threadIdx.x; e < (MAXK + 1) * MAXK; e += NT) g_sh.Hm[e] = cz();
        __syncthreads();
        int m = 0;
        double hn = 0.0, err = 0.0;
        bool found = false;
        auto small_exp = [&](int mm, double sf) -> double {
            const double2 f = cscale(coeff, tau * sf);
            for (int e = threadIdx.x; e < mm * mm; e += NT) {
                const int i = e / mm, j = e % mm;
                g_sh.S[e] = cmul(f, g_sh.Hm[i * MAXK + j]);
            }
            __syncthreads();
            block_expm(g_sh.S, g_sh.F, mm);
            return hn * cabsv(g_sh.F[(mm - 1) * mm]);
        };
        for (int j = 0; j < kdim; ++j) {
            // V_j into LDS for the H-action (static register indexing: unrolled select)
#pragma unroll
            for (int i = 0; i <= KM; ++i)
                if (i == j)
#pragma unroll
                    for (int e = 0; e < SMALL_E; ++e) {
                        const int q = threadIdx.x + e * NT;
                        if (q < n) L.v[q] = vb[i][e];
                    }
            __syncthreads();
            if (bond) {
                PROFILE(PROF_HBOND, small_heff_bond());
            } else {
                PROFILE(PROF_HEFF, small_heff_node());
            }
            ++napply;
            sumj += j + 1;
            const uint64_t ot0 = prof_now(c);
            double2 wv[SMALL_E];
#pragma unroll
            for (int e = 0; e < SMALL_E; ++e) {
                const int q = threadIdx.x + e * NT;
                wv[e] = q < n ? L.w[q] : cz();
            }
            // CGS2 in registers, the coefficients through workgroup sums
#pragma unroll
            for (int pass = 0; pass < 2; ++pass) {
                double p[2 * KM];
#pragma unroll
                for (int i = 0; i < 2 * KM; ++i) p[i] = 0.0;
#pragma unroll
                for (int i = 0; i < KM; ++i)
                    if (i <= j)
#pragma unroll
                        for (int e = 0; e < SMALL_E; ++e) {
                            const double2 pr = cmulc(vb[i][e], wv[e]);
                            p[2 * i] += pr.x;
                            p[2 * i + 1] += pr.y;
                        }
                block_sum_lds<2 * KM>(p, g_sh.hb[pass], 2 * (j + 1));
                const double* h = g_sh.hb[pass];
#pragma unroll
                for (int i = 0; i < KM; ++i)
                    if (i <= j)
#pragma unroll
                        for (int e = 0; e < SMALL_E; ++e) wv[e] = csub(wv[e], cmul(cmk(h[2 * i], h[2 * i + 1]), vb[i][e]));
            }
            double nn[1] = {0.0}, nno[1];
#pragma unroll
            for (int e = 0; e < SMALL_E; ++e) nn[0] += cabs2(wv[e]);
            block_sum<1>(nn, nno);
            hn = sqrt(nno[0]);
            // column j of the Hessenberg matrix and its subdiagonal entry, one barrier
            if (static_cast<int>(threadIdx.x) <= j) {
                const int i = threadIdx.x;
                g_sh.Hm[i * MAXK + j] = cadd(cadd(g_sh.Hm[i * MAXK + j], cmk(g_sh.hb[0][2 * i], g_sh.hb[0][2 * i + 1])),
                                           cmk(g_sh.hb[1][2 * i], g_sh.hb[1][2 * i + 1]));
            }
            if (j + 1 < kdim && threadIdx.x == 0) g_sh.Hm[(j + 1) * MAXK + j] = cmk(hn, 0.0);
            __syncthreads();
            prof_add(c, PROF_ORTH, ot0);
            if (!isfinite(hn)) {
                c.fail(FAIL_KRYLOV_NONFINITE);
                ok = false;
                break;
            }
            m = j + 1;
            double lead = hn;
            for (int i = 0; i < j; ++i)
                lead *= cabsv(coeff) * fabs(tau * s_step) * cabsv(g_sh.Hm[(i + 1) * MAXK + i]) / (i + 1);
            bool next_vec = j + 1 < kdim;
            if (!(lead > 1e3 * tol && j + 1 < kdim)) {
                PROFILE(PROF_EXPM, err = small_exp(m, s_step));
                if (err <= tol) {
                    found = true;
                    next_vec = false;
                }
            }
            if (next_vec) {
#pragma unroll
                for (int i = 1; i <= KM; ++i)
                    if (i == j + 1)
#pragma unroll
                        for (int e = 0; e < SMALL_E; ++e) vb[i][e] = cscale(wv[e], 1.0 / hn);
            }
            if (found) break;
        }
        if (!ok) break;
        for (int k = 0; !found && k < 8; ++k) {
            s_step *= step_factor(tol, err, m, 0.0625, 0.9);
            PROFILE(PROF_EXPM, err = small_exp(m, s_step));
            found = err <= tol;
        }
        for (int k = 0; !found && k < 60; ++k) {
            s_step *= 0.5;
            PROFILE(PROF_EXPM, err = small_exp(m, s_step));
            found = err <= tol;
        }
        if (!found) {
            c.fail(FAIL_KRYLOV_STEP);
            ok = false;
            break;
        }
#pragma unroll
        for (int e = 0; e < SMALL_E; ++e) {
            double2 sacc = cz();
#pragma unroll
            for (int i = 0; i < KM; ++i)
                if (i < m) sacc = cadd(sacc, cmul(g_sh.F[i * m], vb[i][e]));
            xe[e] = cscale(sacc, beta);
        }
        __syncthreads();
        s_done += s_step;
        s_next = s_step * step_factor(tol, err, m, 1.0, 2.0);
    }
#pragma unroll
    for (int e = 0; e < SMALL_E; ++e) {
        const int q = threadIdx.x + e * NT;
        if (q < n) x[q] = xe[e];
    }
    __syncthreads();
    return napply;
}

__device__ void count_small(const Ctx& c, uint64_t t0) {
    prof_add(c, PROF_SMALL, t0);
    if (c.I.prof && c.writer()) c.I.prof[PROF_NSMALL] += 1;
}

__device__ int64_t run_small(Ctx& c, int n_op, bool bond, double2* x, int n, double2 coeff, double tau, int64_t& sumj) {
    if (c.I.kdim <= 12) return expv_small<12>(c, n_op, bond, x, n, coeff, tau, sumj);
    return expv_small<MAXK>(c, n_op, bond, x, n, coeff, tau, sumj);
}

#include "expand.hpp"

// ---------------------------------------------------------------- sweeps
struct SweepState {
    double2 coeff;
    double sub_dt;
    double exp_dt;   // the engine's dt() seen by the spawning test of the expansion
    int64_t nh;
};

__device__ void node_update(Ctx& c, SweepState& st, int n) {
    const DNode& nd = c.node(n);
    const int size = c.shape(n).size;
    c.where = n;
    c.bond = 0;
    int64_t na = 0, sumj = 0, hticks = 0;
    const uint64_t pt0 = prof_now(c);
    if (small_eligible(c, n, false, size)) {
        // one workgroup (the leader, other members wait at the barrier that follows)
        on_team(c, 0, [&](Ctx& cc) { na = run_small(cc, n, false, cc.base + nd.A, size, st.coeff, st.sub_dt / 2.0, sumj); });
        count_small(c, pt0);
    } else {
        on_team(c, size, [&](Ctx& cc) {
            na = expv(
                cc,
                [&](const double2* v, double2* res) {
                    const uint64_t h0 = prof_now(c);
                    heff_node(cc, n, v, res);
                    const uint64_t dh = prof_now(c) - h0;
                    prof_add(c, PROF_HEFF, h0);
                    hticks += static_cast<int64_t>(dh);
                },
                cc.base + nd.A, size, st.coeff, st.sub_dt / 2.0, sumj);
        });
    }
    prof_add(c, PROF_EXPV_NODE, pt0);
    st.nh += na;
    if (c.writer()) {
        c.I.node_stats[NODE_STATS * n] += na;
        c.I.node_stats[NODE_STATS * n + 1] += sumj;
        if (c.I.prof) {
            c.I.node_stats[NODE_STATS * n + 4] += static_cast<int64_t>(prof_now(c) - pt0);
            c.I.node_stats[NODE_STATS * n + 6] += hticks;
        }
    }
}

__device__ void bond_update(Ctx& c, SweepState& st, int n, int ops) {
    const int r = c.hr(ops);
    c.where = ops;
    c.bond = 1;
    int64_t na = 0, sumj = 0, hticks = 0;
    const uint64_t pt0 = prof_now(c);
    if (small_eligible(c, ops, true, static_cast<int64_t>(r) * r)) {
        on_team(c, 0, [&](Ctx& cc) {
            na = run_small(cc, ops, true, cc.base + cc.node(n).R, r * r, st.coeff, -st.sub_dt / 2.0, sumj);
        });
        count_small(c, pt0);
    } else {
        // a bond propagation spreads over the team from half the element count of other
        // operations: its H-action is r x r x r products, more work per entry
        on_team(c, 2 * static_cast<int64_t>(r) * r, [&](Ctx& cc) {
            na = expv(
                cc,
                [&](const double2* v, double2* res) {
                    const uint64_t h0 = prof_now(c);
                    heff_bond(cc, ops, v, res);
                    const uint64_t dh = prof_now(c) - h0;
                    prof_add(c, PROF_HBOND, h0);
                    hticks += static_cast<int64_t>(dh);
                },
                cc.base + cc.node(n).R, static_cast<int64_t>(r) * r, st.coeff, -st.sub_dt / 2.0, sumj);
        });
    }
    prof_add(c, PROF_EXPV_BOND, pt0);
    if (c.writer()) {
        c.I.node_stats[NODE_STATS * ops + 2] += na;
        c.I.node_stats[NODE_STATS * ops + 3] += sumj;
        if (c.I.prof) {
            c.I.node_stats[NODE_STATS * ops + 5] += static_cast<int64_t>(prof_now(c) - pt0);
            c.I.node_stats[NODE_STATS * ops + 7] += hticks;
        }
    }
}

// With expand, every root-to-leaf factorisation is followed by the downward expansion of that
// bond and every leaf-to-root factorisation by the upward one (tdvp_algorithm_engine.hpp:373,
// 479).  Returns whether a bond grew (the same on every member).
__device__ bool forward(Ctx& c, SweepState& st, bool evolve, bool expand, unsigned char* visits) {
    bool grew = false;
    for (int e = threadIdx.x; e < c.I.nnodes; e += NT) visits[e] = 0;
    __syncthreads();
    for (int s = 0; s < c.I.tour_len; ++s) {
        if (c.failed()) return grew;
        const int id = c.I.tour[s];
        const DNode& nd = c.node(id);
        const int tv = visits[id];
        __syncthreads();
        if (threadIdx.x == 0) visits[id] = static_cast<unsigned char>(tv + 1);
        __syncthreads();
        if (tv != nd.nmodes) {
            if (tv == 0 && nd.parent >= 0) PROFILE(PROF_GAUGE, apply_from_parent(c, id));
            if (nd.leaf < 0) {
                PROFILE(PROF_QR, r2l(c, id, tv));
                if (expand && expand_on_leader(c, nd.children[tv],
                                               [&](Ctx& cc) { return expand_down_body(cc, id, tv, st.exp_dt); }))
                    grew = true;
                PROFILE(PROF_MF, update_mf(c, nd.children[tv]));
            }
        } else {
            if (evolve) node_update(c, st, id);
            if (nd.parent >= 0) {
                PROFILE(PROF_QR, l2r(c, id));
                PROFILE(PROF_SPF, update_spf(c, id));
                if (evolve) bond_update(c, st, id, id);
                PROFILE(PROF_GAUGE, apply_to_parent(c, id));
            }
        }
    }
    return grew;
}

__device__ bool backward(Ctx& c, SweepState& st, bool evolve, bool expand, unsigned char* visits) {
    bool grew = false;
    for (int e = threadIdx.x; e < c.I.nnodes; e += NT) visits[e] = 0;
    __syncthreads();
    for (int s = c.I.tour_len - 1; s >= 0; --s) {
        if (c.failed()) return grew;
        const int id = c.I.tour[s];
        const DNode& nd = c.node(id);
        const int tv = visits[id];
        __syncthreads();
        if (threadIdx.x == 0) visits[id] = static_cast<unsigned char>(tv + 1);
        __syncthreads();
        if (tv != nd.nmodes) {
            const int mode = nd.nmodes - (tv + 1);
            if (tv == 0) {
                if (nd.parent >= 0) {
                    if (evolve) bond_update(c, st, nd.parent, id);
                    PROFILE(PROF_GAUGE, apply_from_parent(c, id));
                }
                if (evolve) node_update(c, st, id);
            }
            if (nd.leaf < 0) {
                PROFILE(PROF_QR, r2l(c, id, mode));
                PROFILE(PROF_MF, update_mf(c, nd.children[mode]));
            }
        } else if (nd.parent >= 0) {
            PROFILE(PROF_QR, l2r(c, id));
            if (expand && expand_on_leader(c, id, [&](Ctx& cc) { return expand_up_body(cc, id, st.exp_dt); }))
                grew = true;
            PROFILE(PROF_GAUGE, apply_to_parent(c, id));
            PROFILE(PROF_SPF, update_spf(c, id));
        }
    }
    return grew;
}

// <psi|O|psi> for O on the leaf node `leaf` (path contraction to the root, the other
// subtrees orthonormal as after every sweep) and <psi|psi> = |A_root|^2; uses W1..W4
__device__ void path_expect(Ctx& c, int leaf, int64_t op_off, double2& val, double& nrm) {
    Team& T = *c.T;
    const DNode& ln = c.node(leaf);
    const Shape ls = c.shape(leaf);
    double2 *HA = c.W(1), *M = c.W(2), *Tm = c.W(3), *M2 = c.W(4);
    DK k;
    k.nu = ln.leaf;
    k.kind = 0;
    k.dim = ls.dimen;
    k.nnz = 0;
    k.mat = op_off;
    k.idx = 0;
    t_prim_apply(T, k, c.base, c.I.ints, c.base + ln.A, HA, ls.hrank);
    t_adjoint_mul(T, c.base + ln.A, HA, ls.dimen, ls.hrank, M, cmk(1.0, 0.0), false);
    int n = leaf;
    while (c.node(n).parent >= 0) {
        const DNode& nd = c.node(n);
        const DNode& pn = c.node(nd.parent);
        const Shape ps = c.shape(nd.parent);
        int64_t L, R;
        mode_view(ps, nd.child_id, L, R);
        t_mode_contract(T, M, L, ps.dims[nd.child_id], R, c.base + pn.A, Tm, cmk(1.0, 0.0), false);
        t_adjoint_mul(T, c.base + pn.A, Tm, ps.dimen, ps.hrank, M2, cmk(1.0, 0.0), false);
        double2* tmp = M;
        M = M2;
        M2 = tmp;
        n = nd.parent;
    }
    val = ldc(M);
    const DNode& root = c.node(0);
    nrm = team_norm2(T, c.base + root.A, c.shape(0).size);
}

// every observable's <psi|O_j|psi> and <psi|psi> into the step's trajectory record (the
// driver's mel(mops[j], 0, A) after each step, src/spin_boson.cpp:86-90)
__device__ void step_done(Ctx& c, int64_t step_index) {
    if (c.I.nobs > 0 && c.I.traj != nullptr) {
        const int no = c.I.nobs;
        double* rec = c.I.traj + step_index * (2 * no + 1);
        double nrm = 0.0;
        for (int j = 0; j < no; ++j) {
            double2 v;
            path_expect(c, c.I.obs_leaves[j], c.I.obs_ops[j], v, nrm);
            if (c.writer()) {
                rec[2 * j] = v.x;
                rec[2 * j + 1] = v.y;
            }
        }
        if (c.writer()) rec[2 * no] = nrm;
    }
    if (c.writer()) c.I.counters[1] += 1;
    __syncthreads();
}

__device__ __forceinline__ int xcd_id() {
    // XCC_ID hardware register, bits [3:0]
    return __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20);
}

// plan == null: workgroup b runs instance b alone.  Otherwise every workgroup takes the
// next slot of the XCD it runs on (xcd_fill: one counter per 128-byte line) and joins
// the team that slot names; words[i] are the team words of instance i.
__global__ __launch_bounds__(NT) void tdvp_run_kernel(const DInst* insts, RunArgs a_in, const TeamPlan* plan,
                                                      TeamWords* words, unsigned* xcd_fill) {
    // visit count of every node in the current sweep (at most 1 + MAXC)
    __shared__ unsigned char visits[MAX_NODES];
    __shared__ int slot_info[3];
    if (plan == nullptr) {
        if (threadIdx.x == 0) {
            slot_info[0] = blockIdx.x;
            slot_info[1] = 0;
            slot_info[2] = 1;
        }
    } else if (threadIdx.x == 0) {
        const int xcc = xcd_id();
        int inst = -1, rank = 0, size = 1;
        if (xcc < plan->nxcd) {
            const unsigned s =
                __hip_atomic_fetch_add(&xcd_fill[32 * xcc], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (s < static_cast<unsigned>(plan->count[xcc])) {
                const TeamSlot& t = plan->slot[xcc][s];
                inst = t.inst;
                rank = t.rank;
                size = t.size;
            }
        }
        slot_info[0] = inst;
        slot_info[1] = rank;
        slot_info[2] = size;
    }
    if (threadIdx.x == 0) g_sh.fail = 0;
    __syncthreads();
    const int inst = slot_info[0], rank = slot_info[1], size = slot_info[2];
    if (inst < 0) return;
    const DInst& I = insts[inst];
    RunArgs a = a_in;
    if (I.own != nullptr) {
        a = *I.own;
        a.team_min = a_in.team_min;
        a.small_max = a_in.small_max;
    }
    Team T;
    T.w = size > 1 ? &words[inst] : nullptr;
    T.rank = rank;
    T.size = size;
    T.gen = 0;
    T.pbuf = 0;
    T.part = reinterpret_cast<double*>(I.base + I.team_part);
    T.tiles = I.base + I.team_tiles;
    // roll call: nothing is touched unless the whole team has assembled, and then by every
    // member (a team that did not form leaves its instance untouched)
    if (!T.roll_call()) {
        if (threadIdx.x == 0)
            __hip_atomic_store(I.status, static_cast<int>(FAIL_TEAM_FORM), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    Ctx c(I, &T, a.team_min);
    c.small = &g_small;
    c.small_max = a.small_max;
    for (int e = threadIdx.x; e < I.nnodes; e += NT) g_hr[e] = static_cast<unsigned char>(I.hr[e]);
    __syncthreads();
    const uint64_t prof_t0 = prof_now(c);
    SweepState st;
    st.coeff = a.coeff;
    st.nh = 0;
    int64_t step_index = a.step0;
    // tdvp_algorithm_base::prepare_evolution (:257-321): the bottom-up single-particle pass,
    // then sweep pairs without propagation
    if (a.prepare_spf)
        for (int n = I.nnodes - 1; n >= 0; --n)
            if (I.nodes[n].parent >= 0) update_spf(c, n);
    const bool expand = I.expand != 0;
    bool grew = false;
    for (int pi = 0; pi < a.prepare && !c.failed(); ++pi) {
        // the engine's dt is the full step during prepare_evolution
        st.sub_dt = a.nladder > 0 ? a.ladder_dt[0] : a.dt;
        st.exp_dt = a.dt;
        grew = forward(c, st, false, expand, visits);
        grew = backward(c, st, false, expand, visits) || grew;
        if (expand && !grew) break;
    }
    if (a.prepare > 0 && expand && c.writer()) I.exp[EXP_GREW] = grew ? 1 : 0;
    for (int l = 0; l < a.nladder && !c.failed(); ++l) {
        st.sub_dt = a.ladder_dt[l];
        st.exp_dt = a.ladder_dt[l];
        forward(c, st, true, expand, visits);
        backward(c, st, true, expand, visits);
    }
    if (a.nladder > 0 && !c.failed()) step_done(c, step_index++);
    for (int s = 0; s < a.nregular && !c.failed(); ++s) {
        st.sub_dt = a.dt;
        st.exp_dt = a.dt;
        forward(c, st, true, expand, visits);
        backward(c, st, true, expand, visits);
        step_done(c, step_index++);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        // the first member to report a failure records it
        int expected = 0;
        if (g_sh.fail != 0 && __hip_atomic_compare_exchange_strong(I.status, &expected, g_sh.fail, __ATOMIC_RELAXED,
                                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            I.status[1] = static_cast<int>(step_index);
        if (rank == 0) {
            I.counters[0] += st.nh;
            if (I.prof) I.prof[PROF_TOTAL] += static_cast<int64_t>(wall_clock64() - prof_t0);
        }
    }
}

// <psi|O|psi> on one instance (O stored at op_off); out: re, im, norm2
__global__ __launch_bounds__(NT) void expect_kernel(const DInst* inst, int leaf, int64_t op_off, double* out) {
    if (threadIdx.x == 0) g_sh.fail = 0;
    __syncthreads();
    Team T;
    T.w = nullptr;
    T.rank = 0;
    T.size = 1;
    T.gen = 0;
    T.pbuf = 0;
    T.part = nullptr;
    T.tiles = nullptr;
    Ctx c(*inst, &T, 0);
    for (int e = threadIdx.x; e < inst->nnodes; e += NT) g_hr[e] = static_cast<unsigned char>(inst->hr[e]);
    __syncthreads();
    double2 v;
    double nrm;
    path_expect(c, leaf, op_off, v, nrm);
    if (threadIdx.x == 0) {
        out[0] = v.x;
        out[1] = v.y;
        out[2] = nrm;
    }
}

}  // namespace ttnd

// ---------------------------------------------------------------- host launchers
// Team launches start `grid` workgroups (one per CU); plan == null runs n instances one
// workgroup each.
namespace ttnd {

// Leaf transition densities of pairs of instances of one tree, one workgroup per pair (the
// costate gradient's matrix elements):  rho_m(j, l) = <bra|(|j><l|)_m|ket> for every leaf
// and rhoS_m, the same with S (the ket's leaf node `ins` operator, d x d) inserted, by full
// contraction of the two trees: bottom-up transfer matrices T = Abra^H (children) Aket (TS
// along the insertion leaf's path), then top-down environments E (ES: S outside the
// subtree), then rho = Abra^* E Aket^T on every leaf.  No orthonormality is assumed.
// work: per pair 4 * nnodes * rr + 3 * maxsize entries (rr >= largest hrank^2); out: per
// pair, rho then rhoS (d x d each) at out_off[node] of every leaf node.
__global__ __launch_bounds__(NT) void densities_kernel(const DInst* kets, const DInst* bras, int ins, const double2* S,
                                                       double2* work, int64_t work_stride, int64_t rr,
                                                       const int64_t* out_off, double2* out, int64_t out_stride) {
    __shared__ unsigned char onpath[MAX_NODES], esdef[MAX_NODES];
    const int pair = blockIdx.x;
    const DInst& K = kets[pair];
    const DInst& B = bras[pair];
    if (threadIdx.x == 0) g_sh.fail = 0;
    Team T;
    T.w = nullptr;
    T.rank = 0;
    T.size = 1;
    T.gen = 0;
    T.pbuf = 0;
    T.part = reinterpret_cast<double*>(K.base + K.team_part);
    T.tiles = K.base + K.team_tiles;
    Ctx c(K, &T, 0);
    const int n = K.nnodes;
    for (int e = threadIdx.x; e < n; e += NT) {
        g_hr[e] = static_cast<unsigned char>(K.hr[e]);
        onpath[e] = 0;
        esdef[e] = 0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = ins; q >= 0; q = K.nodes[q].parent) onpath[q] = 1;
        // ES exists below the insertion path (pre-order: parents first)
        for (int i = 1; i < n; ++i) {
            const int p = K.nodes[i].parent;
            esdef[i] = !onpath[i] && (onpath[p] || esdef[p]);
        }
    }
    __syncthreads();
    const double2 one = cmk(1.0, 0.0);
    double2* wk = work + static_cast<int64_t>(pair) * work_stride;
    double2* y0 = wk + 4 * static_cast<int64_t>(n) * rr;
    double2* y1 = y0 + K.maxsize;
    auto TM = [&](int i) { return wk + static_cast<int64_t>(i) * rr; };
    auto TSM = [&](int i) { return wk + static_cast<int64_t>(n + i) * rr; };
    auto EM = [&](int i) { return wk + static_cast<int64_t>(2 * n + i) * rr; };
    auto ESM = [&](int i) { return wk + static_cast<int64_t>(3 * n + i) * rr; };
    // bottom-up transfer matrices
    for (int i = n - 1; i >= 0; --i) {
        const DNode& nd = K.nodes[i];
        const Shape sh = c.shape(i);
        const double2* Ak = K.base + nd.A;
        const double2* Ab = B.base + nd.A;
        if (nd.leaf >= 0) {
            t_adjoint_mul(T, Ab, Ak, sh.dimen, sh.hrank, TM(i), one, false);
            if (onpath[i]) {
                const int d = static_cast<int>(sh.dimen), r = sh.hrank;
                for (int64_t e = threadIdx.x; e < sh.size; e += NT) {
                    const int a = static_cast<int>(e / r), col = static_cast<int>(e % r);
                    double2 acc = cz();
                    for (int b = 0; b < d; ++b) acc = cadd(acc, cmul(S[a * d + b], Ak[b * r + col]));
                    y0[e] = acc;
                }
                T.sync();
                t_adjoint_mul(T, Ab, y0, sh.dimen, sh.hrank, TSM(i), one, false);
            }
            continue;
        }
        for (int pass = 0; pass < (onpath[i] ? 2 : 1); ++pass) {
            const double2* src = Ak;
            for (int k = 0; k < nd.nmodes; ++k) {
                const int ch = nd.children[k];
                int64_t L, R;
                mode_view(sh, k, L, R);
                double2* dst = (k % 2 == 0) ? y0 : y1;
                t_mode_contract(T, (pass == 1 && onpath[ch]) ? TSM(ch) : TM(ch), L, sh.dims[k], R, src, dst, one,
                                false);
                src = dst;
            }
            t_adjoint_mul(T, Ab, src, sh.dimen, sh.hrank, pass == 1 ? TSM(i) : TM(i), one, false);
        }
    }
    // top-down environments; the root's is the identity on its (dummy) upper index
    {
        const int r0 = c.hr(0);
        for (int e = threadIdx.x; e < r0 * r0; e += NT) EM(0)[e] = (e / r0 == e % r0) ? one : cz();
        T.sync();
    }
    for (int p = 0; p < n; ++p) {
        const DNode& nd = K.nodes[p];
        if (nd.leaf >= 0) continue;
        const Shape sh = c.shape(p);
        const double2* Ak = K.base + nd.A;
        const double2* Ab = B.base + nd.A;
        for (int k = 0; k < nd.nmodes; ++k) {
            const int ch = nd.children[k];
            for (int which = 0; which < 2; ++which) {
                if (which == 1 && !esdef[ch]) continue;
                const double2* Ep = (which == 1 && !onpath[p]) ? ESM(p) : EM(p);
                const double2* src = Ak;
                int flip = 0;
                for (int s2 = 0; s2 < nd.nmodes; ++s2) {
                    if (s2 == k) continue;
                    const int sc = nd.children[s2];
                    int64_t L, R;
                    mode_view(sh, s2, L, R);
                    double2* dst = flip ? y1 : y0;
                    flip ^= 1;
                    t_mode_contract(T, (which == 1 && onpath[p] && onpath[sc]) ? TSM(sc) : TM(sc), L, sh.dims[s2], R,
                                    src, dst, one, false);
                    src = dst;
                }
                double2* ye = (src == y0) ? y1 : y0;
                t_mul_transposed(T, src, sh.dimen, sh.hrank, Ep, ye, one, false);
                int64_t L, R;
                mode_view(sh, k, L, R);
                t_outer_contract(T, Ab, ye, L, sh.dims[k], R, which == 1 ? ESM(ch) : EM(ch), one, false);
            }
        }
    }
    // leaf densities
    double2* o = out + static_cast<int64_t>(pair) * out_stride;
    for (int i = 0; i < n; ++i) {
        const DNode& nd = K.nodes[i];
        if (nd.leaf < 0) continue;
        const Shape sh = c.shape(i);
        const int d = static_cast<int>(sh.dimen);
        const double2* Ak = K.base + nd.A;
        const double2* Ab = B.base + nd.A;
        double2* dst = o + out_off[i];
        t_mul_transposed(T, Ak, sh.dimen, sh.hrank, EM(i), y0, one, false);
        t_outer_contract(T, Ab, y0, 1, d, sh.hrank, dst, one, false);
        if (esdef[i]) {
            t_mul_transposed(T, Ak, sh.dimen, sh.hrank, ESM(i), y0, one, false);
            t_outer_contract(T, Ab, y0, 1, d, sh.hrank, dst + d * d, one, false);
        } else {
            for (int e = threadIdx.x; e < d * d; e += NT) dst[d * d + e] = cz();
            T.sync();
        }
    }
}

}  // namespace ttnd

extern "C" hipError_t ttnd_launch_densities(const ttnd::DInst* kets, const ttnd::DInst* bras, int npairs, int ins,
                                            const double2* S, double2* work, int64_t work_stride, int64_t rr,
                                            const int64_t* out_off, double2* out, int64_t out_stride, hipStream_t s) {
    hipLaunchKernelGGL(ttnd::densities_kernel, dim3(npairs), dim3(ttnd::NT), 0, s, kets, bras, ins, S, work,
                       work_stride, rr, out_off, out, out_stride);
    return hipGetLastError();
}

extern "C" hipError_t ttnd_launch_run(const ttnd::DInst* d_insts, int n, const ttnd::RunArgs* a,
                                      const ttnd::TeamPlan* d_plan, ttnd::TeamWords* d_words, unsigned* d_fill,
                                      int grid, hipStream_t s) {
    if (d_plan == nullptr)
        hipLaunchKernelGGL(ttnd::tdvp_run_kernel, dim3(n), dim3(ttnd::NT), 0, s, d_insts, *a, d_plan, d_words, d_fill);
    else
        hipLaunchKernelGGL(ttnd::tdvp_run_kernel, dim3(grid), dim3(ttnd::NT), 0, s, d_insts, *a, d_plan, d_words,
                           d_fill);
    return hipGetLastError();
}

extern "C" hipError_t ttnd_launch_expect(const ttnd::DInst* d_inst, int leaf, int64_t op_off, double* d_out,
                                         hipStream_t s) {
    hipLaunchKernelGGL(ttnd::expect_kernel, dim3(1), dim3(ttnd::NT), 0, s, d_inst, leaf, op_off, d_out);
    return hipGetLastError();
}

// Workgroups of tdvp_run_kernel that fit on one CU at once (the team launch needs every
// workgroup resident).
extern "C" hipError_t ttnd_run_occupancy(int* blocks_per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, ttnd::tdvp_run_kernel, ttnd::NT, 0);
}

// p[0 .. n) = 0: clears the team words before a team launch (a kernel, not a memset, so the
// sequence can be captured in a graph and replayed)
namespace ttnd {
__global__ __launch_bounds__(256) void zero_u32_kernel(unsigned* p, int64_t n) {
    for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * 256) p[i] = 0u;
}
}  // namespace ttnd
extern "C" hipError_t ttnd_launch_zero_u32(unsigned* p, int64_t n, hipStream_t s) {
    const int64_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(ttnd::zero_u32_kernel, dim3(static_cast<unsigned>(blocks < 1024 ? blocks : 1024)), dim3(256), 0,
                       s, p, n);
    return hipGetLastError();
}

// XCD of every workgroup of an n-workgroup launch (placement census)
namespace ttnd {
__global__ __launch_bounds__(NT) void xcd_census_kernel(int* out) {
    if (threadIdx.x == 0) out[blockIdx.x] = xcd_id();
}
}  // namespace ttnd
extern "C" hipError_t ttnd_launch_census(int* d_out, int n, hipStream_t s) {
    hipLaunchKernelGGL(ttnd::xcd_census_kernel, dim3(n), dim3(ttnd::NT), 0, s, d_out);
    return hipGetLastError();
}
