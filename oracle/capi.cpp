// ORACLE — test infrastructure only (see oracle.hpp).  C-ABI of the CPU restatement.
#include "orc_capi.h"

#include "../ttns_lib_amd/csrc/host/marshal.hpp"
#include "../ttns_lib_amd/csrc/host/tables.hpp"
#include "oracle.hpp"

#include <cmath>
#include <memory>
#include <string>

using namespace ttnh;

struct orc_ctx {
    std::string err;
};
struct orc_tree {
    orc_ctx* ctx;
    Topology topo;
};
struct orc_sop {
    orc_ctx* ctx;
    SOP sop;
};
struct orc_integ {
    orc_ctx* ctx;
    std::unique_ptr<orc::Engine> eng;
};

namespace {
template <class F>
int guard(orc_ctx* ctx, F&& f) {
    try {
        f();
        return TTN_OK;
    } catch (const Error& e) {
        if (ctx) ctx->err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        if (ctx) ctx->err = e.what();
        return TTN_ERR_INVALID;
    }
}
}  // namespace

static std::vector<cplx> cvec(const double* p, i64 n) {
    std::vector<cplx> v(n);
    for (i64 i = 0; i < n; ++i) v[i] = cget(p, i);
    return v;
}
static void cput(const std::vector<cplx>& v, double* p) {
    for (size_t i = 0; i < v.size(); ++i) {
        p[2 * i] = v[i].real();
        p[2 * i + 1] = v[i].imag();
    }
}

extern "C" {

int orc_ctx_create(int, orc_ctx** out) {
    *out = new orc_ctx();
    return TTN_OK;
}
int orc_ctx_destroy(orc_ctx* ctx) {
    delete ctx;
    return TTN_OK;
}
const char* orc_last_error(const orc_ctx* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

int orc_tree_create(orc_ctx* ctx, int64_t n, const int64_t* parent, const int64_t* hrank, const int64_t* cap_hrank,
                    const int64_t* leaf_dim, orc_tree** out) {
    return guard(ctx, [&] {
        auto t = std::make_unique<orc_tree>();
        t->ctx = ctx;
        t->topo = topology_from_arrays(static_cast<int>(n), parent, hrank, cap_hrank, leaf_dim);
        *out = t.release();
    });
}
int orc_tree_destroy(orc_tree* t) {
    delete t;
    return TTN_OK;
}
int orc_tree_node(const orc_tree* t, int64_t node, int64_t* nchild, int64_t* rows, int64_t* hrank, int64_t* leaf) {
    return guard(t->ctx, [&] {
        check(node >= 0 && node < t->topo.size(), ERR_INVALID, "node out of range");
        const Node& nd = t->topo.nodes[node];
        if (nchild) *nchild = static_cast<int64_t>(nd.children.size());
        if (rows) *rows = nd.dimen();
        if (hrank) *hrank = nd.hrank;
        if (leaf) *leaf = nd.leaf_index;
    });
}
int orc_tree_capacities(const orc_tree* t, int64_t* caps) {
    return guard(t->ctx, [&] {
        check(caps != nullptr, ERR_INVALID, "null output");
        for (int i = 0; i < t->topo.size(); ++i) caps[i] = t->topo.nodes[i].cap_hrank;
    });
}
int orc_tree_size(const orc_tree* t, int64_t* n, int64_t* m) {
    if (n) *n = t->topo.size();
    if (m) *m = t->topo.nmodes();
    return TTN_OK;
}
int orc_traversal(const orc_tree* t, int64_t* len, int64_t* order) {
    return guard(t->ctx, [&] {
        auto tour = euler_tour(t->topo);
        *len = static_cast<int64_t>(tour.size());
        if (order)
            for (size_t i = 0; i < tour.size(); ++i) order[i] = tour[i];
    });
}

int orc_sop_create(orc_ctx* ctx, int64_t nmodes, const int64_t* mode_dims, int64_t nterms, const double* coeff,
                   int64_t nops, const int32_t* op_mode, const int32_t* op_kind, const int64_t* op_coff,
                   const double* cdata, const int64_t* op_ioff, const int64_t* idata, const int64_t* op_roff,
                   const int64_t* r, orc_sop** out) {
    return guard(ctx, [&] {
        auto s = std::make_unique<orc_sop>();
        s->ctx = ctx;
        s->sop = sop_from_arrays(nmodes, mode_dims, nterms, coeff, nops, op_mode, op_kind, op_coff, cdata, op_ioff,
                                 idata, op_roff, r);
        *out = s.release();
    });
}
int orc_sop_destroy(orc_sop* s) {
    delete s;
    return TTN_OK;
}
int orc_sop_compiled_tables(const orc_tree* t, const orc_sop* s, int64_t* len, int64_t* buf) {
    return guard(t->ctx, [&] {
        auto v = serialise_tables(compile_sop(t->topo, s->sop));
        if (buf) {
            check(*len >= static_cast<int64_t>(v.size()), ERR_CAPACITY, "table buffer too small");
            std::copy(v.begin(), v.end(), buf);
        }
        *len = static_cast<int64_t>(v.size());
    });
}

int orc_integrator_create(orc_ctx* ctx, const orc_tree* t, const orc_sop* s, const ttn_integ_params* p,
                          orc_integ** out) {
    return guard(ctx, [&] {
        check(p->mode == TTN_MODE_FIXED_STEP || p->mode == TTN_MODE_SUBSPACE_EXPANSION, ERR_INVALID,
              "unknown integrator mode");
        check(p->krylov_dim >= 1 && p->krylov_dim <= 64, ERR_INVALID, "krylov_dim must be in [1, 64]");
        check(p->dt > 0.0, ERR_INVALID, "dt must be positive");
        orc::Params pp;
        pp.krylov_dim = static_cast<int>(p->krylov_dim);
        pp.krylov_tol = p->krylov_tol;
        pp.dt = p->dt;
        pp.expansion = p->mode == TTN_MODE_SUBSPACE_EXPANSION;
        pp.spawning_threshold = p->spawning_threshold;
        pp.unoccupied_threshold = p->unoccupied_threshold;
        pp.minimum_unoccupied = p->minimum_unoccupied;
        pp.neigenvalues = p->neigenvalues;
        if (p->eigenvalue_tolerance > 0.0) pp.eigenvalue_tolerance = p->eigenvalue_tolerance;
        pp.only_apply_when_no_unoccupied = p->only_apply_when_no_unoccupied != 0;
        pp.rng_seed = p->rng_seed != 0 ? p->rng_seed : 5489u;
        check(!pp.expansion || pp.neigenvalues >= 1, ERR_INVALID, "neigenvalues must be positive");
        check(pp.minimum_unoccupied >= 0, ERR_INVALID, "minimum_unoccupied must not be negative");
        auto in = std::make_unique<orc_integ>();
        in->ctx = ctx;
        in->eng = std::make_unique<orc::Engine>(t->topo, s->sop, pp);
        *out = in.release();
    });
}
int orc_integrator_destroy(orc_integ* in) {
    delete in;
    return TTN_OK;
}

int orc_state_set(orc_integ* in, int64_t node, const double* data, int64_t rows, int64_t cols) {
    return guard(in->ctx, [&] {
        const auto& topo = in->eng->topo();
        check(node >= 0 && node < topo.size(), ERR_INVALID, "node out of range");
        const Node& nd = topo.nodes[node];
        check(rows == nd.dimen() && cols == nd.hrank, ERR_INVALID, "state_set: shape mismatch");
        auto& A = in->eng->A(static_cast<int>(node));
        for (int64_t i = 0; i < rows * cols; ++i) A[i] = cget(data, i);
    });
}
int orc_state_get(orc_integ* in, int64_t node, double* data, int64_t rows, int64_t cols) {
    return guard(in->ctx, [&] {
        const auto& topo = in->eng->topo();
        check(node >= 0 && node < topo.size(), ERR_INVALID, "node out of range");
        const Node& nd = topo.nodes[node];
        check(rows == nd.dimen() && cols == nd.hrank, ERR_INVALID, "state_get: shape mismatch");
        const auto& A = in->eng->A(static_cast<int>(node));
        for (int64_t i = 0; i < rows * cols; ++i) {
            data[2 * i] = A[i].real();
            data[2 * i + 1] = A[i].imag();
        }
    });
}
int orc_step(orc_integ* in, double cr, double ci, int64_t nsteps) {
    return guard(in->ctx, [&] {
        for (int64_t s = 0; s < nsteps; ++s) in->eng->step(cplx(cr, ci));
    });
}
int orc_time(const orc_integ* in, double* t) {
    *t = in->eng->time();
    return TTN_OK;
}
int orc_counters(const orc_integ* in, int64_t* nh, int64_t* ns) {
    if (nh) *nh = in->eng->nh_applications();
    if (ns) *ns = in->eng->nsteps();
    return TTN_OK;
}
int orc_expansion_counters(const orc_integ* in, int64_t* n1, int64_t* n2) {
    if (n1) *n1 = in->eng->n_onesite();
    if (n2) *n2 = in->eng->n_twosite();
    return TTN_OK;
}
int orc_hranks(const orc_integ* in, int64_t* hr) {
    const auto& topo = in->eng->topo();
    for (int n = 0; n < topo.size(); ++n) hr[n] = topo.nodes[n].hrank;
    return TTN_OK;
}
int orc_state_shape(const orc_integ* in, int64_t node, int64_t* rows, int64_t* cols) {
    return guard(in->ctx, [&] {
        const auto& topo = in->eng->topo();
        check(node >= 0 && node < topo.size(), ERR_INVALID, "node out of range");
        if (rows) *rows = topo.nodes[node].dimen();
        if (cols) *cols = topo.nodes[node].hrank;
    });
}
int orc_set_dt(orc_integ* in, double dt) {
    return guard(in->ctx, [&] {
        check(dt > 0.0, ERR_INVALID, "dt must be positive");
        in->eng->set_dt(dt);
    });
}
int orc_get_dt(const orc_integ* in, double* dt) {
    *dt = in->eng->dt();
    return TTN_OK;
}
int orc_reset_nsteps(orc_integ* in) {
    in->eng->reset_nsteps();
    return TTN_OK;
}
int orc_set_expansion_params(orc_integ* in, const ttn_integ_params* p) {
    return guard(in->ctx, [&] {
        check(p != nullptr, ERR_INVALID, "null parameters");
        check(in->eng->params().expansion, ERR_INVALID, "expansion parameters need a subspace-expansion integrator");
        check(p->neigenvalues >= 1, ERR_INVALID, "neigenvalues must be positive");
        check(p->minimum_unoccupied >= 0, ERR_INVALID, "minimum_unoccupied must not be negative");
        orc::Params q = in->eng->params();
        q.spawning_threshold = p->spawning_threshold;
        q.unoccupied_threshold = p->unoccupied_threshold;
        q.minimum_unoccupied = p->minimum_unoccupied;
        q.neigenvalues = p->neigenvalues;
        q.eigenvalue_tolerance = p->eigenvalue_tolerance > 0.0 ? p->eigenvalue_tolerance : 1e-6;
        q.only_apply_when_no_unoccupied = p->only_apply_when_no_unoccupied != 0;
        in->eng->set_params(q);
    });
}
int orc_set_coefficients(orc_integ* in, int64_t nterms, const double* coeff) {
    return guard(in->ctx, [&] {
        check(coeff != nullptr && nterms == in->eng->nterms(), ERR_INVALID,
              "set_coefficients: term count differs from the operator's");
        std::vector<cplx> c(static_cast<size_t>(nterms));
        for (int64_t r = 0; r < nterms; ++r) {
            check(std::isfinite(coeff[2 * r]) && std::isfinite(coeff[2 * r + 1]), ERR_INVALID,
                  "set_coefficients: non-finite coefficient");
            c[r] = cplx(coeff[2 * r], coeff[2 * r + 1]);
        }
        in->eng->set_coefficients(c);
    });
}
int orc_set_expansion_log(orc_integ* in, int64_t cap) {
    in->eng->set_log(cap);
    return TTN_OK;
}
int orc_expansion_log(const orc_integ* in, int64_t* len, double* out) {
    const auto& lg = in->eng->log();
    const int64_t n = static_cast<int64_t>(lg.size() / 20);
    if (out) std::copy(lg.begin(), lg.begin() + 20 * std::min(n, *len), out);
    *len = n;
    return TTN_OK;
}
int orc_expansion_draws(const orc_integ* in, int64_t* n) {
    *n = in->eng->normals_used();
    return TTN_OK;
}
int orc_normals_used(const orc_integ* in, int64_t* n) {
    *n = in->eng->normals_used();
    return TTN_OK;
}
int orc_singular_values(int64_t n, const double* M, double* sv) {
    return guard(nullptr, [&] {
        auto m = cvec(M, n * n);
        auto s = orc::singular_values(m.data(), n);
        for (i64 i = 0; i < n; ++i) sv[i] = s[i];
    });
}
int orc_tridiag_eigen(int64_t n, const double* a, const double* b, double* w, double* Z) {
    return guard(nullptr, [&] {
        std::vector<double> av(a, a + n), bv(b, b + (n > 0 ? n - 1 : 0)), wv, Zv;
        orc::tridiag_eigen(av, bv, wv, Zv);
        std::copy(wv.begin(), wv.end(), w);
        std::copy(Zv.begin(), Zv.end(), Z);
    });
}

int orc_expect(orc_integ* in, int64_t mode, const double* op, double* out) {
    return guard(in->ctx, [&] {
        const auto& topo = in->eng->topo();
        check(mode >= 0 && mode < topo.nmodes(), ERR_INVALID, "mode out of range");
        const i64 d = topo.mode_dims[mode];
        std::vector<cplx> o(d * d);
        for (i64 i = 0; i < d * d; ++i) o[i] = cget(op, i);
        const cplx v = in->eng->expect(static_cast<int>(mode), o.data());
        out[0] = v.real();
        out[1] = v.imag();
    });
}
int orc_norm2(orc_integ* in, double* out) {
    return guard(in->ctx, [&] { *out = in->eng->expect(-1, nullptr).real(); });
}

int orc_sb_problem(orc_ctx* ctx, const ttn_sb_params* p, orc_tree** tree, orc_sop** sop) {
    return guard(ctx, [&] {
        SBProblem pr = sb_build(sb_from_c(p));
        auto t = std::make_unique<orc_tree>();
        t->ctx = ctx;
        t->topo = pr.topo;
        auto s = std::make_unique<orc_sop>();
        s->ctx = ctx;
        s->sop = pr.sop;
        *tree = t.release();
        *sop = s.release();
    });
}
int orc_sb_initial_state(orc_integ* in, const ttn_sb_params* p) {
    return guard(in->ctx, [&] {
        SBParams sp = sb_from_c(p);
        const auto& topo = in->eng->topo();
        auto st = sb_initial_state(topo, sp.psi0, sp.seed);
        for (int n = 0; n < topo.size(); ++n) in->eng->A(n) = st[n];
    });
}

int orc_sb_adjoint_run(orc_integ* in, const ttn_sb_params* p, int64_t nsteps, int64_t every, int64_t nquad,
                       int64_t mode, const double* op, double* value, double* dg, double* dw) {
    return guard(in->ctx, [&] {
        const auto& topo = in->eng->topo();
        check(mode >= 0 && mode < topo.nmodes(), ERR_INVALID, "mode out of range");
        const i64 d = topo.mode_dims[mode];
        std::vector<cplx> o(d * d);
        for (i64 i = 0; i < d * d; ++i) o[i] = cget(op, i);
        const SBParams sp = sb_from_c(p);
        std::vector<double> g, w;
        const double J = orc::sb_adjoint(*in->eng, sp, nsteps, every, static_cast<int>(nquad), static_cast<int>(mode),
                                         o.data(), g, w);
        std::copy(g.begin(), g.end(), dg);
        std::copy(w.begin(), w.end(), dw);
        if (value) *value = J;
    });
}
int orc_transition_densities(orc_integ* ket, orc_integ* bra, int64_t ins, const double* S, int64_t len, double* out) {
    return guard(ket->ctx, [&] {
        const auto& topo = ket->eng->topo();
        check(ins >= 0 && ins < topo.nmodes(), ERR_INVALID, "mode out of range");
        const i64 d = topo.mode_dims[ins];
        std::vector<cplx> s(d * d);
        for (i64 i = 0; i < d * d; ++i) s[i] = cget(S, i);
        std::vector<std::vector<cplx>> rho, rhoS;
        orc::transition_densities(*ket->eng, *bra->eng, static_cast<int>(ins), s.data(), rho, rhoS);
        // per leaf mode: rho (d x d) then rhoS (d x d, zeros on the insertion leaf)
        i64 at = 0;
        for (int m = 0; m < topo.nmodes(); ++m) {
            const i64 dm = topo.mode_dims[m];
            check(at + 2 * dm * dm <= len, ERR_INVALID, "transition densities: output too small");
            for (i64 q = 0; q < dm * dm; ++q) {
                const cplx a = rho[m][q], b = rhoS[m].empty() ? cplx(0.0, 0.0) : rhoS[m][q];
                out[2 * (at + q)] = a.real();
                out[2 * (at + q) + 1] = a.imag();
                out[2 * (at + dm * dm + q)] = b.real();
                out[2 * (at + dm * dm + q) + 1] = b.imag();
            }
            at += 2 * dm * dm;
        }
    });
}

int orc_sb_partition_modes(orc_ctx* ctx, int64_t nmodes, const double* wk, double ecut, double wc, int64_t nmax,
                           int64_t nmax4, int64_t nmax2, int64_t nmax3, int64_t* nblocks, int64_t* partition) {
    return guard(ctx, [&] { sb_partition_c(nmodes, wk, ecut, wc, nmax, nmax4, nmax2, nmax3, nblocks, partition); });
}
int orc_sb_mode_block(orc_ctx* ctx, int64_t nmodes, const double* w, const double* g, double ecut, int64_t nmax,
                      int64_t* nstates, int64_t* nnz, int64_t* occupations, double* h0, int64_t* hc_rowptr,
                      int64_t* hc_colind, double* hc_vals) {
    return guard(ctx, [&] {
        sb_mode_block_c(nmodes, w, g, ecut, nmax, nstates, nnz, occupations, h0, hc_rowptr, hc_colind, hc_vals);
    });
}

// ---- building blocks ----
int orc_heff_node(orc_integ* in, int64_t node, const double* v, double* out) {
    return guard(in->ctx, [&] {
        const i64 sz = in->eng->topo().nodes.at(node).size();
        auto x = cvec(v, sz);
        std::vector<cplx> y(sz);
        in->eng->heff_node(static_cast<int>(node), x.data(), y.data());
        cput(y, out);
    });
}
int orc_heff_bond(orc_integ* in, int64_t node, const double* v, double* out) {
    return guard(in->ctx, [&] {
        const i64 r = in->eng->topo().nodes.at(node).hrank;
        auto x = cvec(v, r * r);
        std::vector<cplx> y(r * r);
        in->eng->heff_bond(static_cast<int>(node), x.data(), y.data());
        cput(y, out);
    });
}
int orc_update_spf(orc_integ* in, int64_t node) {
    return guard(in->ctx, [&] { in->eng->update_spf(static_cast<int>(node)); });
}
int orc_update_mf(orc_integ* in, int64_t node) {
    return guard(in->ctx, [&] { in->eng->update_mf(static_cast<int>(node)); });
}
int orc_env_get(orc_integ* in, int64_t node, int64_t term, int32_t which_mf, double* out) {
    return guard(in->ctx, [&] {
        const auto& m = which_mf ? in->eng->mf(static_cast<int>(node), static_cast<int>(term))
                                 : in->eng->spf(static_cast<int>(node), static_cast<int>(term));
        check(!m.empty(), ERR_INVALID, "identity environment has no matrix");
        cput(m, out);
    });
}
int orc_env_set(orc_integ* in, int64_t node, int64_t term, int32_t which_mf, const double* src) {
    return guard(in->ctx, [&] {
        auto& m = which_mf ? in->eng->mf_mut(static_cast<int>(node), static_cast<int>(term))
                           : in->eng->spf_mut(static_cast<int>(node), static_cast<int>(term));
        check(!m.empty(), ERR_INVALID, "identity environment has no matrix");
        for (size_t i = 0; i < m.size(); ++i) m[i] = cget(src, static_cast<i64>(i));
    });
}
int orc_node_expv(orc_integ* in, int64_t node, double cr, double ci, double tau, int64_t kdim, double tol,
                  int64_t* napply) {
    return guard(in->ctx, [&] {
        const int n = static_cast<int>(node);
        auto& A = in->eng->A(n);
        const i64 na = orc::expv([&](const cplx* v, cplx* res) { in->eng->heff_node(n, v, res); }, A.data(),
                                 static_cast<i64>(A.size()), cplx(cr, ci), tau, static_cast<int>(kdim), tol);
        if (napply) *napply = na;
    });
}
int orc_expv(int64_t n, const double* Hd, double* x, double cr, double ci, double tau, int64_t kdim, double tol,
             int64_t* napply) {
    return guard(nullptr, [&] {
        auto H = cvec(Hd, n * n);
        auto v = cvec(x, n);
        auto f = [&](const cplx* a, cplx* b) {
            for (i64 i = 0; i < n; ++i) {
                cplx s(0.0, 0.0);
                for (i64 j = 0; j < n; ++j) s += H[i * n + j] * a[j];
                b[i] = s;
            }
        };
        const i64 na = orc::expv(f, v.data(), n, cplx(cr, ci), tau, static_cast<int>(kdim), tol);
        if (napply) *napply = na;
        cput(v, x);
    });
}
int orc_expm(int64_t n, const double* M, double* E) {
    return guard(nullptr, [&] {
        auto m = cvec(M, n * n);
        std::vector<cplx> e(n * n);
        orc::expm(m.data(), static_cast<int>(n), e.data());
        cput(e, E);
    });
}
int orc_qr(int64_t m, int64_t n, const double* A, double* Q, double* R) {
    return guard(nullptr, [&] {
        auto a = cvec(A, m * n);
        std::vector<cplx> q(m * n), r(n * n);
        orc::householder_qr(a.data(), m, n, q.data(), r.data());
        cput(q, Q);
        cput(r, R);
    });
}

}  // extern "C"
