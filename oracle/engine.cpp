// ORACLE — test infrastructure only (see oracle.hpp).  TDVP engine restatement.
#include "oracle.hpp"

#include <cmath>

namespace orc {

using ttnh::check;
using ttnh::ERR_INVALID;
using ttnh::ERR_NUMERIC;
using ttnh::Node;
using ttnh::Term;

Engine::Engine(const ttnh::Topology& topo, const ttnh::SOP& sop, const Params& p)
    : m_topo(topo), m_sop(sop), m_p(p) {
    m_ops = ttnh::compile_sop(m_topo, m_sop);
    m_tour = ttnh::euler_tour(m_topo);
    const int n = m_topo.size();
    m_A.resize(n);
    m_R.resize(n);
    m_spf.resize(n);
    m_mf.resize(n);
    for (int i = 0; i < n; ++i) {
        const Node& nd = m_topo.nodes[i];
        m_A[i].assign(nd.size(), cplx(0.0, 0.0));
        const i64 r = nd.hrank;
        const int nt = m_ops.nodes[i].nterms();
        m_spf[i].resize(nt);
        m_mf[i].resize(nt);
        for (int t = 0; t < nt; ++t) {
            const Term& tm = m_ops.nodes[i].terms[t];
            if (!tm.id_spf) m_spf[i][t].assign(r * r, cplx(0.0, 0.0));
            if (!tm.id_mf) m_mf[i][t].assign(r * r, cplx(0.0, 0.0));
        }
    }
    m_dt = p.dt;
    m_rng.seed(p.rng_seed);
}

// kronecker_product_operator::apply (kronecker_product_operator_helper.hpp:261-315):
// chain of mode contractions with the children's spf matrices.
void Engine::kpo(int n, const std::vector<std::pair<int, int>>& list, const cplx* v, cplx* out,
                 cplx* tmp) const {
    const Node& nd = m_topo.nodes[n];
    const i64 sz = nd.size();
    if (list.empty()) {
        std::copy(v, v + sz, out);
        return;
    }
    // ping-pong so that the last contraction lands in out
    const cplx* src = v;
    const size_t cnt = list.size();
    for (size_t k = 0; k < cnt; ++k) {
        cplx* dst = ((cnt - 1 - k) % 2 == 0) ? out : tmp;
        const int nu = list[k].first, cri = list[k].second;
        i64 L = 1, R = nd.hrank;
        for (int c = 0; c < nu; ++c) L *= nd.dims[c];
        for (int c = nu + 1; c < nd.nmodes(); ++c) R *= nd.dims[c];
        const int ch = nd.children[nu];
        mode_contract(m_spf[ch][cri].data(), L, nd.dims[nu], R, src, dst);
        src = dst;
    }
}

// coefficient_evolution_leaf / coefficient_evolution_branch
// (projector_splitting_evolution_functions.hpp:65-258)
void Engine::heff_node(int n, const cplx* v, cplx* res) const {
    const Node& nd = m_topo.nodes[n];
    const i64 sz = nd.size(), r = nd.hrank, D = nd.dimen();
    std::vector<cplx> t1(sz), t2(sz), t3(sz), t4(sz);
    for (i64 i = 0; i < sz; ++i) res[i] = cplx(0.0, 0.0);
    const auto& terms = m_ops.nodes[n].terms;
    for (size_t ind = 0; ind < terms.size(); ++ind) {
        const Term& tm = terms[ind];
        if (tm.id_spf) {
            cplx coeff(0.0, 0.0);
            if (tm.nspf_terms() > 1)
                for (const cplx& a : tm.accum) coeff += a;
            else
                coeff = tm.coeff;
            if (tm.id_mf) {
                for (i64 i = 0; i < sz; ++i) res[i] += coeff * v[i];
            } else {
                mul_transposed(v, D, r, m_mf[n][ind].data(), t1.data());
                for (i64 i = 0; i < sz; ++i) res[i] += coeff * t1[i];
            }
            continue;
        }
        // t3 = sum_i c_i K_i v
        for (i64 i = 0; i < sz; ++i) t3[i] = cplx(0.0, 0.0);
        for (int i = 0; i < tm.nspf_terms(); ++i) {
            const cplx c = (i == 0) ? (tm.nspf_terms() > 1 ? tm.accum[0] : tm.coeff) : tm.accum[i];
            if (nd.is_leaf()) {
                const auto& idx = tm.spf_index[i][0];
                m_sop.ops[idx.first][idx.second].op.apply(v, t2.data(), r);
            } else {
                kpo(n, tm.spf_index[i], v, t2.data(), t1.data());
            }
            for (i64 q = 0; q < sz; ++q) t3[q] += c * t2[q];
        }
        if (tm.id_mf) {
            for (i64 q = 0; q < sz; ++q) res[q] += t3[q];
        } else {
            mul_transposed(t3.data(), D, r, m_mf[n][ind].data(), t4.data());
            for (i64 q = 0; q < sz; ++q) res[q] += t4[q];
        }
    }
}

// full_hamiltonian_action (projector_splitting_evolution_functions.hpp:7-62): R is hrank x hrank
void Engine::heff_bond(int n, const cplx* v, cplx* res) const {
    const i64 r = m_topo.nodes[n].hrank, sz = r * r;
    std::vector<cplx> t1(sz), t2(sz);
    for (i64 i = 0; i < sz; ++i) res[i] = cplx(0.0, 0.0);
    const auto& terms = m_ops.nodes[n].terms;
    for (size_t ind = 0; ind < terms.size(); ++ind) {
        const Term& tm = terms[ind];
        const cplx* x = v;
        if (!tm.id_spf) {
            mode_contract(m_spf[n][ind].data(), 1, r, r, v, t1.data());
            x = t1.data();
        }
        if (!tm.id_mf) {
            mul_transposed(x, r, r, m_mf[n][ind].data(), t2.data());
            x = t2.data();
        }
        for (i64 i = 0; i < sz; ++i) res[i] += tm.coeff * x[i];
    }
}

// single_particle_operator_engine::evaluate (single_particle_operator.hpp:26-103)
void Engine::update_spf(int n) {
    const Node& nd = m_topo.nodes[n];
    const i64 sz = nd.size(), r = nd.hrank, D = nd.dimen();
    std::vector<cplx> HA(sz), tmp(sz), M(r * r);
    const auto& terms = m_ops.nodes[n].terms;
    const cplx* a = m_A[n].data();
    for (size_t ind = 0; ind < terms.size(); ++ind) {
        const Term& tm = terms[ind];
        if (tm.id_spf) continue;
        auto& out = m_spf[n][ind];
        out.assign(r * r, cplx(0.0, 0.0));
        for (int i = 0; i < tm.nspf_terms(); ++i) {
            if (nd.is_leaf()) {
                const auto& idx = tm.spf_index[i][0];
                m_sop.ops[idx.first][idx.second].op.apply(a, HA.data(), r);
            } else {
                kpo(n, tm.spf_index[i], a, HA.data(), tmp.data());
            }
            adjoint_mul(a, HA.data(), D, r, M.data());
            if (tm.nspf_terms() == 1)
                out = M;
            else
                for (i64 q = 0; q < r * r; ++q) out[q] += tm.accum[i] * M[q];
        }
    }
}

// mean_field_operator_engine::kron_prod (mean_field_operator.hpp:138-184): sibling spf
// matrices contracted into the parent's tensor
void Engine::kron_siblings(int p, const std::vector<std::pair<int, int>>& sib, const cplx* Ap, cplx* out,
                           cplx* tmp) const {
    kpo(p, sib, Ap, out, tmp);
}

// mean_field_operator_engine::evaluate (mean_field_operator.hpp:37-135) for child n
void Engine::update_mf(int n) {
    const Node& nd = m_topo.nodes[n];
    if (nd.is_root()) return;
    const int p = nd.parent;
    const Node& pn = m_topo.nodes[p];
    const i64 psz = pn.size(), pr = pn.hrank, pD = pn.dimen();
    const int mode = nd.child_id;
    i64 L = 1, R = pr;
    for (int c = 0; c < mode; ++c) L *= pn.dims[c];
    for (int c = mode + 1; c < pn.nmodes(); ++c) R *= pn.dims[c];
    const i64 d = pn.dims[mode];
    std::vector<cplx> HA(psz), tmp(psz), M(d * d);
    const auto& terms = m_ops.nodes[n].terms;
    const auto& pterms = m_ops.nodes[p].terms;
    const cplx* Ap = m_A[p].data();
    for (size_t ind = 0; ind < terms.size(); ++ind) {
        const Term& tm = terms[ind];
        if (tm.id_mf) continue;
        auto& out = m_mf[n][ind];
        out.assign(d * d, cplx(0.0, 0.0));
        for (int it = 0; it < tm.nmf_terms(); ++it) {
            const auto& mi = tm.mf_index[it];
            kron_siblings(p, mi.siblings, Ap, tmp.data(), HA.data());
            if (!pterms[mi.parent_index].id_mf)
                mul_transposed(tmp.data(), pD, pr, m_mf[p][mi.parent_index].data(), HA.data());
            else
                HA = tmp;
            outer_contract(Ap, HA.data(), L, d, R, M.data());
            if (tm.nmf_terms() == 1)
                out = M;
            else
                for (i64 q = 0; q < d * d; ++q) out[q] += tm.accum[it] * M[q];
        }
    }
}

// ---- gauge moves ----
// root_to_leaf_decomposition_engine::evaluate + apply_to_node (:112-204): permute mode
// `mode` to the fastest index, factorise, un-permute, keep R^T.
void Engine::r2l(int n, int mode) {
    const Node& nd = m_topo.nodes[n];
    const i64 sz = nd.size();
    i64 id1 = 1;
    for (int c = 0; c <= mode; ++c) id1 *= nd.dims[c];
    const i64 id2 = sz / id1;
    const i64 d2 = nd.dims[mode], d1 = sz / d2;
    std::vector<cplx> U(sz), Q(sz), Rm(d2 * d2);
    // U (id2 x id1) = A (id1 x id2)^T
    for (i64 i = 0; i < id1; ++i)
        for (i64 j = 0; j < id2; ++j) U[j * id1 + i] = m_A[n][i * id2 + j];
    householder_qr(U.data(), d1, d2, Q.data(), Rm.data());
    for (i64 i = 0; i < id1; ++i)
        for (i64 j = 0; j < id2; ++j) m_A[n][i * id2 + j] = Q[j * id1 + i];
    m_R[n].assign(d2 * d2, cplx(0.0, 0.0));
    for (i64 i = 0; i < d2; ++i)
        for (i64 j = 0; j < d2; ++j) m_R[n][i * d2 + j] = Rm[j * d2 + i];
}

// leaf_to_root_decomposition_engine::evaluate + apply_to_node (:56-101)
void Engine::l2r(int n) {
    const Node& nd = m_topo.nodes[n];
    const i64 D = nd.dimen(), r = nd.hrank;
    std::vector<cplx> Q(D * r);
    m_R[n].assign(r * r, cplx(0.0, 0.0));
    householder_qr(m_A[n].data(), D, r, Q.data(), m_R[n].data());
    m_A[n] = Q;
}

// root_to_leaf_decomposition_engine::apply_from_parent (:208-233): A <- A * R_parent
void Engine::apply_from_parent(int n) {
    const Node& nd = m_topo.nodes[n];
    const i64 D = nd.dimen(), r = nd.hrank;
    const auto& pr = m_R[nd.parent];
    check(static_cast<i64>(pr.size()) == r * r, ERR_INVALID, "parent R has the wrong shape");
    std::vector<cplx> out(D * r, cplx(0.0, 0.0));
    for (i64 x = 0; x < D; ++x)
        for (i64 k = 0; k < r; ++k) {
            const cplx a = m_A[n][x * r + k];
            for (i64 j = 0; j < r; ++j) out[x * r + j] += a * pr[k * r + j];
        }
    m_A[n] = out;
}

// leaf_to_root_decomposition_engine::apply_to_parent (:105-138): parent mode <- R *
void Engine::apply_to_parent(int n) {
    const Node& nd = m_topo.nodes[n];
    const int p = nd.parent;
    const Node& pn = m_topo.nodes[p];
    const int mode = nd.child_id;
    i64 L = 1, R = pn.hrank;
    for (int c = 0; c < mode; ++c) L *= pn.dims[c];
    for (int c = mode + 1; c < pn.nmodes(); ++c) R *= pn.dims[c];
    std::vector<cplx> out(pn.size());
    mode_contract(m_R[n].data(), L, pn.dims[mode], R, m_A[p].data(), out.data());
    m_A[p] = out;
}

// ---- propagators (projector_splitting_engine.hpp:117-165) ----
void Engine::node_update(int n) {
    const i64 sz = m_topo.nodes[n].size();
    m_nh += expv([this, n](const cplx* v, cplx* res) { heff_node(n, v, res); }, m_A[n].data(), sz, m_coeff,
                 m_sub_dt / 2.0, m_p.krylov_dim, m_p.krylov_tol);
}

// R of node `n` evolved backwards with the operators of node `ops`
void Engine::bond_update(int n, int ops) {
    const i64 r = m_topo.nodes[ops].hrank;
    check(static_cast<i64>(m_R[n].size()) == r * r, ERR_INVALID, "R has the wrong shape");
    expv([this, ops](const cplx* v, cplx* res) { heff_bond(ops, v, res); }, m_R[n].data(), r * r, m_coeff,
         -m_sub_dt / 2.0, m_p.krylov_dim, m_p.krylov_tol);
}

// ---- sweeps (tdvp_algorithm_engine.hpp:330-501) ----
// With `expand`, every root-to-leaf factorisation is followed by the downward expansion of
// that bond and every leaf-to-root factorisation by the upward one (:373, :479).  Returns
// whether any bond grew.
bool Engine::forward(bool evolve, bool expand) {
    bool grew = false;
    std::vector<int> visits(m_topo.size(), 0);
    for (int id : m_tour) {
        const Node& nd = m_topo.nodes[id];
        const int tv = visits[id]++;
        if (tv != nd.nmodes()) {
            const int mode = tv;
            if (tv == 0 && !nd.is_root()) apply_from_parent(id);
            if (!nd.is_leaf()) {
                r2l(id, mode);
                if (expand && expand_down(id, mode)) grew = true;
                update_mf(m_topo.nodes[id].children[mode]);
            }
        } else {
            if (evolve) node_update(id);
            if (!nd.is_root()) {
                l2r(id);
                update_spf(id);
                if (evolve) bond_update(id, id);
                apply_to_parent(id);
            }
        }
    }
    return grew;
}

bool Engine::backward(bool evolve, bool expand) {
    bool grew = false;
    std::vector<int> visits(m_topo.size(), 0);
    for (auto it = m_tour.rbegin(); it != m_tour.rend(); ++it) {
        const int id = *it;
        const Node& nd = m_topo.nodes[id];
        const int tv = visits[id]++;
        if (tv != nd.nmodes()) {
            const int mode = nd.nmodes() - (tv + 1);
            if (tv == 0) {
                if (!nd.is_root()) {
                    if (evolve) bond_update(nd.parent, id);
                    apply_from_parent(id);
                }
                if (evolve) node_update(id);
            }
            if (!nd.is_leaf()) {
                r2l(id, mode);
                update_mf(nd.children[mode]);
            }
        } else if (!nd.is_root()) {
            l2r(id);
            if (expand && expand_up(id)) grew = true;
            apply_to_parent(id);
            update_spf(id);
        }
    }
    return grew;
}

bool Engine::sweep(bool evolve, bool expand) {
    const bool f = forward(evolve, expand);
    m_t += m_sub_dt / 2.0;
    const bool b = backward(evolve, expand);
    m_t += m_sub_dt / 2.0;
    return f || b;
}

// tdvp_algorithm_base::prepare_evolution (:257-321): single-particle operators bottom-up, then
// sweeps without propagation until a sweep pair adds no function (one pair for the fixed-step
// integrator, whose expansion never adds)
void Engine::prepare_evolution() {
    for (int n = m_topo.size() - 1; n >= 0; --n)
        if (!m_topo.nodes[n].is_root()) update_spf(n);
    while (sweep(false, m_p.expansion)) {
    }
}

void Engine::step(cplx coeff) {
    m_coeff = coeff;
    if (m_nsteps == 0) {
        double tevolve = 0.0;
        double dt = m_dt / 1e5;
        m_sub_dt = (dt + tevolve <= m_dt ? dt : m_dt - tevolve);
        m_exp_dt = m_dt;  // prepare_evolution runs with the engine's dt at the full step
        prepare_evolution();
        while (tevolve < m_dt) {
            m_sub_dt = (dt + tevolve <= m_dt ? dt : m_dt - tevolve);
            m_exp_dt = m_sub_dt;
            m_t = tevolve;
            sweep(true, m_p.expansion);
            tevolve += dt;
            dt *= 10;
        }
    } else {
        m_sub_dt = m_dt;
        m_exp_dt = m_dt;
        m_t = static_cast<double>(m_nsteps) * m_dt;
        sweep(true, m_p.expansion);
    }
    ++m_nsteps;
}

// matrix_element::operator()(op, mode, psi) (matrix_element.hpp:264-314), computed with
// explicit overlaps at every node.
cplx Engine::expect(int mode, const cplx* op) const {
    const int n = m_topo.size();
    std::vector<std::vector<cplx>> M(n);
    for (int i = n - 1; i >= 0; --i) {
        const Node& nd = m_topo.nodes[i];
        const i64 D = nd.dimen(), r = nd.hrank, sz = nd.size();
        std::vector<cplx> HA(sz), tmp(sz);
        if (nd.is_leaf()) {
            if (nd.leaf_index == mode && op != nullptr) {
                ttnh::PrimOp po;
                po.kind = ttnh::OP_DENSE;
                po.dim = D;
                po.dense.assign(op, op + D * D);
                po.apply(m_A[i].data(), HA.data(), r);
            } else {
                HA = m_A[i];
            }
        } else {
            const cplx* src = m_A[i].data();
            std::vector<cplx> cur(src, src + sz);
            for (int c = 0; c < nd.nmodes(); ++c) {
                i64 L = 1, R = r;
                for (int q = 0; q < c; ++q) L *= nd.dims[q];
                for (int q = c + 1; q < nd.nmodes(); ++q) R *= nd.dims[q];
                mode_contract(M[nd.children[c]].data(), L, nd.dims[c], R, cur.data(), tmp.data());
                cur.swap(tmp);
            }
            HA = cur;
        }
        M[i].assign(r * r, cplx(0.0, 0.0));
        adjoint_mul(m_A[i].data(), HA.data(), D, r, M[i].data());
    }
    return M[0][0];
}

}  // namespace orc
