// ORACLE — test infrastructure only (see oracle.hpp).  Gradient of J = Re<psi(T)|O|psi(T)>
// with respect to the spin-boson couplings g_k and frequencies w_k by the costate
// (adjoint-state) method.  The reference has no adjoint; the forward it differentiates is
// the integrator of engine.cpp (projector_splitting_integrator.hpp:136-169), and g_k, w_k
// enter the leaves' operators as src/mode_data.hpp:239-253, 309-361 builds them.
//
//   dJ/dtheta = 2 int_0^T Im <chi(t)| dH/dtheta |psi(t)> dt,
//   chi(t) = exp(iH(T - t)) O_s psi(T),  O_s = (O + O^H) / 2,
//
// chi is a tree state of the same shape, propagated backwards in time by the same TDVP step
// (coefficient +i).  Each step [t_n, t_n + dt] is integrated by n-point Gauss-Legendre
// quadrature: psi at t_n + s dt is one step of length s dt from psi(t_n), chi one backward
// step of length (1 - s) dt from chi(t_n + dt).  psi is checkpointed every K steps and
// recomputed from the checkpoint inside each interval.  The matrix elements come from the
// leaves' transition densities, full contractions of the two trees.
//
// This differentiates the time evolution the TDVP step approximates, not the step's
// floating-point gauge choices: reverse-mode differentiation through the QR gauge moves
// divides by their pivots, which reach 1e-17..1e-25 of the largest in the first steps from
// a product state, and is unusable in double precision.  Where the tree represents the
// dynamics exactly the two agree to the quadrature and Krylov tolerances.
#include "oracle.hpp"

#include <algorithm>
#include <cmath>

namespace orc {

using ttnh::check;
using ttnh::ERR_INVALID;
using ttnh::Node;

namespace {

// (L, R) of node nd's rank-3 view around child slot k
void slot_view(const Node& nd, int k, i64& L, i64& R) {
    L = 1;
    R = nd.hrank;
    for (int q = 0; q < k; ++q) L *= nd.dims[q];
    for (int q = k + 1; q < nd.nmodes(); ++q) R *= nd.dims[q];
}

// the environment of child slot k of node p: E_c(i, j) = sum conj(Abra(l,i,r)) Y(l,j,r),
// Y = ket tensor with every other slot contracted with its transfer matrix and the hrank
// index with Ep (Y(x, a) = sum_b Yk(x, b) Ep(a, b))
std::vector<cplx> child_env(const Node& nd, int k, const cplx* Aket, const cplx* Abra,
                            const std::vector<const std::vector<cplx>*>& T, const std::vector<cplx>& Ep) {
    const i64 sz = nd.size(), D = nd.dimen(), r = nd.hrank;
    std::vector<cplx> Y(Aket, Aket + sz), tmp(sz);
    for (int s = 0; s < nd.nmodes(); ++s) {
        if (s == k) continue;
        i64 L, R;
        slot_view(nd, s, L, R);
        mode_contract(T[s]->data(), L, nd.dims[s], R, Y.data(), tmp.data());
        Y.swap(tmp);
    }
    mul_transposed(Y.data(), D, r, Ep.data(), tmp.data());
    i64 L, R;
    slot_view(nd, k, L, R);
    const i64 d = nd.dims[k];
    std::vector<cplx> E(d * d);
    outer_contract(Abra, tmp.data(), L, d, R, E.data());
    return E;
}

// rho(j, l) = sum conj(Abra(j, a)) E(a, b) Aket(l, b) on a leaf (d x r)
std::vector<cplx> leaf_density(const cplx* Aket, const cplx* Abra, const std::vector<cplx>& E, i64 d, i64 r) {
    std::vector<cplx> Z(d * r), rho(d * d);
    mul_transposed(Aket, d, r, E.data(), Z.data());
    for (i64 j = 0; j < d; ++j)
        for (i64 l = 0; l < d; ++l) {
            cplx s(0.0, 0.0);
            for (i64 a = 0; a < r; ++a) s += std::conj(Abra[j * r + a]) * Z[l * r + a];
            rho[j * d + l] = s;
        }
    return rho;
}

}  // namespace

void transition_densities(const Engine& ket, const Engine& bra, int ins, const cplx* S,
                          std::vector<std::vector<cplx>>& rho, std::vector<std::vector<cplx>>& rhoS) {
    const ttnh::Topology& tp = ket.topo();
    const int n = tp.size();
    check(bra.topo().size() == n, ERR_INVALID, "transition densities: trees differ");
    // nodes whose subtree holds the insertion leaf
    std::vector<char> onpath(n, 0);
    for (int i = 0; i < n; ++i)
        if (tp.nodes[i].is_leaf() && tp.nodes[i].leaf_index == ins)
            for (int q = i; q >= 0; q = tp.nodes[q].parent) onpath[q] = 1;
    // bottom-up transfer matrices T = Abra^H (children) Aket, TS with S on the insertion leaf
    std::vector<std::vector<cplx>> T(n), TS(n);
    for (int i = n - 1; i >= 0; --i) {
        const Node& nd = tp.nodes[i];
        const i64 D = nd.dimen(), r = nd.hrank, sz = nd.size();
        const cplx* Ak = ket.node(i).data();
        const cplx* Ab = bra.node(i).data();
        check(static_cast<i64>(bra.node(i).size()) == sz, ERR_INVALID, "transition densities: shapes differ");
        T[i].assign(r * r, cplx(0.0, 0.0));
        if (nd.is_leaf()) {
            adjoint_mul(Ab, Ak, D, r, T[i].data());
            if (onpath[i]) {
                std::vector<cplx> SA(sz);
                for (i64 a = 0; a < D; ++a)
                    for (i64 c = 0; c < r; ++c) {
                        cplx s(0.0, 0.0);
                        for (i64 b = 0; b < D; ++b) s += S[a * D + b] * Ak[b * r + c];
                        SA[a * r + c] = s;
                    }
                TS[i].assign(r * r, cplx(0.0, 0.0));
                adjoint_mul(Ab, SA.data(), D, r, TS[i].data());
            }
            continue;
        }
        for (int pass = 0; pass < (onpath[i] ? 2 : 1); ++pass) {
            std::vector<cplx> Y(Ak, Ak + sz), tmp(sz);
            for (int k = 0; k < nd.nmodes(); ++k) {
                const int ch = nd.children[k];
                i64 L, R;
                slot_view(nd, k, L, R);
                const std::vector<cplx>& M = (pass == 1 && onpath[ch]) ? TS[ch] : T[ch];
                mode_contract(M.data(), L, nd.dims[k], R, Y.data(), tmp.data());
                Y.swap(tmp);
            }
            std::vector<cplx>& out = pass == 0 ? T[i] : TS[i];
            out.assign(r * r, cplx(0.0, 0.0));
            adjoint_mul(Ab, Y.data(), D, r, out.data());
        }
    }
    // top-down environments E (nothing inserted) and ES (S inserted outside the subtree)
    std::vector<std::vector<cplx>> E(n), ES(n);
    {
        const i64 r0 = tp.nodes[0].hrank;
        E[0].assign(r0 * r0, cplx(0.0, 0.0));
        for (i64 a = 0; a < r0; ++a) E[0][a * r0 + a] = cplx(1.0, 0.0);
    }
    for (int p = 0; p < n; ++p) {
        const Node& nd = tp.nodes[p];
        if (nd.is_leaf()) continue;
        const cplx* Ak = ket.node(p).data();
        const cplx* Ab = bra.node(p).data();
        std::vector<const std::vector<cplx>*> Tp(nd.nmodes()), TSp(nd.nmodes());
        for (int k = 0; k < nd.nmodes(); ++k) {
            const int ch = nd.children[k];
            Tp[k] = &T[ch];
            TSp[k] = onpath[ch] ? &TS[ch] : &T[ch];
        }
        for (int k = 0; k < nd.nmodes(); ++k) {
            const int c = nd.children[k];
            E[c] = child_env(nd, k, Ak, Ab, Tp, E[p]);
            if (onpath[c]) continue;
            if (onpath[p])
                ES[c] = child_env(nd, k, Ak, Ab, TSp, E[p]);
            else if (!ES[p].empty())
                ES[c] = child_env(nd, k, Ak, Ab, Tp, ES[p]);
        }
    }
    rho.assign(tp.nmodes(), {});
    rhoS.assign(tp.nmodes(), {});
    for (int i = 0; i < n; ++i) {
        const Node& nd = tp.nodes[i];
        if (!nd.is_leaf()) continue;
        const i64 d = nd.dimen(), r = nd.hrank;
        rho[nd.leaf_index] = leaf_density(ket.node(i).data(), bra.node(i).data(), E[i], d, r);
        if (!ES[i].empty()) rhoS[nd.leaf_index] = leaf_density(ket.node(i).data(), bra.node(i).data(), ES[i], d, r);
    }
}

double sb_adjoint(Engine& ket, const ttnh::SBParams& p, i64 nsteps, i64 every, int nquad, int mode, const cplx* O,
                  std::vector<double>& dg, std::vector<double>& dw) {
    const ttnh::Topology& tp = ket.topo();
    check(!ket.params().expansion, ERR_INVALID, "adjoint: the fixed-step integrator only");
    check(ket.nsteps() == 0, ERR_INVALID, "adjoint: the state must not have been stepped");
    check(nsteps >= 1 && every >= 1 && nquad >= 0 && nquad <= 16, ERR_INVALID,
          "adjoint: nsteps and the checkpoint interval must be positive, 0-16 quadrature nodes");
    if (nquad == 0) nquad = ttnh::sb_quadrature_nodes(p, ket.dt());
    check(mode >= 0 && mode < tp.nmodes(), ERR_INVALID, "adjoint: mode out of range");
    check(static_cast<i64>(p.wk.size()) == tp.nmodes() - 1, ERR_INVALID, "adjoint: the model does not match the tree");
    const cplx mi(0.0, -1.0), pi(0.0, 1.0);
    const double dt = ket.dt();
    // forward with checkpoints
    std::vector<Engine::Snapshot> ck;
    for (i64 s = 0; s < nsteps; ++s) {
        if (s % every == 0) ck.push_back(ket.snapshot());
        ket.step(mi);
    }
    const Engine::Snapshot fin = ket.snapshot();
    const i64 dm = tp.mode_dims[mode];
    const double J = ket.expect(mode, O).real();
    // chi(T) = O_s psi(T) on a second engine
    Engine chi(tp, ket.sop(), ket.params());
    std::vector<cplx> Os(dm * dm);
    for (i64 a = 0; a < dm; ++a)
        for (i64 b = 0; b < dm; ++b) Os[a * dm + b] = 0.5 * (O[a * dm + b] + std::conj(O[b * dm + a]));
    for (int i = 0; i < tp.size(); ++i) {
        chi.A(i) = ket.node(i);
        const Node& nd = tp.nodes[i];
        if (nd.is_leaf() && nd.leaf_index == mode) {
            const i64 r = nd.hrank;
            std::vector<cplx>& A = chi.A(i);
            std::vector<cplx> out(A.size());
            for (i64 a = 0; a < dm; ++a)
                for (i64 c = 0; c < r; ++c) {
                    cplx s(0.0, 0.0);
                    for (i64 b = 0; b < dm; ++b) s += Os[a * dm + b] * A[b * r + c];
                    out[a * r + c] = s;
                }
            A = out;
        }
    }
    chi.set_dt(dt);
    // Gauss-Legendre nodes and weights on [0, 1]
    std::vector<double> xs(nquad), ws(nquad);
    {
        const double pi_ = std::acos(-1.0);
        for (int i = 0; i < nquad; ++i) {
            double z = std::cos(pi_ * (i + 0.75) / (nquad + 0.5)), dp = 1.0;
            for (int it = 0; it < 100; ++it) {
                double p0 = 1.0, p1 = z;
                for (int k = 2; k <= nquad; ++k) {
                    const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                    p0 = p1;
                    p1 = p2;
                }
                dp = nquad * (z * p1 - p0) / (z * z - 1.0);
                const double dz = p1 / dp;
                z -= dz;
                if (std::abs(dz) < 1e-16) break;
            }
            xs[i] = 0.5 * (1.0 - z);
            ws[i] = 1.0 / ((1.0 - z * z) * dp * dp);
        }
    }
    const cplx* Sc = p.scoup.data();
    dg.assign(p.wk.size(), 0.0);
    dw.assign(p.wk.size(), 0.0);
    std::vector<std::vector<cplx>> rho, rhoS;
    for (i64 g = static_cast<i64>(ck.size()) - 1; g >= 0; --g) {
        ket.restore(ck[g]);
        const i64 s0 = g * every, s1 = std::min(nsteps, s0 + every);
        std::vector<Engine::Snapshot> starts;
        for (i64 s = s0; s < s1; ++s) {
            starts.push_back(ket.snapshot());
            if (s + 1 < s1) ket.step(mi);
        }
        for (i64 s = s1 - 1; s >= s0; --s) {
            const Engine::Snapshot c1 = chi.snapshot();  // chi(t_{s+1})
            for (int q = 0; q < nquad; ++q) {
                ket.restore(starts[s - s0]);
                ket.set_dt(xs[q] * dt);
                ket.step(mi);
                chi.restore(c1);
                chi.set_dt((1.0 - xs[q]) * dt);
                chi.step(pi);
                transition_densities(ket, chi, 0, Sc, rho, rhoS);
                ttnh::sb_density_gradient(p, rho, rhoS, ws[q] * dt, dg, dw);
            }
            chi.restore(c1);
            chi.set_dt(dt);
            chi.step(pi);
        }
    }
    ket.restore(fin);
    return J;
}

}  // namespace orc
