// ORACLE — test infrastructure only.
//
// CPU restatement, in plain C++ complex128, of the reference's one-site
// projector-splitting TDVP step for tree tensor network states.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only as
// the checker or the timed CPU baseline; the product library never links it.
//
// Structure follows the reference engine:
//   integrator ladder         integrators/projector_splitting_integrator.hpp:136-169
//   sweep (forward/backward)  tdvp_core/tdvp_algorithm_engine.hpp:257-501
//   node / bond updates       tdvp_core/projector_splitting_engine.hpp:117-165
//   H-actions                 tdvp_core/projector_splitting_evolution_functions.hpp:7-258
//   mode-k contractions       tdvp_core/kronecker_product_operator_helper.hpp:261-315
//   environments              tdvp_core/mean_field_operator.hpp:37-184,
//                             tdvp_core/single_particle_operator.hpp:26-103
//   gauge moves               tdvp_core/root_to_leaf_decomposition.hpp:112-233,
//                             tdvp_core/leaf_to_root_decomposition.hpp:56-138
//   observables               observables/matrix_element.hpp:264-314
// The integer bookkeeping (tree, sweep order, compiled SOP tables) comes from the
// shared host sources under ttns_lib_amd/csrc/host.
//
// Two pieces of the reference live in un-vendored dependencies (llindoy/utils_cpp
// krylov_integrator and llindoy/linalg's SVD, no pinned versions) and are restated
// from published methods, identically to the HIP path:
//   * expv: Arnoldi with classical Gram-Schmidt applied twice (CGS2), the Saad (1992)
//     a-posteriori error estimate, sub-steps sized by Expokit's rule (Sidje 1998: the
//     model err(s) ~ s^m with safety factor 0.9, growth by at most 2 per sub-step, halving
//     as the last resort), and the small exponential by Pade(13) scaling and squaring
//     (Higham 2005);
//   * gauge factorisations: Householder QR with LAPACK zgeqrf/zlarfg conventions.
// The reference ships no tests and cannot be built here, so their internals (sub-step
// counts, gauge phases) are unpinned against it; they are pinned against mathematics
// (tests/test_oracle_*.py), and end to end against the trajectories the reference printed for
// its own 512-mode runs (tests/test_ref_outputs.py: within 3e-11 after one step).
#pragma once

#include "../ttns_lib_amd/csrc/host/sop.hpp"
#include "../ttns_lib_amd/csrc/host/spin_boson.hpp"
#include "../ttns_lib_amd/csrc/host/topology.hpp"

#include <functional>
#include <random>

namespace orc {

using ttnh::cplx;
using ttnh::i64;

// ---- dense helpers (row-major) ----
// res(l,i,r) = sum_j op(i,j) A(l,j,r) on the (L,d,R) view
void mode_contract(const cplx* op, i64 L, i64 d, i64 R, const cplx* A, cplx* res);
// res = X * M^T  (X: rows x r, M: r x r)
void mul_transposed(const cplx* X, i64 rows, i64 r, const cplx* M, cplx* res);
// res = A^H B  (A, B: rows x r)
void adjoint_mul(const cplx* A, const cplx* B, i64 rows, i64 r, cplx* res);
// res(i,j) = sum_{l,r} conj(A(l,i,r)) B(l,j,r) on the (L,d,R) view
void outer_contract(const cplx* A, const cplx* B, i64 L, i64 d, i64 R, cplx* res);
// Householder QR (zgeqrf + zungqr): A (m x n) = Q (m x n) R (n x n).  For m < n the
// reference zero-pads (decomposition_engine.hpp:124-158): Q gets n-m zero columns and
// R zero rows.
void householder_qr(const cplx* A, i64 m, i64 n, cplx* Q, cplx* R);
// exp(M) for a small dense n x n matrix (Pade 13, scaling and squaring)
void expm(const cplx* M, int n, cplx* E);

struct KrylovStats {
    i64 napply = 0;
    i64 nsub = 0;
};
// x <- exp(c * tau * H) x, H given by its action; returns the number of H applications.
i64 expv(const std::function<void(const cplx*, cplx*)>& H, cplx* x, i64 n, cplx c, double tau, int kdim,
         double tol);

// ---- subspace expansion helpers (expansion.cpp) ----
// singular values of a dense n x n matrix, descending (one-sided Jacobi, Hestenes 1958)
std::vector<double> singular_values(const cplx* M, i64 n);
// eigenvalues (descending) and eigenvectors (columns of Z, n x n row-major) of a real
// symmetric tridiagonal matrix (diagonal a, off-diagonal b), cyclic Jacobi rotations
void tridiag_eigen(const std::vector<double>& a, const std::vector<double>& b, std::vector<double>& w,
                   std::vector<double>& Z);

// ---- the TDVP engine ----
struct Params {
    int krylov_dim = 12;
    double krylov_tol = 1e-12;
    double dt = 0.05;
    // subspace_expansion_projector_splitting_integrator (projector_splitting_integrator.hpp:225-397)
    bool expansion = false;
    double spawning_threshold = -1.0;       // engine defaults (subspace_expansion_..._engine.hpp:45)
    double unoccupied_threshold = -1.0;
    i64 minimum_unoccupied = 0;
    i64 neigenvalues = 2;
    double eigenvalue_tolerance = 1e-6;
    bool only_apply_when_no_unoccupied = false;
    unsigned rng_seed = 5489u;              // std::mt19937 default seed (two_site_energy_variations.hpp:34)
};

class Engine {
public:
    Engine(const ttnh::Topology& topo, const ttnh::SOP& sop, const Params& p);

    std::vector<cplx>& A(int n) { return m_A[n]; }
    const ttnh::Topology& topo() const { return m_topo; }
    const ttnh::CompiledSOP& compiled() const { return m_ops; }
    const ttnh::SOP& sop() const { return m_sop; }

    // one integrator call (projector_splitting_integrator.hpp:136-169, 281-326)
    void step(cplx coeff);
    double time() const { return m_dt * static_cast<double>(m_nsteps); }
    i64 nsteps() const { return m_nsteps; }
    i64 nh_applications() const { return m_nh; }
    // engine().Nonesite() / Ntwosite() (subspace_expansion_projector_splitting_engine.hpp:513-514)
    i64 n_onesite() const { return m_onesite; }
    i64 n_twosite() const { return m_twosite; }
    void set_dt(double dt) { m_dt = dt; }
    double dt() const { return m_dt; }
    void reset_nsteps() { m_nsteps = 0; }
    const Params& params() const { return m_p; }
    void set_params(const Params& p) { m_p = p; }
    // new SOP coefficients: only the coefficient fields of the compiled tables change
    i64 nterms() const { return m_sop.nterms; }
    void set_coefficients(const std::vector<cplx>& c) {
        m_sop.coeff = c;
        m_ops = ttnh::compile_sop(m_topo, m_sop);
    }
    // normal draws consumed by the expansion so far (the position in the std::mt19937 +
    // std::normal_distribution stream)
    i64 normals_used() const { return m_normals; }
    // log of the two-site expansion attempts: records of 20 doubles (child node, direction 0
    // down / 1 up, Ritz values kept, functions added, 16 Ritz values), at most cap records
    void set_log(i64 cap) {
        m_log_cap = cap;
        m_log.clear();
    }
    const std::vector<double>& log() const { return m_log; }

    // <psi|O_mode|psi> with O acting on one primitive mode (full contraction, no
    // orthonormality assumption) and <psi|psi>
    cplx expect(int mode, const cplx* op) const;

    // building blocks, exposed for tests
    void heff_node(int n, const cplx* v, cplx* res) const;  // coefficient_evolution_{leaf,branch}
    void heff_bond(int n, const cplx* v, cplx* res) const;  // full_hamiltonian_action
    void update_spf(int n);                                 // single_particle_operator_engine
    void update_mf(int n);                                  // mean_field_operator_engine
    const std::vector<cplx>& spf(int n, int t) const { return m_spf[n][t]; }
    const std::vector<cplx>& mf(int n, int t) const { return m_mf[n][t]; }
    std::vector<cplx>& mf_mut(int n, int t) { return m_mf[n][t]; }
    std::vector<cplx>& spf_mut(int n, int t) { return m_spf[n][t]; }
    const std::vector<cplx>& node(int n) const { return m_A[n]; }

    // the whole integrator state (node tensors, gauge and environment matrices, counters, dt)
    struct Snapshot {
        std::vector<std::vector<cplx>> A, R;
        std::vector<std::vector<std::vector<cplx>>> spf, mf;
        i64 nsteps = 0, nh = 0;
        double dt = 0.0, t = 0.0;
    };
    Snapshot snapshot() const { return {m_A, m_R, m_spf, m_mf, m_nsteps, m_nh, m_dt, m_t}; }
    void restore(const Snapshot& s) {
        m_A = s.A;
        m_R = s.R;
        m_spf = s.spf;
        m_mf = s.mf;
        m_nsteps = s.nsteps;
        m_nh = s.nh;
        m_dt = s.dt;
        m_t = s.t;
    }

private:
    void kpo(int n, const std::vector<std::pair<int, int>>& list, const cplx* v, cplx* out, cplx* tmp) const;
    void kron_siblings(int n, const std::vector<std::pair<int, int>>& sib, const cplx* Ap, cplx* out,
                       cplx* tmp) const;
    void prepare_evolution();
    bool sweep(bool evolve, bool expand);
    bool forward(bool evolve, bool expand);
    bool backward(bool evolve, bool expand);
    void node_update(int n);
    void bond_update(int n, int with_ops_of);
    void r2l(int n, int mode);
    void l2r(int n);
    void apply_from_parent(int n);
    void apply_to_parent(int n);

    // subspace expansion (subspace_expansion_projector_splitting_engine.hpp:194-510,
    // two_site_energy_variations.hpp:49-562)
    bool expand_down(int p, int mode);
    bool expand_up(int c);
    int n_unoccupied(const std::vector<cplx>& R, i64 n) const;
    void trial_vector(int n, int mode, std::vector<cplx>& x);
    void two_site_lower(int c, int ind, const std::vector<cplx>& r, std::vector<cplx>& out) const;
    void two_site_upper(int c, int ind, std::vector<cplx>& out) const;
    std::vector<int> two_site_terms(int c) const;
    void eigs(const std::function<void(const cplx*, cplx*)>& F, std::vector<cplx> x, int kd, double tol,
              std::vector<double>& S, std::vector<std::vector<cplx>>& V) const;
    void grow(int c, i64 nadd);   // child c's bond dimension += nadd (shapes only)
    void log_attempt(int c, int dir, const std::vector<double>& S, i64 nadd);

    ttnh::Topology m_topo;
    ttnh::SOP m_sop;
    ttnh::CompiledSOP m_ops;
    std::vector<int> m_tour;
    Params m_p;
    std::vector<std::vector<cplx>> m_A, m_R;
    std::vector<std::vector<std::vector<cplx>>> m_spf, m_mf;
    double m_dt = 0.0, m_sub_dt = 0.0, m_t = 0.0;
    double m_exp_dt = 0.0;    // the engine's dt() seen by the spawning criterion
    cplx m_coeff = cplx(0.0, -1.0);
    i64 m_nsteps = 0, m_nh = 0;
    i64 m_onesite = 0, m_twosite = 0, m_normals = 0;
    std::vector<double> m_log;
    i64 m_log_cap = 0;
    std::mt19937 m_rng;
    std::normal_distribution<double> m_ndist{0.0, 1.0};
};

// ---- gradients by the costate method (adjoint.cpp; the reference has no adjoint) ----
// rho[m] = <bra|(|j><l|)_m|ket> for every leaf mode m (d_m x d_m, row-major), and rhoS[m] the same
// with S (d x d) inserted on leaf mode `ins` (rhoS[ins] empty).  Full contractions of the two
// trees; no orthonormality is assumed.
void transition_densities(const Engine& ket, const Engine& bra, int ins, const cplx* S,
                          std::vector<std::vector<cplx>>& rho, std::vector<std::vector<cplx>>& rhoS);

// J = Re<psi(T)|O|psi(T)> after nsteps real-time integrator calls from a state that has not
// been stepped, and dJ/dg_k, dJ/dw_k of the spin-boson model p (adjoint.cpp has the method).
double sb_adjoint(Engine& ket, const ttnh::SBParams& p, i64 nsteps, i64 every, int nquad, int mode, const cplx* O,
                  std::vector<double>& dg, std::vector<double>& dw);

}  // namespace orc
