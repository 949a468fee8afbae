// Spin-boson problem construction: tree topology, SOP Hamiltonian and initial state,
// following the reference's driver (src/spin_boson.cpp:418-649) and harmonic mode
// operators (src/mode_data.hpp:136-361).  Each leaf holds a block of consecutive bath modes
// (one mode per leaf unless block sizes are given).
//
//   H = Hsys (x) 1  +  sum_k  w_k n_k  +  sum_k Scoup (x) g_k (a_k + a_k^dagger)
//
// Terms (sop index r): 0 = Hsys on mode 0; 1+2b = H0 of block b on mode b+1;
// 2+2b = Scoup on mode 0 times the CSR coupling Hc of block b on mode b+1.
#pragma once

#include "common.hpp"
#include "modes.hpp"
#include "sop.hpp"
#include "topology.hpp"

#include <cstdint>

namespace ttnh {

struct SBParams {
    std::vector<double> wk, gk;  // bath frequencies and couplings
    std::vector<i64> boson_dim;  // Fock truncation per mode (one mode per leaf)
    // blocks of consecutive modes, one block per leaf, with the cutoffs
    // min(block_nmax[b], floor(ecut / |w_k|)) (src/spin_boson.cpp:430-441); replaces
    // boson_dim when not empty
    std::vector<i64> block_sizes, block_nmax;
    double ecut = 3000.0;
    i64 nspf = 8, nspf_lower = 8;           // bond dims at the top / bottom bath levels
    i64 nspf_cap = 8, nspf_lower_cap = 8;   // bond-dim capacities
    i64 nhilb = 2;
    std::vector<cplx> hsys;   // nhilb x nhilb row-major
    std::vector<cplx> scoup;  // nhilb x nhilb row-major
    std::vector<cplx> psi0;   // nhilb
    uint64_t seed = 0;
};

struct SBProblem {
    Topology topo;
    SOP sop;
    std::vector<std::vector<cplx>> state;  // per node, row-major dimen x hrank
};

// The leaves' mode blocks: one mode per leaf with boson_dim, or the given block sizes.
std::vector<ModeBlock> sb_blocks(const SBParams& p);

// Spin-boson topology (spin_boson.cpp:447-489): root(1) -> [system node(nhilb) ->
// primitive(nhilb)], [bath node(nhilb) -> balanced binary subtree over the leaf dimensions]
// with level-dependent bond dimensions, sanitised.
Topology sb_topology(const SBParams& p, const std::vector<i64>& leaf_dims);

SBProblem sb_build(const SBParams& p);

// Random orthonormal leaf functions and identity interior nodes, drawn with
// std::mt19937(seed) in pre-order (spin_boson.cpp:578-649).
std::vector<std::vector<cplx>> sb_initial_state(const Topology& t, const std::vector<cplx>& psi0,
                                                uint64_t seed);

// Adds weight * 2 Im <chi|dH/dtheta|psi> to dg (theta = g_k) and dw (theta = w_k) from the
// leaves' transition densities rho_b(j, l) = <chi|(|j><l|)_b|psi> (rho) and the same with the
// coupling operator Scoup inserted on the system leaf (rhoS), both indexed by leaf mode (1 + b;
// d_b x d_b row-major).  dH/dw_k is the occupation of mode k in its block's H0, dH/dg_k is
// Scoup times the lowering + raising part of mode k in its block's Hc (sqrt(n), sqrt(n + 1);
// src/mode_data.hpp:239-253, 309-361); the Fock cutoffs are held fixed.
// Gauss-Legendre nodes per step for the costate gradient's step integrals: the integrand
// <chi|dH|psi> oscillates at most at phi = 1.5 dt (max |w_k| + 2 |Hsys|_F) per step, and the
// n-point rule's error on exp(i phi s) is about (phi / 2)^(2n) / (2n)!; the smallest n in
// [4, 16] that brings this below 1e-13.
int sb_quadrature_nodes(const SBParams& p, double dt);

void sb_density_gradient(const SBParams& p, const std::vector<std::vector<cplx>>& rho,
                         const std::vector<std::vector<cplx>>& rhoS, double weight, std::vector<double>& dg,
                         std::vector<double>& dw);

}  // namespace ttnh
