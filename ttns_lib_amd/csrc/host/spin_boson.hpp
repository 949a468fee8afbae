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

}  // namespace ttnh
