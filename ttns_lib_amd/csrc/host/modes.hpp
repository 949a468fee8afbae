// Harmonic bath modes combined on one leaf, and the partition of a bath into such blocks,
// as the reference's spin-boson driver builds them (src/mode_data.hpp, called from
// src/spin_boson.cpp:418-441 and :515-538).
//
//   * Fock cutoff of mode k: nper_k = min(nmax, floor(ecut / |w_k|))             (:146, :788)
//   * block states: the occupations (n_0, ..., n_{N-1}) with mode 0 most significant, i.e.
//     lexicographic order (populate_state_vector, :797-812): state s has n_k = (s / stride_k)
//     mod nper_k with stride_{N-1} = 1
//   * H0 = diag(sum_k w_k n_k)                                                    (:239-253)
//   * Hc row s: the lowering entries sqrt(n_k) g_k in mode order, then the raising entries
//     sqrt(n_k + 1) g_k in reverse mode order                                     (:309-361)
//   * partition_modes: greedy blocks of consecutive modes                         (:818-886)
//
// The block has prod_k nper_k states: the states the reference's enumeration writes.  Its
// state table is sized nper_0^N (get_nstates_in_cut, :785-792), which is the same number
// whenever the block's modes share one cutoff (always for a block of one mode); for a block
// whose modes have different cutoffs the reference leaves the extra rows unwritten, and this
// build does not reproduce them.
#pragma once

#include "common.hpp"
#include "sop.hpp"

#include <vector>

namespace ttnh {

struct ModeBlock {
    std::vector<double> w, g;
    std::vector<i64> nper;    // Fock states kept per mode
    std::vector<i64> stride;  // state index stride of each mode's occupation
    i64 nstates = 0;
    i64 occ(i64 s, size_t k) const { return (s / stride[k]) % nper[k]; }
};

// min(nmax, floor(ecut / |w|)) (mode_data.hpp:146)
i64 mode_cutoff(double w, double ecut, i64 nmax);

// The block of modes w[0..n) with per-mode cutoffs (construct_basis_topology,
// mode_data.hpp:136-228).
ModeBlock mode_block(const std::vector<double>& w, const std::vector<double>& g, double ecut, i64 nmax);

// A block of one mode with an explicit Fock truncation d.
ModeBlock single_mode(double w, double g, i64 d);

PrimOp block_h0(const ModeBlock& b);
PrimOp block_hc(const ModeBlock& b);

// Sizes of consecutive blocks (mode_data.hpp:818-886).  wk sorted by magnitude; a block
// starting at mode c grows while nper(w_c)^size <= (w_c > wc ? nmax4 : nmax), with the cutoff
// taken at (w_c > wc ? nmax3 : nmax2); a block of one mode is kept even above that bound.
std::vector<i64> partition_modes(const std::vector<double>& wk, double ecut, double wc, i64 nmax, i64 nmax4,
                                 i64 nmax2, i64 nmax3);

}  // namespace ttnh
