// Spin-boson driver written against the reference's C++ API (include/ttns_amd.hpp), the
// way the reference's own driver sets up and runs a trajectory (src/spin_boson.cpp:447-649
// for the tree, operator binding and initial state, :41-107 for the step loop, whose calls
// run() below makes: tdvp(A, H, -i), tdvp.t(), tdvp.mel()(op, 0, A), tdvp.engine().Nonesite()
// / Ntwosite(), tdvp.nh_applications() and every non-root node's hrank()).  Used by the tests
// as the drop-in check: its rows must equal the oracle's on the same model.
//
// usage: sb_driver nmodes boson_dim nspf nsteps dt alpha wc seed krylov_dim
//                  [ecut high_frequency_bound target_dimension maximum_dimension_high_frequency]
//                  [--capacity C] [--spawn S] [--unoccupied U] [--min-unoccupied M]
//                  [--dt2 DT --dt2-at K] [--coeff0 F --coeff0-at K] [--reset-at K] [--clear-at K]
//                  [--checkpoint-at K --restore-at J]
// Rows: t <sigma_z> Nonesite Ntwosite nh_applications/size hranks...  The first row is t = 0
// (no counters).  --capacity lets the bonds grow to C (the expansion integrator's tree
// capacity); --dt2 sets tdvp.dt() = DT after K steps; --coeff0 multiplies the coefficient of
// term 0 (the system Hamiltonian) by F after K steps; --reset-at calls reset_nsteps() after K
// steps; --clear-at calls clear(), restores the parameters and initialise()s again after K
// steps (each of these with the state the run has reached); --checkpoint-at takes a checkpoint
// after K steps and --restore-at goes back to it after J > K steps (the rows after J then
// repeat those after K).
// Bath: w_k = wc * 4 k / n (k = 1..n), g_k = sqrt(J(w_k) dw / pi), J(w) = pi/2 alpha w e^{-w/wc}.
// The modes are split into leaves with partition_modes and each leaf's operators come from
// harmonic_mode_combination, as in the reference driver (src/spin_boson.cpp:418-441,
// 515-538); boson_dim is the maximum dimension.  Without the optional arguments every leaf
// holds one mode (target dimension = boson_dim, no high-frequency modes).
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <iostream>
#include <map>
#include <random>
#include <string>
#include <vector>

#include "ttns_amd.hpp"

using namespace ttns;
using complex_type = std::complex<double>;
using backend_type = blas_backend;

template <class integ_type, class op_type>
void run(httensor<complex_type, backend_type>& A, integ_type& tdvp, sop_operator<complex_type, backend_type>& H,
         const op_type& mop, int nsteps, const std::function<void(int)>& between) {
    auto row_hranks = [&]() {
        for (auto& c : A)
            if (!c.is_root()) std::printf(" %lld", static_cast<long long>(c.hrank()));
        std::printf("\n");
    };
    std::printf("%.17g %.17g", 0.0, std::real(tdvp.mel()(mop, 0, A)));
    row_hranks();
    for (int s = 0; s < nsteps; ++s) {
        tdvp(A, H, complex_type(0, -1));
        std::printf("%.17g %.17g %zu %zu %.17g", tdvp.t(), std::real(tdvp.mel()(mop, 0, A)),
                    static_cast<size_t>(tdvp.engine().Nonesite()), static_cast<size_t>(tdvp.engine().Ntwosite()),
                    static_cast<double>(tdvp.nh_applications()) / static_cast<double>(A.size()));
        row_hranks();
        between(s + 1);
    }
}

int main(int argc, char** argv) {
    std::vector<char*> pos;
    std::map<std::string, double> opt;
    for (int i = 1; i < argc; ++i) {
        if (std::strncmp(argv[i], "--", 2) == 0 && i + 1 < argc) {
            opt[argv[i] + 2] = std::atof(argv[i + 1]);
            ++i;
        } else {
            pos.push_back(argv[i]);
        }
    }
    auto get = [&](const char* k, double d) { return opt.count(k) ? opt[k] : d; };
    argc = static_cast<int>(pos.size()) + 1;
    pos.insert(pos.begin(), argv[0]);
    argv = pos.data();
    if (argc < 10) {
        std::cerr << "usage: sb_driver nmodes boson_dim nspf nsteps dt alpha wc seed krylov_dim"
                     " [ecut high_frequency_bound target_dimension maximum_dimension_high_frequency]\n";
        return 2;
    }
    const size_type nblocks_arg = std::strtoul(argv[1], nullptr, 10);
    const size_type nmax = std::strtoul(argv[2], nullptr, 10);
    const size_type nspf = std::strtoul(argv[3], nullptr, 10), nspf_lower = nspf;
    const int nsteps = std::atoi(argv[4]);
    const double dt = std::atof(argv[5]), alpha = std::atof(argv[6]), wc = std::atof(argv[7]);
    const unsigned seed = static_cast<unsigned>(std::strtoul(argv[8], nullptr, 10));
    const size_type krylov_dim = std::strtoul(argv[9], nullptr, 10);
    const bool blocks = argc >= 14;
    const double ecut = blocks ? std::atof(argv[10]) : 3000.0;
    const double hfb = blocks ? std::atof(argv[11]) : 1e300;
    const size_type ntarget = blocks ? std::strtoul(argv[12], nullptr, 10) : nmax;
    const size_type nmaxlargefreq = blocks ? std::strtoul(argv[13], nullptr, 10) : nmax;
    try {
        const size_type nhilb = 2;
        const size_type N = nblocks_arg;
        std::vector<double> wk(N), gk(N);
        for (size_type k = 0; k < N; ++k) wk[k] = wc * 4.0 * static_cast<double>(k + 1) / N;
        const double dw = N > 1 ? wk[1] - wk[0] : wk[0];
        for (size_type k = 0; k < N; ++k) {
            const double J = 0.5 * std::acos(-1.0) * alpha * wk[k] * std::exp(-wk[k] / wc);
            gk[k] = std::sqrt(J * dw / std::acos(-1.0));
        }
        std::vector<size_t> partitions;
        partition_modes(wk, ecut, 2 * hfb, partitions, ntarget, 250, nmax, nmaxlargefreq);
        const size_type nblocks = partitions.size();
        std::vector<size_type> mode_dimensions(nblocks + 1), nmodes(nblocks), nskip(nblocks);
        mode_dimensions[0] = nhilb;
        harmonic_mode_combination<double> mode_combination;
        for (size_type i = 0, counter = 0; i < nblocks; ++i) {
            nskip[i] = counter;
            const size_type nmaxt = wk[counter] > 2 * hfb ? nmaxlargefreq : nmax;
            mode_combination.set_wk(wk.begin() + counter, wk.begin() + counter + partitions[i]);
            mode_combination.set_gk(gk.begin() + counter, gk.begin() + counter + partitions[i]);
            nmodes[i] = mode_dimensions[i + 1] = mode_combination.construct_basis_topology(ecut, nmaxt);
            counter += partitions[i];
        }

        // topology: root -> [system node -> primitive], [bath node -> balanced binary subtree]
        auto build = [&](size_type top, size_type bottom) {
            ntree<size_type> t{};
            t.insert(1);
            t().insert(mode_dimensions[0]);
            t()[0].insert(mode_dimensions[0]);
            t().insert(mode_dimensions[0]);
            const size_type nlevels = static_cast<size_type>(std::log2(nmodes.size()));
            ntree_builder<size_type>::htucker_subtree(t()[1], nmodes, 2, [top, bottom, nlevels](size_type l) {
                size_type ret = 0;
                if (l >= nlevels)
                    ret = bottom;
                else if (l == 0)
                    ret = top;
                else {
                    const double rmax = std::log2(top), rmin = std::log2(bottom);
                    ret = static_cast<size_type>(std::pow(2.0, ((nlevels - l) * (rmax - rmin)) / nlevels + rmin));
                }
                return ret;
            });
            ntree_builder<size_type>::sanitise_tree(t);
            return t;
        };
        const size_type cap = static_cast<size_type>(get("capacity", static_cast<double>(nspf)));
        ntree<size_type> topology = build(nspf, nspf_lower), capacity = build(cap, cap);
        httensor<complex_type, backend_type> A(topology, capacity);

        // Hamiltonian: Hsys on mode 0, Scoup on mode 0 for terms 2+2i, H0_i and Hc_i on mode i+1
        matrix<complex_type> Hsys(2, 2), Scoup(2, 2), Sz(2, 2);
        Hsys(0, 1) = Hsys(1, 0) = 1.0;
        Scoup(0, 0) = 1.0;
        Scoup(1, 1) = -1.0;
        Sz = Scoup;
        std::vector<size_type> terms_per_mode(nblocks + 1, 2);
        sop_operator<complex_type, backend_type> H(1 + 2 * nblocks, mode_dimensions, terms_per_mode);
        H.bind(ops::dense_matrix_operator<complex_type, backend_type>{Hsys}, {0}, 0);
        std::vector<size_type> sigma_z_indices(nblocks);
        for (size_type i = 0; i < nblocks; ++i) sigma_z_indices[i] = 1 + 2 * i + 1;
        H.bind(ops::dense_matrix_operator<complex_type, backend_type>{Scoup}, sigma_z_indices, 0);
        for (size_type i = 0; i < nblocks; ++i) {
            const size_type nmaxt = wk[nskip[i]] > 2 * hfb ? nmaxlargefreq : nmax;
            mode_combination.set_wk(wk.begin() + nskip[i], wk.begin() + nskip[i] + partitions[i]);
            mode_combination.set_gk(gk.begin() + nskip[i], gk.begin() + nskip[i] + partitions[i]);
            mode_combination.construct_basis_topology(ecut, nmaxt);
            diagonal_matrix<complex_type> H0;
            mode_combination.H0(H0);
            H.bind(ops::diagonal_matrix_operator<complex_type, backend_type>(H0), {1 + 2 * i}, i + 1);
            csr_matrix<complex_type> Hc;
            mode_combination.Hc(Hc);
            H.bind(ops::sparse_matrix_operator<complex_type, backend_type>(Hc), {1 + 2 * i + 1}, i + 1);
        }

        // initial state: spin up x vacuum, random orthonormal unoccupied functions
        std::mt19937 rng(seed);
        for (auto& c : A) {
            if (c.is_leaf()) {
                const size_type D = c().shape(0), r = c().shape(1);
                matrix<complex_type> ct(r, D);
                if (c.leaf_index() == 0) {
                    const complex_type psi0[2] = {1.0, 0.0};
                    for (size_type j = 0; j < D; ++j) ct(0, j) = psi0[j];
                } else {
                    ct(0, 0) = 1.0;
                }
                std::uniform_real_distribution<double> dist(0, 2.0 * std::acos(-1.0));
                std::normal_distribution<double> length_dist(0, 1);
                for (size_type i = 1; i < r; ++i) {
                    bool vector_generated = false;
                    while (!vector_generated) {
                        for (size_type j = 0; j < D; ++j) {
                            const double theta = dist(rng);
                            ct(i, j) = length_dist(rng) * complex_type(std::cos(theta), std::sin(theta));
                        }
                        double nrm = 0.0;
                        for (size_type j = 0; j < D; ++j) nrm += std::norm(ct(i, j));
                        for (size_type j = 0; j < D; ++j) ct(i, j) /= std::sqrt(nrm);
                        for (size_type k = 0; k < i; ++k) {
                            complex_type ov(0.0, 0.0);
                            for (size_type j = 0; j < D; ++j) ov += std::conj(ct(k, j)) * ct(i, j);
                            for (size_type j = 0; j < D; ++j) ct(i, j) -= ov * ct(k, j);
                        }
                        double nn = 0.0;
                        for (size_type j = 0; j < D; ++j) nn += std::norm(ct(i, j));
                        nn = std::sqrt(nn);
                        if (nn > 1e-12) {
                            for (size_type j = 0; j < D; ++j) ct(i, j) /= nn;
                            vector_generated = true;
                        }
                    }
                }
                for (size_type j = 0; j < D; ++j)
                    for (size_type i = 0; i < r; ++i) c()(j, i) = ct(i, j);
            } else if (!c.is_root()) {
                for (size_type i = 0; i < c().size(0); ++i)
                    for (size_type j = 0; j < c().size(1); ++j) c()(i, j) = (i == j ? 1.0 : 0.0);
            } else {
                for (size_type i = 0; i < c().size(0); ++i)
                    for (size_type j = 0; j < c().size(1); ++j) c()(i, j) = ((i == j) && (i == 0) ? 1.0 : 0.0);
            }
        }
        A.set_is_orthogonalised();

        subspace_expansion_projector_splitting_integrator<complex_type, backend_type> tdvp(A, H, krylov_dim, 1e-12, 1);
        auto knobs = [&]() {
            tdvp.dt() = dt;
            tdvp.spawning_threshold() = get("spawn", 1e-7);
            tdvp.unoccupied_threshold() = get("unoccupied", 1e-4);
            tdvp.minimum_unoccupied() = static_cast<size_type>(get("min-unoccupied", 0.0));
        };
        knobs();
        ops::dense_matrix_operator<complex_type, backend_type> mop(Sz);
        const int dt2_at = static_cast<int>(get("dt2-at", -1.0)), c0_at = static_cast<int>(get("coeff0-at", -1.0));
        const int reset_at = static_cast<int>(get("reset-at", -1.0)), clear_at = static_cast<int>(get("clear-at", -1.0));
        const int ck_at = static_cast<int>(get("checkpoint-at", -1.0)), rs_at = static_cast<int>(get("restore-at", -1.0));
        std::vector<char> ck;
        run(A, tdvp, H, mop, nsteps, [&](int done) {
            if (done == dt2_at) tdvp.set_dt(get("dt2", dt));
            if (done == c0_at) H.coeff(0) *= get("coeff0", 1.0);
            if (done == reset_at) tdvp.reset_nsteps();
            if (done == ck_at) ck = tdvp.checkpoint(A);
            if (done == rs_at) tdvp.restore(ck, A);
            if (done == clear_at) {
                const double dtnow = tdvp.dt();
                tdvp.clear();
                tdvp.krylov_dim() = krylov_dim;
                knobs();
                tdvp.dt() = dtnow;
                tdvp.initialise(A, H);
            }
        });
    } catch (const std::exception& ex) {
        std::cerr << ex.what() << std::endl;
        return 1;
    }
    return 0;
}
