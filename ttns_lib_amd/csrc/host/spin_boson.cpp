#include "spin_boson.hpp"

#include <algorithm>
#include <cmath>
#include <random>

namespace ttnh {

std::vector<ModeBlock> sb_blocks(const SBParams& p) {
    const size_t n = p.wk.size();
    check(n >= 1 && p.gk.size() == n, ERR_INVALID, "spin-boson: wk and gk must have equal, nonzero length");
    std::vector<ModeBlock> out;
    if (p.block_sizes.empty()) {
        check(p.boson_dim.size() == n, ERR_INVALID, "wk, gk and boson_dim must have equal length");
        for (size_t i = 0; i < n; ++i) out.push_back(single_mode(p.wk[i], p.gk[i], p.boson_dim[i]));
        return out;
    }
    check(p.block_nmax.size() == p.block_sizes.size(), ERR_INVALID, "spin-boson: one cutoff cap per block");
    size_t at = 0;
    for (size_t b = 0; b < p.block_sizes.size(); ++b) {
        const i64 m = p.block_sizes[b];
        check(m >= 1 && at + static_cast<size_t>(m) <= n, ERR_INVALID, "spin-boson: block sizes must cover the modes");
        out.push_back(mode_block(std::vector<double>(p.wk.begin() + at, p.wk.begin() + at + m),
                                 std::vector<double>(p.gk.begin() + at, p.gk.begin() + at + m), p.ecut,
                                 p.block_nmax[b]));
        at += static_cast<size_t>(m);
    }
    check(at == n, ERR_INVALID, "spin-boson: block sizes must add up to the number of modes");
    return out;
}

Topology sb_topology(const SBParams& p, const std::vector<i64>& leaf_dims) {
    const size_t nblocks = leaf_dims.size();
    check(nblocks >= 1, ERR_INVALID, "spin-boson model needs at least one bath mode");
    const size_t nlevels = static_cast<size_t>(std::log2(static_cast<double>(nblocks)));
    auto level_fn = [nlevels](i64 top, i64 bottom) {
        return [nlevels, top, bottom](size_t l) -> i64 {
            if (l >= nlevels) return bottom;
            if (l == 0) return top;
            const double rmax = std::log2(static_cast<double>(top));
            const double rmin = std::log2(static_cast<double>(bottom));
            return static_cast<i64>(
                std::pow(2.0, ((nlevels - l) * static_cast<double>(rmax - rmin)) / nlevels + rmin));
        };
    };
    auto build = [&](i64 top, i64 bottom) {
        NTree t(1);
        t.insert(p.nhilb).insert(p.nhilb);
        NTree& bath = t.insert(p.nhilb);
        htucker_subtree(bath, leaf_dims, 2, level_fn(top, bottom));
        sanitise_tree(t);
        return t;
    };
    NTree topo = build(p.nspf, p.nspf_lower);
    NTree cap = build(std::max(p.nspf, p.nspf_cap), std::max(p.nspf_lower, p.nspf_lower_cap));
    return topology_from_ntree(topo, cap);
}

std::vector<std::vector<cplx>> sb_initial_state(const Topology& t, const std::vector<cplx>& psi0,
                                                uint64_t seed) {
    std::mt19937 rng(static_cast<std::mt19937::result_type>(seed));
    std::vector<std::vector<cplx>> A(t.size());
    for (int n = 0; n < t.size(); ++n) {
        const Node& nd = t.nodes[n];
        const i64 D = nd.dimen(), r = nd.hrank;
        A[n].assign(D * r, cplx(0.0, 0.0));
        if (nd.is_leaf()) {
            // ct: r x D, row i = i-th single-particle function
            std::vector<cplx> ct(r * D, cplx(0.0, 0.0));
            if (nd.leaf_index == 0) {
                check(static_cast<i64>(psi0.size()) == D, ERR_INVALID, "psi0 has the wrong size");
                for (i64 j = 0; j < D; ++j) ct[j] = psi0[j];
            } else {
                ct[0] = cplx(1.0, 0.0);
            }
            std::uniform_real_distribution<double> dist(0, 2.0 * std::acos(-1.0));
            std::normal_distribution<double> length_dist(0, 1);
            for (i64 i = 1; i < r; ++i) {
                bool generated = false;
                while (!generated) {
                    for (i64 j = 0; j < D; ++j) {
                        const double theta = dist(rng);
                        const double len = length_dist(rng);
                        ct[i * D + j] = len * cplx(std::cos(theta), std::sin(theta));
                    }
                    double nrm = 0.0;
                    for (i64 j = 0; j < D; ++j) nrm += std::norm(ct[i * D + j]);
                    nrm = std::sqrt(nrm);
                    for (i64 j = 0; j < D; ++j) ct[i * D + j] /= nrm;
                    for (i64 k = 0; k < i; ++k) {
                        cplx ov(0.0, 0.0);
                        for (i64 j = 0; j < D; ++j) ov += std::conj(ct[k * D + j]) * ct[i * D + j];
                        for (i64 j = 0; j < D; ++j) ct[i * D + j] -= ov * ct[k * D + j];
                    }
                    double nn = 0.0;
                    for (i64 j = 0; j < D; ++j) nn += std::norm(ct[i * D + j]);
                    nn = std::sqrt(nn);
                    if (nn > 1e-12) {
                        for (i64 j = 0; j < D; ++j) ct[i * D + j] /= nn;
                        generated = true;
                    }
                }
            }
            for (i64 j = 0; j < D; ++j)
                for (i64 i = 0; i < r; ++i) A[n][j * r + i] = ct[i * D + j];
        } else if (!nd.is_root()) {
            for (i64 i = 0; i < std::min(D, r); ++i) A[n][i * r + i] = cplx(1.0, 0.0);
        } else {
            A[n][0] = cplx(1.0, 0.0);
        }
    }
    return A;
}

SBProblem sb_build(const SBParams& p) {
    const std::vector<ModeBlock> blocks = sb_blocks(p);
    const size_t nb = blocks.size();
    std::vector<i64> leaf_dims(nb);
    for (size_t i = 0; i < nb; ++i) leaf_dims[i] = blocks[i].nstates;
    check(static_cast<i64>(p.hsys.size()) == p.nhilb * p.nhilb &&
              static_cast<i64>(p.scoup.size()) == p.nhilb * p.nhilb,
          ERR_INVALID, "Hsys / Scoup must be nhilb x nhilb");
    SBProblem prob;
    prob.topo = sb_topology(p, leaf_dims);

    std::vector<i64> mode_dims(nb + 1);
    mode_dims[0] = p.nhilb;
    for (size_t i = 0; i < nb; ++i) mode_dims[i + 1] = leaf_dims[i];
    SOP H(static_cast<int>(1 + 2 * nb), mode_dims);

    PrimOp hs;
    hs.kind = OP_DENSE;
    hs.dim = p.nhilb;
    hs.dense = p.hsys;
    H.bind(hs, {0}, 0);
    PrimOp sc = hs;
    sc.dense = p.scoup;
    std::vector<int> zr(nb);
    for (size_t i = 0; i < nb; ++i) zr[i] = static_cast<int>(2 * i + 2);
    H.bind(sc, zr, 0);
    for (size_t i = 0; i < nb; ++i) {
        H.bind(block_h0(blocks[i]), {static_cast<int>(1 + 2 * i)}, static_cast<int>(i + 1));
        H.bind(block_hc(blocks[i]), {static_cast<int>(2 + 2 * i)}, static_cast<int>(i + 1));
    }
    prob.sop = std::move(H);
    prob.state = sb_initial_state(prob.topo, p.psi0, p.seed);
    return prob;
}

int sb_quadrature_nodes(const SBParams& p, double dt) {
    double wmax = 0.0, hs = 0.0;
    for (double w : p.wk) wmax = std::max(wmax, std::abs(w));
    for (const cplx& h : p.hsys) hs += std::norm(h);
    const double half = 0.75 * std::abs(dt) * (wmax + 2.0 * std::sqrt(hs));
    for (int n = 4; n < 16; ++n) {
        // log of (phi/2)^(2n) / (2n)!
        const double lg = 2.0 * n * std::log(std::max(half, 1e-300)) - std::lgamma(2.0 * n + 1.0);
        if (lg < std::log(1e-13)) return n;
    }
    return 16;
}

void sb_density_gradient(const SBParams& p, const std::vector<std::vector<cplx>>& rho,
                         const std::vector<std::vector<cplx>>& rhoS, double weight, std::vector<double>& dg,
                         std::vector<double>& dw) {
    const std::vector<ModeBlock> blocks = sb_blocks(p);
    check(rho.size() == blocks.size() + 1 && rhoS.size() == blocks.size() + 1, ERR_INVALID,
          "sb gradient: one density per leaf mode expected");
    if (dg.size() != p.wk.size()) dg.assign(p.wk.size(), 0.0);
    if (dw.size() != p.wk.size()) dw.assign(p.wk.size(), 0.0);
    size_t first = 0;  // global index of the block's first mode
    for (size_t b = 0; b < blocks.size(); ++b) {
        const ModeBlock& B = blocks[b];
        const i64 d = B.nstates;
        const std::vector<cplx>& r1 = rho[b + 1];
        const std::vector<cplx>& rS = rhoS[b + 1];
        check(static_cast<i64>(r1.size()) == d * d && static_cast<i64>(rS.size()) == d * d, ERR_INVALID,
              "sb gradient: density of the wrong size");
        const size_t N = B.w.size();
        // H0[s] = sum_k w_k occ(s, k): <chi|dH0/dw_k|psi> = sum_s occ(s, k) rho(s, s)
        for (i64 s = 0; s < d; ++s)
            for (size_t k = 0; k < N; ++k)
                dw[first + k] += weight * 2.0 * r1[s * d + s].imag() * static_cast<double>(B.occ(s, k));
        // Hc in the order block_hc writes it: lowering in mode order, then raising in reverse
        for (i64 s = 0; s < d; ++s) {
            for (size_t k = 0; k < N; ++k) {
                const i64 n = B.occ(s, k);
                if (n > 0)
                    dg[first + k] += weight * 2.0 * rS[s * d + (s - B.stride[k])].imag() * std::sqrt(static_cast<double>(n));
            }
            for (size_t k = N; k-- > 0;) {
                const i64 n = B.occ(s, k);
                if (n + 1 < B.nper[k])
                    dg[first + k] += weight * 2.0 * rS[s * d + (s + B.stride[k])].imag() *
                                     std::sqrt(static_cast<double>(n) + 1.0);
            }
        }
        first += N;
    }
}

}  // namespace ttnh
