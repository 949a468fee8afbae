// Algorithmic floating-point operation counts of the TDVP step (8 real flops per complex
// multiply-add; copies and permutations not counted), driven by the compiled term tables
// and by the H-application counts measured during a run.
#pragma once

#include <algorithm>

#include "sop.hpp"
#include "topology.hpp"

namespace ttnh {

struct FlopModel {
    std::vector<double> heff_node;  // one node H-action (coefficient_evolution_{leaf,branch})
    std::vector<double> heff_bond;  // one bond H-action (full_hamiltonian_action) with node n's operators
    std::vector<double> node_size;  // Krylov vector length of the node propagation
    std::vector<double> bond_size;  // Krylov vector length of the bond propagation (hrank^2)
    double sweep_pair = 0.0;        // forward + backward sweep without propagations: QR gauges,
                                    // gauge applications, mean-field and single-particle updates
    double prepare_spf = 0.0;       // prepare_evolution's bottom-up single-particle pass
};

inline double qr_flops(double m, double n) {
    if (m < n) std::swap(m, n);
    // Householder factorisation plus explicit formation of the thin Q
    return 2.0 * 8.0 * (2.0 * m * n * n - 2.0 * n * n * n / 3.0);
}

inline FlopModel flop_model(const Topology& t, const CompiledSOP& ops, const SOP& sop) {
    const int n = t.size();
    FlopModel f;
    f.heff_node.assign(n, 0.0);
    f.heff_bond.assign(n, 0.0);
    f.node_size.assign(n, 0.0);
    f.bond_size.assign(n, 0.0);
    auto chain = [&](const Node& nd, const std::vector<std::pair<int, int>>& list) {
        double s = 0.0;
        const double size = static_cast<double>(nd.size());
        if (nd.is_leaf()) {
            for (const auto& pr : list) {
                const PrimOp& op = sop.ops[pr.first][pr.second].op;
                const double d = static_cast<double>(op.dim), r = static_cast<double>(nd.hrank);
                if (op.kind == OP_DENSE) s += 8.0 * d * d * r;
                if (op.kind == OP_DIAG) s += 8.0 * d * r;
                if (op.kind == OP_CSR) s += 8.0 * static_cast<double>(op.nnz()) * r;
            }
        } else {
            for (const auto& pr : list) s += 8.0 * size * static_cast<double>(nd.dims[pr.first]);
        }
        return s;
    };
    std::vector<double> spf_cost(n, 0.0), mf_cost(n, 0.0);
    for (int i = 0; i < n; ++i) {
        const Node& nd = t.nodes[i];
        const double size = static_cast<double>(nd.size()), r = static_cast<double>(nd.hrank);
        const double D = static_cast<double>(nd.dimen());
        f.node_size[i] = size;
        f.bond_size[i] = r * r;
        for (const Term& tm : ops.nodes[i].terms) {
            double h = 0.0;
            if (!tm.id_spf)
                for (const auto& s : tm.spf_index) h += chain(nd, s);
            if (!tm.id_mf) h += 8.0 * size * r;
            f.heff_node[i] += h;
            f.heff_bond[i] += (tm.id_spf ? 0.0 : 8.0 * r * r * r) + (tm.id_mf ? 0.0 : 8.0 * r * r * r);
            if (!tm.id_spf)
                for (const auto& s : tm.spf_index) spf_cost[i] += chain(nd, s) + 8.0 * D * r * r;
        }
        if (!nd.is_root()) {
            const Node& pn = t.nodes[nd.parent];
            const double psize = static_cast<double>(pn.size()), pr = static_cast<double>(pn.hrank);
            const double dc = static_cast<double>(pn.dims[nd.child_id]);
            for (const Term& tm : ops.nodes[i].terms) {
                if (tm.id_mf) continue;
                for (const MfIndex& m : tm.mf_index) {
                    double s = 8.0 * psize * dc;  // outer contraction over the parent's other indices
                    for (const auto& sb : m.siblings) s += 8.0 * psize * static_cast<double>(pn.dims[sb.first]);
                    if (!ops.nodes[nd.parent].terms[m.parent_index].id_mf) s += 8.0 * psize * pr;
                    mf_cost[i] += s;
                }
            }
        }
    }
    double sweep = 0.0;
    for (int i = 0; i < n; ++i) {
        const Node& nd = t.nodes[i];
        const double size = static_cast<double>(nd.size()), r = static_cast<double>(nd.hrank);
        const double D = static_cast<double>(nd.dimen());
        if (!nd.is_leaf())
            for (int c = 0; c < nd.nmodes(); ++c) {
                const double dc = static_cast<double>(nd.dims[c]);
                sweep += qr_flops(size / dc, dc) + mf_cost[nd.children[c]];
            }
        if (!nd.is_root()) {
            const Node& pn = t.nodes[nd.parent];
            sweep += qr_flops(D, r) + spf_cost[i] + 8.0 * D * r * r +
                     8.0 * static_cast<double>(pn.size()) * static_cast<double>(pn.dims[nd.child_id]);
            f.prepare_spf += spf_cost[i];
        }
    }
    f.sweep_pair = 2.0 * sweep;
    return f;
}

}  // namespace ttnh
