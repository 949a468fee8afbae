// Sum-of-products (SOP) operators and their compilation into per-node term tables.
//
//  * SOP: H = sum_r coeff_r prod_nu op_{r,nu}; operators are bound per mode with the
//    list of term indices r they belong to (reference operators/sop_operator.hpp:234-520).
//  * compile_sop(): the set-intersection compression of the reference's
//    operator_container (tdvp_core/operator_container.hpp:43-96,139-784) and
//    operator_compression (tdvp_core/operator_compression.hpp:23-489).  It produces,
//    per tree node, "common spf", "common mf" and "standard" terms with the child
//    (spf) and parent/sibling (mf) index lists and accumulation coefficients that the
//    H-action and environment kernels consume.  The integer tables are reproduced
//    exactly, including the order in which the reference visits the nodes.
#pragma once

#include "common.hpp"
#include "topology.hpp"

#include <utility>

namespace ttnh {

enum OpKind : int { OP_DENSE = 0, OP_DIAG = 1, OP_CSR = 2, OP_IDENTITY = 3 };

// A primitive single-mode operator (reference operators/matrix_operators.hpp):
// dense (row-major d x d), diagonal (d entries) or CSR (rowptr d+1, colind/vals nnz).
struct PrimOp {
    int kind = OP_DENSE;
    i64 dim = 0;
    std::vector<cplx> dense;
    std::vector<cplx> diag;
    std::vector<i64> rowptr, colind;
    std::vector<cplx> vals;

    bool is_identity() const { return kind == OP_IDENTITY; }
    i64 nnz() const { return kind == OP_CSR ? static_cast<i64>(vals.size()) : 0; }
    // y = op * x with x, y row-major (dim x ncols).
    void apply(const cplx* x, cplx* y, i64 ncols) const;
};

struct ModeTerm {
    PrimOp op;
    std::vector<int> r;  // sorted term indices this operator belongs to
};

struct SOP {
    int nterms = 0;
    std::vector<i64> mode_dims;
    std::vector<cplx> coeff;                   // one per term, default 1
    std::vector<std::vector<ModeTerm>> ops;    // per mode, in bind order

    SOP() = default;
    SOP(int nt, const std::vector<i64>& dims)
        : nterms(nt), mode_dims(dims), coeff(nt, cplx(1.0, 0.0)), ops(dims.size()) {}

    // reference sop_operator.hpp:355-392: identity operators are not stored.
    void bind(const PrimOp& op, std::vector<int> r, int nu);
    bool is_identity(int r, int nu) const;   // sop_operator.hpp:394-401
    int index(int r, int nu) const;          // sop_operator.hpp:403-412
    std::vector<int> ridentity(int nu) const;  // sop_operator.hpp:452-469
    int nmodes() const { return static_cast<int>(mode_dims.size()); }
};

// ---- compiled term tables (reference ttn_nodes/operator_node.hpp:23-361) ----
struct MfIndex {
    int parent_index = 0;
    std::vector<std::pair<int, int>> siblings;  // (child slot nu of the parent, term index)
};

struct Term {
    bool id_spf = false;
    bool id_mf = false;
    cplx coeff = cplx(0.0, 0.0);
    std::vector<cplx> accum;
    // spf sub-terms: each a list of (nu, cri); at a leaf node nu is the primitive mode
    // and cri the operator index within that mode, at a branch node nu is the child slot
    // and cri the child's term index.
    std::vector<std::vector<std::pair<int, int>>> spf_index;
    std::vector<MfIndex> mf_index;

    int nspf_terms() const { return static_cast<int>(spf_index.size()); }
    int nmf_terms() const { return static_cast<int>(mf_index.size()); }
};

struct NodeOps {
    int ncommon_spf = 0, ncommon_mf = 0, nstandard = 0;
    std::vector<Term> terms;
    int nterms() const { return static_cast<int>(terms.size()); }
};

struct CompiledSOP {
    std::vector<NodeOps> nodes;  // indexed like Topology::nodes
};

CompiledSOP compile_sop(const Topology& topo, const SOP& sop);

}  // namespace ttnh
