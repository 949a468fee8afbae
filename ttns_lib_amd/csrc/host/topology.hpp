// Tree topology of a tree tensor network state (TTNS) and the TDVP sweep order.
//
// Layout contract (reference httensor.hpp:128-288, ttn_nodes/httensor_node.hpp:20-285):
//  * one node per non-leaf vertex of the topology tree, numbered in pre-order;
//  * node n stores a row-major matrix of shape (prod_c dims[c]) x hrank, child 0
//    the slowest index, so the rank-3 view of mode c is
//    (prod_{c'<c} dims[c'], dims[c], prod_{c'>c} dims[c'] * hrank);
//  * a node whose single child is a primitive degree of freedom is a "leaf node":
//    dims = {primitive dimension}, leaf_index = mode number (leaves counted in
//    pre-order).
#pragma once

#include "common.hpp"

#include <functional>

namespace ttnh {

struct Node {
    int parent = -1;        // -1 for the root
    int child_id = 0;       // position among the parent's children
    std::vector<int> children;
    std::vector<i64> dims;      // mode dimensions (children's hranks, or the primitive dim)
    std::vector<i64> cap_dims;  // capacity of each mode dimension
    i64 hrank = 1;
    i64 cap_hrank = 1;
    int leaf_index = -1;    // primitive mode index for leaf nodes, -1 otherwise
    int level = 0;

    bool is_leaf() const { return children.empty(); }
    bool is_root() const { return parent < 0; }
    int nmodes() const { return static_cast<int>(dims.size()); }
    i64 dimen() const { i64 d = 1; for (i64 x : dims) d *= x; return d; }
    i64 cap_dimen() const { i64 d = 1; for (i64 x : cap_dims) d *= x; return d; }
    i64 size() const { return dimen() * hrank; }
    i64 capacity() const { return cap_dimen() * cap_hrank; }
};

struct Topology {
    std::vector<Node> nodes;        // pre-order
    std::vector<int> leaf_node;     // mode index -> node id
    std::vector<i64> mode_dims;     // primitive dimension of each mode

    int size() const { return static_cast<int>(nodes.size()); }
    int nmodes() const { return static_cast<int>(leaf_node.size()); }
    // every bond at its capacity: the subspace expansion has no room to add functions
    // (subspace_expansion_projector_splitting_engine.hpp:214 and :378 return at capacity)
    bool at_capacity() const {
        for (const Node& nd : nodes) {
            if (nd.hrank < nd.cap_hrank) return false;
            for (size_t c = 0; c < nd.dims.size(); ++c)
                if (nd.dims[c] < nd.cap_dims[c]) return false;
        }
        return true;
    }
};

// Build a topology from flat pre-order arrays (the C-ABI form):
//   parent[i]    parent node id, -1 for node 0 only, parent[i] < i;
//   hrank[i]     bond dimension of node i to its parent (1 for the root);
//   cap_hrank[i] capacity of that bond dimension (>= hrank[i]);
//   leaf_dim[i]  > 0 marks a leaf node carrying a primitive mode of that
//                dimension, 0 marks a branch node.
// Validates pre-order consistency and the rank constraints the reference asserts
// (httensor.hpp:231: hrank <= prod dims).
Topology topology_from_arrays(int n, const i64* parent, const i64* hrank, const i64* cap_hrank,
                              const i64* leaf_dim);

// Euler-tour sweep order of the one-site TDVP sweep
// (reference tdvp_core/tdvp_tree_traversal.hpp:37-87): length sum_n (leaf ? 2 : 1 + nchildren).
std::vector<int> euler_tour(const Topology& t);

// ----- n-ary topology trees as built by the reference's spin-boson driver -----
struct NTree {
    i64 value = 0;
    std::vector<NTree> children;
    NTree() = default;
    explicit NTree(i64 v) : value(v) {}
    NTree& insert(i64 v) { children.emplace_back(v); return children.back(); }
    bool is_leaf() const { return children.empty(); }
};

// Balanced degree-ary subtree with one leaf value per entry of hb, values of the
// internal vertices from fl(level) (reference tree/ntree_builder.hpp:59-138, 251-261).
void htucker_subtree(NTree& root, const std::vector<i64>& hb, size_t degree,
                     const std::function<i64(size_t)>& fl);

// Caps every internal value at the product of its children's values, post-order
// (reference tree/ntree_builder.hpp:215-233).
void sanitise_tree(NTree& t);

// Flatten a (topology, capacity) pair of n-ary trees into a Topology
// (reference httensor.hpp:160-283).
Topology topology_from_ntree(const NTree& tree, const NTree& capacity);

// Flatten back to the C-ABI arrays (inverse of topology_from_arrays).
void topology_to_arrays(const Topology& t, std::vector<i64>& parent, std::vector<i64>& hrank,
                        std::vector<i64>& cap_hrank, std::vector<i64>& leaf_dim);

}  // namespace ttnh
