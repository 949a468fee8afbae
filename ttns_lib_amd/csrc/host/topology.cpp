#include "topology.hpp"

#include <algorithm>

namespace ttnh {

Topology topology_from_arrays(int n, const i64* parent, const i64* hrank, const i64* cap_hrank,
                              const i64* leaf_dim) {
    check(n >= 1, ERR_INVALID, "tree must have at least one node");
    check(parent[0] == -1, ERR_INVALID, "node 0 must be the root (parent -1)");
    Topology t;
    t.nodes.resize(n);
    // Pre-order check: the parent of node i must lie on the current root->(i-1) path.
    std::vector<int> path;
    for (int i = 0; i < n; ++i) {
        Node& nd = t.nodes[i];
        nd.parent = static_cast<int>(parent[i]);
        nd.hrank = hrank[i];
        nd.cap_hrank = cap_hrank[i];
        check(nd.hrank >= 1 && nd.cap_hrank >= nd.hrank, ERR_INVALID,
              "node " + std::to_string(i) + ": need 1 <= hrank <= cap_hrank");
        if (i > 0) {
            check(parent[i] >= 0 && parent[i] < i, ERR_INVALID, "parent ids must precede children");
            while (!path.empty() && path.back() != parent[i]) path.pop_back();
            check(!path.empty(), ERR_INVALID, "node order is not a pre-order of the tree");
            Node& p = t.nodes[nd.parent];
            check(p.leaf_index < 0 && leaf_dim[nd.parent] == 0, ERR_INVALID,
                  "a leaf node cannot have children");
            nd.child_id = static_cast<int>(p.children.size());
            nd.level = p.level + 1;
            p.children.push_back(i);
        }
        path.push_back(i);
    }
    // dims and leaf indices
    int nleaf = 0;
    for (int i = 0; i < n; ++i) {
        Node& nd = t.nodes[i];
        if (leaf_dim[i] > 0) {
            check(nd.children.empty(), ERR_INVALID, "leaf node with children");
            nd.dims = {leaf_dim[i]};
            nd.cap_dims = {leaf_dim[i]};
            nd.leaf_index = nleaf++;
            t.leaf_node.push_back(i);
            t.mode_dims.push_back(leaf_dim[i]);
        } else {
            check(!nd.children.empty(), ERR_INVALID,
                  "node " + std::to_string(i) + " has neither children nor a primitive dimension");
            for (int c : nd.children) {
                nd.dims.push_back(t.nodes[c].hrank);
                nd.cap_dims.push_back(t.nodes[c].cap_hrank);
            }
        }
    }
    for (int i = 0; i < n; ++i) {
        const Node& nd = t.nodes[i];
        check(nd.hrank <= nd.dimen(), ERR_INVALID,
              "node " + std::to_string(i) + ": hrank exceeds the product of its mode dimensions");
        check(nd.cap_hrank <= nd.cap_dimen(), ERR_INVALID,
              "node " + std::to_string(i) + ": capacity hrank exceeds the capacity mode product");
    }
    check(t.nodes[0].hrank == 1, ERR_INVALID, "the root must have hrank 1");
    return t;
}

std::vector<int> euler_tour(const Topology& t) {
    const int n = t.size();
    size_t len = 0;
    for (const Node& nd : t.nodes) len += nd.is_leaf() ? 2 : 1 + nd.children.size();
    std::vector<int> order(len);
    std::vector<size_t> visits(n, 0);
    int cur = 0;
    for (size_t i = 0; i < len; ++i) {
        order[i] = cur;
        const Node& nd = t.nodes[cur];
        size_t tv = visits[cur];
        int next = cur;
        if (tv == 0 && nd.is_leaf()) {
            // a leaf is visited twice in a row
        } else if (tv < nd.children.size()) {
            next = nd.children[tv];
        } else if (!nd.is_root()) {
            next = nd.parent;
        } else {
            check(i + 1 == len, ERR_INVALID, "inconsistent traversal");
        }
        ++visits[cur];
        cur = next;
    }
    return order;
}

// ---------------------------------------------------------------------------------------------
namespace {

using Path = std::vector<size_t>;

NTree& at(NTree& root, const Path& p) {
    NTree* nd = &root;
    for (size_t i : p) nd = &nd->children[i];
    return *nd;
}

void balanced_subtree_internal(size_t degree, NTree& node, size_t nadd,
                               const std::function<i64(size_t)>& fl, std::vector<Path>& linds,
                               size_t level, size_t count) {
    if (nadd < degree) {
        if (nadd == 1) return;
        for (size_t i = 0; i < nadd; ++i) {
            linds[count + i].push_back(node.children.size());
            node.insert(fl(level));
        }
        return;
    }
    size_t r = nadd % degree;
    for (size_t i = 0; i < degree; ++i) {
        size_t nchild = nadd / degree + (i < r ? 1 : 0);
        for (size_t j = 0; j < nchild; ++j) linds[count + j].push_back(node.children.size());
        node.insert(fl(level));
        balanced_subtree_internal(degree, node.children.back(), nchild, fl, linds, level + 1, count);
        count += nchild;
    }
}

void balanced_subtree(NTree& root, size_t nleaves, size_t degree,
                      const std::function<i64(size_t)>& fl, std::vector<Path>& linds) {
    check(degree > 1 && nleaves > 0, ERR_INVALID, "balanced subtree needs degree>1 and leaves");
    linds.assign(nleaves, Path{});
    if (nleaves < degree) {
        for (size_t i = 0; i < nleaves; ++i) {
            linds[i] = Path{root.children.size()};
            root.insert(fl(0));
        }
        return;
    }
    size_t r = nleaves % degree;
    size_t count = 0;
    for (size_t i = 0; i < degree; ++i) {
        size_t nchild = nleaves / degree + (i < r ? 1 : 0);
        for (size_t j = 0; j < nchild; ++j) linds[count + j].push_back(root.children.size());
        root.insert(fl(0));
        balanced_subtree_internal(degree, root.children.back(), nchild, fl, linds, 1, count);
        count += nchild;
    }
}

void sanitise_rec(NTree& t) {
    for (NTree& c : t.children) sanitise_rec(c);
    if (!t.is_leaf()) {
        i64 s = 1;
        for (const NTree& c : t.children) s *= c.value;
        if (s < t.value) t.value = s;
    }
}

void flatten(const NTree& tr, const NTree& cap, int parent, int child_id, int level, Topology& t,
             int& leaf_counter) {
    check(tr.children.size() == cap.children.size(), ERR_INVALID,
          "capacity and topology trees differ in structure");
    check(tr.value <= cap.value, ERR_INVALID, "capacity is less than size");
    if (tr.is_leaf()) return;  // primitive dimension, recorded by its parent
    int id = t.size();
    t.nodes.emplace_back();
    {
        Node& nd = t.nodes.back();
        nd.parent = parent;
        nd.child_id = child_id;
        nd.level = level;
        nd.hrank = tr.value;
        nd.cap_hrank = cap.value;
        for (size_t c = 0; c < tr.children.size(); ++c) {
            nd.dims.push_back(tr.children[c].value);
            nd.cap_dims.push_back(cap.children[c].value);
        }
    }
    if (parent >= 0) t.nodes[parent].children.push_back(id);
    if (tr.children.size() == 1 && tr.children[0].is_leaf()) {
        t.nodes[id].leaf_index = leaf_counter++;
        t.leaf_node.push_back(id);
        t.mode_dims.push_back(tr.children[0].value);
        return;
    }
    int cid = 0;
    for (size_t c = 0; c < tr.children.size(); ++c) {
        check(!tr.children[c].is_leaf(), ERR_INVALID,
              "a primitive leaf must be the only child of its parent");
        flatten(tr.children[c], cap.children[c], id, cid++, level + 1, t, leaf_counter);
    }
}

}  // namespace

void htucker_subtree(NTree& root, const std::vector<i64>& hb, size_t degree,
                     const std::function<i64(size_t)>& fl) {
    std::vector<Path> linds;
    balanced_subtree(root, hb.size(), degree, fl, linds);
    for (size_t i = 0; i < linds.size(); ++i) at(root, linds[i]).insert(hb[i]);
}

void sanitise_tree(NTree& t) { sanitise_rec(t); }

Topology topology_from_ntree(const NTree& tree, const NTree& capacity) {
    Topology t;
    int leaf_counter = 0;
    flatten(tree, capacity, -1, 0, 0, t, leaf_counter);
    for (const Node& nd : t.nodes) {
        check(nd.hrank <= nd.dimen(), ERR_INVALID, "hrank exceeds the product of mode dimensions");
    }
    return t;
}

void topology_to_arrays(const Topology& t, std::vector<i64>& parent, std::vector<i64>& hrank,
                        std::vector<i64>& cap_hrank, std::vector<i64>& leaf_dim) {
    const int n = t.size();
    parent.resize(n);
    hrank.resize(n);
    cap_hrank.resize(n);
    leaf_dim.resize(n);
    for (int i = 0; i < n; ++i) {
        const Node& nd = t.nodes[i];
        parent[i] = nd.parent;
        hrank[i] = nd.hrank;
        cap_hrank[i] = nd.cap_hrank;
        leaf_dim[i] = nd.is_leaf() ? nd.dims[0] : 0;
    }
}

}  // namespace ttnh
