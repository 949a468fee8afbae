#include "sop.hpp"

#include <algorithm>
#include <iterator>

namespace ttnh {

void PrimOp::apply(const cplx* x, cplx* y, i64 ncols) const {
    switch (kind) {
        case OP_DENSE:
            for (i64 i = 0; i < dim; ++i)
                for (i64 c = 0; c < ncols; ++c) {
                    cplx s(0.0, 0.0);
                    for (i64 j = 0; j < dim; ++j) s += dense[i * dim + j] * x[j * ncols + c];
                    y[i * ncols + c] = s;
                }
            break;
        case OP_DIAG:
            for (i64 i = 0; i < dim; ++i)
                for (i64 c = 0; c < ncols; ++c) y[i * ncols + c] = diag[i] * x[i * ncols + c];
            break;
        case OP_CSR:
            for (i64 i = 0; i < dim; ++i)
                for (i64 c = 0; c < ncols; ++c) {
                    cplx s(0.0, 0.0);
                    for (i64 k = rowptr[i]; k < rowptr[i + 1]; ++k) s += vals[k] * x[colind[k] * ncols + c];
                    y[i * ncols + c] = s;
                }
            break;
        default:
            for (i64 i = 0; i < dim * ncols; ++i) y[i] = x[i];
    }
}

void SOP::bind(const PrimOp& op, std::vector<int> r, int nu) {
    check(nu >= 0 && nu < nmodes(), ERR_INVALID, "bind: mode index out of range");
    for (int ri : r) check(ri >= 0 && ri < nterms, ERR_INVALID, "bind: term index out of range");
    // no r index may already be bound on this mode
    for (const ModeTerm& mt : ops[nu])
        for (int ri : mt.r)
            check(std::find(r.begin(), r.end(), ri) == r.end(), ERR_INVALID,
                  "bind: term " + std::to_string(ri) + " is already bound on mode " + std::to_string(nu));
    if (op.is_identity()) return;
    check(op.dim == mode_dims[nu], ERR_INVALID, "bind: operator does not have the mode's dimension");
    std::sort(r.begin(), r.end());
    ops[nu].push_back(ModeTerm{op, std::move(r)});
}

bool SOP::is_identity(int r, int nu) const {
    for (const ModeTerm& mt : ops[nu])
        if (std::find(mt.r.begin(), mt.r.end(), r) != mt.r.end()) return false;
    return true;
}

int SOP::index(int r, int nu) const {
    int count = 0;
    for (const ModeTerm& mt : ops[nu]) {
        if (std::find(mt.r.begin(), mt.r.end(), r) != mt.r.end()) return count;
        ++count;
    }
    return count;
}

std::vector<int> SOP::ridentity(int nu) const {
    std::vector<int> ret;
    for (int i = 0; i < nterms; ++i) {
        bool ins = true;
        for (const ModeTerm& mt : ops[nu]) ins = ins && (std::find(mt.r.begin(), mt.r.end(), i) == mt.r.end());
        if (ins) ret.push_back(i);
    }
    return ret;
}

// ---------------------------------------------------------------------------------------------
namespace {

using IVec = std::vector<int>;

struct OpIndex {
    bool id;
    IVec r;  // sorted
};
using InfoList = std::vector<OpIndex>;

IVec intersect(const IVec& a, const IVec& b) {
    IVec out;
    std::set_intersection(a.begin(), a.end(), b.begin(), b.end(), std::back_inserter(out));
    return out;
}

// Combine the r-sets of one element from each list (odometer order, last list fastest),
// keeping every non-empty running intersection.  This is the shared enumeration of
// operator_compression.hpp:300-370 ("compression" form, with the add_one_bodies
// switch) and operator_container.hpp:707-784 ("container" form: the first set is
// intersected with riv, hits are reported with the element indices).
struct Odometer {
    std::vector<size_t> len;
    std::vector<size_t> it;
    std::vector<bool> ru;
    std::vector<IVec> uni;

    explicit Odometer(std::vector<size_t> lens)
        : len(std::move(lens)), it(len.size(), 0), ru(len.size(), true), uni(len.size()) {}

    bool done() const { return it[0] == len[0]; }

    // operator_compression.hpp:261-297 / operator_container.hpp:668-704
    void advance(size_t u) {
        ++it[u];
        ru[u] = true;
        for (size_t i = u + 1; i < len.size(); ++i) {
            it[i] = 0;
            ru[i] = true;
            uni[i].clear();
        }
        while (it[u] == len[u]) {
            if (u != 0) {
                it[u] = 0;
                uni[u].clear();
                --u;
                ++it[u];
                ru[u] = true;
            } else {
                it[u] = len[u];
                return;
            }
        }
    }

    // running intersection of level i with level i-1 (identical branch structure in both forms)
    static void combine(const IVec& prev, const IVec& cur, bool one_bodies, IVec& out) {
        out.clear();
        if (prev.size() > 1 && cur.size() > 1) {
            std::set_intersection(cur.begin(), cur.end(), prev.begin(), prev.end(), std::back_inserter(out));
        } else if (prev.size() == 1 && cur.size() == 1 && one_bodies) {
            if (prev.back() == cur.back()) out.push_back(prev.back());
        } else if (prev.size() == 1 && one_bodies) {
            if (std::binary_search(cur.begin(), cur.end(), prev.back())) out.push_back(prev.back());
        } else if (cur.size() == 1 && one_bodies) {
            if (std::binary_search(prev.begin(), prev.end(), cur.back())) out.push_back(cur.back());
        }
    }
};

// compression form: lists of OpIndex; result OpIndex{all identity, intersection}
void compress_iterate(const std::vector<const InfoList*>& lists, size_t maxterms, InfoList& res,
                      bool add_one_bodies) {
    std::vector<size_t> lens;
    for (auto* l : lists) lens.push_back(l->size());
    Odometer od(lens);
    const size_t nch = lists.size();
    while (!od.done()) {
        bool all_id = (*lists[0])[od.it[0]].id;
        if (od.ru[0]) {
            od.uni[0] = (*lists[0])[od.it[0]].r;
            od.ru[0] = false;
        }
        bool keep = true;
        for (size_t i = 1; i < nch && keep; ++i) {
            all_id = all_id && (*lists[i])[od.it[i]].id;
            if (od.ru[i]) {
                Odometer::combine(od.uni[i - 1], (*lists[i])[od.it[i]].r, add_one_bodies, od.uni[i]);
                od.ru[i] = false;
            }
            if (od.uni[i].empty()) {
                keep = false;
                od.advance(i);
            }
        }
        if (keep) {
            if (add_one_bodies || od.uni.back().size() > 1) {
                res.push_back(OpIndex{all_id, od.uni.back()});
                if (res.size() == maxterms) return;
            }
            od.advance(nch - 1);
        }
    }
}

// container form: lists of r-sets (a node's term r-sets); hits are (intersection, indices)
void container_iterate(const std::vector<const std::vector<IVec>*>& lists, const IVec& riv,
                       std::vector<IVec>& rres, std::vector<std::vector<int>>& nids,
                       size_t nterms_to_find = 0) {
    std::vector<size_t> lens;
    for (auto* l : lists) lens.push_back(l->size());
    Odometer od(lens);
    const size_t nch = lists.size();
    while (!od.done()) {
        if (od.ru[0]) {
            od.uni[0] = intersect((*lists[0])[od.it[0]], riv);
            od.ru[0] = false;
        }
        bool keep = true;
        for (size_t i = 1; i < nch && keep; ++i) {
            if (od.ru[i]) {
                Odometer::combine(od.uni[i - 1], (*lists[i])[od.it[i]], true, od.uni[i]);
                od.ru[i] = false;
            }
            if (od.uni[i].empty()) {
                keep = false;
                od.advance(i);
            }
        }
        if (keep) {
            if (!od.uni.back().empty()) {
                rres.push_back(od.uni.back());
                nids.emplace_back(od.it.begin(), od.it.end());
                if (nterms_to_find > 0 && rres.size() == nterms_to_find) return;
            }
            od.advance(nch - 1);
        }
    }
}

// operator_compression.hpp:373-428
void spf_node_info(const SOP& op, const Topology& topo, int n, std::vector<InfoList>& spf) {
    const Node& nd = topo.nodes[n];
    InfoList p;
    if (nd.is_leaf()) {
        const int nu = nd.leaf_index;
        for (const ModeTerm& mt : op.ops[nu])
            if (mt.r.size() > 1) p.push_back(OpIndex{mt.op.is_identity(), mt.r});
        IVec rid = op.ridentity(nu);
        if (rid.size() > 1) p.push_back(OpIndex{true, rid});
    } else {
        bool compute = true;
        for (int c : nd.children)
            if (spf[c].empty()) compute = false;
        if (compute) {
            std::vector<const InfoList*> lists;
            for (int c : nd.children) lists.push_back(&spf[c]);
            compress_iterate(lists, static_cast<size_t>(op.nterms), p, false);
        }
    }
    spf[n] = std::move(p);
}

// operator_compression.hpp:430-489
void mf_node_info(const Topology& topo, int n, int nterms, const std::vector<InfoList>& spf,
                  std::vector<InfoList>& mf) {
    const Node& nd = topo.nodes[n];
    InfoList p;
    if (nd.is_root()) {
        IVec v(nterms);
        for (int i = 0; i < nterms; ++i) v[i] = i;
        p.push_back(OpIndex{true, v});
    } else {
        const Node& par = topo.nodes[nd.parent];
        bool compute = !mf[nd.parent].empty();
        for (int c : par.children)
            if (c != n && spf[c].empty()) compute = false;
        if (compute) {
            std::vector<const InfoList*> lists{&mf[nd.parent]};
            for (int c : par.children)
                if (c != n) lists.push_back(&spf[c]);
            compress_iterate(lists, static_cast<size_t>(nterms), p, false);
        }
    }
    mf[n] = std::move(p);
}

int find_index(const std::vector<IVec>& rn, int r) {
    for (size_t i = 0; i < rn.size(); ++i)
        if (std::find(rn[i].begin(), rn[i].end(), r) != rn[i].end()) return static_cast<int>(i);
    fail(ERR_INVALID, "SOP compilation: r index not found");
}

// operator_container.hpp:373-444 and the three upward initialisers (:139-371)
void spf_node_indices(const Topology& topo, const SOP& op, int n, const InfoList& spf,
                      const InfoList& mf, std::vector<std::vector<IVec>>& rindex,
                      std::vector<NodeOps>& res) {
    const int nterms = op.nterms;
    const Node& nd = topo.nodes[n];
    // remove_common_terms
    std::vector<char> removed(nterms, 0);
    for (const OpIndex& s : spf)
        for (int ri : s.r) removed[ri] = 1;
    IVec r_rem_spf;
    for (int i = 0; i < nterms; ++i)
        if (!removed[i]) r_rem_spf.push_back(i);
    for (const OpIndex& m : mf)
        for (int ri : m.r) removed[ri] = 1;
    IVec r_rem;
    for (int i = 0; i < nterms; ++i)
        if (!removed[i]) r_rem.push_back(i);

    NodeOps& no = res[n];
    no.ncommon_spf = static_cast<int>(spf.size());
    no.ncommon_mf = static_cast<int>(mf.size());
    no.nstandard = static_cast<int>(r_rem.size());
    no.terms.assign(no.ncommon_spf + no.ncommon_mf + no.nstandard, Term{});
    std::vector<IVec>& rn = rindex[n];
    rn.assign(no.terms.size(), IVec{});

    // common spf terms (:139-197)
    int count = 0;
    for (const OpIndex& s : spf) {
        rn[count] = s.r;
        Term& t = no.terms[count];
        if (nd.is_leaf()) {
            const int nu = nd.leaf_index;
            const int cind = op.index(s.r[0], nu);
            if (!s.id) t.spf_index = {{std::make_pair(nu, cind)}};
        } else {
            std::vector<std::pair<int, int>> list;
            for (size_t c = 0; c < nd.children.size(); ++c) {
                const int ch = nd.children[c];
                const int crind = find_index(rindex[ch], s.r[0]);
                if (!res[ch].terms[crind].id_spf) list.emplace_back(static_cast<int>(c), crind);
            }
            if (!list.empty()) t.spf_index = {list};
        }
        t.id_spf = s.id;
        t.id_mf = false;
        ++count;
    }

    // common mf terms (:199-311)
    const int skip = no.ncommon_spf;
    count = 0;
    for (const OpIndex& m : mf) {
        IVec rnot = intersect(m.r, r_rem_spf);
        rn[count + skip] = rnot;
        std::vector<std::vector<std::pair<int, int>>> spfop;
        std::vector<cplx> coeff;
        if (nd.is_leaf()) {
            const int nu = nd.leaf_index;
            coeff.resize(rnot.size());
            for (size_t k = 0; k < rnot.size(); ++k) {
                coeff[k] = op.coeff[rnot[k]];
                if (!op.is_identity(rnot[k], nu))
                    spfop.push_back({std::make_pair(nu, op.index(rnot[k], nu))});
                else
                    spfop.push_back({});
            }
        } else {
            std::vector<const std::vector<IVec>*> lists;
            for (int ch : nd.children) lists.push_back(&rindex[ch]);
            std::vector<IVec> rres;
            std::vector<std::vector<int>> nids;
            container_iterate(lists, rnot, rres, nids);
            coeff.assign(rres.size(), cplx(0.0, 0.0));
            for (size_t k = 0; k < rres.size(); ++k) {
                coeff[k] = rres[k].size() == 1 ? op.coeff[rres[k][0]] : cplx(1.0, 0.0);
                std::vector<std::pair<int, int>> sparse;
                for (size_t c = 0; c < nids[k].size(); ++c)
                    if (!res[nd.children[c]].terms[nids[k][c]].id_spf)
                        sparse.emplace_back(static_cast<int>(c), nids[k][c]);
                spfop.push_back(sparse);
            }
        }
        check(!spfop.empty(), ERR_INVALID, "SOP compilation: common mean-field term without spf contribution");
        Term& t = no.terms[count + skip];
        t.accum = coeff;
        t.coeff = cplx(1.0, 0.0);
        t.id_spf = false;
        t.id_mf = m.id;
        t.spf_index = spfop;
        ++count;
    }

    // standard terms (:313-371)
    const int skip2 = no.ncommon_spf + no.ncommon_mf;
    count = 0;
    for (int r : r_rem) {
        rn[count + skip2] = IVec{r};
        Term& t = no.terms[count + skip2];
        t.accum = {cplx(1.0, 0.0)};
        if (nd.is_leaf()) {
            const int nu = nd.leaf_index;
            if (!op.is_identity(r, nu)) {
                t.spf_index = {{std::make_pair(nu, op.index(r, nu))}};
                t.id_spf = false;
            } else {
                t.id_spf = true;
            }
        } else {
            std::vector<std::pair<int, int>> list;
            for (size_t c = 0; c < nd.children.size(); ++c) {
                const int ch = nd.children[c];
                const int crind = find_index(rindex[ch], r);
                if (!res[ch].terms[crind].id_spf) list.emplace_back(static_cast<int>(c), crind);
            }
            if (!list.empty()) {
                t.spf_index = {list};
                t.id_spf = false;
            } else {
                t.id_spf = true;
            }
        }
        t.coeff = op.coeff[r];
        ++count;
    }
}

// iterator lists for the downward pass: the parent's term r-sets then every sibling's
std::vector<const std::vector<IVec>*> parent_sibling_lists(const Topology& topo, int n,
                                                          const std::vector<std::vector<IVec>>& rindex) {
    const Node& nd = topo.nodes[n];
    const Node& par = topo.nodes[nd.parent];
    std::vector<const std::vector<IVec>*> lists{&rindex[nd.parent]};
    for (int c : par.children)
        if (c != n) lists.push_back(&rindex[c]);
    return lists;
}

std::vector<std::pair<int, int>> sibling_indexing(const Topology& topo, int n, const std::vector<int>& ni,
                                                 const std::vector<NodeOps>& res, bool* any) {
    const Node& nd = topo.nodes[n];
    const Node& par = topo.nodes[nd.parent];
    std::vector<std::pair<int, int>> indexing;
    size_t mcount = 1;
    for (size_t nu = 0; nu < par.children.size(); ++nu) {
        const int c = par.children[nu];
        if (c == n) continue;
        if (!res[c].terms[ni[mcount]].id_spf) {
            indexing.emplace_back(static_cast<int>(nu), ni[mcount]);
            if (any) *any = true;
        }
        ++mcount;
    }
    return indexing;
}

// operator_container.hpp:648-664 and the three downward initialisers (:448-646)
void mf_node_indices(const Topology& topo, const SOP& op, int n, const std::vector<std::vector<IVec>>& rindex,
                     std::vector<NodeOps>& res) {
    const Node& nd = topo.nodes[n];
    NodeOps& no = res[n];
    if (nd.is_root()) {
        for (Term& t : no.terms) t.id_mf = true;
        return;
    }
    const auto lists = parent_sibling_lists(topo, n, rindex);
    // common spf terms: one mean-field contribution per matching (parent, siblings) combination
    for (int i = 0; i < no.ncommon_spf; ++i) {
        std::vector<IVec> rres;
        std::vector<std::vector<int>> nids;
        container_iterate(lists, rindex[n][i], rres, nids);
        Term& t = no.terms[i];
        t.mf_index.assign(rres.size(), MfIndex{});
        t.accum.assign(rres.size(), cplx(0.0, 0.0));
        t.coeff = cplx(1.0, 0.0);
        for (size_t k = 0; k < rres.size(); ++k) {
            t.accum[k] = rres[k].size() == 1 ? op.coeff[rres[k][0]] : cplx(1.0, 0.0);
            t.mf_index[k].parent_index = nids[k][0];
            t.mf_index[k].siblings = sibling_indexing(topo, n, nids[k], res, nullptr);
        }
    }
    // common mf terms: a single (parent, siblings) combination found from the first r value
    const int skip = no.ncommon_spf;
    for (int i = 0; i < no.ncommon_mf; ++i) {
        const IVec& rt = rindex[n][i + skip];
        check(!rt.empty(), ERR_INVALID, "SOP compilation: empty common mean-field term");
        std::vector<IVec> rres;
        std::vector<std::vector<int>> nids;
        container_iterate(lists, IVec{rt.front()}, rres, nids, 1);
        check(!nids.empty(), ERR_INVALID, "SOP compilation: no parent term for a common mean-field term");
        Term& t = no.terms[i + skip];
        t.mf_index.assign(1, MfIndex{});
        t.mf_index[0].parent_index = nids[0][0];
        t.mf_index[0].siblings = sibling_indexing(topo, n, nids[0], res, nullptr);
    }
    // standard terms
    const int skip2 = no.ncommon_spf + no.ncommon_mf;
    for (int i = 0; i < no.nstandard; ++i) {
        const IVec& rt = rindex[n][i + skip2];
        std::vector<IVec> rres;
        std::vector<std::vector<int>> nids;
        container_iterate(lists, rt, rres, nids, 1);
        check(!nids.empty(), ERR_INVALID, "SOP compilation: no parent term for a standard term");
        Term& t = no.terms[i + skip2];
        // the parent's flag as it stands when this node is processed (children are
        // processed before their parent, operator_container.hpp:84)
        bool nonid_sib = false;
        const bool parent_id = res[nd.parent].terms[nids[0][0]].id_mf;
        t.mf_index.assign(1, MfIndex{});
        t.mf_index[0].parent_index = nids[0][0];
        t.mf_index[0].siblings = sibling_indexing(topo, n, nids[0], res, &nonid_sib);
        t.id_mf = parent_id && !nonid_sib;
    }
}

}  // namespace

CompiledSOP compile_sop(const Topology& topo, const SOP& op) {
    check(op.nmodes() == topo.nmodes(), ERR_INVALID, "SOP and tree have a different number of modes");
    for (int m = 0; m < topo.nmodes(); ++m)
        check(op.mode_dims[m] == topo.mode_dims[m], ERR_INVALID,
              "SOP mode " + std::to_string(m) + " dimension differs from the tree's");
    const int n = topo.size();
    std::vector<InfoList> spf(n), mf(n);
    for (int i = n - 1; i >= 0; --i) spf_node_info(op, topo, i, spf);
    for (int i = 0; i < n; ++i) mf_node_info(topo, i, op.nterms, spf, mf);

    CompiledSOP out;
    out.nodes.resize(n);
    std::vector<std::vector<IVec>> rindex(n);
    for (int i = n - 1; i >= 0; --i) spf_node_indices(topo, op, i, spf[i], mf[i], rindex, out.nodes);
    for (int i = n - 1; i >= 0; --i) mf_node_indices(topo, op, i, rindex, out.nodes);
    return out;
}

}  // namespace ttnh
