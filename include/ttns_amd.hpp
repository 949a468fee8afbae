// ttns_amd.hpp - header-only C++ drop-in for the reference's TDVP integrator API on top of
// the C ABI (ttn_capi.h), so a driver written against llindoy/ttns_lib builds against the
// MI355X library by swapping includes.
//
// Reference interfaces mirrored (paths relative to the reference repository):
//   ntree<T>, ntree_builder<T>::htucker_subtree / sanitise_tree
//                                     include/ttns_lib/tree/ntree.hpp, tree/ntree_builder.hpp:59-138,215-261
//   httensor<T, backend>(topology, capacity), node()(), hrank(), is_leaf(), leaf_index()
//                                     include/ttns_lib/httensor.hpp:128-288, ttn_nodes/httensor_node.hpp:20-285
//   ops::dense_matrix_operator, ops::diagonal_matrix_operator, ops::sparse_matrix_operator
//                                     include/ttns_lib/operators/matrix_operators.hpp:14-474
//   sop_operator<T, backend>(nterms, mode_dims).bind(op, r_indices, mode)
//                                     include/ttns_lib/operators/sop_operator.hpp:234-392
//   fixed_step_ / subspace_expansion_projector_splitting_integrator<T, backend>(A, H, krylov_dim,
//       krylov_tol, nthreads): initialise(), clear(), reset_nsteps(), dt(), set_dt(), t(),
//       operator()(A, H, coeff), nh_applications(), mel()(op, mode, A), the expansion knobs and
//       engine().Nonesite() / Ntwosite()
//                                     include/ttns_lib/integrators/projector_splitting_integrator.hpp:18-397,
//                                     include/ttns_lib/observables/matrix_element.hpp:264-314
//
// Semantics kept: node matrices are row-major (prod child dims) x hrank with child 0
// slowest, nodes in pre-order; operator() advances A in place by one step of dt (the first
// call runs the prepare pass and the six-sub-step ladder) and returns dt; failures throw a
// std::runtime_error whose text starts with the reference's "Failed to apply projector
// splitting integrator.".  All arithmetic runs on the GPU; A is copied to the device when
// the integrator is first used or after the caller changed it, and copied back only when the
// caller next reads a node matrix (bond dimensions are updated after every step).  Only
// complex<double> and dense / diagonal / CSR leaf operators exist.
#pragma once

#include <cmath>
#include <complex>
#include <cstdint>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "ttn_capi.h"

namespace ttns {

using size_type = std::size_t;

struct blas_backend {};  // backend tags: accepted, all work runs on the GPU
struct hip_backend {};

namespace detail {
inline void check(int st, ttn_ctx* ctx, const char* what) {
    if (st != TTN_OK)
        throw std::runtime_error(std::string("Failed to apply projector splitting integrator.\n") + what + ": " +
                                 (ctx ? ttn_last_error(ctx) : "invalid arguments"));
}
}  // namespace detail

// ------------------------------------------------------------------ dense containers
template <typename T>
class matrix {
  public:
    matrix() = default;
    matrix(size_type r, size_type c) : m_r(r), m_c(c), m_d(r * c) {}
    void resize(size_type r, size_type c) {
        m_r = r;
        m_c = c;
        m_d.assign(r * c, T(0));
    }
    T& operator()(size_type i, size_type j) { return m_d[i * m_c + j]; }
    const T& operator()(size_type i, size_type j) const { return m_d[i * m_c + j]; }
    size_type size(int dim) const { return dim == 0 ? m_r : m_c; }
    size_type shape(int dim) const { return size(dim); }
    T* data() { return m_d.data(); }
    const T* data() const { return m_d.data(); }
    void fill_zeros() { std::fill(m_d.begin(), m_d.end(), T(0)); }

  private:
    size_type m_r = 0, m_c = 0;
    std::vector<T> m_d;
};

template <typename T>
struct diagonal_matrix {
    std::vector<T> diag;
};

// CSR matrix: rowptr (n + 1), colind and values (nnz)
template <typename T>
struct csr_matrix {
    size_type n = 0;
    std::vector<int64_t> rowptr, colind;
    std::vector<T> values;
};

// ------------------------------------------------------------------ trees
template <typename T>
class ntree_node {
  public:
    ntree_node() = default;
    explicit ntree_node(const T& v) : m_value(v) {}
    ntree_node& insert(const T& v) {
        m_children.emplace_back(v);
        return m_children.back();
    }
    ntree_node& operator[](size_type i) { return m_children[i]; }
    const ntree_node& operator[](size_type i) const { return m_children[i]; }
    T& value() { return m_value; }
    const T& value() const { return m_value; }
    size_type size() const { return m_children.size(); }
    bool is_leaf() const { return m_children.empty(); }

  private:
    T m_value{};
    std::vector<ntree_node> m_children;
};

template <typename T>
class ntree {
  public:
    ntree& insert(const T& v) {
        m_root.assign(1, ntree_node<T>(v));
        return *this;
    }
    ntree_node<T>& operator()() { return m_root.at(0); }
    const ntree_node<T>& operator()() const { return m_root.at(0); }
    bool empty() const { return m_root.empty(); }

  private:
    std::vector<ntree_node<T>> m_root;
};

template <typename T>
struct ntree_builder {
    // Balanced degree-ary subtree below `root` with one leaf of value hb[i] per entry and
    // internal values fl(level) (the reference's htucker_subtree).
    template <typename F>
    static void htucker_subtree(ntree_node<T>& root, const std::vector<T>& hb, size_type degree, F&& fl) {
        if (degree < 2 || hb.empty()) throw std::invalid_argument("htucker_subtree needs degree > 1 and leaves");
        std::vector<std::vector<size_type>> paths(hb.size());
        balanced(root, hb.size(), degree, fl, paths, 0, 0);
        for (size_type i = 0; i < hb.size(); ++i) {
            ntree_node<T>* nd = &root;
            for (size_type k : paths[i]) nd = &(*nd)[k];
            nd->insert(hb[i]);
        }
    }
    // caps every internal value at the product of its children's values (post-order)
    static void sanitise_tree(ntree<T>& t) { sanitise(t()); }

  private:
    template <typename F>
    static void balanced(ntree_node<T>& node, size_type nadd, size_type degree, F& fl,
                         std::vector<std::vector<size_type>>& paths, size_type level, size_type count) {
        if (nadd < degree) {
            if (level > 0 && nadd == 1) return;
            for (size_type i = 0; i < nadd; ++i) {
                paths[count + i].push_back(node.size());
                node.insert(fl(level));
            }
            return;
        }
        const size_type r = nadd % degree;
        for (size_type i = 0; i < degree; ++i) {
            const size_type nchild = nadd / degree + (i < r ? 1 : 0);
            for (size_type j = 0; j < nchild; ++j) paths[count + j].push_back(node.size());
            node.insert(fl(level));
            balanced(node[node.size() - 1], nchild, degree, fl, paths, level + 1, count);
            count += nchild;
        }
    }
    static void sanitise(ntree_node<T>& t) {
        for (size_type c = 0; c < t.size(); ++c) sanitise(t[c]);
        if (!t.is_leaf()) {
            T s = 1;
            for (size_type c = 0; c < t.size(); ++c) s *= t[c].value();
            if (s < t.value()) t.value() = s;
        }
    }
};

// ------------------------------------------------------------------ TTNS storage
// After an integrator step the device holds the newer state: the tensor records that (and the
// bond dimensions, which the expansion integrator grows) without copying, and the first access
// of a node matrix copies every node back.  Writing through a node (non-const access) marks the
// host copy newer, and the integrator copies it to the device before its next step.
template <typename T, typename backend = blas_backend>
class httensor {
  public:
    class node {
      public:
        bool is_leaf() const { return m_leaf >= 0; }
        bool is_root() const { return m_parent < 0; }
        int64_t leaf_index() const { return m_leaf; }
        int64_t hrank() const { return m_hrank; }
        int64_t parent() const { return m_parent; }
        matrix<T>& operator()() {
            m_owner->pull();
            ++m_owner->m_version;
            return m_A;
        }
        const matrix<T>& operator()() const {
            m_owner->pull();
            return m_A;
        }

      private:
        friend class httensor;
        int64_t m_parent = -1, m_hrank = 1, m_leaf = -1;
        matrix<T> m_A;
        httensor* m_owner = nullptr;
    };

    httensor(const ntree<size_type>& topology, const ntree<size_type>& capacity) {
        int64_t leaves = 0;
        flatten(topology(), capacity(), -1, leaves);
        for (node& n : m_nodes) n.m_owner = this;
    }
    httensor(const httensor&) = delete;
    httensor& operator=(const httensor&) = delete;

    size_type size() const { return m_nodes.size(); }
    node& operator[](size_type i) { return m_nodes[i]; }
    const node& operator[](size_type i) const { return m_nodes[i]; }
    typename std::vector<node>::iterator begin() { return m_nodes.begin(); }
    typename std::vector<node>::iterator end() { return m_nodes.end(); }
    typename std::vector<node>::const_iterator begin() const { return m_nodes.begin(); }
    typename std::vector<node>::const_iterator end() const { return m_nodes.end(); }
    void set_is_orthogonalised() {}

    // flat pre-order arrays of the C ABI (ttn_tree_create); hranks() are the current ones
    const std::vector<int64_t>& parent() const { return m_par; }
    const std::vector<int64_t>& hranks() const { return m_hr; }
    const std::vector<int64_t>& cap_hranks() const { return m_cap; }
    const std::vector<int64_t>& leaf_dims() const { return m_leafdim; }
    uint64_t version() const { return m_version; }

    // ---- used by the integrators
    // the device now holds the newer state, produced by `source`; `pull(A)` copies it into the
    // node matrices (through raw()) when a node matrix is next accessed
    void mark_device_newer(const void* source, std::function<void(httensor&)> pull,
                           const std::vector<int64_t>& hranks) {
        m_source = source;
        m_pull = std::move(pull);
        for (size_type i = 0; i < m_nodes.size(); ++i) m_nodes[i].m_hrank = m_hr[i] = hranks[i];
    }
    // the state the device holds came from `source` and the host copy has not changed since
    bool device_current(const void* source, uint64_t version) const {
        return m_source == source && version == m_version;
    }
    // forget `source` (its device state is about to go away): copy its newer state first
    void detach(const void* source) {
        if (m_source != source) return;
        pull();
        m_source = nullptr;
    }
    matrix<T>& raw(size_type i) { return m_nodes[i].m_A; }
    void pull() const {
        if (!m_pull) return;
        auto f = std::move(m_pull);
        m_pull = nullptr;
        f(const_cast<httensor&>(*this));
    }
    std::weak_ptr<int> alive() const { return m_alive; }

  private:
    void flatten(const ntree_node<size_type>& t, const ntree_node<size_type>& c, int64_t parent, int64_t& leaves) {
        if (t.size() != c.size()) throw std::invalid_argument("capacity and topology trees differ in structure");
        if (t.is_leaf()) return;
        const int64_t id = static_cast<int64_t>(m_nodes.size());
        m_nodes.emplace_back();
        node& n = m_nodes.back();
        n.m_parent = parent;
        n.m_hrank = static_cast<int64_t>(t.value());
        int64_t rows = 1;
        for (size_type k = 0; k < t.size(); ++k) rows *= static_cast<int64_t>(t[k].value());
        const bool leafnode = t.size() == 1 && t[0].is_leaf();
        m_par.push_back(parent);
        m_hr.push_back(static_cast<int64_t>(t.value()));
        m_cap.push_back(static_cast<int64_t>(c.value()));
        m_leafdim.push_back(leafnode ? static_cast<int64_t>(t[0].value()) : 0);
        if (leafnode) n.m_leaf = leaves++;
        n.m_A.resize(static_cast<size_type>(rows), t.value());
        if (leafnode) return;
        for (size_type k = 0; k < t.size(); ++k) flatten(t[k], c[k], id, leaves);
    }
    std::vector<node> m_nodes;
    std::vector<int64_t> m_par, m_hr, m_cap, m_leafdim;
    uint64_t m_version = 0;
    const void* m_source = nullptr;
    mutable std::function<void(httensor&)> m_pull;
    std::shared_ptr<int> m_alive = std::make_shared<int>(0);
};

// ------------------------------------------------------------------ operators
namespace ops {
template <typename T, typename backend = blas_backend>
struct dense_matrix_operator {
    dense_matrix_operator() = default;
    explicit dense_matrix_operator(const matrix<T>& m) : mat(m) {}
    matrix<T> mat;
};
template <typename T, typename backend = blas_backend>
struct diagonal_matrix_operator {
    explicit diagonal_matrix_operator(const diagonal_matrix<T>& d) : mat(d) {}
    diagonal_matrix<T> mat;
};
template <typename T, typename backend = blas_backend>
struct sparse_matrix_operator {
    explicit sparse_matrix_operator(const csr_matrix<T>& m) : mat(m) {}
    csr_matrix<T> mat;
};
}  // namespace ops

// ------------------------------------------------------------------ bath modes of the driver
// The spin-boson driver's mode handling (src/mode_data.hpp): partition_modes (:818-886) and
// harmonic_mode_combination (:21-361) with the calls the driver makes (set_wk / set_gk over
// iterator ranges, construct_basis_topology, H0, Hc; src/spin_boson.cpp:418-441, 515-538).
template <typename T>
bool partition_modes(const std::vector<T>& wk, T ecut, T wc, std::vector<size_t>& partition, size_t nmax,
                     size_t nmax4, size_t nmax2, size_t nmax3) {
    std::vector<double> w(wk.begin(), wk.end());
    std::vector<int64_t> part(w.size());
    int64_t nb = 0;
    detail::check(ttn_sb_partition_modes(nullptr, static_cast<int64_t>(w.size()), w.data(), ecut, wc,
                                         static_cast<int64_t>(nmax), static_cast<int64_t>(nmax4),
                                         static_cast<int64_t>(nmax2), static_cast<int64_t>(nmax3), &nb, part.data()),
                  nullptr, "partition_modes");
    partition.assign(part.begin(), part.begin() + nb);
    return true;
}

template <typename T>
class harmonic_mode_combination {
  public:
    template <class It>
    void set_wk(It a, It b) { m_w.assign(a, b); }
    template <class It>
    void set_gk(It a, It b) { m_g.assign(a, b); }
    // number of states of the block (the leaf dimension)
    size_t construct_basis_topology(T ecut, size_t nmax) {
        if (m_g.size() != m_w.size()) m_g.assign(m_w.size(), T(0));
        m_ecut = ecut;
        m_nmax = nmax;
        int64_t ns = 0, nnz = 0;
        detail::check(ttn_sb_mode_block(nullptr, static_cast<int64_t>(m_w.size()), m_w.data(), m_g.data(), ecut,
                                        static_cast<int64_t>(nmax), &ns, &nnz, nullptr, nullptr, nullptr, nullptr,
                                        nullptr),
                      nullptr, "construct_basis_topology");
        return static_cast<size_t>(ns);
    }
    template <typename U>
    void H0(diagonal_matrix<U>& H) const {
        std::vector<double> d;
        tables<std::complex<double>>(&d, nullptr);
        H.diag.assign(d.begin(), d.end());
    }
    template <typename U>
    void Hc(csr_matrix<U>& H) const {
        tables(nullptr, &H);
    }

  private:
    template <typename U>
    void tables(std::vector<double>* h0, csr_matrix<U>* hc) const {
        int64_t ns = 0, nnz = 0;
        const int64_t n = static_cast<int64_t>(m_w.size());
        detail::check(ttn_sb_mode_block(nullptr, n, m_w.data(), m_g.data(), m_ecut, static_cast<int64_t>(m_nmax), &ns,
                                        &nnz, nullptr, nullptr, nullptr, nullptr, nullptr),
                      nullptr, "harmonic_mode_combination");
        std::vector<double> d(ns), vals(nnz);
        std::vector<int64_t> rp(ns + 1), ci(nnz);
        detail::check(ttn_sb_mode_block(nullptr, n, m_w.data(), m_g.data(), m_ecut, static_cast<int64_t>(m_nmax), &ns,
                                        &nnz, nullptr, d.data(), rp.data(), ci.data(), vals.data()),
                      nullptr, "harmonic_mode_combination");
        if (h0) *h0 = d;
        if (hc) {
            hc->n = static_cast<size_type>(ns);
            hc->rowptr = rp;
            hc->colind = ci;
            hc->values.assign(vals.begin(), vals.end());
        }
    }
    std::vector<double> m_w, m_g;
    T m_ecut = T(0);
    size_t m_nmax = 0;
};

template <typename T, typename backend = blas_backend>
class sop_operator {
  public:
    sop_operator(size_type nterms, const std::vector<size_type>& mode_dims)
        : m_nterms(nterms), m_dims(mode_dims.begin(), mode_dims.end()) {}
    // the reference's third argument (expected terms per mode) only reserves storage
    sop_operator(size_type nterms, const std::vector<size_type>& mode_dims, const std::vector<size_type>&)
        : sop_operator(nterms, mode_dims) {}

    void bind(const ops::dense_matrix_operator<T, backend>& op, const std::vector<size_type>& r, size_type mode) {
        add(mode, TTN_OP_DENSE, op.mat.data(), op.mat.size(0) * op.mat.size(1), {}, r);
    }
    void bind(const ops::diagonal_matrix_operator<T, backend>& op, const std::vector<size_type>& r, size_type mode) {
        add(mode, TTN_OP_DIAG, op.mat.diag.data(), op.mat.diag.size(), {}, r);
    }
    void bind(const ops::sparse_matrix_operator<T, backend>& op, const std::vector<size_type>& r, size_type mode) {
        std::vector<int64_t> idx(op.mat.rowptr);
        idx.insert(idx.end(), op.mat.colind.begin(), op.mat.colind.end());
        add(mode, TTN_OP_CSR, op.mat.values.data(), op.mat.values.size(), idx, r);
    }
    // coeff() / coeff(r) (sop_operator.hpp:471-473): read by the integrator on every step
    std::vector<T>& coeff() { return m_coeff_store(); }
    T& coeff(size_type r) { return m_coeff_store().at(r); }
    T coeff(size_type r) const { return r < m_coeff.size() ? m_coeff[r] : T(1.0); }
    size_type nterms() const { return m_nterms; }
    // interleaved re/im of every coefficient (1 where none was set)
    std::vector<double> coefficients() const {
        std::vector<double> c(2 * m_nterms);
        for (size_type t = 0; t < m_nterms; ++t) {
            const T v = coeff(t);
            c[2 * t] = v.real();
            c[2 * t + 1] = v.imag();
        }
        return c;
    }

    // C-ABI handle on a context (rebuilt lazily by the integrator)
    ttn_sop* create(ttn_ctx* ctx) const {
        const std::vector<double> coeff = coefficients();
        std::vector<int32_t> mode(m_mode.begin(), m_mode.end()), kind(m_kind.begin(), m_kind.end());
        ttn_sop* s = nullptr;
        detail::check(ttn_sop_create(ctx, static_cast<int64_t>(m_dims.size()), m_dims.data(),
                                     static_cast<int64_t>(m_nterms), coeff.data(), static_cast<int64_t>(mode.size()),
                                     mode.data(), kind.data(), m_coff.data(), m_cdata.data(), m_ioff.data(),
                                     m_idata.data(), m_roff.data(), m_r.data(), &s),
                      ctx, "ttn_sop_create");
        return s;
    }

  private:
    std::vector<T>& m_coeff_store() {
        if (m_coeff.empty()) m_coeff.assign(m_nterms, T(1.0));
        return m_coeff;
    }
    void add(size_type mode, int kind, const T* data, size_type n, const std::vector<int64_t>& idx,
             const std::vector<size_type>& r) {
        if (mode >= m_dims.size()) throw std::invalid_argument("sop_operator::bind: mode out of range");
        m_mode.push_back(static_cast<int32_t>(mode));
        m_kind.push_back(kind);
        for (size_type i = 0; i < n; ++i) {
            m_cdata.push_back(data[i].real());
            m_cdata.push_back(data[i].imag());
        }
        m_coff.push_back(m_coff.back() + static_cast<int64_t>(n));
        m_idata.insert(m_idata.end(), idx.begin(), idx.end());
        m_ioff.push_back(static_cast<int64_t>(m_idata.size()));
        for (size_type t : r) m_r.push_back(static_cast<int64_t>(t));
        m_roff.push_back(static_cast<int64_t>(m_r.size()));
    }
    size_type m_nterms;
    std::vector<int64_t> m_dims;
    std::vector<T> m_coeff;
    std::vector<int32_t> m_mode, m_kind;
    std::vector<double> m_cdata;
    std::vector<int64_t> m_coff{0}, m_idata, m_ioff{0}, m_r, m_roff{0};
};

// ------------------------------------------------------------------ integrator
template <typename T, typename backend = blas_backend>
class fixed_step_projector_splitting_integrator;
template <typename T, typename backend>
class subspace_expansion_projector_splitting_integrator;

template <typename T, typename backend = blas_backend>
class matrix_element {
  public:
    // <psi|O|psi> for a dense operator on primitive mode `mode` (matrix_element.hpp:264-314)
    T operator()(const ops::dense_matrix_operator<T, backend>& op, size_type mode, httensor<T, backend>& A) {
        return m_integ->expect(op, mode, A);
    }

  private:
    friend class fixed_step_projector_splitting_integrator<T, backend>;
    fixed_step_projector_splitting_integrator<T, backend>* m_integ = nullptr;
};

// The integrators (projector_splitting_integrator.hpp:18-397).  The reference's state lives in
// the integrator object: dt(), krylov_dim(), krylov_tolerance(), nthreads() and the expansion
// knobs are plain members the caller assigns, and every call reads them.  Here the device
// integrator is created on first use (a step or mel()) with the values set by then; dt, the
// operator's coefficients and the expansion knobs are forwarded again whenever they changed
// since the last call.  krylov_dim / krylov_tolerance take effect at initialise().
template <typename T, typename backend>
class fixed_step_projector_splitting_integrator {
  public:
    using real_type = double;
    fixed_step_projector_splitting_integrator() { m_mel.m_integ = this; }
    fixed_step_projector_splitting_integrator(httensor<T, backend>& A, sop_operator<T, backend>& H,
                                              size_type krylov_dim = 4, real_type krylov_tol = 1e-12,
                                              size_type nthreads = 1)
        : m_kdim(krylov_dim), m_tol(krylov_tol), m_nthreads(nthreads) {
        m_mel.m_integ = this;
        initialise(A, H);
    }
    virtual ~fixed_step_projector_splitting_integrator() {
        release();
        if (m_ctx) ttn_ctx_destroy(m_ctx);
    }
    fixed_step_projector_splitting_integrator(const fixed_step_projector_splitting_integrator&) = delete;
    fixed_step_projector_splitting_integrator& operator=(const fixed_step_projector_splitting_integrator&) = delete;

    // initialise(A, op) (:100-116): binds the tree (its current bond dimensions and capacities)
    // and the operator; the previous device state, if any, is copied into its tensor first
    void initialise(httensor<T, backend>& A, sop_operator<T, backend>& H) {
        release();
        if (!m_ctx) detail::check(ttn_ctx_create(0, &m_ctx), m_ctx, "ttn_ctx_create");
        const auto& p = A.parent();
        detail::check(ttn_tree_create(m_ctx, static_cast<int64_t>(p.size()), p.data(), A.hranks().data(),
                                      A.cap_hranks().data(), A.leaf_dims().data(), &m_tree),
                      m_ctx, "ttn_tree_create");
        m_sop = H.create(m_ctx);
        m_coeff_sent = H.coefficients();
    }
    // clear() (:118-137): unbinds and returns to the default parameters
    void clear() {
        release();
        m_kdim = 4;
        m_tol = 1e-12;
        m_nthreads = 1;
        m_dt = 0.0;
    }
    // reset_nsteps() (:33): t() restarts at 0 and the next call is a first step again
    // (prepare_evolution and the sub-step ladder)
    void reset_nsteps() {
        if (m_integ) detail::check(ttn_reset_nsteps(m_integ), m_ctx, "ttn_reset_nsteps");
    }

    real_type& dt() { return m_dt; }
    const real_type& dt() const { return m_dt; }
    void set_dt(real_type dt) { m_dt = dt; }
    // t() = dt() * nsteps, with the current dt() (:188, 362)
    real_type t() const { return m_dt * static_cast<real_type>(nsteps()); }
    size_type& krylov_dim() { return m_kdim; }
    const size_type& krylov_dim() const { return m_kdim; }
    real_type& krylov_tolerance() { return m_tol; }
    const real_type& krylov_tolerance() const { return m_tol; }
    size_type& nthreads() { return m_nthreads; }
    const size_type& nthreads() const { return m_nthreads; }
    size_type nh_applications() const {
        int64_t nh = 0, ns = 0;
        if (m_integ) detail::check(ttn_counters(m_integ, &nh, &ns), m_ctx, "ttn_counters");
        return static_cast<size_type>(nh);
    }
    matrix_element<T, backend>& mel() { return m_mel; }

    // one call A <- TDVP(A, H, coeff) of length dt() (coeff = -i for real time); returns dt().
    // The first call after construction or reset_nsteps() runs prepare_evolution and the
    // sub-step ladder (:145-168, 284-312).
    real_type operator()(httensor<T, backend>& A, sop_operator<T, backend>& H, T coeff) {
        if (!(m_dt > 0.0)) throw std::runtime_error("Failed to apply projector splitting integrator.\ndt must be positive");
        bind(A);
        if (m_dt != m_dt_sent) {
            detail::check(ttn_set_dt(m_integ, m_dt), m_ctx, "ttn_set_dt");
            m_dt_sent = m_dt;
        }
        const std::vector<double> c = H.coefficients();
        if (c != m_coeff_sent) {
            detail::check(ttn_set_coefficients(m_integ, static_cast<int64_t>(H.nterms()), c.data()), m_ctx,
                          "ttn_set_coefficients");
            m_coeff_sent = c;
        }
        forward_parameters();
        detail::check(ttn_step(m_integ, coeff.real(), coeff.imag(), 1), m_ctx, "ttn_step");
        std::vector<int64_t> hr(A.size());
        detail::check(ttn_hranks(m_integ, hr.data()), m_ctx, "ttn_hranks");
        A.mark_device_newer(this, [this](httensor<T, backend>& B) { copy_from_device(B); }, hr);
        m_version = A.version();
        return m_dt;
    }

    // checkpoint / restart in this library's own format (ttn_checkpoint_save / _load; the
    // reference's cereal archives, projector_splitting_integrator.hpp:50-68, 198-216): the
    // integrator's state after the last call, and A taken back to that state
    std::vector<char> checkpoint(httensor<T, backend>& A) {
        bind(A);
        int64_t n = 0;
        detail::check(ttn_checkpoint_save(m_integ, &n, nullptr), m_ctx, "ttn_checkpoint_save");
        std::vector<char> buf(static_cast<size_t>(n));
        detail::check(ttn_checkpoint_save(m_integ, &n, buf.data()), m_ctx, "ttn_checkpoint_save");
        buf.resize(static_cast<size_t>(n));
        return buf;
    }
    void restore(const std::vector<char>& buf, httensor<T, backend>& A) {
        ensure();
        if (m_bound && !m_alive.expired()) m_bound->detach(this);
        detail::check(ttn_checkpoint_load(m_integ, buf.data(), static_cast<int64_t>(buf.size())), m_ctx,
                      "ttn_checkpoint_load");
        double dt = 0.0;
        detail::check(ttn_get_dt(m_integ, &dt), m_ctx, "ttn_get_dt");
        m_dt = m_dt_sent = dt;
        std::vector<int64_t> hr(A.size());
        detail::check(ttn_hranks(m_integ, hr.data()), m_ctx, "ttn_hranks");
        A.pull();
        A.mark_device_newer(this, [this](httensor<T, backend>& B) { copy_from_device(B); }, hr);
        m_bound = &A;
        m_alive = A.alive();
        m_version = A.version();
    }

    T expect(const ops::dense_matrix_operator<T, backend>& op, size_type mode, httensor<T, backend>& A) {
        bind(A);
        double out[2] = {0.0, 0.0};
        detail::check(ttn_expect(m_integ, static_cast<int64_t>(mode),
                                 reinterpret_cast<const double*>(op.mat.data()), out),
                      m_ctx, "ttn_expect");
        return T(out[0], out[1]);
    }

  protected:
    // the expansion integrator forwards its knobs here
    virtual void forward_parameters() {}
    ttn_integ_params params() const {
        ttn_integ_params p{};
        p.krylov_dim = static_cast<int64_t>(m_kdim);
        p.krylov_tol = m_tol;
        p.dt = m_dt > 0.0 ? m_dt : 1.0;  // forwarded again before every step
        p.mode = m_mode;
        p.spawning_threshold = m_spawning_threshold;
        p.unoccupied_threshold = m_unoccupied_threshold;
        p.minimum_unoccupied = static_cast<int64_t>(m_minimum_unoccupied);
        p.neigenvalues = static_cast<int64_t>(m_neigenvalues);
        p.only_apply_when_no_unoccupied = m_only_unoccupied ? 1 : 0;
        return p;
    }
    ttn_integ* handle() const { return m_integ; }
    ttn_ctx* context() const { return m_ctx; }

    int m_mode = TTN_MODE_FIXED_STEP;
    real_type m_spawning_threshold = -1.0;
    real_type m_unoccupied_threshold = -1.0;
    size_type m_minimum_unoccupied = 0;
    size_type m_neigenvalues = 2;
    bool m_only_unoccupied = false;

  private:
    int64_t nsteps() const {
        int64_t nh = 0, ns = 0;
        if (m_integ) detail::check(ttn_counters(m_integ, &nh, &ns), m_ctx, "ttn_counters");
        return ns;
    }
    void ensure() {
        if (m_integ) return;
        if (!m_tree) throw std::runtime_error("Failed to apply projector splitting integrator.\nnot initialised");
        const ttn_integ_params p = params();
        detail::check(ttn_integrator_create(m_ctx, m_tree, m_sop, &p, &m_integ), m_ctx, "ttn_integrator_create");
        m_dt_sent = p.dt;
    }
    // make the device hold A: copy A up unless the device state is A's current one
    void bind(httensor<T, backend>& A) {
        ensure();
        if (A.device_current(this, m_version)) return;
        // the tensor bound so far keeps this device's newer state; A's own newer state (from
        // another integrator) is copied into A before A goes up
        if (m_bound && m_bound != &A && !m_alive.expired()) m_bound->detach(this);
        A.pull();
        for (size_type i = 0; i < A.size(); ++i) {
            const matrix<T>& m = A.raw(i);
            detail::check(ttn_state_set(m_integ, static_cast<int64_t>(i), reinterpret_cast<const double*>(m.data()),
                                        static_cast<int64_t>(m.size(0)), static_cast<int64_t>(m.size(1))),
                          m_ctx, "ttn_state_set");
        }
        m_bound = &A;
        m_alive = A.alive();
        A.mark_device_newer(this, nullptr, A.hranks());
        m_version = A.version();
    }
    void copy_from_device(httensor<T, backend>& A) {
        for (size_type i = 0; i < A.size(); ++i) {
            int64_t rows = 0, cols = 0;
            detail::check(ttn_state_shape(m_integ, static_cast<int64_t>(i), &rows, &cols), m_ctx, "ttn_state_shape");
            matrix<T>& m = A.raw(i);
            if (m.size(0) != static_cast<size_type>(rows) || m.size(1) != static_cast<size_type>(cols))
                m.resize(static_cast<size_type>(rows), static_cast<size_type>(cols));
            detail::check(ttn_state_get(m_integ, static_cast<int64_t>(i), reinterpret_cast<double*>(m.data()), rows, cols),
                          m_ctx, "ttn_state_get");
        }
    }
    // the device state is about to go: the bound tensor gets it first
    void release() {
        if (m_bound && !m_alive.expired()) m_bound->detach(this);
        m_bound = nullptr;
        m_alive.reset();
        if (m_integ) ttn_integrator_destroy(m_integ);
        if (m_sop) ttn_sop_destroy(m_sop);
        if (m_tree) ttn_tree_destroy(m_tree);
        m_integ = nullptr;
        m_sop = nullptr;
        m_tree = nullptr;
    }

    ttn_ctx* m_ctx = nullptr;
    ttn_tree* m_tree = nullptr;
    ttn_sop* m_sop = nullptr;
    ttn_integ* m_integ = nullptr;
    size_type m_kdim = 4;
    real_type m_tol = 1e-12;
    size_type m_nthreads = 1;
    real_type m_dt = 0.0, m_dt_sent = 0.0;
    std::vector<double> m_coeff_sent;
    httensor<T, backend>* m_bound = nullptr;
    std::weak_ptr<int> m_alive;
    uint64_t m_version = 0;
    matrix_element<T, backend> m_mel;
};

// engine() of the expansion integrator: the counters the reference's driver prints
// (subspace_expansion_projector_splitting_engine.hpp:513-514, src/spin_boson.cpp:91)
class expansion_engine_view {
  public:
    size_type Nonesite() const { return counters().first; }
    size_type Ntwosite() const { return counters().second; }

  private:
    template <typename, typename>
    friend class subspace_expansion_projector_splitting_integrator;
    std::pair<size_type, size_type> counters() const {
        int64_t n1 = 0, n2 = 0;
        if (m_integ && *m_integ) detail::check(ttn_expansion_counters(*m_integ, &n1, &n2), nullptr, "ttn_expansion_counters");
        return {static_cast<size_type>(n1), static_cast<size_type>(n2)};
    }
    ttn_integ* const* m_integ = nullptr;
};

// subspace_expansion_projector_splitting_integrator (projector_splitting_integrator.hpp:229-397):
// the bonds grow up to the capacities of the tree's second ntree; the knobs are those the
// driver sets (src/spin_boson.cpp:652-658)
template <typename T, typename backend = blas_backend>
class subspace_expansion_projector_splitting_integrator : public fixed_step_projector_splitting_integrator<T, backend> {
    using base = fixed_step_projector_splitting_integrator<T, backend>;

  public:
    using real_type = typename base::real_type;
    subspace_expansion_projector_splitting_integrator(httensor<T, backend>& A, sop_operator<T, backend>& H,
                                                      size_type krylov_dim = 4, real_type krylov_tol = 1e-12,
                                                      size_type nthreads = 1)
        : base(A, H, krylov_dim, krylov_tol, nthreads) {
        this->m_mode = TTN_MODE_SUBSPACE_EXPANSION;
    }
    real_type& spawning_threshold() { return this->m_spawning_threshold; }
    const real_type& spawning_threshold() const { return this->m_spawning_threshold; }
    real_type& unoccupied_threshold() { return this->m_unoccupied_threshold; }
    const real_type& unoccupied_threshold() const { return this->m_unoccupied_threshold; }
    size_type& minimum_unoccupied() { return this->m_minimum_unoccupied; }
    const size_type& minimum_unoccupied() const { return this->m_minimum_unoccupied; }
    size_type& neigenvalues() { return this->m_neigenvalues; }
    const size_type& neigenvalues() const { return this->m_neigenvalues; }
    bool& only_apply_when_no_unoccupied() { return this->m_only_unoccupied; }
    const bool& only_apply_when_no_unoccupied() const { return this->m_only_unoccupied; }

    const expansion_engine_view& engine() const {
        m_view.m_integ = &m_handle;
        m_handle = this->handle();
        return m_view;
    }

  protected:
    void forward_parameters() override {
        const ttn_integ_params p = this->params();
        const std::vector<double> now = {p.spawning_threshold, p.unoccupied_threshold,
                                         static_cast<double>(p.minimum_unoccupied), static_cast<double>(p.neigenvalues),
                                         static_cast<double>(p.only_apply_when_no_unoccupied)};
        if (now == m_sent) return;
        detail::check(ttn_set_expansion_params(this->handle(), &p), this->context(), "ttn_set_expansion_params");
        m_sent = now;
    }

  private:
    mutable expansion_engine_view m_view;
    mutable ttn_integ* m_handle = nullptr;
    std::vector<double> m_sent;
};

}  // namespace ttns
