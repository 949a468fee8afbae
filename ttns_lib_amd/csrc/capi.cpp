// C ABI of the MI355X TDVP library (include/ttn_capi.h).
//
// Host side of the drop-in: owns the tree layout, the per-instance HBM arena holding
// every node tensor, gauge matrix and environment, the compiled term tables flattened
// for the device, and the sweep schedule.  Each ttn_step / ttn_batch_step is a single
// launch of the persistent sweep kernel (device/sweep.hip).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <string>

#include "../../include/ttn_capi.h"
#include "device/desc.hpp"
#include "host/flops.hpp"
#include "host/marshal.hpp"
#include "host/tables.hpp"

extern "C" hipError_t ttnd_launch_run(const ttnd::DInst* d_insts, int n, const ttnd::RunArgs* a, hipStream_t s);
extern "C" hipError_t ttnd_launch_expect(const ttnd::DInst* d_inst, int leaf, int64_t op_off, double* d_out,
                                         hipStream_t s);

using namespace ttnh;

namespace {

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) fail(ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    void alloc(size_t count) {
        release();
        if (count == 0) count = 1;
        hip_check(hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)), "hipMalloc");
        n = count;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DevBuf() { release(); }
};

int64_t align16(int64_t x) { return (x + 15) & ~int64_t(15); }

}  // namespace

struct ttn_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
};
struct ttn_tree {
    ttn_ctx* ctx;
    Topology topo;
};
struct ttn_sop {
    ttn_ctx* ctx;
    SOP sop;
};

struct ttn_integ {
    ttn_ctx* ctx;
    Topology topo;
    SOP sop;
    CompiledSOP ops;
    ttn_integ_params p{};
    ttnd::DInst inst{};         // host copy (device pointers inside)
    DevBuf<double2> arena;
    DevBuf<char> tables;
    DevBuf<ttnd::DInst> d_inst;
    DevBuf<int64_t> counters;
    DevBuf<int> status;
    DevBuf<double> traj;
    DevBuf<double> scratch;
    DevBuf<int64_t> node_stats;
    DevBuf<int64_t> prof;
    FlopModel fm;
    int64_t sweep_pairs = 0;   // forward+backward sweeps run (with or without propagation)
    int64_t prepares = 0;
    int64_t arena_size = 0;
    std::vector<int64_t> node_A;
    int64_t obs_op = -1;
    int64_t tmp_op = -1;
    int obs_leaf = -1;
    int64_t obs_dim = 0;
    int64_t nsteps = 0;
    double last_ms = 0.0;
    std::vector<double> trajectory;  // (re <O>, <1>) per completed step
};

struct ttn_batch {
    ttn_ctx* ctx;
    std::vector<ttn_integ*> integs;
    DevBuf<ttnd::DInst> d_insts;
    double last_ms = 0.0;
};

namespace {

template <class F>
int guard(ttn_ctx* ctx, F&& f) {
    try {
        f();
        return TTN_OK;
    } catch (const Error& e) {
        if (ctx) ctx->err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        if (ctx) ctx->err = e.what();
        return TTN_ERR_INVALID;
    }
}

// Flatten topology + compiled tables into device descriptors and lay out the arena.
void build_instance(ttn_integ* in) {
    const Topology& t = in->topo;
    const CompiledSOP& ops = in->ops;
    const int n = t.size();
    check(n <= 1024, ERR_INVALID, "device path supports at most 1024 nodes");
    const int kdim = static_cast<int>(in->p.krylov_dim);
    check(kdim >= 1 && kdim <= ttnd::MAXK, ERR_INVALID, "krylov_dim must be in [1, 16] on the device path");

    int64_t off = 0;
    auto alloc = [&off](int64_t cnt) {
        const int64_t o = align16(off);
        off = o + std::max<int64_t>(cnt, 1);
        return o;
    };
    std::vector<std::pair<int64_t, std::vector<cplx>>> init;  // arena initial data

    std::vector<ttnd::DNode> dn(n);
    int64_t maxsize = 1;
    in->node_A.assign(n, 0);
    for (int i = 0; i < n; ++i) {
        const Node& nd = t.nodes[i];
        check(nd.nmodes() <= ttnd::MAXC, ERR_INVALID, "device path supports at most 4 children per node");
        check(nd.hrank <= 64, ERR_INVALID, "device path supports bond dimensions up to 64");
        for (i64 d : nd.dims) check(d <= 64 || nd.is_leaf(), ERR_INVALID, "device path supports mode dims up to 64");
        ttnd::DNode& d = dn[i];
        d.parent = nd.parent;
        d.child_id = nd.child_id;
        d.nchild = static_cast<int>(nd.children.size());
        d.leaf = nd.leaf_index;
        d.nmodes = nd.nmodes();
        d.hrank = static_cast<int>(nd.hrank);
        d.dimen = static_cast<int>(nd.dimen());
        d.size = static_cast<int>(nd.size());
        for (int c = 0; c < ttnd::MAXC; ++c) {
            d.dims[c] = c < nd.nmodes() ? static_cast<int>(nd.dims[c]) : 1;
            d.children[c] = c < static_cast<int>(nd.children.size()) ? nd.children[c] : -1;
        }
        d.A = alloc(nd.size());
        in->node_A[i] = d.A;
        int64_t rmax = nd.hrank;
        if (!nd.is_leaf())
            for (i64 x : nd.dims) rmax = std::max<int64_t>(rmax, x);
        d.R = alloc(rmax * rmax);
        maxsize = std::max<int64_t>(maxsize, nd.size());
    }
    // environment matrices
    std::vector<std::vector<int64_t>> spf_off(n), mf_off(n);
    for (int i = 0; i < n; ++i) {
        const i64 r = t.nodes[i].hrank;
        const auto& terms = ops.nodes[i].terms;
        spf_off[i].assign(terms.size(), -1);
        mf_off[i].assign(terms.size(), -1);
        for (size_t k = 0; k < terms.size(); ++k) {
            if (!terms[k].id_spf) spf_off[i][k] = alloc(r * r);
            if (!terms[k].id_mf) mf_off[i][k] = alloc(r * r);
        }
    }
    // primitive operators (uploaded once per (mode, index))
    std::vector<int64_t> ints;
    std::map<std::pair<int, int>, ttnd::DK> prim;
    auto prim_dk = [&](int mode, int cind) {
        auto key = std::make_pair(mode, cind);
        auto it = prim.find(key);
        if (it != prim.end()) return it->second;
        const PrimOp& op = in->sop.ops[mode][cind].op;
        ttnd::DK k{};
        k.nu = mode;
        k.kind = op.kind;
        k.dim = static_cast<int>(op.dim);
        k.nnz = static_cast<int>(op.nnz());
        k.idx = -1;
        std::vector<cplx> data;
        if (op.kind == OP_DENSE)
            data = op.dense;
        else if (op.kind == OP_DIAG)
            data = op.diag;
        else if (op.kind == OP_CSR) {
            data = op.vals;
            k.idx = static_cast<int64_t>(ints.size());
            ints.insert(ints.end(), op.rowptr.begin(), op.rowptr.end());
            ints.insert(ints.end(), op.colind.begin(), op.colind.end());
        } else {
            fail(ERR_INVALID, "identity primitive operators are not stored");
        }
        k.mat = alloc(static_cast<int64_t>(data.size()));
        init.emplace_back(k.mat, data);
        prim[key] = k;
        return k;
    };
    // term tables
    std::vector<ttnd::DTerm> terms;
    std::vector<ttnd::DSub> subs;
    std::vector<ttnd::DK> ks;
    std::vector<ttnd::DMf> mfs;
    for (int i = 0; i < n; ++i) {
        const Node& nd = t.nodes[i];
        const auto& tl = ops.nodes[i].terms;
        dn[i].term0 = static_cast<int>(terms.size());
        dn[i].nterms = static_cast<int>(tl.size());
        for (size_t k = 0; k < tl.size(); ++k) {
            const Term& tm = tl[k];
            ttnd::DTerm d{};
            d.id_spf = tm.id_spf;
            d.id_mf = tm.id_mf;
            d.cre = tm.coeff.real();
            d.cim = tm.coeff.imag();
            cplx ic = tm.coeff;
            if (tm.nspf_terms() > 1) {
                ic = cplx(0.0, 0.0);
                for (const cplx& a : tm.accum) ic += a;
            }
            d.ire = ic.real();
            d.iim = ic.imag();
            d.spf = spf_off[i][k];
            d.mf = mf_off[i][k];
            d.sub0 = static_cast<int>(subs.size());
            d.nsub = tm.nspf_terms();
            for (int s = 0; s < tm.nspf_terms(); ++s) {
                ttnd::DSub ds{};
                const cplx a = (tm.nspf_terms() > 1) ? tm.accum.at(s) : cplx(0.0, 0.0);
                ds.are = a.real();
                ds.aim = a.imag();
                ds.k0 = static_cast<int>(ks.size());
                ds.nk = static_cast<int>(tm.spf_index[s].size());
                for (const auto& pr : tm.spf_index[s]) {
                    if (nd.is_leaf()) {
                        ks.push_back(prim_dk(pr.first, pr.second));
                    } else {
                        const int ch = nd.children[pr.first];
                        ttnd::DK dk{};
                        dk.nu = pr.first;
                        dk.kind = 0;
                        dk.dim = static_cast<int>(nd.dims[pr.first]);
                        dk.mat = spf_off[ch][pr.second];
                        check(dk.mat >= 0, ERR_INVALID, "spf chain references an identity term");
                        ks.push_back(dk);
                    }
                }
                subs.push_back(ds);
            }
            d.mf0 = static_cast<int>(mfs.size());
            d.nmf = tm.nmf_terms();
            for (int m = 0; m < tm.nmf_terms(); ++m) {
                const MfIndex& mi = tm.mf_index[m];
                ttnd::DMf dm{};
                const cplx a = (tm.nmf_terms() > 1) ? tm.accum.at(m) : cplx(0.0, 0.0);
                dm.are = a.real();
                dm.aim = a.imag();
                const int p = nd.parent;
                dm.pmf = ops.nodes[p].terms[mi.parent_index].id_mf ? -1 : mf_off[p][mi.parent_index];
                dm.k0 = static_cast<int>(ks.size());
                dm.nk = static_cast<int>(mi.siblings.size());
                for (const auto& sb : mi.siblings) {
                    const int sib = t.nodes[p].children[sb.first];
                    ttnd::DK dk{};
                    dk.nu = sb.first;
                    dk.kind = 0;
                    dk.dim = static_cast<int>(t.nodes[p].dims[sb.first]);
                    dk.mat = spf_off[sib][sb.second];
                    check(dk.mat >= 0, ERR_INVALID, "mean-field chain references an identity term");
                    ks.push_back(dk);
                }
                mfs.push_back(dm);
            }
            terms.push_back(d);
        }
    }
    // observable slot and work buffers
    int64_t maxdim = 1;
    for (i64 d : t.mode_dims) maxdim = std::max<int64_t>(maxdim, d);
    in->obs_op = alloc(maxdim * maxdim);
    in->tmp_op = alloc(maxdim * maxdim);
    ttnd::DInst& I = in->inst;
    for (int k = 0; k < ttnd::NWORK; ++k) I.work[k] = alloc(maxsize);
    I.krylov = alloc((kdim + 1) * maxsize);
    I.maxsize = maxsize;
    in->arena_size = align16(off);

    // device allocations
    in->arena.alloc(in->arena_size);
    hip_check(hipMemsetAsync(in->arena.p, 0, in->arena_size * sizeof(double2), in->ctx->stream), "memset arena");
    for (auto& e : init)
        hip_check(hipMemcpyAsync(in->arena.p + e.first, e.second.data(), e.second.size() * sizeof(double2),
                                 hipMemcpyHostToDevice, in->ctx->stream),
                  "upload operators");
    const auto tour = euler_tour(t);
    std::vector<int> tour_i(tour.begin(), tour.end());
    // pack tables
    std::vector<char> blob;
    auto pack = [&blob](const void* src, size_t bytes) {
        size_t o = (blob.size() + 255) & ~size_t(255);
        blob.resize(o + std::max<size_t>(bytes, 1));
        if (bytes) std::memcpy(blob.data() + o, src, bytes);
        return o;
    };
    const size_t o_nodes = pack(dn.data(), dn.size() * sizeof(ttnd::DNode));
    const size_t o_terms = pack(terms.data(), terms.size() * sizeof(ttnd::DTerm));
    const size_t o_subs = pack(subs.data(), subs.size() * sizeof(ttnd::DSub));
    const size_t o_ks = pack(ks.data(), ks.size() * sizeof(ttnd::DK));
    const size_t o_mfs = pack(mfs.data(), mfs.size() * sizeof(ttnd::DMf));
    const size_t o_tour = pack(tour_i.data(), tour_i.size() * sizeof(int));
    const size_t o_ints = pack(ints.data(), ints.size() * sizeof(int64_t));
    in->tables.alloc(blob.size());
    hip_check(hipMemcpyAsync(in->tables.p, blob.data(), blob.size(), hipMemcpyHostToDevice, in->ctx->stream),
              "upload tables");
    in->counters.alloc(2);
    hip_check(hipMemsetAsync(in->counters.p, 0, 2 * sizeof(int64_t), in->ctx->stream), "memset counters");
    in->status.alloc(2);
    hip_check(hipMemsetAsync(in->status.p, 0, 2 * sizeof(int), in->ctx->stream), "memset status");
    in->scratch.alloc(4);
    in->node_stats.alloc(4 * n);
    hip_check(hipMemsetAsync(in->node_stats.p, 0, 4 * n * sizeof(int64_t), in->ctx->stream), "memset stats");
    in->fm = flop_model(t, ops, in->sop);
    in->prof.alloc(ttnd::PROF_NSLOTS);
    hip_check(hipMemsetAsync(in->prof.p, 0, ttnd::PROF_NSLOTS * sizeof(int64_t), in->ctx->stream), "memset prof");

    char* tb = in->tables.p;
    I.base = in->arena.p;
    I.nodes = reinterpret_cast<const ttnd::DNode*>(tb + o_nodes);
    I.terms = reinterpret_cast<const ttnd::DTerm*>(tb + o_terms);
    I.subs = reinterpret_cast<const ttnd::DSub*>(tb + o_subs);
    I.ks = reinterpret_cast<const ttnd::DK*>(tb + o_ks);
    I.mfs = reinterpret_cast<const ttnd::DMf*>(tb + o_mfs);
    I.tour = reinterpret_cast<const int*>(tb + o_tour);
    I.ints = reinterpret_cast<const int64_t*>(tb + o_ints);
    I.nnodes = n;
    I.tour_len = static_cast<int>(tour_i.size());
    I.kdim = kdim;
    I.tol = in->p.krylov_tol;
    I.obs_op = in->obs_op;
    I.obs_leaf = -1;
    I.traj = nullptr;
    I.counters = in->counters.p;
    I.node_stats = in->node_stats.p;
    I.prof = in->prof.p;
    I.status = in->status.p;
    in->d_inst.alloc(1);
    hip_check(hipMemcpyAsync(in->d_inst.p, &I, sizeof(I), hipMemcpyHostToDevice, in->ctx->stream), "upload inst");
    hip_check(hipStreamSynchronize(in->ctx->stream), "sync");
}

// first-step ladder of the reference integrator (projector_splitting_integrator.hpp:138-160)
int ladder(double dtfull, double* out) {
    int cnt = 0;
    double tevolve = 0.0, dt = dtfull / 1e5;
    while (tevolve < dtfull && cnt < 8) {
        out[cnt++] = (dt + tevolve <= dtfull ? dt : dtfull - tevolve);
        tevolve += dt;
        dt *= 10;
    }
    return cnt;
}

void make_args(const ttn_integ* in, double cr, double ci, int64_t nsteps, ttnd::RunArgs& a) {
    std::memset(&a, 0, sizeof(a));
    a.coeff = make_double2(cr, ci);
    a.dt = in->p.dt;
    a.step0 = 0;
    if (in->nsteps == 0) {
        a.prepare = 1;
        a.nladder = ladder(in->p.dt, a.ladder_dt);
        a.nregular = static_cast<int>(nsteps - 1);
    } else {
        a.nregular = static_cast<int>(nsteps);
    }
}

void prepare_traj(ttn_integ* in, int64_t nsteps) {
    if (in->obs_leaf < 0) {
        in->inst.traj = nullptr;
        return;
    }
    if (static_cast<int64_t>(in->traj.n) < 2 * nsteps) in->traj.alloc(2 * nsteps);
    in->inst.traj = in->traj.p;
    in->inst.obs_leaf = in->obs_leaf;
}

void account(ttn_integ* in, const ttnd::RunArgs& a) {
    if (a.prepare) {
        ++in->prepares;
        ++in->sweep_pairs;
    }
    in->sweep_pairs += a.nladder + a.nregular;
}

// Message for a failure record of the sweep kernel (reason | bond << 4 | node << 5).  The
// record is sticky: an instance whose sweep failed keeps failing until it is recreated.
std::string failure_message(const int st[2]) {
    static const char* reasons[] = {"", "the vector to propagate has a non-finite norm",
                                    "a Krylov basis vector has a non-finite norm",
                                    "no Krylov sub-step met the tolerance"};
    const int reason = st[0] & 15, bond = (st[0] >> 4) & 1, node = st[0] >> 5;
    std::string m = "TDVP sweep failed at step " + std::to_string(st[1]) + ", " +
                    (bond ? "bond propagation above node " : "node propagation of node ") + std::to_string(node) +
                    ": " + (reason < 4 ? reasons[reason] : "unknown reason");
    return m;
}

void finish(ttn_integ* in, int64_t nsteps) {
    int st[2] = {0, 0};
    hip_check(hipMemcpy(st, in->status.p, sizeof(st), hipMemcpyDeviceToHost), "status");
    if (in->obs_leaf >= 0 && nsteps > 0) {
        std::vector<double> buf(2 * nsteps);
        hip_check(hipMemcpy(buf.data(), in->traj.p, buf.size() * sizeof(double), hipMemcpyDeviceToHost), "traj");
        in->trajectory.insert(in->trajectory.end(), buf.begin(), buf.end());
    }
    if (st[0] != 0) fail(ERR_NUMERIC, failure_message(st));
    in->nsteps += nsteps;
}

}  // namespace

extern "C" {

int ttn_version(int* major, int* minor) {
    *major = 0;
    *minor = 1;
    return TTN_OK;
}

int ttn_ctx_create(int device, ttn_ctx** out) {
    auto c = std::make_unique<ttn_ctx>();
    c->device = device;
    int rc = guard(c.get(), [&] {
        int ndev = 0;
        hip_check(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
        check(device >= 0 && device < ndev, ERR_DEVICE, "no such GPU");
        hip_check(hipSetDevice(device), "hipSetDevice");
        hip_check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
        hip_check(hipEventCreate(&c->ev0), "hipEventCreate");
        hip_check(hipEventCreate(&c->ev1), "hipEventCreate");
    });
    if (rc != TTN_OK) {
        static std::string last;
        last = c->err;
        *out = nullptr;
        return rc;
    }
    *out = c.release();
    return TTN_OK;
}

int ttn_ctx_destroy(ttn_ctx* c) {
    if (!c) return TTN_OK;
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return TTN_OK;
}

const char* ttn_last_error(const ttn_ctx* c) { return c ? c->err.c_str() : "no context (device unavailable?)"; }

int ttn_tree_create(ttn_ctx* ctx, int64_t n, const int64_t* parent, const int64_t* hrank, const int64_t* cap_hrank,
                    const int64_t* leaf_dim, ttn_tree** out) {
    return guard(ctx, [&] {
        auto t = std::make_unique<ttn_tree>();
        t->ctx = ctx;
        t->topo = topology_from_arrays(static_cast<int>(n), parent, hrank, cap_hrank, leaf_dim);
        *out = t.release();
    });
}
int ttn_tree_destroy(ttn_tree* t) {
    delete t;
    return TTN_OK;
}
int ttn_tree_node(const ttn_tree* t, int64_t node, int64_t* nchild, int64_t* rows, int64_t* hrank, int64_t* leaf) {
    return guard(t->ctx, [&] {
        check(node >= 0 && node < t->topo.size(), ERR_INVALID, "node out of range");
        const Node& nd = t->topo.nodes[node];
        if (nchild) *nchild = static_cast<int64_t>(nd.children.size());
        if (rows) *rows = nd.dimen();
        if (hrank) *hrank = nd.hrank;
        if (leaf) *leaf = nd.leaf_index;
    });
}
int ttn_tree_size(const ttn_tree* t, int64_t* n, int64_t* m) {
    if (n) *n = t->topo.size();
    if (m) *m = t->topo.nmodes();
    return TTN_OK;
}
int ttn_traversal(const ttn_tree* t, int64_t* len, int64_t* order) {
    return guard(t->ctx, [&] {
        auto tour = euler_tour(t->topo);
        *len = static_cast<int64_t>(tour.size());
        if (order)
            for (size_t i = 0; i < tour.size(); ++i) order[i] = tour[i];
    });
}

int ttn_sop_create(ttn_ctx* ctx, int64_t nmodes, const int64_t* mode_dims, int64_t nterms, const double* coeff,
                   int64_t nops, const int32_t* op_mode, const int32_t* op_kind, const int64_t* op_coff,
                   const double* cdata, const int64_t* op_ioff, const int64_t* idata, const int64_t* op_roff,
                   const int64_t* r, ttn_sop** out) {
    return guard(ctx, [&] {
        auto s = std::make_unique<ttn_sop>();
        s->ctx = ctx;
        s->sop = sop_from_arrays(nmodes, mode_dims, nterms, coeff, nops, op_mode, op_kind, op_coff, cdata, op_ioff,
                                 idata, op_roff, r);
        *out = s.release();
    });
}
int ttn_sop_destroy(ttn_sop* s) {
    delete s;
    return TTN_OK;
}
int ttn_sop_compiled_tables(const ttn_tree* t, const ttn_sop* s, int64_t* len, int64_t* buf) {
    return guard(t->ctx, [&] {
        auto v = serialise_tables(compile_sop(t->topo, s->sop));
        if (buf) {
            check(*len >= static_cast<int64_t>(v.size()), ERR_CAPACITY, "table buffer too small");
            std::copy(v.begin(), v.end(), buf);
        }
        *len = static_cast<int64_t>(v.size());
    });
}

int ttn_integrator_create(ttn_ctx* ctx, const ttn_tree* t, const ttn_sop* s, const ttn_integ_params* p,
                          ttn_integ** out) {
    return guard(ctx, [&] {
        check(p != nullptr, ERR_INVALID, "null parameters");
        check(p->mode == TTN_MODE_FIXED_STEP, ERR_INVALID,
              "subspace expansion is not available on the device path yet; use TTN_MODE_FIXED_STEP");
        check(p->dt > 0.0, ERR_INVALID, "dt must be positive");
        auto in = std::make_unique<ttn_integ>();
        in->ctx = ctx;
        in->topo = t->topo;
        in->sop = s->sop;
        in->p = *p;
        in->ops = compile_sop(in->topo, in->sop);
        hip_check(hipSetDevice(ctx->device), "hipSetDevice");
        build_instance(in.get());
        *out = in.release();
    });
}
int ttn_integrator_destroy(ttn_integ* in) {
    delete in;
    return TTN_OK;
}

int ttn_state_set(ttn_integ* in, int64_t node, const double* data, int64_t rows, int64_t cols) {
    return guard(in->ctx, [&] {
        check(node >= 0 && node < in->topo.size(), ERR_INVALID, "node out of range");
        const Node& nd = in->topo.nodes[node];
        check(rows == nd.dimen() && cols == nd.hrank, ERR_INVALID, "state_set: shape mismatch");
        hip_check(hipMemcpy(in->arena.p + in->node_A[node], data, rows * cols * sizeof(double2), hipMemcpyHostToDevice),
                  "state_set");
    });
}
int ttn_state_get(ttn_integ* in, int64_t node, double* data, int64_t rows, int64_t cols) {
    return guard(in->ctx, [&] {
        check(node >= 0 && node < in->topo.size(), ERR_INVALID, "node out of range");
        const Node& nd = in->topo.nodes[node];
        check(rows == nd.dimen() && cols == nd.hrank, ERR_INVALID, "state_get: shape mismatch");
        hip_check(hipStreamSynchronize(in->ctx->stream), "sync");
        hip_check(hipMemcpy(data, in->arena.p + in->node_A[node], rows * cols * sizeof(double2), hipMemcpyDeviceToHost),
                  "state_get");
    });
}

int ttn_set_observable(ttn_integ* in, int64_t mode, const double* op) {
    return guard(in->ctx, [&] {
        check(mode >= 0 && mode < in->topo.nmodes(), ERR_INVALID, "mode out of range");
        const i64 d = in->topo.mode_dims[mode];
        hip_check(hipMemcpy(in->arena.p + in->obs_op, op, d * d * sizeof(double2), hipMemcpyHostToDevice),
                  "observable");
        in->obs_leaf = in->topo.leaf_node[mode];
        in->obs_dim = d;
    });
}

int ttn_trajectory(const ttn_integ* in, int64_t* len, double* out) {
    const int64_t n = static_cast<int64_t>(in->trajectory.size() / 2);
    if (out) {
        const int64_t m = std::min(*len, n);
        std::copy(in->trajectory.begin(), in->trajectory.begin() + 2 * m, out);
    }
    *len = n;
    return TTN_OK;
}

int ttn_step(ttn_integ* in, double cr, double ci, int64_t nsteps) {
    return guard(in->ctx, [&] {
        if (nsteps <= 0) return;
        ttnd::RunArgs a;
        make_args(in, cr, ci, nsteps, a);
        prepare_traj(in, nsteps);
        hipStream_t s = in->ctx->stream;
        hip_check(hipMemcpyAsync(in->d_inst.p, &in->inst, sizeof(in->inst), hipMemcpyHostToDevice, s), "inst");
        hip_check(hipEventRecord(in->ctx->ev0, s), "event");
        hip_check(ttnd_launch_run(in->d_inst.p, 1, &a, s), "launch tdvp_run_kernel");
        hip_check(hipEventRecord(in->ctx->ev1, s), "event");
        hip_check(hipEventSynchronize(in->ctx->ev1), "sync");
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, in->ctx->ev0, in->ctx->ev1), "elapsed");
        in->last_ms = ms;
        account(in, a);
        finish(in, nsteps);
    });
}

int ttn_time(const ttn_integ* in, double* t) {
    *t = in->p.dt * static_cast<double>(in->nsteps);
    return TTN_OK;
}
int ttn_counters(const ttn_integ* in, int64_t* nh, int64_t* ns) {
    return guard(in->ctx, [&] {
        int64_t c[2];
        hip_check(hipMemcpy(c, in->counters.p, sizeof(c), hipMemcpyDeviceToHost), "counters");
        if (nh) *nh = c[0];
        if (ns) *ns = in->nsteps;
    });
}

static void run_expect(ttn_integ* in, int leaf, int64_t op_off, double* out3) {
    hip_check(ttnd_launch_expect(in->d_inst.p, leaf, op_off, in->scratch.p, in->ctx->stream), "launch expect");
    hip_check(hipStreamSynchronize(in->ctx->stream), "sync");
    hip_check(hipMemcpy(out3, in->scratch.p, 3 * sizeof(double), hipMemcpyDeviceToHost), "expect");
}

int ttn_expect(ttn_integ* in, int64_t mode, const double* op, double* out) {
    return guard(in->ctx, [&] {
        check(mode >= 0 && mode < in->topo.nmodes(), ERR_INVALID, "mode out of range");
        const i64 d = in->topo.mode_dims[mode];
        hip_check(hipMemcpy(in->arena.p + in->tmp_op, op, d * d * sizeof(double2), hipMemcpyHostToDevice), "op");
        double o[3];
        run_expect(in, in->topo.leaf_node[mode], in->tmp_op, o);
        out[0] = o[0];
        out[1] = o[1];
    });
}

int ttn_norm2(ttn_integ* in, double* out) {
    return guard(in->ctx, [&] {
        const i64 d = in->topo.mode_dims[0];
        std::vector<double> eye(2 * d * d, 0.0);
        for (i64 i = 0; i < d; ++i) eye[2 * (i * d + i)] = 1.0;
        hip_check(hipMemcpy(in->arena.p + in->tmp_op, eye.data(), d * d * sizeof(double2), hipMemcpyHostToDevice),
                  "op");
        double o[3];
        run_expect(in, in->topo.leaf_node[0], in->tmp_op, o);
        *out = o[0];
    });
}

int ttn_sb_problem(ttn_ctx* ctx, const ttn_sb_params* p, ttn_tree** tree, ttn_sop** sop) {
    return guard(ctx, [&] {
        SBProblem pr = sb_build(sb_from_c(p));
        auto t = std::make_unique<ttn_tree>();
        t->ctx = ctx;
        t->topo = pr.topo;
        auto s = std::make_unique<ttn_sop>();
        s->ctx = ctx;
        s->sop = pr.sop;
        *tree = t.release();
        *sop = s.release();
    });
}

int ttn_sb_initial_state(ttn_integ* in, const ttn_sb_params* p) {
    return guard(in->ctx, [&] {
        SBParams sp = sb_from_c(p);
        auto st = sb_initial_state(in->topo, sp.psi0, sp.seed);
        for (int n = 0; n < in->topo.size(); ++n)
            hip_check(hipMemcpy(in->arena.p + in->node_A[n], st[n].data(), st[n].size() * sizeof(double2),
                                hipMemcpyHostToDevice),
                      "initial state");
    });
}

int ttn_batch_create(ttn_ctx* ctx, int64_t n, ttn_integ* const* integs, ttn_batch** out) {
    return guard(ctx, [&] {
        check(n >= 1, ERR_INVALID, "empty batch");
        auto b = std::make_unique<ttn_batch>();
        b->ctx = ctx;
        b->integs.assign(integs, integs + n);
        b->d_insts.alloc(n);
        *out = b.release();
    });
}
int ttn_batch_destroy(ttn_batch* b) {
    delete b;
    return TTN_OK;
}

int ttn_batch_step(ttn_batch* b, double cr, double ci, int64_t nsteps) {
    return guard(b->ctx, [&] {
        if (nsteps <= 0) return;
        const ttn_integ* first = b->integs[0];
        for (ttn_integ* in : b->integs) {
            check(in->nsteps == first->nsteps, ERR_INVALID, "batched instances must be at the same step");
            check(in->p.dt == first->p.dt, ERR_INVALID, "batched instances must share dt");
        }
        ttnd::RunArgs a;
        make_args(first, cr, ci, nsteps, a);
        std::vector<ttnd::DInst> insts;
        for (ttn_integ* in : b->integs) {
            prepare_traj(in, nsteps);
            insts.push_back(in->inst);
        }
        hipStream_t s = b->ctx->stream;
        hip_check(hipMemcpyAsync(b->d_insts.p, insts.data(), insts.size() * sizeof(ttnd::DInst), hipMemcpyHostToDevice,
                                 s),
                  "insts");
        hip_check(hipEventRecord(b->ctx->ev0, s), "event");
        hip_check(ttnd_launch_run(b->d_insts.p, static_cast<int>(insts.size()), &a, s), "launch tdvp_run_kernel");
        hip_check(hipEventRecord(b->ctx->ev1, s), "event");
        hip_check(hipEventSynchronize(b->ctx->ev1), "sync");
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, b->ctx->ev0, b->ctx->ev1), "elapsed");
        b->last_ms = ms;
        // every instance is accounted before the first failure is reported
        std::string err;
        for (size_t k = 0; k < b->integs.size(); ++k) {
            ttn_integ* in = b->integs[k];
            account(in, a);
            try {
                finish(in, nsteps);
            } catch (const Error& e) {
                if (err.empty()) err = "batch instance " + std::to_string(k) + ": " + e.what();
            }
        }
        if (!err.empty()) fail(ERR_NUMERIC, err);
    });
}

int ttn_flops(const ttn_integ* in, double* flops) {
    return guard(in->ctx, [&] {
        const int n = in->topo.size();
        std::vector<int64_t> st(4 * n);
        hip_check(hipMemcpy(st.data(), in->node_stats.p, st.size() * sizeof(int64_t), hipMemcpyDeviceToHost),
                  "stats");
        const FlopModel& f = in->fm;
        double tot = static_cast<double>(in->sweep_pairs) * f.sweep_pair + static_cast<double>(in->prepares) * f.prepare_spf;
        for (int i = 0; i < n; ++i) {
            tot += static_cast<double>(st[4 * i]) * f.heff_node[i] + 32.0 * f.node_size[i] * static_cast<double>(st[4 * i + 1]);
            tot += static_cast<double>(st[4 * i + 2]) * f.heff_bond[i] +
                   32.0 * f.bond_size[i] * static_cast<double>(st[4 * i + 3]);
        }
        *flops = tot;
    });
}

int ttn_profile(const ttn_integ* in, int64_t* slots) {
    return guard(in->ctx, [&] {
        hip_check(hipMemcpy(slots, in->prof.p, ttnd::PROF_NSLOTS * sizeof(int64_t), hipMemcpyDeviceToHost), "prof");
    });
}

int ttn_last_step_ms(const ttn_integ* in, double* ms) {
    *ms = in->last_ms;
    return TTN_OK;
}
int ttn_batch_last_step_ms(const ttn_batch* b, double* ms) {
    *ms = b->last_ms;
    return TTN_OK;
}

}  // extern "C"
