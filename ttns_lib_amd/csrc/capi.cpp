// C ABI of the MI355X TDVP library (include/ttn_capi.h).
//
// Host side of the drop-in: owns the tree layout, the per-instance HBM arena holding
// every node tensor, gauge matrix and environment, the compiled term tables flattened
// for the device, and the sweep schedule.  Each ttn_step / ttn_batch_step is a single
// launch of the persistent sweep kernel (device/sweep.hip).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <random>
#include <sstream>
#include <string>

#include "../../include/ttn_capi.h"
#include "device/desc.hpp"
#include "runtime.hpp"
#include "host/flops.hpp"
#include "host/marshal.hpp"
#include "host/tables.hpp"

extern "C" hipError_t ttnd_launch_run(const ttnd::DInst* d_insts, int n, const ttnd::RunArgs* a,
                                      const ttnd::TeamPlan* d_plan, ttnd::TeamWords* d_words, unsigned* d_fill,
                                      int grid, hipStream_t s);
extern "C" hipError_t ttnd_launch_expect(const ttnd::DInst* d_inst, int leaf, int64_t op_off, double* d_out,
                                         hipStream_t s);
extern "C" hipError_t ttnd_run_occupancy(int* blocks_per_cu);
extern "C" hipError_t ttnd_launch_census(int* d_out, int n, hipStream_t s);
extern "C" hipError_t ttnd_launch_zero_u32(unsigned* p, int64_t n, hipStream_t s);
extern "C" hipError_t ttnd_launch_densities(const ttnd::DInst* kets, const ttnd::DInst* bras, int npairs, int ins,
                                            const double2* S, double2* work, int64_t work_stride, int64_t rr,
                                            const int64_t* out_off, double2* out, int64_t out_stride, hipStream_t s);

using namespace ttnh;
using namespace ttnrt;

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
    DevBuf<int> hr;                 // live bond dimension of every node
    // subspace expansion: the normal stream of the trial vectors (std::mt19937 +
    // std::normal_distribution, the generator the reference's two_site_variations holds),
    // generated here and uploaded ahead of each launch; pending holds the draws not yet consumed
    bool expand = false;
    std::mt19937 rng;
    std::normal_distribution<double> ndist{0.0, 1.0};
    std::vector<double> pending;
    int64_t normals_used = 0;
    DevBuf<double> d_rng;
    DevBuf<int64_t> d_exp;          // ExpWord words
    DevBuf<double> d_exp_log;       // expansion log (ttn_set_expansion_log)
    // host copies of the coefficient-bearing tables and their byte offsets in `tables`
    // (ttn_set_coefficients patches and re-uploads them)
    std::vector<ttnd::DTerm> h_terms;
    std::vector<ttnd::DSub> h_subs;
    std::vector<ttnd::DMf> h_mfs;
    size_t o_terms = 0, o_subs = 0, o_mfs = 0;
    // arena regions that carry state from one step to the next besides the node tensors
    // (gauge factors, single-particle and mean-field matrices): what a checkpoint stores
    std::vector<std::pair<int64_t, int64_t>> carried;
    FlopModel fm;
    int64_t sweep_pairs = 0;   // forward+backward sweeps run (with or without propagation)
    int64_t prepares = 0;
    int64_t arena_size = 0;
    std::vector<int64_t> node_A;
    int64_t obs_slot[ttnd::MAX_OBS] = {};  // arena offsets of the observable operators
    int64_t tmp_op = -1;
    int nobs = 0;                   // observables recorded after every step
    int64_t nsteps = 0;
    int64_t n_onesite = 0, n_twosite = 0;  // expansion attempts (engine().Nonesite/Ntwosite)
    double last_ms = 0.0;
    std::vector<double> trajectory;  // per completed step: re, im of every <O_j>, then <1>
};

// A captured sweep launch (zero fill of the team words + tdvp_run_kernel), replayed while
// the launch arguments and the team plan stay the same.
struct SweepGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ttnd::RunArgs args{};
    ttnd::TeamPlan plan{};
    bool teams = false;
    int n = 0;
    void reset() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    ~SweepGraph() { reset(); }
};

struct ttn_batch {
    ttn_ctx* ctx;
    bool use_graph = true;         // replay captured launches (ttn_batch_set_graph)
    SweepGraph graph;
    int64_t graph_replays = 0, graph_captures = 0;
    std::vector<ttn_integ*> integs;
    DevBuf<ttnd::DInst> d_insts;
    double last_ms = 0.0;
    std::vector<double> cost;      // per-instance cost of one sweep pair on one workgroup: modelled
                                   // flops before the first launch, then measured device time
    std::vector<int> team;         // team sizes of the last launch
};

namespace {

// The coefficients of the compiled tables (a term's own factor, the sum of its accumulation
// coefficients, accum_coeff(0), and the per-sub-term and per-mean-field accumulation
// coefficients): the only fields of the tables that depend on the SOP's coefficients.
void term_coeffs(const Term& tm, ttnd::DTerm& d) {
    d.cre = tm.coeff.real();
    d.cim = tm.coeff.imag();
    cplx ic = tm.coeff;
    if (tm.nspf_terms() > 1) {
        ic = cplx(0.0, 0.0);
        for (const cplx& a : tm.accum) ic += a;
    }
    d.ire = ic.real();
    d.iim = ic.imag();
    d.a0re = tm.accum.empty() ? 0.0 : tm.accum[0].real();
    d.a0im = tm.accum.empty() ? 0.0 : tm.accum[0].imag();
}
void sub_coeff(const Term& tm, int s, ttnd::DSub& ds) {
    const cplx a = (tm.nspf_terms() > 1) ? tm.accum.at(s) : cplx(0.0, 0.0);
    ds.are = a.real();
    ds.aim = a.imag();
}
void mf_coeff(const Term& tm, int m, ttnd::DMf& dm) {
    const cplx a = (tm.nmf_terms() > 1) ? tm.accum.at(m) : cplx(0.0, 0.0);
    dm.are = a.real();
    dm.aim = a.imag();
}

// Flatten topology + compiled tables into device descriptors and lay out the arena.
void build_instance(ttn_integ* in) {
    const Topology& t = in->topo;
    const CompiledSOP& ops = in->ops;
    const int n = t.size();
    check(n <= ttnd::MAX_NODES, ERR_INVALID, "device path supports at most 4096 nodes");
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
    // storage at capacity: a bond grows inside a launch without moving anything
    for (int i = 0; i < n; ++i) {
        const Node& nd = t.nodes[i];
        check(nd.nmodes() <= ttnd::MAXC, ERR_INVALID, "device path supports at most 4 children per node");
        check(nd.cap_hrank <= 64, ERR_INVALID, "device path supports bond dimensions (and capacities) up to 64");
        for (size_t c = 0; c < nd.dims.size(); ++c)
            check(nd.cap_dims[c] <= 64 || nd.is_leaf(), ERR_INVALID, "device path supports mode dims up to 64");
        check(nd.capacity() < (int64_t(1) << 31), ERR_INVALID, "node tensor too large for the device path");
        ttnd::DNode& d = dn[i];
        d.parent = nd.parent;
        d.child_id = nd.child_id;
        d.nchild = static_cast<int>(nd.children.size());
        d.leaf = nd.leaf_index;
        d.nmodes = nd.nmodes();
        d.leafdim = nd.is_leaf() ? static_cast<int>(nd.dims[0]) : 0;
        d.cap_hrank = static_cast<int>(nd.cap_hrank);
        for (int c = 0; c < ttnd::MAXC; ++c) d.children[c] = c < static_cast<int>(nd.children.size()) ? nd.children[c] : -1;
        d.A = alloc(nd.capacity());
        in->node_A[i] = d.A;
        int64_t rmax = nd.cap_hrank;
        if (!nd.is_leaf())
            for (i64 x : nd.cap_dims) rmax = std::max<int64_t>(rmax, x);
        d.R = alloc(rmax * rmax);
        in->carried.emplace_back(d.R, rmax * rmax);
        maxsize = std::max<int64_t>(maxsize, nd.capacity());
    }
    // environment matrices (capacity hrank squared)
    std::vector<std::vector<int64_t>> spf_off(n), mf_off(n);
    for (int i = 0; i < n; ++i) {
        const i64 r = t.nodes[i].cap_hrank;
        const auto& terms = ops.nodes[i].terms;
        spf_off[i].assign(terms.size(), -1);
        mf_off[i].assign(terms.size(), -1);
        for (size_t k = 0; k < terms.size(); ++k) {
            if (!terms[k].id_spf) {
                spf_off[i][k] = alloc(r * r);
                in->carried.emplace_back(spf_off[i][k], r * r);
            }
            if (!terms[k].id_mf) {
                mf_off[i][k] = alloc(r * r);
                in->carried.emplace_back(mf_off[i][k], r * r);
            }
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
            term_coeffs(tm, d);
            d.spf = spf_off[i][k];
            d.mf = mf_off[i][k];
            d.sub0 = static_cast<int>(subs.size());
            d.nsub = tm.nspf_terms();
            for (int s = 0; s < tm.nspf_terms(); ++s) {
                ttnd::DSub ds{};
                sub_coeff(tm, s, ds);
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
                        dk.dim = static_cast<int>(nd.cap_dims[pr.first]);
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
                mf_coeff(tm, m, dm);
                const int p = nd.parent;
                dm.pmf = ops.nodes[p].terms[mi.parent_index].id_mf ? -1 : mf_off[p][mi.parent_index];
                dm.k0 = static_cast<int>(ks.size());
                dm.nk = static_cast<int>(mi.siblings.size());
                for (const auto& sb : mi.siblings) {
                    const int sib = t.nodes[p].children[sb.first];
                    ttnd::DK dk{};
                    dk.nu = sb.first;
                    dk.kind = 0;
                    dk.dim = static_cast<int>(t.nodes[p].cap_dims[sb.first]);
                    dk.mat = spf_off[sib][sb.second];
                    check(dk.mat >= 0, ERR_INVALID, "mean-field chain references an identity term");
                    ks.push_back(dk);
                }
                mfs.push_back(dm);
            }
            terms.push_back(d);
        }
    }
    // expansion: two-site factors and vectors
    int nt2 = 0;
    for (int i = 0; i < n; ++i) {
        int k = 0;
        for (const Term& tm : ops.nodes[i].terms)
            if (!tm.id_spf && !tm.id_mf) ++k;
        nt2 = std::max(nt2, k);
    }
    // observable slot and work buffers
    int64_t maxdim = 1;
    for (i64 d : t.mode_dims) maxdim = std::max<int64_t>(maxdim, d);
    for (int j = 0; j < ttnd::MAX_OBS; ++j) in->obs_slot[j] = alloc(maxdim * maxdim);
    in->tmp_op = alloc(maxdim * maxdim);
    ttnd::DInst& I = in->inst;
    for (int k = 0; k < ttnd::NWORK; ++k) I.work[k] = alloc(maxsize);
    I.krylov = alloc((kdim + 1) * maxsize);
    I.maxsize = maxsize;
    // team scratch: partial sums (doubles, two buffers) and partial products per member
    I.team_part = alloc(ttnd::TEAM_MAX * ttnd::PART_K);
    I.team_tiles = alloc(static_cast<int64_t>(ttnd::TEAM_MAX) * ttnd::TILE_MAX);
    in->expand = in->p.mode == TTN_MODE_SUBSPACE_EXPANSION;
    if (in->expand) {
        check(nt2 <= 8, ERR_INVALID, "device expansion supports at most 8 two-site terms per node");
        I.nt2 = std::max(nt2, 1);
        I.e_op1 = alloc(I.nt2 * maxsize);
        I.e_op2 = alloc(I.nt2 * maxsize);
        I.e_vec = alloc(2 * maxsize);
    }
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
    in->h_terms = terms;
    in->h_subs = subs;
    in->h_mfs = mfs;
    in->o_terms = o_terms;
    in->o_subs = o_subs;
    in->o_mfs = o_mfs;
    in->tables.alloc(blob.size());
    hip_check(hipMemcpyAsync(in->tables.p, blob.data(), blob.size(), hipMemcpyHostToDevice, in->ctx->stream),
              "upload tables");
    in->counters.alloc(2);
    hip_check(hipMemsetAsync(in->counters.p, 0, 2 * sizeof(int64_t), in->ctx->stream), "memset counters");
    // live bond dimensions
    std::vector<int> hr(n);
    for (int i = 0; i < n; ++i) hr[i] = static_cast<int>(t.nodes[i].hrank);
    in->hr.alloc(n);
    hip_check(hipMemcpyAsync(in->hr.p, hr.data(), n * sizeof(int), hipMemcpyHostToDevice, in->ctx->stream), "hranks");
    in->status.alloc(2);
    hip_check(hipMemsetAsync(in->status.p, 0, 2 * sizeof(int), in->ctx->stream), "memset status");
    in->scratch.alloc(4);
    in->node_stats.alloc(ttnd::NODE_STATS * n);
    hip_check(hipMemsetAsync(in->node_stats.p, 0, ttnd::NODE_STATS * n * sizeof(int64_t), in->ctx->stream),
              "memset stats");
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
    for (int j = 0; j < ttnd::MAX_OBS; ++j) {
        I.obs_ops[j] = in->obs_slot[j];
        I.obs_leaves[j] = -1;
    }
    I.nobs = 0;
    I.traj = nullptr;
    I.counters = in->counters.p;
    I.node_stats = in->node_stats.p;
    I.prof = in->prof.p;
    I.status = in->status.p;
    I.hr = in->hr.p;
    in->d_exp.alloc(ttnd::EXP_NWORDS);
    hip_check(hipMemsetAsync(in->d_exp.p, 0, ttnd::EXP_NWORDS * sizeof(int64_t), in->ctx->stream), "memset exp");
    I.exp = in->d_exp.p;
    I.expand = in->expand ? 1 : 0;
    I.neig = static_cast<int>(in->p.neigenvalues);
    I.min_unocc = static_cast<int>(in->p.minimum_unoccupied);
    I.only_unocc = in->p.only_apply_when_no_unoccupied != 0 ? 1 : 0;
    I.spawn = in->p.spawning_threshold;
    I.unocc_thr = in->p.unoccupied_threshold;
    I.eig_tol = in->p.eigenvalue_tolerance > 0.0 ? in->p.eigenvalue_tolerance : 1e-6;
    in->rng.seed(in->p.rng_seed != 0 ? in->p.rng_seed : 5489u);
    I.rng = nullptr;
    I.rng_n = 0;
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
    a.team_min = in->ctx->team_min;
    a.small_max = in->ctx->small_max;
    if (in->nsteps == 0) {
        a.prepare = 1;
        a.prepare_spf = 1;
        a.nladder = ladder(in->p.dt, a.ladder_dt);
        a.nregular = static_cast<int>(nsteps - 1);
    } else {
        a.nregular = static_cast<int>(nsteps);
    }
}

int64_t traj_stride(const ttn_integ* in) { return 2 * in->nobs + 1; }

void prepare_traj(ttn_integ* in, int64_t nsteps) {
    in->inst.nobs = in->nobs;
    if (in->nobs == 0) {
        in->inst.traj = nullptr;
        return;
    }
    const int64_t need = traj_stride(in) * nsteps;
    if (static_cast<int64_t>(in->traj.n) < need) in->traj.alloc(need);
    in->inst.traj = in->traj.p;
}

// the trajectory records of the first `steps` steps of the last launch(es)
void keep_traj(ttn_integ* in, int64_t steps) {
    if (in->nobs == 0 || steps <= 0) return;
    std::vector<double> buf(traj_stride(in) * steps);
    hip_check(hipMemcpy(buf.data(), in->traj.p, buf.size() * sizeof(double), hipMemcpyDeviceToHost), "traj");
    in->trajectory.insert(in->trajectory.end(), buf.begin(), buf.end());
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
    static const char* reasons[] = {"",
                                    "the vector to propagate has a non-finite norm",
                                    "a Krylov basis vector has a non-finite norm",
                                    "no Krylov sub-step met the tolerance",
                                    "a team barrier timed out",
                                    "the team's workgroups did not assemble",
                                    "the expansion needed more random draws than were uploaded",
                                    "the expansion found no orthogonal trial vector or a non-finite value"};
    const int reason = st[0] & 15, bond = (st[0] >> 4) & 1, node = st[0] >> 5;
    if (reason == ttnd::FAIL_TEAM_TIMEOUT || reason == ttnd::FAIL_TEAM_FORM)
        return std::string("TDVP sweep failed at step ") + std::to_string(st[1]) + ": " + reasons[reason];
    if (reason == ttnd::FAIL_RNG_EXHAUSTED || reason == ttnd::FAIL_EXPANSION)
        return std::string("TDVP sweep failed at step ") + std::to_string(st[1]) + ", bond above node " +
               std::to_string(node) + ": " + reasons[reason];
    std::string m = "TDVP sweep failed at step " + std::to_string(st[1]) + ", " +
                    (bond ? "bond propagation above node " : "node propagation of node ") + std::to_string(node) +
                    ": " + (reason < 8 ? reasons[reason] : "unknown reason");
    return m;
}

void finish(ttn_integ* in, int64_t nsteps) {
    int st[2] = {0, 0};
    hip_check(hipMemcpy(st, in->status.p, sizeof(st), hipMemcpyDeviceToHost), "status");
    keep_traj(in, nsteps);
    if (st[0] != 0) fail(ERR_NUMERIC, failure_message(st));
    in->nsteps += nsteps;
}

// ---------------------------------------------------------------- expansion launches
// Upper bound on the normal draws one sweep pair of the expansion integrator can take: a bond
// below its capacity makes at most one downward and one upward attempt per sweep pair (at
// capacity an attempt returns before drawing), each trial vector takes two draws per entry of
// the complement (sized here at capacity), a two-site attempt draws one vector and a random
// one-site attempt up to minimum_unoccupied, and redrawn vectors are covered by a factor 2.
int64_t rng_bound(const ttn_integ* in) {
    const Topology& t = in->topo;
    int64_t tot = 0;
    const int64_t vecs = 1 + std::max<int64_t>(in->p.minimum_unoccupied, 0);
    for (int n = 1; n < t.size(); ++n) {
        const Node& nd = t.nodes[n];
        if (nd.hrank >= nd.cap_hrank) continue;
        const Node& pn = t.nodes[nd.parent];
        const int64_t m_down = pn.capacity() / pn.cap_dims[nd.child_id];
        const int64_t m_up = nd.cap_dimen();
        tot += 2 * 2 * (m_down + m_up) * vecs;
    }
    return tot;
}

// Makes at least `need` normal draws available to the next launch: the draws the last launch
// consumed are dropped, the stream is continued on the host and the pending draws uploaded.
void top_up_rng(ttn_integ* in, int64_t need) {
    int64_t pos = 0;
    hip_check(hipMemcpy(&pos, in->d_exp.p + ttnd::EXP_RNG_POS, sizeof(pos), hipMemcpyDeviceToHost), "rng cursor");
    check(pos >= 0 && pos <= static_cast<int64_t>(in->pending.size()), ERR_DEVICE, "random stream cursor out of range");
    in->pending.erase(in->pending.begin(), in->pending.begin() + pos);
    in->normals_used += pos;
    while (static_cast<int64_t>(in->pending.size()) < need) in->pending.push_back(in->ndist(in->rng));
    const int64_t n = static_cast<int64_t>(in->pending.size());
    if (static_cast<int64_t>(in->d_rng.n) < n) in->d_rng.alloc(static_cast<size_t>(std::max<int64_t>(n, 1)));
    if (n > 0)
        hip_check(hipMemcpy(in->d_rng.p, in->pending.data(), n * sizeof(double), hipMemcpyHostToDevice), "rng");
    const int64_t zero = 0;
    hip_check(hipMemcpy(in->d_exp.p + ttnd::EXP_RNG_POS, &zero, sizeof(zero), hipMemcpyHostToDevice), "rng cursor");
    in->inst.rng = in->d_rng.p;
    in->inst.rng_n = n;
}

// Live bond dimensions and expansion counters of the last launch into the host copy of the tree
// (the node shapes ttn_state_get/set and the flop model use).  Returns the device's grew flag.
bool sync_shapes(ttn_integ* in) {
    const int n = in->topo.size();
    std::vector<int> hr(n);
    hip_check(hipMemcpy(hr.data(), in->hr.p, n * sizeof(int), hipMemcpyDeviceToHost), "hranks");
    bool changed = false;
    for (int i = 0; i < n; ++i) {
        Node& nd = in->topo.nodes[i];
        if (nd.hrank == hr[i]) continue;
        nd.hrank = hr[i];
        if (nd.parent >= 0) in->topo.nodes[nd.parent].dims[nd.child_id] = hr[i];
        changed = true;
    }
    if (changed) in->fm = flop_model(in->topo, in->ops, in->sop);
    int64_t w[ttnd::EXP_NWORDS];
    hip_check(hipMemcpy(w, in->d_exp.p, sizeof(w), hipMemcpyDeviceToHost), "expansion words");
    in->n_onesite = w[ttnd::EXP_ONESITE];
    in->n_twosite = w[ttnd::EXP_TWOSITE];
    return w[ttnd::EXP_GREW] != 0;
}

// ---------------------------------------------------------------- teams
// Placement census: which XCD each workgroup of a resident grid lands on.  Teams are
// enabled only when every workgroup of the grid is resident at once and the grid spreads
// evenly over at most MAX_XCD XCDs.
void team_census(ttn_ctx* c) {
    hipDeviceProp_t prop;
    hip_check(hipGetDeviceProperties(&prop, c->device), "hipGetDeviceProperties");
    int occ = 0;
    hip_check(ttnd_run_occupancy(&occ), "occupancy");
    c->teams_ok = false;
    if (occ < 1 || prop.multiProcessorCount < 2) return;
    const int grid = prop.multiProcessorCount * occ;
    DevBuf<int> ids;
    ids.alloc(grid);
    hip_check(ttnd_launch_census(ids.p, grid, c->stream), "census");
    std::vector<int> h(grid);
    hip_check(hipMemcpyAsync(h.data(), ids.p, grid * sizeof(int), hipMemcpyDeviceToHost, c->stream), "census");
    hip_check(hipStreamSynchronize(c->stream), "sync");
    int nx = 0;
    for (int x : h) {
        if (x < 0 || x >= ttnd::MAX_XCD) return;
        nx = std::max(nx, x + 1);
    }
    std::vector<int> cnt(nx, 0);
    for (int x : h) ++cnt[x];
    for (int x = 0; x < nx; ++x)
        if (cnt[x] != grid / nx || grid % nx != 0) return;
    c->grid = grid;
    c->nxcd = nx;
    c->xcd_cap = std::min(grid / nx, ttnd::XCD_SLOTS);
    c->teams_ok = c->xcd_cap >= 2;
}

// largest team the context can place (one XCD, TEAM_MAX members)
int max_team(const ttn_ctx* c) { return c->teams_ok ? std::min(c->xcd_cap, ttnd::TEAM_MAX) : 1; }

// First-fit-decreasing placement of the teams on the XCDs.  Returns false when they do
// not fit.
bool make_plan(const ttn_ctx* c, const std::vector<int>& size, ttnd::TeamPlan& plan) {
    std::memset(&plan, 0, sizeof(plan));
    plan.nxcd = c->nxcd;
    std::vector<int> order(size.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = static_cast<int>(i);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return size[a] > size[b]; });
    for (int i : order) {
        int x = 0;
        while (x < c->nxcd && plan.count[x] + size[i] > c->xcd_cap) ++x;
        if (x == c->nxcd) return false;
        for (int r = 0; r < size[i]; ++r) plan.slot[x][plan.count[x] + r] = ttnd::TeamSlot{i, r, size[i], 0};
        plan.count[x] += size[i];
    }
    return true;
}

// Launch the sweep kernel over n instances with the given team sizes (size 1 everywhere
// runs one workgroup per instance without a plan).  A team that did not assemble leaves
// its instance untouched with FAIL_TEAM_FORM; such instances are run again one workgroup
// each, and teams are switched off for the context.  With a graph cache (gb) the zero fill
// of the team words and the kernel launch are captured once and replayed as long as the
// launch arguments and the team plan are unchanged (the instance descriptors and the plan
// are uploaded before, outside the graph).
void launch(ttn_ctx* c, ttnd::DInst* d_insts, const std::vector<ttnd::DInst>& insts, const ttnd::RunArgs& a,
            std::vector<int> size, ttn_batch* gb = nullptr) {
    const int n = static_cast<int>(insts.size());
    hipStream_t s = c->stream;
    int maxsize = 1;
    for (int& t : size) {
        t = std::max(1, std::min(t, max_team(c)));
        maxsize = std::max(maxsize, t);
    }
    ttnd::TeamPlan plan{};
    const bool teams = c->teams_ok && maxsize > 1 && make_plan(c, size, plan);
    const int64_t head = 32 * ttnd::MAX_XCD;  // slot counters, one 128-byte line each
    const int64_t words = static_cast<int64_t>(sizeof(ttnd::TeamWords) / sizeof(unsigned));
    if (teams) {
        if (c->team_words_n < n) {
            if (gb) gb->graph.reset();  // the captured launch points at the old buffer
            c->team_words.alloc(head + words * n);
            c->team_words_n = n;
        }
        if (c->plan.p == nullptr) c->plan.alloc(1);
        hip_check(hipMemcpyAsync(c->plan.p, &plan, sizeof(plan), hipMemcpyHostToDevice, s), "plan");
    }
    // the work itself: a fill kernel zeroes the team words (no memset node, so the sequence
    // can live in a graph), then the sweep
    auto enqueue = [&] {
        if (!teams) {
            hip_check(ttnd_launch_run(d_insts, n, &a, nullptr, nullptr, nullptr, 0, s), "launch tdvp_run_kernel");
            return;
        }
        hip_check(ttnd_launch_zero_u32(c->team_words.p, head + words * n, s), "zero team words");
        hip_check(ttnd_launch_run(d_insts, n, &a, c->plan.p, reinterpret_cast<ttnd::TeamWords*>(c->team_words.p + head),
                                  c->team_words.p, c->grid, s),
                  "launch tdvp_run_kernel");
    };
    if (gb && gb->use_graph) {
        SweepGraph& g = gb->graph;
        const bool same = g.exec && g.n == n && g.teams == teams && std::memcmp(&g.args, &a, sizeof(a)) == 0 &&
                          (!teams || std::memcmp(&g.plan, &plan, sizeof(plan)) == 0);
        if (!same) {
            g.reset();
            hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
            enqueue();
            hip_check(hipStreamEndCapture(s, &g.graph), "end capture");
            hip_check(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0), "instantiate graph");
            g.args = a;
            g.plan = plan;
            g.teams = teams;
            g.n = n;
            ++gb->graph_captures;
        }
        hip_check(hipEventRecord(c->ev0, s), "event");
        hip_check(hipGraphLaunch(g.exec, s), "graph launch");
        ++gb->graph_replays;
    } else {
        hip_check(hipEventRecord(c->ev0, s), "event");
        enqueue();
    }
    hip_check(hipEventRecord(c->ev1, s), "event");
    hip_check(hipEventSynchronize(c->ev1), "sync");
    if (!teams) return;
    // instances whose team did not assemble: run them again, one workgroup each
    std::vector<ttnd::DInst> again;
    for (const ttnd::DInst& I : insts) {
        int st = 0;
        hip_check(hipMemcpy(&st, I.status, sizeof(int), hipMemcpyDeviceToHost), "status");
        if ((st & 15) == ttnd::FAIL_TEAM_FORM) {
            hip_check(hipMemset(I.status, 0, 2 * sizeof(int)), "reset status");
            again.push_back(I);
        }
    }
    if (again.empty()) return;
    c->teams_ok = false;
    DevBuf<ttnd::DInst> d_again;
    d_again.alloc(again.size());
    hip_check(hipMemcpyAsync(d_again.p, again.data(), again.size() * sizeof(ttnd::DInst), hipMemcpyHostToDevice, s),
              "insts");
    hip_check(ttnd_launch_run(d_again.p, static_cast<int>(again.size()), &a, nullptr, nullptr, nullptr, 0, s),
              "launch tdvp_run_kernel");
    hip_check(hipEventRecord(c->ev1, s), "event");
    hip_check(hipEventSynchronize(c->ev1), "sync");
}

// Algorithmic flops executed by an instance so far (the ttn_flops model)
double measured_flops(const ttn_integ* in) {
    const int n = in->topo.size();
    constexpr int S = ttnd::NODE_STATS;
    std::vector<int64_t> st(S * n);
    hip_check(hipMemcpy(st.data(), in->node_stats.p, st.size() * sizeof(int64_t), hipMemcpyDeviceToHost), "stats");
    const FlopModel& f = in->fm;
    double tot = static_cast<double>(in->sweep_pairs) * f.sweep_pair + static_cast<double>(in->prepares) * f.prepare_spf;
    for (int i = 0; i < n; ++i) {
        tot += static_cast<double>(st[S * i]) * f.heff_node[i] + 32.0 * f.node_size[i] * static_cast<double>(st[S * i + 1]);
        tot += static_cast<double>(st[S * i + 2]) * f.heff_bond[i] + 32.0 * f.bond_size[i] * static_cast<double>(st[S * i + 3]);
    }
    return tot;
}

// Device wall-clock ticks the instance's team leader has spent in sweep launches
int64_t device_ticks(const ttn_integ* in) {
    int64_t v = 0;
    hip_check(hipMemcpy(&v, in->prof.p + ttnd::PROF_TOTAL, sizeof(int64_t), hipMemcpyDeviceToHost), "prof");
    return v;
}

// Team size of an instance stepped on its own: one workgroup per 1024 entries of its
// largest node tensor, at most one XCD (the largest nodes' Krylov iterations keep getting faster
// up to a whole XCD).
int auto_team(const ttn_integ* in) {
    if (in->p.team_size > 0) return static_cast<int>(in->p.team_size);
    const int64_t t = in->inst.maxsize / 1024;
    return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(t, max_team(in->ctx))));
}

// Static cost model of one sweep pair (flops), used before anything is measured: every
// propagation taken as six H applications with an average Krylov size of three.
double static_cost(const ttn_integ* in) {
    const FlopModel& f = in->fm;
    double c = f.sweep_pair;
    for (size_t i = 0; i < f.heff_node.size(); ++i)
        c += 6.0 * (f.heff_node[i] + 3.0 * 32.0 * f.node_size[i]) + 6.0 * (f.heff_bond[i] + 3.0 * 32.0 * f.bond_size[i]);
    return c;
}

// Share of an instance's work in operations below team_min, which its leader runs alone.
double serial_share(const ttn_integ* in) {
    const FlopModel& f = in->fm;
    double all = 0.0, small = 0.0;
    for (size_t i = 0; i < f.heff_node.size(); ++i) {
        const double wn = f.heff_node[i] + 3.0 * 32.0 * f.node_size[i];
        const double wb = f.heff_bond[i] + 3.0 * 32.0 * f.bond_size[i];
        all += wn + wb;
        if (f.node_size[i] < in->ctx->team_min) small += wn;
        if (f.bond_size[i] < in->ctx->team_min) small += wb;
    }
    return all > 0.0 ? small / all : 1.0;
}

// Team sizes of a batch (the autotuner): starting from one workgroup per instance, the
// next workgroup always goes to the instance with the longest predicted time
// cost * (s + (1 - s) / t), s its serial share, until the XCDs are full; sizes given
// explicitly by an integrator's team_size are kept.
std::vector<int> batch_teams(const ttn_batch* b) {
    const ttn_ctx* c = b->ctx;
    const size_t n = b->integs.size();
    std::vector<int> t(n, 1);
    if (!c->teams_ok) return t;
    std::vector<double> s(n);
    std::vector<bool> fixed(n, false);
    int used = 0;
    for (size_t i = 0; i < n; ++i) {
        s[i] = serial_share(b->integs[i]);
        if (b->integs[i]->p.team_size > 0) {
            t[i] = static_cast<int>(std::min<int64_t>(b->integs[i]->p.team_size, max_team(c)));
            fixed[i] = true;
        }
        used += t[i];
    }
    const int cap = c->nxcd * c->xcd_cap;
    auto time = [&](size_t i, int ti) { return b->cost[i] * (s[i] + (1.0 - s[i]) / ti); };
    // an instance takes another workgroup only while that cuts its predicted time by at
    // least team_gain (every member adds barrier traffic)
    auto grows = [&](size_t i) { return time(i, t[i] + 1) < (1.0 - c->team_gain) * time(i, t[i]); };
    while (used < cap) {
        size_t best = n;
        for (size_t i = 0; i < n; ++i)
            if (!fixed[i] && t[i] < max_team(c) && grows(i) && (best == n || time(i, t[i]) > time(best, t[best])))
                best = i;
        if (best == n) break;
        ++t[best];
        ++used;
    }
    // keep the placement feasible: shrink the largest team until first-fit succeeds
    ttnd::TeamPlan plan{};
    while (!make_plan(c, t, plan)) {
        size_t big = 0;
        for (size_t i = 1; i < n; ++i)
            if (t[i] > t[big]) big = i;
        if (t[big] == 1) break;
        --t[big];
    }
    return t;
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
        team_census(c.get());
        ttnrt_noalloc_setup(c.get());
        // tuning knobs: TTN_TEAM_MIN (elements below which an operation runs on a team's
        // leader alone), TTN_TEAMS=0 switches teams off, TTN_SMALL_MAX (largest propagation
        // run in one workgroup's LDS, 0: none), TTN_TEAM_GAIN (smallest predicted relative gain
        // for which the team-size autotuner adds a member)
        if (const char* e = std::getenv("TTN_TEAM_MIN")) c->team_min = std::max<int64_t>(1, std::atoll(e));
        if (const char* e = std::getenv("TTN_TEAM_GAIN")) c->team_gain = std::max(0.0, std::atof(e));
        if (const char* e = std::getenv("TTN_SMALL_MAX")) c->small_max = std::max<int64_t>(0, std::atoll(e));
        if (const char* e = std::getenv("TTN_TEAMS"))
            if (std::atoi(e) == 0) c->teams_ok = false;
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
int ttn_tree_capacities(const ttn_tree* t, int64_t* caps) {
    return guard(t->ctx, [&] {
        check(caps != nullptr, ERR_INVALID, "null output");
        for (int i = 0; i < t->topo.size(); ++i) caps[i] = t->topo.nodes[i].cap_hrank;
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
        check(p->mode == TTN_MODE_FIXED_STEP || p->mode == TTN_MODE_SUBSPACE_EXPANSION, ERR_INVALID,
              "unknown integrator mode");
        check(p->dt > 0.0, ERR_INVALID, "dt must be positive");
        check(p->mode == TTN_MODE_FIXED_STEP || p->neigenvalues >= 1, ERR_INVALID, "neigenvalues must be positive");
        check(p->minimum_unoccupied >= 0, ERR_INVALID, "minimum_unoccupied must not be negative");
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

int ttn_set_observables(ttn_integ* in, int64_t nops, const int64_t* modes, const double* ops) {
    return guard(in->ctx, [&] {
        check(nops >= 0 && nops <= ttnd::MAX_OBS, ERR_INVALID, "at most 8 observables are recorded on the device");
        check(nops == 0 || (modes != nullptr && ops != nullptr), ERR_INVALID, "null observables");
        int64_t at = 0;
        for (int64_t j = 0; j < nops; ++j) {
            check(modes[j] >= 0 && modes[j] < in->topo.nmodes(), ERR_INVALID, "mode out of range");
            const i64 d = in->topo.mode_dims[modes[j]];
            hip_check(hipMemcpy(in->arena.p + in->obs_slot[j], ops + 2 * at, d * d * sizeof(double2),
                                hipMemcpyHostToDevice),
                      "observable");
            in->inst.obs_leaves[j] = in->topo.leaf_node[modes[j]];
            at += d * d;
        }
        in->nobs = static_cast<int>(nops);
        in->trajectory.clear();  // records of another observable set
    });
}

int ttn_set_observable(ttn_integ* in, int64_t mode, const double* op) {
    return ttn_set_observables(in, 1, &mode, op);
}

int ttn_observables(const ttn_integ* in, int64_t* len, double* out) {
    const int64_t stride = traj_stride(in);
    const int64_t n = in->nobs == 0 ? 0 : static_cast<int64_t>(in->trajectory.size()) / stride;
    if (out) {
        const int64_t m = std::min(*len, n);
        std::copy(in->trajectory.begin(), in->trajectory.begin() + stride * m, out);
    }
    *len = n;
    return TTN_OK;
}

int ttn_trajectory(const ttn_integ* in, int64_t* len, double* out) {
    const int64_t stride = traj_stride(in);
    const int64_t n = in->nobs == 0 ? 0 : static_cast<int64_t>(in->trajectory.size()) / stride;
    if (out) {
        const int64_t m = std::min(*len, n);
        for (int64_t k = 0; k < m; ++k) {
            out[2 * k] = in->trajectory[stride * k];
            out[2 * k + 1] = in->trajectory[stride * k + stride - 1];
        }
    }
    *len = n;
    return TTN_OK;
}

// The expansion integrator's steps: prepare_evolution one sweep pair per launch until a pair
// adds nothing, the first-step ladder, then the regular steps in launches short enough for the
// normal draws uploaded ahead of them (one launch once every bond is at capacity).
void step_expanding(ttn_integ* in, double cr, double ci, int64_t nsteps) {
    constexpr int64_t RNG_BUDGET = int64_t(1) << 23;  // draws per launch (64 MiB)
    prepare_traj(in, nsteps);
    hipStream_t s = in->ctx->stream;
    int64_t done = 0;
    double ms_total = 0.0;
    int st[2] = {0, 0};
    auto run = [&](ttnd::RunArgs& a, int64_t pairs) {
        top_up_rng(in, pairs * rng_bound(in));
        hip_check(hipMemcpyAsync(in->d_inst.p, &in->inst, sizeof(in->inst), hipMemcpyHostToDevice, s), "inst");
        launch(in->ctx, in->d_inst.p, {in->inst}, a, {auto_team(in)});
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, in->ctx->ev0, in->ctx->ev1), "elapsed");
        ms_total += ms;
        account(in, a);
        const bool grew = sync_shapes(in);
        hip_check(hipMemcpy(st, in->status.p, sizeof(st), hipMemcpyDeviceToHost), "status");
        return grew;
    };
    ttnd::RunArgs a;
    make_args(in, cr, ci, 1, a);
    if (in->nsteps == 0) {
        for (int it = 0; st[0] == 0; ++it) {
            check(it < 1000000, ERR_NUMERIC, "prepare_evolution: the expansion did not settle");
            ttnd::RunArgs b = a;
            b.prepare = 1;
            b.prepare_spf = it == 0 ? 1 : 0;
            b.nladder = 0;
            b.nregular = 0;
            if (!run(b, 1)) break;
        }
        if (st[0] == 0) {
            ttnd::RunArgs b = a;
            b.prepare = 0;
            b.prepare_spf = 0;
            b.nregular = 0;
            run(b, b.nladder);
            if (st[0] == 0) done = 1;
        }
    }
    while (st[0] == 0 && done < nsteps) {
        const int64_t bound = rng_bound(in);
        const int64_t chunk =
            bound == 0 ? nsteps - done : std::max<int64_t>(1, std::min<int64_t>(RNG_BUDGET / bound, nsteps - done));
        ttnd::RunArgs b = a;
        b.prepare = 0;
        b.prepare_spf = 0;
        b.nladder = 0;
        b.nregular = static_cast<int>(chunk);
        b.step0 = done;
        run(b, chunk);
        if (st[0] == 0) done += chunk;
    }
    in->last_ms = ms_total;
    if (st[0] != 0) {
        // the steps completed before the failure are kept; the failing one is not
        keep_traj(in, done);
        in->nsteps += done;
        fail(ERR_NUMERIC, failure_message(st));
    }
    finish(in, nsteps);
}

int ttn_step(ttn_integ* in, double cr, double ci, int64_t nsteps) {
    return guard(in->ctx, [&] {
        if (nsteps <= 0) return;
        if (in->expand) {
            step_expanding(in, cr, ci, nsteps);
            return;
        }
        ttnd::RunArgs a;
        make_args(in, cr, ci, nsteps, a);
        prepare_traj(in, nsteps);
        hipStream_t s = in->ctx->stream;
        hip_check(hipMemcpyAsync(in->d_inst.p, &in->inst, sizeof(in->inst), hipMemcpyHostToDevice, s), "inst");
        launch(in->ctx, in->d_inst.p, {in->inst}, a, {auto_team(in)});
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

int ttn_expansion_counters(const ttn_integ* in, int64_t* n1, int64_t* n2) {
    if (n1) *n1 = in->n_onesite;
    if (n2) *n2 = in->n_twosite;
    return TTN_OK;
}
int ttn_hranks(const ttn_integ* in, int64_t* hr) {
    for (int n = 0; n < in->topo.size(); ++n) hr[n] = in->topo.nodes[n].hrank;
    return TTN_OK;
}
int ttn_state_shape(const ttn_integ* in, int64_t node, int64_t* rows, int64_t* cols) {
    return guard(in->ctx, [&] {
        check(node >= 0 && node < in->topo.size(), ERR_INVALID, "node out of range");
        if (rows) *rows = in->topo.nodes[node].dimen();
        if (cols) *cols = in->topo.nodes[node].hrank;
    });
}
int ttn_set_dt(ttn_integ* in, double dt) {
    return guard(in->ctx, [&] {
        check(dt > 0.0 && std::isfinite(dt), ERR_INVALID, "dt must be positive");
        in->p.dt = dt;
    });
}
int ttn_get_dt(const ttn_integ* in, double* dt) {
    *dt = in->p.dt;
    return TTN_OK;
}
int ttn_set_expansion_log(ttn_integ* in, int64_t capacity) {
    return guard(in->ctx, [&] {
        check(capacity >= 0, ERR_INVALID, "log capacity must not be negative");
        if (capacity == 0) {
            in->d_exp_log.release();
            in->inst.exp_log = nullptr;
        } else {
            in->d_exp_log.alloc(static_cast<size_t>(capacity * ttnd::EXP_LOG_REC));
            in->inst.exp_log = in->d_exp_log.p;
        }
        in->inst.exp_log_cap = capacity;
        const int64_t zero = 0;
        hip_check(hipMemcpy(in->d_exp.p + ttnd::EXP_LOG_N, &zero, sizeof(zero), hipMemcpyHostToDevice), "log");
    });
}
int ttn_expansion_log(const ttn_integ* in, int64_t* len, double* out) {
    return guard(in->ctx, [&] {
        int64_t n = 0;
        hip_check(hipMemcpy(&n, in->d_exp.p + ttnd::EXP_LOG_N, sizeof(n), hipMemcpyDeviceToHost), "log");
        n = std::min(n, in->inst.exp_log_cap);
        if (out && n > 0)
            hip_check(hipMemcpy(out, in->d_exp_log.p, std::min(n, *len) * ttnd::EXP_LOG_REC * sizeof(double),
                                hipMemcpyDeviceToHost),
                      "log");
        *len = n;
    });
}
int ttn_expansion_draws(const ttn_integ* in, int64_t* n) {
    return guard(in->ctx, [&] {
        int64_t pos = 0;
        hip_check(hipMemcpy(&pos, in->d_exp.p + ttnd::EXP_RNG_POS, sizeof(pos), hipMemcpyDeviceToHost), "rng cursor");
        *n = in->normals_used + pos;
    });
}
int ttn_reset_nsteps(ttn_integ* in) {
    in->nsteps = 0;
    return TTN_OK;
}

// ---------------------------------------------------------------- checkpoints
// Layout (native byte order, for resuming with the same library build): the 8-byte tag
// "TTNCKPT1", a fixed block of int64 fields (nodes, terms, tree/operator fingerprint, steps,
// H applications, sweeps completed, Nonesite, Ntwosite, normal draws consumed, pending draws,
// length of the generator text, sweep pairs, prepare passes), dt, the bond dimension of every
// node, the std::mt19937 / std::normal_distribution state as text, the pending normal draws,
// every node tensor at its current shape and the carried matrices (gauge factors,
// single-particle and mean-field matrices) at capacity.
extern "C++" {
namespace {
constexpr char CKPT_TAG[8] = {'T', 'T', 'N', 'C', 'K', 'P', 'T', '1'};
constexpr int CKPT_NFIELDS = 13;
enum CkptField { CK_NODES, CK_TERMS, CK_FPRINT, CK_NSTEPS, CK_NH, CK_SWEEPS, CK_ONESITE, CK_TWOSITE,
                 CK_NORMALS, CK_PENDING, CK_RNGLEN, CK_PAIRS, CK_PREPARES };

uint64_t fingerprint(const ttn_integ* in) {
    uint64_t h = 1469598103934665603ull;
    auto mix = [&h](int64_t v) {
        for (int b = 0; b < 8; ++b) {
            h ^= static_cast<uint64_t>((v >> (8 * b)) & 0xff);
            h *= 1099511628211ull;
        }
    };
    for (const Node& nd : in->topo.nodes) {
        mix(nd.parent);
        mix(nd.cap_hrank);
        mix(nd.leaf_index);
        mix(nd.is_leaf() ? nd.dims[0] : -1);
    }
    mix(in->sop.nterms);
    mix(in->p.krylov_dim);
    mix(in->p.mode);
    return h;
}

std::string rng_text(const ttn_integ* in) {
    std::ostringstream o;
    o << in->rng << ' ' << in->ndist;
    return o.str();
}

int64_t node_live(const Node& nd) { return nd.dimen() * nd.hrank; }

// bytes of the checkpoint of `in` with `pending` pending draws and generator text `rt`
int64_t ckpt_bytes(const ttn_integ* in, int64_t pending, int64_t rt) {
    int64_t b = 8 + 8 * CKPT_NFIELDS + 8 + 8 * in->topo.size() + rt + 8 * pending;
    for (const Node& nd : in->topo.nodes) b += 16 * node_live(nd);
    for (const auto& r : in->carried) b += 16 * r.second;
    return b;
}
}  // namespace
}  // extern "C++"

int ttn_checkpoint_save(ttn_integ* in, int64_t* bytes, void* buf) {
    return guard(in->ctx, [&] {
        check(bytes != nullptr, ERR_INVALID, "checkpoint: null size");
        int st[2] = {0, 0};
        hip_check(hipMemcpy(st, in->status.p, sizeof(st), hipMemcpyDeviceToHost), "status");
        check(st[0] == 0, ERR_INVALID, "checkpoint: the integrator's last step failed");
        int64_t pos = 0;
        hip_check(hipMemcpy(&pos, in->d_exp.p + ttnd::EXP_RNG_POS, sizeof(pos), hipMemcpyDeviceToHost), "rng cursor");
        check(pos >= 0 && pos <= static_cast<int64_t>(in->pending.size()), ERR_DEVICE, "random stream cursor out of range");
        const std::string rt = rng_text(in);
        const int64_t npend = static_cast<int64_t>(in->pending.size()) - pos;
        const int64_t need = ckpt_bytes(in, npend, static_cast<int64_t>(rt.size()));
        if (buf == nullptr) {
            *bytes = need;
            return;
        }
        check(*bytes >= need, ERR_INVALID, "checkpoint: buffer too small");
        int64_t c[2];
        hip_check(hipMemcpy(c, in->counters.p, sizeof(c), hipMemcpyDeviceToHost), "counters");
        char* w = static_cast<char*>(buf);
        auto put = [&w](const void* src, size_t n) {
            std::memcpy(w, src, n);
            w += n;
        };
        put(CKPT_TAG, 8);
        int64_t f[CKPT_NFIELDS];
        f[CK_NODES] = in->topo.size();
        f[CK_TERMS] = in->sop.nterms;
        f[CK_FPRINT] = static_cast<int64_t>(fingerprint(in));
        f[CK_NSTEPS] = in->nsteps;
        f[CK_NH] = c[0];
        f[CK_SWEEPS] = c[1];
        f[CK_ONESITE] = in->n_onesite;
        f[CK_TWOSITE] = in->n_twosite;
        f[CK_NORMALS] = in->normals_used + pos;
        f[CK_PENDING] = npend;
        f[CK_RNGLEN] = static_cast<int64_t>(rt.size());
        f[CK_PAIRS] = in->sweep_pairs;
        f[CK_PREPARES] = in->prepares;
        put(f, sizeof(f));
        put(&in->p.dt, 8);
        for (const Node& nd : in->topo.nodes) put(&nd.hrank, 8);
        put(rt.data(), rt.size());
        if (npend > 0) put(in->pending.data() + pos, 8 * npend);
        for (int i = 0; i < in->topo.size(); ++i) {
            const int64_t n = node_live(in->topo.nodes[i]);
            hip_check(hipMemcpy(w, in->arena.p + in->node_A[i], 16 * n, hipMemcpyDeviceToHost), "checkpoint");
            w += 16 * n;
        }
        for (const auto& r : in->carried) {
            hip_check(hipMemcpy(w, in->arena.p + r.first, 16 * r.second, hipMemcpyDeviceToHost), "checkpoint");
            w += 16 * r.second;
        }
        *bytes = need;
    });
}

int ttn_checkpoint_load(ttn_integ* in, const void* buf, int64_t bytes) {
    return guard(in->ctx, [&] {
        check(buf != nullptr && bytes >= 8 + 8 * CKPT_NFIELDS + 8, ERR_INVALID, "checkpoint: truncated");
        const char* r = static_cast<const char*>(buf);
        const char* end = r + bytes;
        auto get = [&r, end](void* dst, size_t n) {
            check(r + n <= end, ERR_INVALID, "checkpoint: truncated");
            std::memcpy(dst, r, n);
            r += n;
        };
        char tag[8];
        get(tag, 8);
        check(std::memcmp(tag, CKPT_TAG, 8) == 0, ERR_INVALID, "checkpoint: not a checkpoint of this library");
        int64_t f[CKPT_NFIELDS];
        get(f, sizeof(f));
        check(f[CK_NODES] == in->topo.size() && f[CK_TERMS] == in->sop.nterms &&
                  f[CK_FPRINT] == static_cast<int64_t>(fingerprint(in)),
              ERR_INVALID, "checkpoint: written for another tree, operator or integrator mode");
        check(f[CK_PENDING] >= 0 && f[CK_RNGLEN] >= 0 && f[CK_NSTEPS] >= 0, ERR_INVALID, "checkpoint: corrupt header");
        double dt = 0.0;
        get(&dt, 8);
        check(dt > 0.0 && std::isfinite(dt), ERR_INVALID, "checkpoint: corrupt dt");
        const int n = in->topo.size();
        std::vector<int64_t> hr(n);
        get(hr.data(), 8 * n);
        for (int i = 0; i < n; ++i) {
            const Node& nd = in->topo.nodes[i];
            check(hr[i] >= 1 && hr[i] <= nd.cap_hrank, ERR_INVALID, "checkpoint: bond dimension beyond capacity");
        }
        std::string rt(static_cast<size_t>(f[CK_RNGLEN]), '\0');
        get(rt.data(), rt.size());
        std::vector<double> pend(static_cast<size_t>(f[CK_PENDING]));
        get(pend.data(), 8 * pend.size());
        // shapes first: the node tensors are stored at their checkpointed shapes
        for (int i = 0; i < n; ++i) {
            Node& nd = in->topo.nodes[i];
            nd.hrank = hr[i];
            if (nd.parent >= 0) in->topo.nodes[nd.parent].dims[nd.child_id] = hr[i];
        }
        check(bytes == ckpt_bytes(in, f[CK_PENDING], f[CK_RNGLEN]), ERR_INVALID,
              "checkpoint: size does not match its header");
        for (int i = 0; i < n; ++i) {
            const int64_t cnt = node_live(in->topo.nodes[i]);
            hip_check(hipMemcpy(in->arena.p + in->node_A[i], r, 16 * cnt, hipMemcpyHostToDevice), "restore");
            r += 16 * cnt;
        }
        for (const auto& reg : in->carried) {
            hip_check(hipMemcpy(in->arena.p + reg.first, r, 16 * reg.second, hipMemcpyHostToDevice), "restore");
            r += 16 * reg.second;
        }
        std::vector<int> hri(hr.begin(), hr.end());
        hip_check(hipMemcpy(in->hr.p, hri.data(), n * sizeof(int), hipMemcpyHostToDevice), "restore hranks");
        const int64_t c[2] = {f[CK_NH], f[CK_SWEEPS]};
        hip_check(hipMemcpy(in->counters.p, c, sizeof(c), hipMemcpyHostToDevice), "restore counters");
        int64_t w[ttnd::EXP_NWORDS] = {};
        w[ttnd::EXP_ONESITE] = f[CK_ONESITE];
        w[ttnd::EXP_TWOSITE] = f[CK_TWOSITE];
        hip_check(hipMemcpy(in->d_exp.p, w, sizeof(w), hipMemcpyHostToDevice), "restore expansion words");
        const int zero[2] = {0, 0};
        hip_check(hipMemcpy(in->status.p, zero, sizeof(zero), hipMemcpyHostToDevice), "restore status");
        std::istringstream is(rt);
        is >> in->rng >> in->ndist;
        check(!is.fail(), ERR_INVALID, "checkpoint: corrupt generator state");
        in->pending = std::move(pend);
        in->normals_used = f[CK_NORMALS];
        in->n_onesite = f[CK_ONESITE];
        in->n_twosite = f[CK_TWOSITE];
        in->nsteps = f[CK_NSTEPS];
        in->sweep_pairs = f[CK_PAIRS];
        in->prepares = f[CK_PREPARES];
        in->p.dt = dt;
        in->inst.exp_log_cap = 0;
        in->inst.exp_log = nullptr;
        in->fm = flop_model(in->topo, in->ops, in->sop);
    });
}

int ttn_set_expansion_params(ttn_integ* in, const ttn_integ_params* p) {
    return guard(in->ctx, [&] {
        check(p != nullptr, ERR_INVALID, "null parameters");
        check(in->expand, ERR_INVALID, "expansion parameters need a subspace-expansion integrator");
        check(p->neigenvalues >= 1, ERR_INVALID, "neigenvalues must be positive");
        check(p->minimum_unoccupied >= 0, ERR_INVALID, "minimum_unoccupied must not be negative");
        in->p.spawning_threshold = p->spawning_threshold;
        in->p.unoccupied_threshold = p->unoccupied_threshold;
        in->p.minimum_unoccupied = p->minimum_unoccupied;
        in->p.neigenvalues = p->neigenvalues;
        in->p.eigenvalue_tolerance = p->eigenvalue_tolerance;
        in->p.only_apply_when_no_unoccupied = p->only_apply_when_no_unoccupied;
        ttnd::DInst& I = in->inst;  // uploaded with every launch
        I.neig = static_cast<int>(p->neigenvalues);
        I.min_unocc = static_cast<int>(p->minimum_unoccupied);
        I.only_unocc = p->only_apply_when_no_unoccupied != 0 ? 1 : 0;
        I.spawn = p->spawning_threshold;
        I.unocc_thr = p->unoccupied_threshold;
        I.eig_tol = p->eigenvalue_tolerance > 0.0 ? p->eigenvalue_tolerance : 1e-6;
    });
}

int ttn_set_coefficients(ttn_integ* in, int64_t nterms, const double* coeff) {
    return guard(in->ctx, [&] {
        check(nterms == in->sop.nterms, ERR_INVALID, "set_coefficients: term count differs from the operator's");
        check(coeff != nullptr, ERR_INVALID, "set_coefficients: null coefficients");
        for (int64_t r = 0; r < nterms; ++r) {
            check(std::isfinite(coeff[2 * r]) && std::isfinite(coeff[2 * r + 1]), ERR_INVALID,
                  "set_coefficients: non-finite coefficient");
            in->sop.coeff[r] = cplx(coeff[2 * r], coeff[2 * r + 1]);
        }
        // coefficients are copied into the compiled tables, never used to shape them: the
        // recompiled tables differ from the uploaded ones in the coefficient fields only
        in->ops = compile_sop(in->topo, in->sop);
        size_t ti = 0, si = 0, mi = 0;
        for (int i = 0; i < in->topo.size(); ++i)
            for (const Term& tm : in->ops.nodes[i].terms) {
                check(ti < in->h_terms.size(), ERR_INVALID, "set_coefficients: compiled tables changed shape");
                term_coeffs(tm, in->h_terms[ti++]);
                for (int q = 0; q < tm.nspf_terms(); ++q) sub_coeff(tm, q, in->h_subs.at(si++));
                for (int q = 0; q < tm.nmf_terms(); ++q) mf_coeff(tm, q, in->h_mfs.at(mi++));
            }
        check(ti == in->h_terms.size() && si == in->h_subs.size() && mi == in->h_mfs.size(), ERR_INVALID,
              "set_coefficients: compiled tables changed shape");
        hipStream_t s = in->ctx->stream;
        hip_check(hipStreamSynchronize(s), "sync");
        auto up = [&](size_t off, const void* src, size_t bytes) {
            if (bytes) hip_check(hipMemcpy(in->tables.p + off, src, bytes, hipMemcpyHostToDevice), "coefficients");
        };
        up(in->o_terms, in->h_terms.data(), in->h_terms.size() * sizeof(ttnd::DTerm));
        up(in->o_subs, in->h_subs.data(), in->h_subs.size() * sizeof(ttnd::DSub));
        up(in->o_mfs, in->h_mfs.data(), in->h_mfs.size() * sizeof(ttnd::DMf));
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

}  // extern "C"

// ---------------------------------------------------------------- costate gradient
namespace {

// A device copy of an integrator's whole arena (node tensors, gauge and environment
// matrices, operators) with its step count and dt: instances built from the same tree,
// operator and parameters share the arena layout, so a copy moves a state between them.
struct ArenaSnap {
    DevBuf<double2> a;
    int64_t nsteps = 0;
    double dt = 0.0;
};

void snap_take(const ttn_integ* in, ArenaSnap& s) {
    if (static_cast<int64_t>(s.a.n) < in->arena_size) s.a.alloc(in->arena_size);
    hip_check(hipMemcpyAsync(s.a.p, in->arena.p, in->arena_size * sizeof(double2), hipMemcpyDeviceToDevice,
                             in->ctx->stream),
              "snapshot");
    s.nsteps = in->nsteps;
    s.dt = in->p.dt;
}

void snap_put(ttn_integ* in, const ArenaSnap& s) {
    hip_check(hipMemcpyAsync(in->arena.p, s.a.p, in->arena_size * sizeof(double2), hipMemcpyDeviceToDevice,
                             in->ctx->stream),
              "restore");
    in->nsteps = s.nsteps;
    in->p.dt = s.dt;
}

void copy_state(ttn_integ* dst, const ttn_integ* src) {
    check(dst->arena_size == src->arena_size, ERR_INVALID, "adjoint: instances with different layouts");
    hip_check(hipMemcpyAsync(dst->arena.p, src->arena.p, src->arena_size * sizeof(double2), hipMemcpyDeviceToDevice,
                             src->ctx->stream),
              "copy state");
    dst->nsteps = src->nsteps;
    dst->p.dt = src->p.dt;
}

// a second integrator on the same tree, operator (current coefficients) and parameters
std::unique_ptr<ttn_integ> twin(const ttn_integ* in) {
    auto t = std::make_unique<ttn_integ>();
    t->ctx = in->ctx;
    t->topo = in->topo;
    t->sop = in->sop;
    t->p = in->p;
    t->ops = compile_sop(t->topo, t->sop);
    build_instance(t.get());
    return t;
}

// n integrator calls of one instance in one launch (ttn_step without the trajectory bookkeeping)
void step_n(ttn_integ* in, double ci, int64_t n) {
    ttnd::RunArgs a;
    make_args(in, 0.0, ci, n, a);
    prepare_traj(in, n);
    hip_check(hipMemcpyAsync(in->d_inst.p, &in->inst, sizeof(in->inst), hipMemcpyHostToDevice, in->ctx->stream),
              "inst");
    launch(in->ctx, in->d_inst.p, {in->inst}, a, {auto_team(in)});
    account(in, a);
    finish(in, n);
}

// one call of every instance, each with its own coefficient (-i forward, +i backward in time)
// and dt, in one persistent launch (teams sized by the batch autotuner)
void step_each(ttn_ctx* ctx, const std::vector<ttn_integ*>& ins, const std::vector<double>& ci,
               DevBuf<ttnd::RunArgs>& d_args, DevBuf<ttnd::DInst>& d_insts) {
    const size_t n = ins.size();
    std::vector<ttnd::RunArgs> args(n);
    for (size_t k = 0; k < n; ++k) make_args(ins[k], 0.0, ci[k], 1, args[k]);
    if (d_args.n < n) d_args.alloc(n);
    if (d_insts.n < n) d_insts.alloc(n);
    hipStream_t s = ctx->stream;
    hip_check(hipMemcpyAsync(d_args.p, args.data(), n * sizeof(ttnd::RunArgs), hipMemcpyHostToDevice, s), "args");
    std::vector<ttnd::DInst> insts;
    for (size_t k = 0; k < n; ++k) {
        prepare_traj(ins[k], 1);
        ttnd::DInst I = ins[k]->inst;
        I.own = d_args.p + k;
        insts.push_back(I);
    }
    hip_check(hipMemcpyAsync(d_insts.p, insts.data(), n * sizeof(ttnd::DInst), hipMemcpyHostToDevice, s), "insts");
    ttn_batch b;
    b.ctx = ctx;
    b.integs = ins;
    for (ttn_integ* in : ins) b.cost.push_back(static_cost(in));
    launch(ctx, d_insts.p, insts, args[0], batch_teams(&b));
    std::string err;
    for (size_t k = 0; k < n; ++k) {
        account(ins[k], args[k]);
        try {
            finish(ins[k], 1);
        } catch (const Error& e) {
            if (err.empty()) err = e.what();
        }
    }
    if (!err.empty()) fail(ERR_NUMERIC, "adjoint: " + err);
}

// Gauss-Legendre nodes and weights on [0, 1] (Newton's method on P_n)
void gauss_legendre01(int q, std::vector<double>& x, std::vector<double>& w) {
    x.assign(q, 0.0);
    w.assign(q, 0.0);
    const double pi = std::acos(-1.0);
    for (int i = 0; i < q; ++i) {
        double z = std::cos(pi * (i + 0.75) / (q + 0.5)), dp = 1.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= q; ++k) {
                const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = q * (z * p1 - p0) / (z * z - 1.0);
            const double dz = p1 / dp;
            z -= dz;
            if (std::abs(dz) < 1e-16) break;
        }
        x[i] = 0.5 * (1.0 - z);
        w[i] = 1.0 / ((1.0 - z * z) * dp * dp);
    }
}

}  // namespace

extern "C" {

int ttn_sb_adjoint_run(ttn_integ* in, const ttn_sb_params* p, int64_t nsteps, int64_t every, int64_t nquad,
                       int64_t mode, const double* op, double* value, double* dg, double* dw) {
    return guard(in->ctx, [&] {
        const Topology& tp = in->topo;
        check(!in->expand, ERR_INVALID, "adjoint: the fixed-step integrator only");
        check(in->nsteps == 0, ERR_INVALID, "adjoint: the state must not have been stepped");
        check(nsteps >= 1 && every >= 1 && nquad >= 0 && nquad <= 16, ERR_INVALID,
              "adjoint: nsteps and the checkpoint interval must be positive, 0-16 quadrature nodes");
        check(mode >= 0 && mode < tp.nmodes(), ERR_INVALID, "adjoint: mode out of range");
        const SBParams sp = sb_from_c(p);
        check(static_cast<int64_t>(sb_blocks(sp).size()) == tp.nmodes() - 1, ERR_INVALID,
              "adjoint: the model's leaves do not match the tree");
        hip_check(hipSetDevice(in->ctx->device), "hipSetDevice");
        hipStream_t st = in->ctx->stream;
        const double dt = in->p.dt;
        const int Q = nquad > 0 ? static_cast<int>(nquad) : sb_quadrature_nodes(sp, dt);
        // forward: one launch per checkpoint interval
        std::vector<std::unique_ptr<ArenaSnap>> ck;
        for (int64_t s0 = 0; s0 < nsteps; s0 += every) {
            ck.push_back(std::make_unique<ArenaSnap>());
            snap_take(in, *ck.back());
            step_n(in, -1.0, std::min(every, nsteps - s0));
        }
        ArenaSnap fin;
        snap_take(in, fin);
        // J and chi(T) = O_s psi(T)
        const int64_t dm = tp.mode_dims[mode];
        std::vector<cplx> O(dm * dm), Os(dm * dm);
        for (int64_t i = 0; i < dm * dm; ++i) O[i] = cget(op, i);
        for (int64_t a = 0; a < dm; ++a)
            for (int64_t b = 0; b < dm; ++b) Os[a * dm + b] = 0.5 * (O[a * dm + b] + std::conj(O[b * dm + a]));
        hip_check(hipMemcpy(in->arena.p + in->tmp_op, O.data(), dm * dm * sizeof(double2), hipMemcpyHostToDevice), "op");
        double e3[3];
        run_expect(in, tp.leaf_node[mode], in->tmp_op, e3);
        auto chi = twin(in);
        copy_state(chi.get(), in);
        {
            const int ln = tp.leaf_node[mode];
            const int64_t r = tp.nodes[ln].hrank;
            std::vector<cplx> A(dm * r), B(dm * r);
            hip_check(hipStreamSynchronize(st), "sync");
            hip_check(hipMemcpy(A.data(), chi->arena.p + chi->node_A[ln], A.size() * sizeof(double2),
                                hipMemcpyDeviceToHost),
                      "leaf");
            for (int64_t a = 0; a < dm; ++a)
                for (int64_t c = 0; c < r; ++c) {
                    cplx s(0.0, 0.0);
                    for (int64_t b = 0; b < dm; ++b) s += Os[a * dm + b] * A[b * r + c];
                    B[a * r + c] = s;
                }
            hip_check(hipMemcpy(chi->arena.p + chi->node_A[ln], B.data(), B.size() * sizeof(double2),
                                hipMemcpyHostToDevice),
                      "leaf");
        }
        chi->nsteps = 0;
        chi->p.dt = dt;
        // one launch per step: the 2Q instances of the step's quadrature (psi at t_n + x_q dt,
        // chi at the same times) and chi's own full step back to t_n
        std::vector<std::unique_ptr<ttn_integ>> W;
        std::vector<ttn_integ*> wp;
        for (int q = 0; q < 2 * Q; ++q) {
            W.push_back(twin(in));
            wp.push_back(W.back().get());
        }
        wp.push_back(chi.get());
        std::vector<double> xs, ws;
        gauss_legendre01(Q, xs, ws);
        std::vector<double> ci(2 * Q + 1, 1.0);
        for (int q = 0; q < Q; ++q) ci[q] = -1.0;
        // transition densities: pairs (W[q], W[Q + q]), S = Scoup on the system leaf
        const int64_t d0 = tp.mode_dims[0];
        check(static_cast<int64_t>(sp.scoup.size()) == d0 * d0, ERR_INVALID, "adjoint: Scoup has the wrong size");
        DevBuf<double2> dS;
        dS.alloc(d0 * d0);
        hip_check(hipMemcpy(dS.p, sp.scoup.data(), d0 * d0 * sizeof(double2), hipMemcpyHostToDevice), "Scoup");
        int64_t rr = 1;
        for (const Node& nd : tp.nodes) rr = std::max<int64_t>(rr, nd.hrank * nd.hrank);
        const int64_t work_stride = 4 * static_cast<int64_t>(tp.size()) * rr + 3 * in->inst.maxsize;
        DevBuf<double2> dwork;
        dwork.alloc(work_stride * Q);
        std::vector<int64_t> off(tp.size(), -1), moff(tp.nmodes());
        int64_t out_stride = 0;
        for (int m = 0; m < tp.nmodes(); ++m) {
            moff[m] = out_stride;
            off[tp.leaf_node[m]] = out_stride;
            out_stride += 2 * tp.mode_dims[m] * tp.mode_dims[m];
        }
        DevBuf<int64_t> doff;
        doff.alloc(tp.size());
        hip_check(hipMemcpy(doff.p, off.data(), off.size() * sizeof(int64_t), hipMemcpyHostToDevice), "offsets");
        DevBuf<double2> dout;
        dout.alloc(out_stride * Q);
        std::vector<cplx> hout(out_stride * Q);
        DevBuf<ttnd::RunArgs> d_args;
        DevBuf<ttnd::DInst> d_insts, d_pairs;
        d_pairs.alloc(2 * Q);
        // (the pairs read the quadrature instances only; chi has moved on by then)
        std::vector<double> g(sp.wk.size(), 0.0), w(sp.wk.size(), 0.0);
        std::vector<std::vector<cplx>> rho(tp.nmodes()), rhoS(tp.nmodes());
        std::vector<std::unique_ptr<ArenaSnap>> starts;
        for (int64_t gi = static_cast<int64_t>(ck.size()) - 1; gi >= 0; --gi) {
            snap_put(in, *ck[gi]);
            const int64_t s0 = gi * every, s1 = std::min(nsteps, s0 + every);
            for (int64_t s = s0; s < s1; ++s) {
                const size_t k = static_cast<size_t>(s - s0);
                if (starts.size() <= k) starts.push_back(std::make_unique<ArenaSnap>());
                snap_take(in, *starts[k]);
                if (s + 1 < s1) step_n(in, -1.0, 1);
            }
            for (int64_t s = s1 - 1; s >= s0; --s) {
                for (int q = 0; q < Q; ++q) {
                    snap_put(W[q].get(), *starts[s - s0]);
                    W[q]->p.dt = xs[q] * dt;
                    copy_state(W[Q + q].get(), chi.get());  // chi(t_{s+1})
                    W[Q + q]->p.dt = (1.0 - xs[q]) * dt;
                }
                chi->p.dt = dt;
                step_each(in->ctx, wp, ci, d_args, d_insts);  // chi -> chi(t_s)
                std::vector<ttnd::DInst> pr;
                for (int q = 0; q < 2 * Q; ++q) pr.push_back(W[q]->inst);
                hip_check(hipMemcpyAsync(d_pairs.p, pr.data(), pr.size() * sizeof(ttnd::DInst), hipMemcpyHostToDevice,
                                         st),
                          "pairs");
                hip_check(ttnd_launch_densities(d_pairs.p, d_pairs.p + Q, Q, tp.leaf_node[0], dS.p, dwork.p,
                                                work_stride, rr, doff.p, dout.p, out_stride, st),
                          "launch densities_kernel");
                hip_check(hipMemcpyAsync(hout.data(), dout.p, hout.size() * sizeof(double2), hipMemcpyDeviceToHost, st),
                          "densities");
                hip_check(hipStreamSynchronize(st), "sync");
                for (int q = 0; q < Q; ++q) {
                    for (int m = 0; m < tp.nmodes(); ++m) {
                        const int64_t dd = tp.mode_dims[m] * tp.mode_dims[m];
                        const cplx* base = hout.data() + q * out_stride + moff[m];
                        rho[m].assign(base, base + dd);
                        rhoS[m].assign(base + dd, base + 2 * dd);
                    }
                    sb_density_gradient(sp, rho, rhoS, ws[q] * dt, g, w);
                }
            }
        }
        snap_put(in, fin);
        hip_check(hipStreamSynchronize(st), "sync");
        std::copy(g.begin(), g.end(), dg);
        std::copy(w.begin(), w.end(), dw);
        if (value) *value = e3[0];
    });
}

int ttn_sb_partition_modes(ttn_ctx* ctx, int64_t nmodes, const double* wk, double ecut, double wc, int64_t nmax,
                           int64_t nmax4, int64_t nmax2, int64_t nmax3, int64_t* nblocks, int64_t* partition) {
    return guard(ctx, [&] { sb_partition_c(nmodes, wk, ecut, wc, nmax, nmax4, nmax2, nmax3, nblocks, partition); });
}

int ttn_sb_mode_block(ttn_ctx* ctx, int64_t nmodes, const double* w, const double* g, double ecut, int64_t nmax,
                      int64_t* nstates, int64_t* nnz, int64_t* occupations, double* h0, int64_t* hc_rowptr,
                      int64_t* hc_colind, double* hc_vals) {
    return guard(ctx, [&] {
        sb_mode_block_c(nmodes, w, g, ecut, nmax, nstates, nnz, occupations, h0, hc_rowptr, hc_colind, hc_vals);
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
            // growing bonds need the per-instance prepare loop and normal-stream top-ups of
            // ttn_step; at capacity the expansion adds nothing and the step is the fixed one
            check(!in->expand || rng_bound(in) == 0, ERR_INVALID,
                  "batched stepping takes expansion-mode instances only once every bond is at capacity; "
                  "step growing instances with ttn_step");
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
        if (b->cost.empty())
            for (ttn_integ* in : b->integs) b->cost.push_back(static_cost(in));
        b->team = batch_teams(b);
        std::vector<int64_t> t0(b->integs.size());
        for (size_t k = 0; k < b->integs.size(); ++k) t0[k] = device_ticks(b->integs[k]);
        launch(b->ctx, b->d_insts.p, insts, a, b->team, b);
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, b->ctx->ev0, b->ctx->ev1), "elapsed");
        b->last_ms = ms;
        // every instance is accounted before the first failure is reported; the measured
        // device time per sweep pair, brought back to one workgroup with the same team
        // model, becomes the autotuner's cost for the next call
        std::string err;
        const double pairs = static_cast<double>(a.prepare + a.nladder + a.nregular);
        for (size_t k = 0; k < b->integs.size(); ++k) {
            ttn_integ* in = b->integs[k];
            account(in, a);
            const double s = serial_share(in), tk = static_cast<double>(std::max(1, b->team[k]));
            const double ticks = static_cast<double>(device_ticks(in) - t0[k]);
            b->cost[k] = std::max(1.0, ticks / pairs / (s + (1.0 - s) / tk));
            try {
                finish(in, nsteps);
            } catch (const Error& e) {
                if (err.empty()) err = "batch instance " + std::to_string(k) + ": " + e.what();
            }
        }
        if (!err.empty()) fail(ERR_NUMERIC, err);
    });
}

int ttn_batch_set_graph(ttn_batch* b, int enabled) {
    return guard(b->ctx, [&] {
        b->use_graph = enabled != 0;
        if (!b->use_graph) b->graph.reset();
    });
}

int ttn_batch_graph_stats(const ttn_batch* b, int64_t* captures, int64_t* replays) {
    if (captures) *captures = b->graph_captures;
    if (replays) *replays = b->graph_replays;
    return TTN_OK;
}

int ttn_flops(const ttn_integ* in, double* flops) {
    return guard(in->ctx, [&] { *flops = measured_flops(in); });
}

int ttn_profile(const ttn_integ* in, int64_t* slots) {
    return guard(in->ctx, [&] {
        hip_check(hipMemcpy(slots, in->prof.p, ttnd::PROF_NSLOTS * sizeof(int64_t), hipMemcpyDeviceToHost), "prof");
    });
}

int ttn_node_stats(const ttn_integ* in, int64_t* out) {
    return guard(in->ctx, [&] {
        const int n = in->topo.size();
        hip_check(hipMemcpy(out, in->node_stats.p, ttnd::NODE_STATS * n * sizeof(int64_t), hipMemcpyDeviceToHost),
                  "stats");
    });
}

int ttn_ctx_set_tuning(ttn_ctx* c, int64_t team_min, int teams_enabled) {
    return guard(c, [&] {
        check(team_min >= 1, ERR_INVALID, "team_min must be positive");
        c->team_min = team_min;
        if (!teams_enabled)
            c->teams_ok = false;
        else
            team_census(c);
    });
}

int ttn_ctx_set_small(ttn_ctx* c, int64_t small_max) {
    return guard(c, [&] {
        check(small_max >= 0, ERR_INVALID, "small_max must not be negative");
        c->small_max = small_max;
    });
}

int ttn_batch_teams(const ttn_batch* b, int64_t* sizes) {
    for (size_t k = 0; k < b->integs.size(); ++k) sizes[k] = k < b->team.size() ? b->team[k] : 1;
    return TTN_OK;
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
