// Node-level C ABI (include/ttn_capi.h, ttn_node_*): one large branch node tensor on the
// whole GPU.  The effective Hamiltonian is given as product terms coeff_t (O0_t (x) O1_t)
// . MF_t^T over the node's two child modes and its bond (the environments an operator node
// of the reference holds), and is applied with the batched FP64-MFMA GEMMs of
// device/node_gemm.hip in three stages:
//   Y_t = O1_t x_1 A           batch (t, i0):  d1 x r = (d1 x d1)(d1 x r)
//   Z_t = O0_t x_0 Y_t         batch t:        d0 x (d1 r) = (d0 x d0)(d0 x d1 r)
//   out = sum_t Z_t (coeff_t MF_t^T)           one GEMM with K = nterms * r, split over K
// Mean fields of child c and the Krylov propagator reuse the same GEMMs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "device/node_gemm.hpp"
#include "host/smallexp.hpp"
#include "runtime.hpp"

extern "C" hipError_t ttnd_zgemm(const ttnd::GemmDesc* g, int nsplit, int tile, hipStream_t s);
extern "C" hipError_t ttnd_zheff_fused(const ttnd::FusedDesc* f, int ngroups, hipStream_t s);
extern "C" hipError_t ttnd_zmf_fused(const ttnd::MfDesc* f, int nchunks, hipStream_t s);
extern "C" hipError_t ttnd_zdots(const double2* V, int64_t ld, int m, const double2* w, int64_t n, double* partial,
                                 int nblk, double* h, hipStream_t s);
extern "C" hipError_t ttnd_zorth(const double2* V, int64_t ld, int m, const double* h, double2* w, int64_t n,
                                 double* partial, int nblk, hipStream_t s);
extern "C" hipError_t ttnd_zscale(double2* y, const double2* x, double ar, double ai, int64_t n, int nblk,
                                  hipStream_t s);
extern "C" hipError_t ttnd_zcombine(double2* y, const double2* V, int64_t ld, int m, const double* c, double beta,
                                    int64_t n, int nblk, hipStream_t s);
extern "C" hipError_t ttnd_znorm(const double2* x, int64_t n, double* partial, int nblk, hipStream_t s);
extern "C" hipError_t ttnd_zmft(const double2* coeff, const double2* mf, int64_t nt, int64_t r, double2* out,
                                hipStream_t s);

// ---------------------------------------------------------------- allocation-free scratch
// Device memory of ttn_node_heff_noalloc: the mean-field factors coeff_t MF_t^T, the per-term
// intermediates of one chunk of terms and the K-split workspace.  Every GPU the library runs
// on holds its own copy until the process ends.  Host code never passes `scratch` itself to
// a kernel: its device address is taken once per device (ttnrt_noalloc_setup, called when a
// context is created) and only that address is used.
constexpr int64_t NOALLOC_SCRATCH_BYTES = int64_t(64) << 20;
static __device__ __align__(256) double2 scratch[NOALLOC_SCRATCH_BYTES / sizeof(double2)];

using namespace ttnh;
using namespace ttnrt;

struct ttn_node {
    ttn_ctx* ctx;
    int64_t d0, d1, r, nt;
    int64_t D, size;
    DevBuf<double2> op0, op1, mf, mft;  // operators; mft_t = coeff_t MF_t^T stacked (nt r x r)
    DevBuf<double2> A, out, Y, Z, ws, V, w;
    DevBuf<double> part, h;
    int nblk = 1024;                  // blocks of the Krylov vector kernels
    int64_t kcap = 0;                 // Krylov vectors allocated
    double last_ms = 0.0;
    double last_flops = 0.0;
};

namespace {

constexpr int KCHUNK = 16;   // GEMM K chunk
constexpr int TARGET_WG = 512;
// output tile (rows, columns) of every ttnd::GemmTile
constexpr int TILE_DIMS[ttnd::GEMM_TILES][2] = {{32, 32}, {64, 32}, {32, 64}, {64, 64}, {128, 64}, {64, 128}};

int tiles(int64_t M, int64_t N, int tile = ttnd::TILE_64x64) {
    const int64_t tm = TILE_DIMS[tile][0], tn = TILE_DIMS[tile][1];
    return static_cast<int>(((M + tm - 1) / tm) * ((N + tn - 1) / tn));
}

// K split that gives the GEMM about TARGET_WG workgroups
int split_for(int64_t M, int64_t N, int64_t batch, int64_t K, int tile = ttnd::TILE_64x64) {
    const int64_t wg = static_cast<int64_t>(tiles(M, N, tile)) * batch;
    if (wg >= TARGET_WG / 2) return 1;
    const int64_t want = (TARGET_WG + wg - 1) / wg;
    const int64_t chunks = (K + KCHUNK - 1) / KCHUNK;
    return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(want, chunks)));
}

ttnd::GemmDesc desc0() {
    ttnd::GemmDesc g{};
    g.nb0 = g.nb1 = 1;
    g.alpha = make_double2(1.0, 0.0);
    return g;
}

// The buffers one H-action works on: a node's own allocations, or slices of the
// allocation-free scratch (ttn_node_heff_noalloc).
struct NodeView {
    int64_t d0, d1, r, nt, D, size;
    const double2 *op0, *op1, *mft;  // operators: nt d0 x d0, nt d1 x d1, nt r x r (coeff_t MF_t^T)
    double2 *Y, *Z;                  // per-term intermediates, nt x size each
    double2* ws;                     // K-split workspace
    int64_t ws_n;                    // its entries
    bool ws_cap;                     // run unsplit when a split would not fit (else an error)
    hipStream_t s;
    ttn_ctx* ctx;                    // its tile table; new shapes are timed only when alloc_ok
    bool alloc_ok;
};

NodeView view(ttn_node* nd) {
    return NodeView{nd->d0, nd->d1, nd->r, nd->nt, nd->D, nd->size, nd->op0.p, nd->op1.p, nd->mft.p,
                    nd->Y.p, nd->Z.p, nd->ws.p, static_cast<int64_t>(nd->ws.n), false, nd->ctx->stream, nd->ctx,
                    true};
}

int64_t split_need(const ttnd::GemmDesc& g, int ns) { return static_cast<int64_t>(ns) * g.M * g.N * g.nb0 * g.nb1; }

void launch_gemm(const NodeView& v, ttnd::GemmDesc g, int tile, int ns) {
    if (ns > 1) {
        g.W = v.ws;
        g.w_batch = static_cast<int64_t>(g.M) * g.N;
        g.w_split = g.w_batch * g.nb0 * g.nb1;
    }
    hip_check(ttnd_zgemm(&g, ns, tile, v.s), "zgemm");
}

// The autotuner: the first time a context meets a GEMM shape it times every output tile with
// the K split that gives that tile about TARGET_WG workgroups, twice and half that split and
// no split (those whose partial products fit the workspace), into a scratch output, and keeps
// the fastest for the shape.  Every tile sums each output element in the same order, so the
// choice of tile never changes results; a different K split changes their rounding.
// A pair of HIP events for timing work on a stream, released on every path out.
class EventTimer {
public:
    EventTimer() {
        hip_check(hipEventCreate(&a_), "event");
        const hipError_t e = hipEventCreate(&b_);
        if (e != hipSuccess) {
            (void)hipEventDestroy(a_);
            hip_check(e, "event");
        }
    }
    ~EventTimer() {
        (void)hipEventDestroy(a_);
        (void)hipEventDestroy(b_);
    }
    EventTimer(const EventTimer&) = delete;
    EventTimer& operator=(const EventTimer&) = delete;
    // device time of f() in ms
    template <class F>
    float time(hipStream_t s, F&& f) {
        hip_check(hipEventRecord(a_, s), "event");
        f();
        hip_check(hipEventRecord(b_, s), "event");
        hip_check(hipEventSynchronize(b_), "sync");
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, a_, b_), "elapsed");
        return ms;
    }

private:
    hipEvent_t a_ = nullptr, b_ = nullptr;
};

std::array<int64_t, 8> tune_key(const ttnd::GemmDesc& g) {
    return {g.M, g.N, g.K, static_cast<int64_t>(g.nb0) * g.nb1, g.kb, g.a_k == 1 ? 1 : 0, g.b_c == 1 ? 1 : 0,
            g.accumulate};
}

std::array<double, 3> tune(const NodeView& v, const ttnd::GemmDesc& g) {
    ttn_ctx* c = v.ctx;
    int64_t extent = 1 + (g.nb0 - 1) * g.c_b0 + (g.nb1 - 1) * g.c_b1 + (g.M - 1) * g.c_r + (g.N - 1) * g.c_c;
    if (static_cast<int64_t>(c->tune_out.n) < extent) c->tune_out.alloc(extent);
    ttnd::GemmDesc t = g;
    t.C = c->tune_out.p;
    EventTimer timer;
    std::array<double, 3> best = {static_cast<double>(ttnd::TILE_64x64), 1.0, 1e300};
    const int64_t batch = static_cast<int64_t>(g.nb0) * g.nb1;
    const int64_t chunks = (g.K + KCHUNK - 1) / KCHUNK;
    for (int tile = 0; tile < ttnd::GEMM_TILES; ++tile) {
        const int base = split_for(g.M, g.N, batch, g.K, tile);
        int splits[4] = {1, base, std::max(1, base / 2), static_cast<int>(std::min<int64_t>(chunks, 2 * base))};
        for (int si = 0; si < 4; ++si) {
            const int ns = splits[si];
            bool seen = false;
            for (int q = 0; q < si; ++q) seen = seen || splits[q] == ns;
            if (seen || (ns > 1 && split_need(g, ns) > v.ws_n)) continue;
            launch_gemm(v, t, tile, ns);  // warm-up
            float ms = 1e30f;
            for (int rep = 0; rep < 3; ++rep) ms = std::min(ms, timer.time(v.s, [&] { launch_gemm(v, t, tile, ns); }));
            if (ms < best[2]) best = {static_cast<double>(tile), static_cast<double>(ns), static_cast<double>(ms)};
        }
    }
    return best;
}

void run(const NodeView& v, ttnd::GemmDesc g) {
    if (g.kb <= 0) g.kb = g.K;
    int tile = ttnd::TILE_64x64;
    int ns = split_for(g.M, g.N, static_cast<int64_t>(g.nb0) * g.nb1, g.K);
    if (v.ctx && v.ctx->autotune) {
        const auto key = tune_key(g);
        auto it = v.ctx->tune.find(key);
        if (it == v.ctx->tune.end() && v.alloc_ok) it = v.ctx->tune.emplace(key, tune(v, g)).first;
        if (it != v.ctx->tune.end() && (it->second[1] <= 1.0 || split_need(g, static_cast<int>(it->second[1])) <= v.ws_n)) {
            tile = static_cast<int>(it->second[0]);
            ns = static_cast<int>(it->second[1]);
        }
    }
    if (ns > 1) {
        const int64_t need = split_need(g, ns);
        if (need > v.ws_n && v.ws_cap) ns = 1;
        check(need <= v.ws_n || ns == 1, ERR_INVALID, "node workspace too small for the K split");
    }
    launch_gemm(v, g, tile, ns);
}
void run(ttn_node* nd, const ttnd::GemmDesc& g) { run(view(nd), g); }

// Y_t = O1_t x_1 src for every term
void stage_mode1(const NodeView& v, const double2* src) {
    ttnd::GemmDesc g = desc0();
    g.M = static_cast<int>(v.d1);
    g.N = static_cast<int>(v.r);
    g.K = static_cast<int>(v.d1);
    g.nb0 = static_cast<int>(v.d0);
    g.nb1 = static_cast<int>(v.nt);
    g.A = v.op1;
    g.a_b1 = v.d1 * v.d1;
    g.a_r = v.d1;
    g.a_k = 1;
    g.B = src;
    g.b_b0 = v.d1 * v.r;
    g.b_k = v.r;
    g.b_c = 1;
    g.C = v.Y;
    g.c_b0 = v.d1 * v.r;
    g.c_b1 = v.size;
    g.c_r = v.r;
    g.c_c = 1;
    run(v, g);
}

// dst_t = O0_t x_0 src_t (src batch stride src_b: size for Y, 0 for a single tensor)
void stage_mode0(const NodeView& v, const double2* src, int64_t src_b, double2* dst) {
    ttnd::GemmDesc g = desc0();
    g.M = static_cast<int>(v.d0);
    g.N = static_cast<int>(v.d1 * v.r);
    g.K = static_cast<int>(v.d0);
    g.nb0 = static_cast<int>(v.nt);
    g.A = v.op0;
    g.a_b0 = v.d0 * v.d0;
    g.a_r = v.d0;
    g.a_k = 1;
    g.B = src;
    g.b_b0 = src_b;
    g.b_k = v.d1 * v.r;
    g.b_c = 1;
    g.C = dst;
    g.c_b0 = v.size;
    g.c_r = v.d1 * v.r;
    g.c_c = 1;
    run(v, g);
}

// res (=|+=) sum_t Z_t (coeff_t MF_t^T): K = nt r, blocked by term
void stage_mf_sum(const NodeView& v, double2* res, bool accumulate) {
    ttnd::GemmDesc g = desc0();
    g.M = static_cast<int>(v.D);
    g.N = static_cast<int>(v.r);
    g.K = static_cast<int>(v.nt * v.r);
    g.kb = static_cast<int>(v.r);
    g.accumulate = accumulate ? 1 : 0;
    g.A = v.Z;
    g.a_r = v.r;
    g.a_k = 1;
    g.a_kblk = v.size;
    g.B = v.mft;
    g.b_k = v.r;
    g.b_kblk = v.r * v.r;
    g.b_c = 1;
    g.C = res;
    g.c_r = v.r;
    g.c_c = 1;
    run(v, g);
}

// Stages 2 and 3 fused (zheff_fused_kernel): term groups so that d1 x groups fills the GPU at two
// workgroups per CU, fewer when their partial sums would not fit the workspace.  Returns false if
// the shape does not fit the kernel.
bool fused_tail(const NodeView& v, double2* res, bool accumulate) {
    if (v.d0 > 64 || v.r > 64) return false;
    int groups = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(v.nt, (512 + v.d1 - 1) / v.d1)));
    while (groups > 1 && groups * v.D * v.r > v.ws_n) --groups;
    ttnd::FusedDesc f{};
    f.d0 = static_cast<int>(v.d0);
    f.d1 = static_cast<int>(v.d1);
    f.r = static_cast<int>(v.r);
    f.nt = static_cast<int>(v.nt);
    f.accumulate = accumulate ? 1 : 0;
    f.Y = v.Y;
    f.op0 = v.op0;
    f.mft = v.mft;
    f.res = res;
    f.W = v.ws;
    f.w_split = v.D * v.r;
    hip_check(ttnd_zheff_fused(&f, groups, v.s), "fused H-action");
    return true;
}

// Whether this H-action shape runs its mode-0 and mean-field stages fused: the autotuner times
// both forms once per shape (into a scratch result) and keeps the faster; without it the staged
// form runs, unless ttn_ctx_set_heff_form fixed the form.
bool use_fused(const NodeView& v) {
    if (!v.ctx || v.d0 > 64 || v.r > 64 || v.ctx->heff_form == 1) return false;
    if (v.ctx->heff_form == 2) return true;
    if (!v.ctx->autotune) return false;
    const std::array<int64_t, 4> key = {v.d0, v.d1, v.r, v.nt};
    auto it = v.ctx->fuse.find(key);
    if (it == v.ctx->fuse.end()) {
        if (!v.alloc_ok) return false;
        DevBuf<double2> scratch;
        scratch.alloc(v.D * v.r);
        EventTimer timer;
        float best[2] = {1e30f, 1e30f};
        for (int form = 0; form < 2; ++form)
            for (int rep = 0; rep < 4; ++rep) {
                const float ms = timer.time(v.s, [&] {
                    if (form == 0) {
                        stage_mode0(v, v.Y, v.size, v.Z);
                        stage_mf_sum(v, scratch.p, false);
                    } else {
                        fused_tail(v, scratch.p, false);
                    }
                });
                if (rep > 0) best[form] = std::min(best[form], ms);  // the first is a warm-up
            }
        const int form = best[1] < best[0] ? 1 : 0;
        it = v.ctx->fuse.emplace(key, std::array<double, 2>{static_cast<double>(form), best[form]}).first;
    }
    return it->second[0] > 0.5;
}

void heff(const NodeView& v, const double2* x, double2* res) {
    stage_mode1(v, x);
    if (use_fused(v) && fused_tail(v, res, false)) return;
    stage_mode0(v, v.Y, v.size, v.Z);
    stage_mf_sum(v, res, false);
}
void heff(ttn_node* nd, const double2* x, double2* res) { heff(view(nd), x, res); }

// The mean fields of child `child` for every term, res[t] (dc x dc): the other child's operator
// on A (stage 1, batched GEMM), then either two more batched GEMMs (Z_t = Y_t MF_t^T to memory,
// then the long-K product with conj(A)) or the fused kernel that keeps each slab's Z_t in LDS.
void mean_field(ttn_node* nd, int child, double2* res, bool fused) {
    const NodeView v = view(nd);
    if (child == 0)
        stage_mode1(v, nd->A.p);
    else
        stage_mode0(v, nd->A.p, 0, nd->Y.p);
    const int64_t dc = child == 0 ? nd->d0 : nd->d1, ns = child == 0 ? nd->d1 : nd->d0;
    if (fused) {
        ttnd::MfDesc f{};
        f.dc = static_cast<int>(dc);
        f.ns = static_cast<int>(ns);
        f.r = static_cast<int>(nd->r);
        f.nt = static_cast<int>(nd->nt);
        f.ss = child == 0 ? nd->r : nd->d1 * nd->r;
        f.si = child == 0 ? nd->d1 * nd->r : nd->r;
        f.size = nd->size;
        f.Y = nd->Y.p;
        f.A = nd->A.p;
        f.mf = nd->mf.p;
        f.out = res;
        f.W = v.ws;
        f.w_split = nd->nt * dc * dc;
        // chunks of slabs so that chunks x terms fills the GPU at two workgroups per CU, fewer
        // when their partial sums would not fit the workspace
        int chunks = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(ns, (512 + nd->nt - 1) / nd->nt)));
        while (chunks > 1 && chunks * f.w_split > v.ws_n) --chunks;
        hip_check(ttnd_zmf_fused(&f, chunks, v.s), "fused mean field");
        return;
    }
    ttnd::GemmDesc g = desc0();
    g.M = static_cast<int>(nd->D);
    g.N = static_cast<int>(nd->r);
    g.K = static_cast<int>(nd->r);
    g.nb0 = static_cast<int>(nd->nt);
    g.A = nd->Y.p;
    g.a_b0 = nd->size;
    g.a_r = nd->r;
    g.a_k = 1;
    g.B = nd->mf.p;  // B(p, j) = MF_t(j, p)
    g.b_b0 = nd->r * nd->r;
    g.b_k = 1;
    g.b_c = nd->r;
    g.C = nd->Z.p;
    g.c_b0 = nd->size;
    g.c_r = nd->r;
    g.c_c = 1;
    run(v, g);
    // mf_t(i, j) = sum over the other indices of conj(A)(.., i, ..) Z_t(.., j, ..)
    ttnd::GemmDesc m = desc0();
    m.M = m.N = static_cast<int>(dc);
    m.nb0 = static_cast<int>(nd->nt);
    m.conja = 1;
    m.A = nd->A.p;
    m.B = nd->Z.p;
    m.b_b0 = nd->size;
    if (child == 0) {
        m.K = static_cast<int>(nd->d1 * nd->r);
        m.a_r = nd->d1 * nd->r;
        m.a_k = 1;
        m.b_k = 1;
        m.b_c = nd->d1 * nd->r;
    } else {
        m.K = static_cast<int>(nd->d0 * nd->r);
        m.kb = static_cast<int>(nd->r);
        m.a_r = nd->r;
        m.a_k = 1;
        m.a_kblk = nd->d1 * nd->r;
        m.b_k = 1;
        m.b_c = nd->r;
        m.b_kblk = nd->d1 * nd->r;
    }
    m.C = res;
    m.c_b0 = dc * dc;
    m.c_r = dc;
    m.c_c = 1;
    run(v, m);
}

// Whether the mean fields of this child and shape run fused: fixed by ttn_ctx_set_heff_form, or
// timed once per shape (both forms into a scratch result) when the autotuner is on.
bool use_fused_mf(ttn_node* nd, int child) {
    ttn_ctx* ctx = nd->ctx;
    const int64_t dc = child == 0 ? nd->d0 : nd->d1;
    // the fused kernel puts the terms on the grid's second dimension (at most 65,535)
    if (dc > 64 || nd->r > 64 || nd->nt > 65535 || ctx->heff_form == 1) return false;
    if (ctx->heff_form == 2) return true;
    if (!ctx->autotune) return false;
    const std::array<int64_t, 5> key = {child, nd->d0, nd->d1, nd->r, nd->nt};
    auto it = ctx->fuse_mf.find(key);
    if (it == ctx->fuse_mf.end()) {
        DevBuf<double2> scratch;
        scratch.alloc(static_cast<size_t>(nd->nt * dc * dc));
        EventTimer timer;
        float best[2] = {1e30f, 1e30f};
        for (int form = 0; form < 2; ++form)
            for (int rep = 0; rep < 4; ++rep) {
                const float ms = timer.time(ctx->stream, [&] { mean_field(nd, child, scratch.p, form == 1); });
                if (rep > 0) best[form] = std::min(best[form], ms);  // the first is a warm-up
            }
        const int form = best[1] < best[0] ? 1 : 0;
        it = ctx->fuse_mf.emplace(key, std::array<double, 2>{static_cast<double>(form), best[form]}).first;
    }
    return it->second[0] > 0.5;
}

double heff_flops(const ttn_node* nd) {
    return 8.0 * static_cast<double>(nd->nt) * static_cast<double>(nd->size) *
           static_cast<double>(nd->d0 + nd->d1 + nd->r);
}

void ensure_krylov(ttn_node* nd, int64_t kdim) {
    if (nd->kcap >= kdim + 1) return;
    nd->V.alloc(static_cast<size_t>((kdim + 1) * nd->size));
    nd->kcap = kdim + 1;
}

// host sum (fixed order) of nblk partial norms
double sum_partials(ttn_node* nd, int nblk) {
    std::vector<double> p(nblk);
    hip_check(hipMemcpyAsync(p.data(), nd->part.p, nblk * sizeof(double), hipMemcpyDeviceToHost, nd->ctx->stream),
              "partials");
    hip_check(hipStreamSynchronize(nd->ctx->stream), "sync");
    double s = 0.0;
    for (double x : p) s += x;
    return s;
}

double step_factor(double tol, double err, int m, double lo, double hi) {
    if (!(err > 0.0)) return hi;
    const double f = 0.9 * std::pow(tol / err, 1.0 / m);
    return std::min(hi, std::max(lo, f));
}

// A <- exp(coeff tau H) A: Arnoldi with CGS2, Saad estimate, Expokit sub-step sizes, Pade(13)
// small exponential; the same algorithm, step for step, as the sweep kernel's expv.
int64_t node_expv(ttn_node* nd, cplx coeff, double tau, int kdim, double tol) {
    if (tau == 0.0) return 0;
    ensure_krylov(nd, kdim);
    hipStream_t s = nd->ctx->stream;
    const int64_t n = nd->size, ld = nd->size;
    const int nb = nd->nblk;
    double2* V = nd->V.p;
    double2* w = nd->w.p;
    int64_t napply = 0;
    double s_done = 0.0, s_next = 1.0;
    std::vector<cplx> Hm(static_cast<size_t>((kdim + 1) * kdim)), S, F;
    std::vector<double> hv(4 * ttnd::MAXV);
    int guard_iter = 0;
    while (s_done < 1.0 && guard_iter++ < 4096) {
        double s_step = std::min(s_next, 1.0 - s_done);
        hip_check(ttnd_znorm(nd->A.p, n, nd->part.p, nb, s), "norm");
        const double beta = std::sqrt(sum_partials(nd, nb));
        check(std::isfinite(beta), ERR_NUMERIC, "node expv: non-finite vector");
        if (beta == 0.0) return napply;
        hip_check(ttnd_zscale(V, nd->A.p, 1.0 / beta, 0.0, n, nb, s), "scale");
        std::fill(Hm.begin(), Hm.end(), cplx(0.0, 0.0));
        int m = 0;
        double hn = 0.0, err = 0.0;
        bool found = false;
        auto small_exp = [&](int mm, double sf) {
            S.assign(static_cast<size_t>(mm) * mm, cplx(0.0, 0.0));
            F.assign(static_cast<size_t>(mm) * mm, cplx(0.0, 0.0));
            const cplx f = coeff * (tau * sf);
            for (int i = 0; i < mm; ++i)
                for (int j = 0; j < mm; ++j) S[i * mm + j] = f * Hm[i * kdim + j];
            small_expm(S.data(), mm, F.data());
            return hn * std::abs(F[(mm - 1) * mm]);
        };
        for (int j = 0; j < kdim; ++j) {
            heff(nd, V + j * ld, w);
            ++napply;
            // CGS2: both passes on the device (coefficients of pass p at h + 2 MAXV p), then one
            // read-back of both coefficient sets and the partial norms of the final w
            for (int pass = 0; pass < 2; ++pass) {
                double* hp = nd->h.p + 2 * ttnd::MAXV * pass;
                hip_check(ttnd_zdots(V, ld, j + 1, w, n, nd->part.p, nb, hp, s), "dots");
                hip_check(ttnd_zorth(V, ld, j + 1, hp, w, n, nd->part.p, nb, s), "orth");
            }
            hip_check(hipMemcpyAsync(hv.data(), nd->h.p, 4 * ttnd::MAXV * sizeof(double), hipMemcpyDeviceToHost, s),
                      "h");
            const double nn = sum_partials(nd, nb);  // synchronises the stream
            for (int pass = 0; pass < 2; ++pass)
                for (int i = 0; i <= j; ++i)
                    Hm[i * kdim + j] += cplx(hv[2 * ttnd::MAXV * pass + 2 * i], hv[2 * ttnd::MAXV * pass + 2 * i + 1]);
            hn = std::sqrt(nn);
            check(std::isfinite(hn), ERR_NUMERIC, "node expv: non-finite Krylov vector");
            if (j + 1 < kdim) Hm[(j + 1) * kdim + j] = hn;
            m = j + 1;
            double lead = hn;
            for (int i = 0; i < j; ++i)
                lead *= std::abs(coeff) * std::abs(tau * s_step) * std::abs(Hm[(i + 1) * kdim + i]) / (i + 1);
            if (lead > 1e3 * tol && j + 1 < kdim) {
                hip_check(ttnd_zscale(V + (j + 1) * ld, w, 1.0 / hn, 0.0, n, nb, s), "scale");
                continue;
            }
            err = small_exp(m, s_step);
            if (err <= tol) {
                found = true;
                break;
            }
            if (j + 1 < kdim) hip_check(ttnd_zscale(V + (j + 1) * ld, w, 1.0 / hn, 0.0, n, nb, s), "scale");
        }
        for (int k = 0; !found && k < 8; ++k) {
            s_step *= step_factor(tol, err, m, 0.0625, 0.9);
            err = small_exp(m, s_step);
            found = err <= tol;
        }
        for (int k = 0; !found && k < 60; ++k) {
            s_step *= 0.5;
            err = small_exp(m, s_step);
            found = err <= tol;
        }
        check(found, ERR_NUMERIC, "node expv: Krylov step did not reach the tolerance");
        for (int i = 0; i < m; ++i) {
            hv[2 * i] = F[i * m].real();
            hv[2 * i + 1] = F[i * m].imag();
        }
        hip_check(hipMemcpyAsync(nd->h.p, hv.data(), 2 * m * sizeof(double), hipMemcpyHostToDevice, s), "coef");
        hip_check(ttnd_zcombine(nd->A.p, V, ld, m, nd->h.p, beta, n, nb, s), "combine");
        s_done += s_step;
        s_next = s_step * step_factor(tol, err, m, 1.0, 2.0);
    }
    return napply;
}

void upload(const double* src, DevBuf<double2>& dst, int64_t count, hipStream_t s) {
    dst.alloc(static_cast<size_t>(count));
    hip_check(hipMemcpyAsync(dst.p, src, count * sizeof(double2), hipMemcpyHostToDevice, s), "upload");
}

struct Timer {
    ttn_node* nd;
    explicit Timer(ttn_node* n) : nd(n) { hip_check(hipEventRecord(nd->ctx->ev0, nd->ctx->stream), "event"); }
    void stop() {
        hip_check(hipEventRecord(nd->ctx->ev1, nd->ctx->stream), "event");
        hip_check(hipEventSynchronize(nd->ctx->ev1), "sync");
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, nd->ctx->ev0, nd->ctx->ev1), "elapsed");
        nd->last_ms = ms;
    }
};

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[MAX_DEVICES] = {};  // device address of `scratch`, per device

}  // namespace

void ttnrt_noalloc_setup(ttn_ctx* c) {
    check(c->device >= 0 && c->device < MAX_DEVICES, ERR_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[c->device] == nullptr) {
        void* p = nullptr;
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        check(p != nullptr, ERR_DEVICE, "hipGetSymbolAddress(scratch) returned a null address");
        g_scratch_ptr[c->device] = p;
    }
    c->noalloc_scratch = g_scratch_ptr[c->device];
}

extern "C" {

int ttn_node_create(ttn_ctx* ctx, int64_t d0, int64_t d1, int64_t r, int64_t nterms, const double* coeff,
                    const double* op0, const double* op1, const double* mf, ttn_node** out) {
    return guard(ctx, [&] {
        check(ctx != nullptr && out != nullptr, ERR_INVALID, "null argument");
        check(d0 >= 1 && d1 >= 1 && r >= 1 && nterms >= 1, ERR_INVALID, "node dimensions must be positive");
        check(op0 && op1 && mf && coeff, ERR_INVALID, "null operator array");
        check(d0 * d1 * r <= (int64_t(1) << 26), ERR_INVALID, "node tensor too large");
        auto nd = std::make_unique<ttn_node>();
        nd->ctx = ctx;
        nd->d0 = d0;
        nd->d1 = d1;
        nd->r = r;
        nd->nt = nterms;
        nd->D = d0 * d1;
        nd->size = d0 * d1 * r;
        hipStream_t s = ctx->stream;
        upload(op0, nd->op0, nterms * d0 * d0, s);
        upload(op1, nd->op1, nterms * d1 * d1, s);
        upload(mf, nd->mf, nterms * r * r, s);
        // coeff_t MF_t^T, stacked over terms: (p, j) of term t = coeff_t MF_t(j, p)
        std::vector<double2> mft(static_cast<size_t>(nterms * r * r));
        for (int64_t t = 0; t < nterms; ++t) {
            const cplx c(coeff[2 * t], coeff[2 * t + 1]);
            for (int64_t p = 0; p < r; ++p)
                for (int64_t j = 0; j < r; ++j) {
                    const int64_t src = 2 * (t * r * r + j * r + p);
                    const cplx v = c * cplx(mf[src], mf[src + 1]);
                    mft[t * r * r + p * r + j] = make_double2(v.real(), v.imag());
                }
        }
        nd->mft.alloc(mft.size());
        hip_check(hipMemcpyAsync(nd->mft.p, mft.data(), mft.size() * sizeof(double2), hipMemcpyHostToDevice, s),
                  "upload");
        nd->A.alloc(nd->size);
        nd->out.alloc(nd->size);
        nd->w.alloc(nd->size);
        nd->Y.alloc(static_cast<size_t>(nterms * nd->size));
        nd->Z.alloc(static_cast<size_t>(nterms * nd->size));
        // workspace of the K splits: the largest need over every GEMM the node runs
        // (M, N, batch, K of: mode-1 and mode-0 contractions, the mean-field sum, the
        // per-term mean-field products and both mean-field contractions)
        const int64_t shapes[][4] = {{d1, r, d0 * nterms, d1},        {d0, d1 * r, nterms, d0},
                                     {nd->D, r, 1, nterms * r},       {nd->D, r, nterms, r},
                                     {d0, d0, nterms, nd->size / d0}, {d1, d1, nterms, nd->size / d1}};
        int64_t ws = 1;
        for (const auto& g : shapes) {
            const int ns = split_for(g[0], g[1], g[2], g[3]);
            if (ns > 1) ws = std::max<int64_t>(ws, static_cast<int64_t>(ns) * g[0] * g[1] * g[2]);
        }
        nd->ws.alloc(static_cast<size_t>(ws));
        nd->part.alloc(static_cast<size_t>(nd->nblk) * 2 * ttnd::MAXV);
        nd->h.alloc(4 * ttnd::MAXV);
        hip_check(hipMemsetAsync(nd->A.p, 0, nd->size * sizeof(double2), s), "memset");
        hip_check(hipStreamSynchronize(s), "sync");
        *out = nd.release();
    });
}

int ttn_node_destroy(ttn_node* nd) {
    delete nd;
    return TTN_OK;
}

int ttn_node_set(ttn_node* nd, const double* A) {
    return guard(nd->ctx, [&] {
        hip_check(hipMemcpyAsync(nd->A.p, A, nd->size * sizeof(double2), hipMemcpyHostToDevice, nd->ctx->stream),
                  "set");
        hip_check(hipStreamSynchronize(nd->ctx->stream), "sync");
    });
}

int ttn_node_get(ttn_node* nd, double* A) {
    return guard(nd->ctx, [&] {
        hip_check(hipMemcpy(A, nd->A.p, nd->size * sizeof(double2), hipMemcpyDeviceToHost), "get");
    });
}

int ttn_node_get_result(ttn_node* nd, double* out) {
    return guard(nd->ctx, [&] {
        hip_check(hipMemcpy(out, nd->out.p, nd->size * sizeof(double2), hipMemcpyDeviceToHost), "get");
    });
}

int ttn_node_heff(ttn_node* nd, int64_t reps) {
    return guard(nd->ctx, [&] {
        check(reps >= 1, ERR_INVALID, "reps must be positive");
        Timer tm(nd);
        for (int64_t k = 0; k < reps; ++k) heff(nd, nd->A.p, nd->out.p);
        tm.stop();
        nd->last_flops = heff_flops(nd) * static_cast<double>(reps);
    });
}

int ttn_node_mf(ttn_node* nd, int64_t child, int64_t reps, double* out) {
    return guard(nd->ctx, [&] {
        check(child == 0 || child == 1, ERR_INVALID, "child must be 0 or 1");
        check(reps >= 1, ERR_INVALID, "reps must be positive");
        const int64_t dc = child == 0 ? nd->d0 : nd->d1;
        DevBuf<double2> res;
        res.alloc(static_cast<size_t>(nd->nt * dc * dc));
        const bool fused = use_fused_mf(nd, static_cast<int>(child));
        Timer tm(nd);
        for (int64_t k = 0; k < reps; ++k) mean_field(nd, static_cast<int>(child), res.p, fused);
        tm.stop();
        nd->last_flops = static_cast<double>(reps) * 8.0 * static_cast<double>(nd->nt) *
                         static_cast<double>(nd->size) * static_cast<double>((child == 0 ? nd->d1 : nd->d0) + nd->r + dc);
        if (out)
            hip_check(hipMemcpy(out, res.p, nd->nt * dc * dc * sizeof(double2), hipMemcpyDeviceToHost), "mf");
    });
}

int ttn_node_expv(ttn_node* nd, double cr, double ci, double tau, int64_t kdim, double tol, int64_t* napply) {
    return guard(nd->ctx, [&] {
        check(kdim >= 1 && kdim < ttnd::MAXV, ERR_INVALID, "krylov_dim must be in [1, 31]");
        check(tol > 0.0, ERR_INVALID, "tolerance must be positive");
        Timer tm(nd);
        const int64_t na = node_expv(nd, cplx(cr, ci), tau, static_cast<int>(kdim), tol);
        tm.stop();
        if (napply) *napply = na;
        nd->last_flops = static_cast<double>(na) * heff_flops(nd);
    });
}

int ttn_node_heff_noalloc(ttn_ctx* ctx, int64_t d0, int64_t d1, int64_t r, int64_t nterms, const double* coeff,
                          const double* op0, const double* op1, const double* mf, const double* A, double* res,
                          void* stream) {
    return guard(ctx, [&] {
        check(ctx != nullptr, ERR_INVALID, "null context");
        check(d0 >= 1 && d1 >= 1 && r >= 1 && nterms >= 1, ERR_INVALID, "node dimensions must be positive");
        check(coeff && op0 && op1 && mf && A && res, ERR_INVALID, "null device pointer");
        check(ctx->noalloc_scratch != nullptr, ERR_DEVICE, "the context has no scratch address");
        check(d0 <= (int64_t(1) << 20) && d1 <= (int64_t(1) << 20) && r <= (int64_t(1) << 20) &&
                  nterms <= (int64_t(1) << 20),
              ERR_CAPACITY, "ttn_node_heff_noalloc: dimension out of range");
        const int64_t cap = NOALLOC_SCRATCH_BYTES / static_cast<int64_t>(sizeof(double2));
        const int64_t size = d0 * d1 * r, rr = r * r;
        // scratch layout (entries, each slice on a 256-byte boundary): mean-field factors
        // (nterms r^2), Y and Z for a chunk of tc terms, then the rest as K-split workspace
        const int64_t off_y = align16(nterms * rr);
        auto end_of = [&](int64_t t) { return align16(off_y + t * size) + t * size; };
        int64_t tc = std::min<int64_t>(nterms, std::max<int64_t>(0, (cap - off_y) / (2 * size)));
        while (tc > 0 && end_of(tc) > cap) --tc;
        check(tc >= 1, ERR_CAPACITY,
              "ttn_node_heff_noalloc: the node needs " + std::to_string(end_of(1) * 16) +
                  " bytes of scratch (mean-field factors and one term's intermediates); the maximum is " +
                  std::to_string(NOALLOC_SCRATCH_BYTES));
        double2* base = static_cast<double2*>(ctx->noalloc_scratch);
        double2* mft = base;
        double2* Y = base + off_y;
        const int64_t off_z = align16(off_y + tc * size);
        double2* Z = base + off_z;
        const int64_t off_ws = std::min(cap, align16(off_z + tc * size));
        double2* ws = base + off_ws;
        const int64_t ws_n = cap - off_ws;
        hipStream_t s = static_cast<hipStream_t>(stream);
        hip_check(ttnd_zmft(reinterpret_cast<const double2*>(coeff), reinterpret_cast<const double2*>(mf), nterms, r,
                            mft, s),
                  "mean-field factors");
        for (int64_t t0 = 0; t0 < nterms; t0 += tc) {
            const int64_t cnt = std::min(tc, nterms - t0);
            NodeView v{d0, d1, r, cnt, d0 * d1, size,
                       reinterpret_cast<const double2*>(op0) + t0 * d0 * d0,
                       reinterpret_cast<const double2*>(op1) + t0 * d1 * d1, mft + t0 * rr,
                       Y, Z, ws, ws_n, true, s, ctx, false};
            stage_mode1(v, reinterpret_cast<const double2*>(A));
            stage_mode0(v, v.Y, v.size, v.Z);
            stage_mf_sum(v, reinterpret_cast<double2*>(res), t0 > 0);
        }
    });
}

int ttn_ctx_set_autotune(ttn_ctx* ctx, int enabled) {
    return guard(ctx, [&] { ctx->autotune = enabled != 0; });
}

int ttn_ctx_set_heff_form(ttn_ctx* ctx, int form) {
    return guard(ctx, [&] {
        check(form >= 0 && form <= 2, ERR_INVALID, "form must be 0 (autotuned), 1 (staged) or 2 (fused)");
        ctx->heff_form = form;
    });
}

int ttn_autotune_table(const ttn_ctx* ctx, int64_t* n, int64_t* entries, double* ms) {
    if (!ctx || !n) return TTN_ERR_INVALID;
    if (entries || ms) {
        int64_t i = 0;
        for (const auto& kv : ctx->tune) {
            if (i >= *n) break;
            if (entries) {
                for (int q = 0; q < 4; ++q) entries[7 * i + q] = kv.first[q];
                entries[7 * i + 4] = kv.first[4];
                entries[7 * i + 5] = static_cast<int64_t>(kv.second[0]);
                entries[7 * i + 6] = static_cast<int64_t>(kv.second[1]);
            }
            if (ms) ms[i] = kv.second[2];
            ++i;
        }
    }
    *n = static_cast<int64_t>(ctx->tune.size());
    return TTN_OK;
}

int ttn_node_last(const ttn_node* nd, double* ms, double* flops) {
    if (ms) *ms = nd->last_ms;
    if (flops) *flops = nd->last_flops;
    return TTN_OK;
}

}  // extern "C"
