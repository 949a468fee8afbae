// Team-level complex128 primitives for the persistent TDVP sweep.
//
// A team is a set of workgroups that cooperate on one TTNS instance.  All members sit on
// one XCD, so they share its L2: a member publishes its global stores by waiting for
// them to complete (the vector L1 is write-through) and reaches the others through an
// atomic counter in L2; data other members wrote is read with coherent loads (ldc) that
// bypass the L1.  A team of one is a single workgroup and its barrier is __syncthreads().
//
// Every primitive is executed by all threads of all members and spreads rows, or 16x16
// output tiles, over them.  Primitives whose results other threads read end with a team
// barrier, so callers need none of their own between two primitives.  Reductions are
// summed member by member in rank order, so every member obtains bit-identical totals
// and all members take the same control-flow decisions.
#pragma once

#include <hip/hip_runtime.h>

#include "desc.hpp"
#include "zgemm.hpp"

namespace ttnd {

constexpr int NT = 256;          // threads per workgroup (4 wave64)
constexpr int NWAVE = NT / 64;

// ---------------------------------------------------------------- complex helpers
__device__ __forceinline__ double2 cz() { return make_double2(0.0, 0.0); }
__device__ __forceinline__ double2 cmk(double a, double b) { return make_double2(a, b); }
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ __forceinline__ double cabs2(double2 a) { return a.x * a.x + a.y * a.y; }
__device__ __forceinline__ double cabsv(double2 a) { return hypot(a.x, a.y); }
// a / b with Smith's range reduction
__device__ __forceinline__ double2 cdiv(double2 a, double2 b) {
    if (fabs(b.x) >= fabs(b.y)) {
        const double r = b.y / b.x, den = b.x + b.y * r;
        return make_double2((a.x + a.y * r) / den, (a.y - a.x * r) / den);
    }
    const double r = b.x / b.y, den = b.x * r + b.y;
    return make_double2((a.x * r + a.y) / den, (a.y * r - a.x) / den);
}

// ---------------------------------------------------------------- workgroup state
struct Shared {
    double red[NWAVE * 2 * (MAXK + 1)];
    double tsum[2 * MAXK];   // team totals of block reductions
    double2 qtau[64];        // Householder tau (up to 64 columns)
    double2 qw[64];          // Householder trailing dot products
    // Krylov small matrices
    double2 Hm[(MAXK + 1) * MAXK];
    double2 S[MAXK * MAXK], F[MAXK * MAXK];
    double2 E[8][MAXK * MAXK];
    double2 lu_f[MAXK];
    int fail;                // failure record, identical on all members after a barrier
    double hb[2][2 * MAXK];  // Gram-Schmidt coefficients of the two CGS passes
    int expm_sc;             // squarings of the current small exponential
};

// The workgroup's shared area.  Declared once at namespace scope and named directly by the
// routines that use it, so that every access compiles to an LDS instruction (through a
// reference passed between non-inlined functions it would be a flat access).
__shared__ Shared g_sh;

// v from another lane of the same row of 16 by a DPP lane move (a VALU operation, no LDS
// round trip), one 32-bit half at a time
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(static_cast<uint32_t>(b)), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(static_cast<uint32_t>(b >> 32)), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, (static_cast<uint64_t>(static_cast<uint32_t>(hi)) << 32) | static_cast<uint32_t>(lo));
}

// Sum over the wave, every lane receives it.  Butterfly: lane ^ 1 and lane ^ 2 (quad
// permutations), then the other quad of the half-row and the other half of the row (half-row and
// row mirrors: after the previous steps every lane of a quad, then of a half-row, holds the same
// value, so a mirror pairs the same partial sums as lane ^ 4 and lane ^ 8), then lane ^ 16 and
// lane ^ 32 across rows through ds_bpermute.
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_move<0xB1>(v);   // quad_perm [1, 0, 3, 2]
    v += dpp_move<0x4E>(v);   // quad_perm [2, 3, 0, 1]
    v += dpp_move<0x141>(v);  // row_half_mirror
    v += dpp_move<0x140>(v);  // row_mirror
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    return v;
}

// sum of K doubles over the workgroup; every thread receives the totals in out[]
template <int K>
__device__ void block_sum(double (&v)[K], double (&out)[K]) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const double s = wave_sum(v[k]);
        if (lane == 0) g_sh.red[wv * K + k] = s;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < NWAVE; ++w) s += g_sh.red[w * K + k];
        out[k] = s;
    }
    __syncthreads();
}

// sum of the first cnt (uniform, <= K) of K doubles over the workgroup, totals left in the LDS
// array dst[cnt]; the others are skipped
template <int K>
__device__ void block_sum_lds(double (&v)[K], double* dst, int cnt = K) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (k < cnt) {
            const double s = wave_sum(v[k]);
            if (lane == 0) g_sh.red[wv * K + k] = s;
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < cnt; k += NT) {
        double s = 0.0;
        for (int w = 0; w < NWAVE; ++w) s += g_sh.red[w * K + k];
        dst[k] = s;
    }
    __syncthreads();
}

// maximum over the workgroup; every thread receives it
__device__ double block_max(double v) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    if (lane == 0) g_sh.red[wv] = v;
    __syncthreads();
    double m = g_sh.red[0];
#pragma unroll
    for (int w = 1; w < NWAVE; ++w) m = fmax(m, g_sh.red[w]);
    __syncthreads();
    return m;
}

// ---------------------------------------------------------------- teams
struct Team {
    TeamWords* w;       // null for a team of one
    int rank, size;
    unsigned gen;       // barriers completed by this member
    int pbuf;           // which of the two partial-sum buffers the next reduction uses
    double* part;       // 2 x TEAM_MAX x PART_K doubles
    double2* tiles;     // TEAM_MAX x TILE_MAX partial products

    __device__ int64_t tid() const { return static_cast<int64_t>(rank) * NT + threadIdx.x; }
    __device__ int64_t nthr() const { return static_cast<int64_t>(size) * NT; }
    __device__ int64_t gwave() const { return static_cast<int64_t>(rank) * NWAVE + (threadIdx.x >> 6); }
    __device__ int64_t nwave() const { return static_cast<int64_t>(size) * NWAVE; }
    // the leader alone, as a team of one
    __device__ Team single() const {
        Team s = *this;
        s.w = nullptr;
        s.rank = 0;
        s.size = 1;
        return s;
    }

    // Team barrier.  Release: each wave waits for its global stores to complete (they are
    // then in the shared L2).  Arrival on the L2 counter; the last arrival publishes the
    // generation.  No acquire step is needed: shared data is only read with coherent
    // (L1-bypassing) loads, see ldc().
    // A barrier that does not complete within TEAM_TIMEOUT records FAIL_TEAM_TIMEOUT and
    // returns, so a broken team ends instead of hanging.  The team-wide failure record is
    // copied into every member's LDS flag.  A member that has timed out no longer waits,
    // so a broken team drains instead of timing out barrier after barrier.
    __device__ void sync() {
        if (size == 1 || g_sh.fail == FAIL_TEAM_TIMEOUT) {
            // the workgroup barrier alone does not wait for this wave's global stores, and the
            // other waves read them with loads that bypass the L1
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            return;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (threadIdx.x == 0) {
            const unsigned target = gen + 1;
            const unsigned old = __hip_atomic_fetch_add(&w->count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (old + 1 == target * static_cast<unsigned>(size)) {
                __hip_atomic_store(&w->gen, target, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                const uint64_t t0 = wall_clock64();
                while (__hip_atomic_load(&w->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
                    __builtin_amdgcn_s_sleep(1);
                    if (wall_clock64() - t0 > TEAM_TIMEOUT) {
                        g_sh.fail = FAIL_TEAM_TIMEOUT;
                        break;
                    }
                }
            }
            const int f = __hip_atomic_load(&w->fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (f != 0 && g_sh.fail == 0) g_sh.fail = f;
        }
        ++gen;
        __syncthreads();
    }

    // Roll call before anything is touched: decides once, in the outcome word, whether the
    // whole team assembled.  The last arrival claims FORMED (0 -> 1) before it publishes the
    // first generation; a member that gives up waiting claims ABORTED (0 -> 2).  Whichever
    // claim lands first wins, and every member reads the same decided word afterwards, so
    // either all members proceed or none does.  Returns true when the team formed.
    __device__ bool roll_call() {
        if (size == 1) return true;
        if (threadIdx.x == 0) {
            const unsigned old = __hip_atomic_fetch_add(&w->count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (old + 1 == static_cast<unsigned>(size)) {
                int expected = 0;
                __hip_atomic_compare_exchange_strong(&w->outcome, &expected, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&w->gen, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                const uint64_t t0 = wall_clock64();
                while (__hip_atomic_load(&w->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < 1u) {
                    __builtin_amdgcn_s_sleep(1);
                    if (wall_clock64() - t0 > TEAM_TIMEOUT) {
                        int expected = 0;
                        __hip_atomic_compare_exchange_strong(&w->outcome, &expected, 2, __ATOMIC_RELAXED,
                                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        break;
                    }
                }
            }
            // decided here: by the last arrival before it published the generation, or by
            // this member's own claim (or the earlier one that beat it)
            g_sh.fail = __hip_atomic_load(&w->outcome, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 1
                            ? 0
                            : static_cast<int>(FAIL_TEAM_FORM);
        }
        ++gen;
        __syncthreads();
        return g_sh.fail == 0;
    }

    // record a failure on this member and in the team word (read by the others at their
    // next barrier)
    __device__ void fail(int code) {
        g_sh.fail = code;
        if (w != nullptr && threadIdx.x == 0)
            __hip_atomic_store(&w->fail, code, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
};

// Every member holds `count` partial values in the LDS array v (complete: the caller has
// passed a workgroup barrier); afterwards v holds the team totals, summed in rank order.
__device__ void team_reduce_lds(Team& T, double* v, int count) {
    if (T.size == 1) return;
    double* mine = T.part + (static_cast<int64_t>(T.pbuf) * TEAM_MAX + T.rank) * PART_K;
    for (int e = threadIdx.x; e < count; e += NT) mine[e] = v[e];
    T.sync();
    const double* all = T.part + static_cast<int64_t>(T.pbuf) * TEAM_MAX * PART_K;
    for (int e = threadIdx.x; e < count; e += NT) {
        // the members' partials are loaded 16 at a time (independent loads in flight: two
        // round trips for a team of a whole XCD) and added in rank order
        double s = 0.0;
        for (int r0 = 0; r0 < T.size; r0 += 16) {
            double p[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) p[u] = r0 + u < T.size ? ldc(all + (r0 + u) * PART_K + e) : 0.0;
#pragma unroll
            for (int u = 0; u < 16; ++u)
                if (r0 + u < T.size) s += p[u];
        }
        v[e] = s;
    }
    __syncthreads();
    T.pbuf ^= 1;
}

// team sum of K doubles; every thread receives the totals in out[]
template <int K>
__device__ void team_sum(Team& T, double (&v)[K], double (&out)[K]) {
    if (T.size == 1) {
        block_sum<K>(v, out);
        return;
    }
    block_sum_lds<K>(v, g_sh.tsum);
    team_reduce_lds(T, g_sh.tsum, K);
#pragma unroll
    for (int k = 0; k < K; ++k) out[k] = g_sh.tsum[k];
    __syncthreads();
}

// team sum of the first cnt (uniform, <= K) of K doubles, totals left in the LDS array dst[cnt]
template <int K>
__device__ void team_sum_lds(Team& T, double (&v)[K], double* dst, int cnt = K) {
    block_sum_lds<K>(v, dst, cnt);
    team_reduce_lds(T, dst, cnt);
}

__device__ double team_max(Team& T, double v) {
    double m = block_max(v);
    if (T.size == 1) return m;
    if (threadIdx.x == 0) g_sh.tsum[0] = m;
    __syncthreads();
    // maxima are exact, so a second reduction buffer is not needed for determinism
    double* mine = T.part + (static_cast<int64_t>(T.pbuf) * TEAM_MAX + T.rank) * PART_K;
    if (threadIdx.x == 0) mine[0] = m;
    T.sync();
    const double* all = T.part + static_cast<int64_t>(T.pbuf) * TEAM_MAX * PART_K;
    m = ldc(all);
    for (int r0 = 1; r0 < T.size; r0 += 8) {
        double p[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) p[u] = r0 + u < T.size ? ldc(all + (r0 + u) * PART_K) : m;
#pragma unroll
        for (int u = 0; u < 8; ++u) m = fmax(m, p[u]);
    }
    __syncthreads();
    T.pbuf ^= 1;
    return m;
}

__device__ double team_norm2(Team& T, const double2* x, int64_t n) {
    double p[1] = {0.0}, o[1];
    const int64_t st = T.nthr();
    for (int64_t b = T.tid(); b < n; b += 4 * st) {
        double2 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = (b + u * st < n) ? ldc(x + b + u * st) : cz();
#pragma unroll
        for (int u = 0; u < 4; ++u) p[0] += cabs2(v[u]);
    }
    team_sum<1>(T, p, o);
    return o[0];
}

// ---------------------------------------------------------------- elementwise
// One workgroup has a single wave per SIMD, so latency is hidden by issuing the loads
// of UNR consecutive strides before using any of them.
constexpr int UNR = 4;

__device__ void t_zero(Team& T, double2* y, int64_t n) {
    for (int64_t i = T.tid(); i < n; i += T.nthr()) y[i] = cz();
    T.sync();
}
__device__ void t_copy(Team& T, double2* __restrict__ y, const double2* __restrict__ x, int64_t n) {
    const int64_t st = T.nthr();
    for (int64_t b = T.tid(); b < n; b += UNR * st) {
        double2 v[UNR];
#pragma unroll
        for (int u = 0; u < UNR; ++u) v[u] = (b + u * st < n) ? ldc(x + b + u * st) : cz();
#pragma unroll
        for (int u = 0; u < UNR; ++u)
            if (b + u * st < n) y[b + u * st] = v[u];
    }
    T.sync();
}
// y += a * x
__device__ void t_axpy(Team& T, double2* __restrict__ y, double2 a, const double2* __restrict__ x, int64_t n) {
    const int64_t st = T.nthr();
    for (int64_t b = T.tid(); b < n; b += UNR * st) {
        double2 v[UNR], o[UNR];
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            const bool ok = b + u * st < n;
            v[u] = ok ? ldc(x + b + u * st) : cz();
            o[u] = ok ? ldc(y + b + u * st) : cz();
        }
#pragma unroll
        for (int u = 0; u < UNR; ++u)
            if (b + u * st < n) y[b + u * st] = cadd(o[u], cmul(a, v[u]));
    }
    T.sync();
}
// y = a * x
__device__ void t_scale_to(Team& T, double2* __restrict__ y, double2 a, const double2* __restrict__ x, int64_t n) {
    const int64_t st = T.nthr();
    for (int64_t b = T.tid(); b < n; b += UNR * st) {
        double2 v[UNR];
#pragma unroll
        for (int u = 0; u < UNR; ++u) v[u] = (b + u * st < n) ? ldc(x + b + u * st) : cz();
#pragma unroll
        for (int u = 0; u < UNR; ++u)
            if (b + u * st < n) y[b + u * st] = cmul(a, v[u]);
    }
    T.sync();
}

// ---------------------------------------------------------------- contractions (FP64 MFMA)
// dst(l,i,r) (=|+=) alpha * sum_j op(i,j) src(l,j,r) on the (L, d, R) view.
// Reference: contract(op, 1, A3, 1) in kronecker_product_operator_helper.hpp:292-303.
// One 16x16 tile of (i, r) per wave work item; op rows come from L1.
__device__ void t_mode_contract(Team& T, const double2* op, int64_t L, int d, int64_t R, const double2* src,
                                double2* dst, double2 alpha, bool accumulate) {
    const int mt = (d + 15) >> 4;
    const int64_t nt = (R + 15) >> 4;
    const int64_t items = L * mt * nt;
    for (int64_t it = T.gwave(); it < items; it += T.nwave()) {
        const int64_t n = it % nt;
        const int m = static_cast<int>((it / nt) % mt);
        const int64_t l = it / (nt * mt);
        ZAcc acc;
        zacc_zero(acc);
        const int mv = min(16, d - 16 * m);
        const int nv = static_cast<int>(R - 16 * n < 16 ? R - 16 * n : 16);
        const double2* s = src + l * d * R + 16 * n;
        ztile<false>(acc, op + static_cast<int64_t>(16 * m) * d, d, 1, mv, s, R, 1, nv, d);
        ztile_store(acc, dst + l * d * R + static_cast<int64_t>(16 * m) * R + 16 * n, R, 1, mv, nv, alpha,
                    accumulate);
    }
    T.sync();
}

// dst (=|+=) alpha * X M^T, X: rows x r, M: r x r  (v * trans(mf))
__device__ void t_mul_transposed(Team& T, const double2* X, int64_t rows, int r, const double2* M, double2* dst,
                                 double2 alpha, bool accumulate) {
    const int64_t mt = (rows + 15) >> 4;
    const int nt = (r + 15) >> 4;
    const int64_t items = mt * nt;
    for (int64_t it = T.gwave(); it < items; it += T.nwave()) {
        const int n = static_cast<int>(it % nt);
        const int64_t m = it / nt;
        ZAcc acc;
        zacc_zero(acc);
        const int mv = static_cast<int>(rows - 16 * m < 16 ? rows - 16 * m : 16);
        const int nv = min(16, r - 16 * n);
        ztile<false>(acc, X + 16 * m * r, r, 1, mv, M + static_cast<int64_t>(16 * n) * r, 1, r, nv, r);
        ztile_store(acc, dst + 16 * m * r + 16 * n, r, 1, mv, nv, alpha, accumulate);
    }
    T.sync();
}

// dst = X M, X: rows x r, M: r x r
__device__ void t_mul(Team& T, const double2* X, int64_t rows, int r, const double2* M, double2* dst) {
    const int64_t mt = (rows + 15) >> 4;
    const int nt = (r + 15) >> 4;
    const int64_t items = mt * nt;
    for (int64_t it = T.gwave(); it < items; it += T.nwave()) {
        const int n = static_cast<int>(it % nt);
        const int64_t m = it / nt;
        ZAcc acc;
        zacc_zero(acc);
        const int mv = static_cast<int>(rows - 16 * m < 16 ? rows - 16 * m : 16);
        const int nv = min(16, r - 16 * n);
        ztile<false>(acc, X + 16 * m * r, r, 1, mv, M + 16 * n, r, 1, nv, r);
        ztile_store(acc, dst + 16 * m * r + 16 * n, r, 1, mv, nv, make_double2(1.0, 0.0), false);
    }
    T.sync();
}

// out(i,j) (=|+=) alpha * sum_{l,r} conj(A(l,i,r)) B(l,j,r) on the (L, d, R) view with
// element (l, i, r) at (l*d + i)*ldr + r, by one workgroup: a long-K product.  With
// several output tiles the waves take whole tiles; with a single tile (d <= 16) they
// split K and the partial tiles are summed through LDS.  out is d x d.
__device__ void wg_outer_contract(const double2* A, const double2* B, int64_t L, int d, int64_t R,
                                  int64_t ldr, double2* out, double2 alpha, bool accumulate) {
    const int wv = threadIdx.x >> 6;
    const int mt = (d + 15) >> 4;
    const int tiles = mt * mt;
    const bool split = tiles < NWAVE;
    for (int t0 = 0; t0 < tiles; t0 += split ? 1 : NWAVE) {
        const int t = split ? t0 : t0 + wv;
        ZAcc acc;
        zacc_zero(acc);
        const int m = t / mt, n = t % mt;
        const int mv = t < tiles ? min(16, d - 16 * m) : 0;
        const int nv = t < tiles ? min(16, d - 16 * n) : 0;
        if (t < tiles) {
            if (R == 1 && ldr == 1) {
                // element (l, i) at l*d + i: K runs over l
                const int64_t k0 = split ? (L * wv) / NWAVE : 0, k1 = split ? (L * (wv + 1)) / NWAVE : L;
                ztile<true>(acc, A + k0 * d + 16 * m, 1, d, mv, B + k0 * d + 16 * n, d, 1, nv, k1 - k0);
            } else if (!split || L >= NWAVE) {
                for (int64_t l = split ? wv : 0; l < L; l += split ? NWAVE : 1) {
                    const double2* a = A + (l * d + 16 * m) * ldr;
                    const double2* b = B + (l * d + 16 * n) * ldr;
                    ztile<true>(acc, a, ldr, 1, mv, b, 1, ldr, nv, R);
                }
            } else {
                // too few slices: split each slice's K range over the waves instead
                const int64_t k0 = (R * wv) / NWAVE, k1 = (R * (wv + 1)) / NWAVE;
                for (int64_t l = 0; l < L; ++l) {
                    const double2* a = A + (l * d + 16 * m) * ldr + k0;
                    const double2* b = B + (l * d + 16 * n) * ldr + k0;
                    ztile<true>(acc, a, ldr, 1, mv, b, 1, ldr, nv, k1 - k0);
                }
            }
        }
        if (!split) {
            if (t < tiles)
                ztile_store(acc, out + static_cast<int64_t>(16 * m) * d + 16 * n, d, 1, mv, nv, alpha, accumulate);
        } else {
            // sum the four partial tiles in LDS (reuses the Krylov scratch E[0..7])
            double* red = reinterpret_cast<double*>(&g_sh.E[0][0]);
            const int lane = threadIdx.x & 63;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                red[(wv * 64 + lane) * 8 + j] = acc.re[j];
                red[(wv * 64 + lane) * 8 + 4 + j] = acc.im[j];
            }
            __syncthreads();
            if (wv == 0) {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    double re = 0.0, im = 0.0;
                    for (int w = 0; w < NWAVE; ++w) {
                        re += red[(w * 64 + lane) * 8 + j];
                        im += red[(w * 64 + lane) * 8 + 4 + j];
                    }
                    acc.re[j] = re;
                    acc.im[j] = im;
                }
                ztile_store(acc, out + static_cast<int64_t>(16 * m) * d + 16 * n, d, 1, mv, nv, alpha, accumulate);
            }
            __syncthreads();
        }
    }
    __syncthreads();
}

// Team form: the K range (slices l, or the r range when there are fewer slices than
// members) is split over the members, each member writes its partial d x d product to
// its tile buffer, and the partials are summed in rank order.
__device__ void t_outer_contract(Team& T, const double2* A, const double2* B, int64_t L, int d, int64_t R,
                                 double2* out, double2 alpha, bool accumulate) {
    if (T.size == 1) {
        wg_outer_contract(A, B, L, d, R, R, out, alpha, accumulate);
        return;
    }
    double2* mine = T.tiles + static_cast<int64_t>(T.rank) * TILE_MAX;
    if (L >= T.size) {
        const int64_t l0 = L * T.rank / T.size, l1 = L * (T.rank + 1) / T.size;
        wg_outer_contract(A + l0 * d * R, B + l0 * d * R, l1 - l0, d, R, R, mine, make_double2(1.0, 0.0),
                          false);
    } else {
        const int64_t r0 = R * T.rank / T.size, r1 = R * (T.rank + 1) / T.size;
        wg_outer_contract(A + r0, B + r0, L, d, r1 - r0, R, mine, make_double2(1.0, 0.0), false);
    }
    T.sync();
    const int64_t dd = static_cast<int64_t>(d) * d;
    for (int64_t e = T.tid(); e < dd; e += T.nthr()) {
        double2 s = cz();
        for (int r0 = 0; r0 < T.size; r0 += 8) {
            double2 p[8];
#pragma unroll
            for (int u = 0; u < 8; ++u)
                p[u] = r0 + u < T.size ? ldc(T.tiles + static_cast<int64_t>(r0 + u) * TILE_MAX + e) : cz();
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (r0 + u < T.size) s = cadd(s, p[u]);
        }
        const double2 v = cmul(alpha, s);
        out[e] = accumulate ? cadd(ldc(out + e), v) : v;
    }
    T.sync();
}

// out = A^H B for A, B: rows x r
__device__ void t_adjoint_mul(Team& T, const double2* A, const double2* B, int64_t rows, int r, double2* out,
                              double2 alpha, bool accumulate) {
    t_outer_contract(T, A, B, rows, r, 1, out, alpha, accumulate);
}

// ---------------------------------------------------------------- primitive leaf operators
// y (=|+=) alpha * op x, x/y: dim x ncols row-major; kind: 0 dense, 1 diag, 2 csr
__device__ void t_prim_apply(Team& T, const DK& k, const double2* base, const int64_t* ints, const double2* x,
                             double2* y, int ncols, double2 alpha = make_double2(1.0, 0.0), bool accumulate = false) {
    const int d = k.dim;
    const double2* mat = base + k.mat;
    const int64_t total = static_cast<int64_t>(d) * ncols;
    for (int64_t idx = T.tid(); idx < total; idx += T.nthr()) {
        const int i = static_cast<int>(idx / ncols), c = static_cast<int>(idx % ncols);
        double2 acc = cz();
        if (k.kind == 0) {
            for (int j = 0; j < d; ++j) acc = cadd(acc, cmul(mat[i * d + j], ldc(x + static_cast<int64_t>(j) * ncols + c)));
        } else if (k.kind == 1) {
            acc = cmul(mat[i], ldc(x + idx));
        } else {
            const int64_t* rowptr = ints + k.idx;
            const int64_t* colind = rowptr + d + 1;
            for (int64_t q = rowptr[i]; q < rowptr[i + 1]; ++q)
                acc = cadd(acc, cmul(mat[q], ldc(x + colind[q] * ncols + c)));
        }
        acc = cmul(alpha, acc);
        y[idx] = accumulate ? cadd(ldc(y + idx), acc) : acc;
    }
    T.sync();
}

// ---------------------------------------------------------------- safe norms (LAPACK dznrm2/dlapy3 roles)
// The plain sum of squares is exact to rounding whenever it lies in [kNrmTiny, kNrmHuge];
// otherwise the column is rescaled by its largest component first.
constexpr double kNrmTiny = 1e-280, kNrmHuge = 1e280;
// safmin = (smallest normal) / (unit roundoff) = 2^-969, as in zlarfg
constexpr double kSafmin = 2.2250738585072014e-308 / 1.1102230246251565e-16;

// ||A(lo:hi, i)||_2 over the rows this member owns, totalled over the team
__device__ double t_col_nrm2(Team& T, const double2* A, int i, int lo, int hi, int n) {
    double p[1] = {0.0}, o[1];
    for (int rr = lo + threadIdx.x; rr < hi; rr += NT) p[0] += cabs2(ldc(A + static_cast<int64_t>(rr) * n + i));
    team_sum<1>(T, p, o);
    if (o[0] >= kNrmTiny && o[0] <= kNrmHuge) return sqrt(o[0]);
    double mx = 0.0;
    for (int rr = lo + threadIdx.x; rr < hi; rr += NT) {
        const double2 v = ldc(A + static_cast<int64_t>(rr) * n + i);
        mx = fmax(mx, fmax(fabs(v.x), fabs(v.y)));
    }
    const double scale = team_max(T, mx);
    if (scale == 0.0) return 0.0;
    p[0] = 0.0;
    for (int rr = lo + threadIdx.x; rr < hi; rr += NT) {
        const double2 v = ldc(A + static_cast<int64_t>(rr) * n + i);
        p[0] += cabs2(make_double2(v.x / scale, v.y / scale));
    }
    team_sum<1>(T, p, o);
    return scale * sqrt(o[0]);
}

// sqrt(a^2 + b^2 + c^2) scaled by the largest magnitude
__device__ __forceinline__ double lapy3(double a, double b, double c) {
    const double w = fmax(fabs(a), fmax(fabs(b), fabs(c)));
    if (w == 0.0) return 0.0;
    const double x = a / w, y = b / w, z = c / w;
    return w * sqrt(x * x + y * y + z * z);
}

// ---------------------------------------------------------------- Householder QR
// A (m x n, row-major, destroyed: holds the reflectors) = Q (m x n) R (n x n).
// LAPACK zgeqr2/zlarfg/zung2r conventions; for m < n, Q gets zero columns and R zero
// rows beyond m (decomposition_engine.hpp:124-158).  Each member owns a fixed block of
// rows for the whole factorisation, so a column costs two team reductions (its norm and
// the trailing dot products) and forming Q one per reflector.
__device__ void t_householder_qr(Team& T, double2* A, int m, int n, double2* Q, double2* Rout) {
    const int k = m < n ? m : n;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int r0 = static_cast<int>(static_cast<int64_t>(m) * T.rank / T.size);
    const int r1 = static_cast<int>(static_cast<int64_t>(m) * (T.rank + 1) / T.size);
    for (int i = 0; i < k; ++i) {
        const int lo = r0 > i + 1 ? r0 : i + 1;  // own rows strictly below the diagonal
        const bool own_i = i >= r0 && i < r1;
        // reflector of column i (zlarfg): beta from a scaled norm; tiny columns are
        // rescaled by 1/safmin until |beta| >= safmin
        double xnorm = t_col_nrm2(T, A, i, lo, r1, n);
        const double2 alpha = ldc(A + static_cast<int64_t>(i) * n + i);  // published by the barrier above
        double ar = alpha.x, ai = alpha.y;
        double2 tau = cz(), scal = cz();
        double beta = ar;
        const bool refl = !(xnorm == 0.0 && ai == 0.0);
        int knt = 0;
        if (refl) {
            beta = -copysign(lapy3(ar, ai, xnorm), ar);
            if (fabs(beta) < kSafmin) {
                do {
                    ++knt;
                    for (int rr = lo + threadIdx.x; rr < r1; rr += NT) {
                        double2* v = A + static_cast<int64_t>(rr) * n + i;
                        *v = cscale(ldc(v), 1.0 / kSafmin);
                    }
                    beta *= 1.0 / kSafmin;
                    ar *= 1.0 / kSafmin;
                    ai *= 1.0 / kSafmin;
                } while (fabs(beta) < kSafmin && knt < 20);
                __syncthreads();
                xnorm = t_col_nrm2(T, A, i, lo, r1, n);
                beta = -copysign(lapy3(ar, ai, xnorm), ar);
            }
            tau = make_double2((beta - ar) / beta, -ai / beta);
            scal = cdiv(make_double2(1.0, 0.0), make_double2(ar - beta, ai));
        }
        if (threadIdx.x == 0) g_sh.qtau[i] = tau;
        if (tau.x != 0.0 || tau.y != 0.0) {
            for (int rr = lo + threadIdx.x; rr < r1; rr += NT) {
                double2* v = A + static_cast<int64_t>(rr) * n + i;
                *v = cmul(ldc(v), scal);
            }
            __syncthreads();
            // w_c = v^H A(i:m, c) with v_i = 1, for c > i: one wave per column over the
            // member's rows, then a team reduction
            for (int c = i + 1 + wv; c < n; c += NWAVE) {
                double2 acc = cz();
                for (int rr = lo + lane; rr < r1; rr += 64)
                    acc = cadd(acc, cmulc(ldc(A + static_cast<int64_t>(rr) * n + i), ldc(A + static_cast<int64_t>(rr) * n + c)));
                acc.x = wave_sum(acc.x);
                acc.y = wave_sum(acc.y);
                if (lane == 0) g_sh.qw[c] = own_i ? cadd(acc, ldc(A + static_cast<int64_t>(i) * n + c)) : acc;
            }
            __syncthreads();
            team_reduce_lds(T, reinterpret_cast<double*>(&g_sh.qw[i + 1]), 2 * (n - i - 1));
            const double2 tc = make_double2(tau.x, -tau.y);
            const int ncol = n - i - 1;
            const int rs = r0 > i ? r0 : i;
            const int64_t tot = static_cast<int64_t>(r1 > rs ? r1 - rs : 0) * ncol;
            for (int64_t e = threadIdx.x; e < tot; e += NT) {
                const int rr = rs + static_cast<int>(e / ncol);
                const int c = i + 1 + static_cast<int>(e % ncol);
                const double2 v = rr == i ? make_double2(1.0, 0.0) : ldc(A + static_cast<int64_t>(rr) * n + i);
                double2* a = A + static_cast<int64_t>(rr) * n + c;
                *a = csub(ldc(a), cmul(tc, cmul(v, g_sh.qw[c])));
            }
        }
        __syncthreads();
        if (refl && own_i) {
            for (int j = 0; j < knt; ++j) beta *= kSafmin;
            if (threadIdx.x == 0) A[static_cast<int64_t>(i) * n + i] = make_double2(beta, 0.0);
        }
        __syncthreads();
    }
    T.sync();
    // R
    for (int64_t e = T.tid(); e < static_cast<int64_t>(n) * n; e += T.nthr()) {
        const int i = static_cast<int>(e / n), j = static_cast<int>(e % n);
        Rout[e] = (i < k && j >= i) ? ldc(A + static_cast<int64_t>(i) * n + j) : cz();
    }
    // Q = H_0 ... H_{k-1} [I_k; 0] on the member's rows
    for (int64_t e = threadIdx.x; e < static_cast<int64_t>(r1 - r0) * n; e += NT) {
        const int i = r0 + static_cast<int>(e / n), j = static_cast<int>(e % n);
        Q[static_cast<int64_t>(i) * n + j] = (i == j && j < k) ? make_double2(1.0, 0.0) : cz();
    }
    __syncthreads();
    for (int i = k - 1; i >= 0; --i) {
        const double2 tau = g_sh.qtau[i];
        if (tau.x == 0.0 && tau.y == 0.0) continue;
        const int lo = r0 > i + 1 ? r0 : i + 1;
        const bool own_i = i >= r0 && i < r1;
        for (int c = wv; c < k; c += NWAVE) {
            double2 acc = cz();
            for (int rr = lo + lane; rr < r1; rr += 64)
                acc = cadd(acc, cmulc(ldc(A + static_cast<int64_t>(rr) * n + i), ldc(Q + static_cast<int64_t>(rr) * n + c)));
            acc.x = wave_sum(acc.x);
            acc.y = wave_sum(acc.y);
            if (lane == 0) g_sh.qw[c] = own_i ? cadd(acc, ldc(Q + static_cast<int64_t>(i) * n + c)) : acc;
        }
        __syncthreads();
        team_reduce_lds(T, reinterpret_cast<double*>(&g_sh.qw[0]), 2 * k);
        const int rs = r0 > i ? r0 : i;
        const int64_t tot = static_cast<int64_t>(r1 > rs ? r1 - rs : 0) * k;
        for (int64_t e = threadIdx.x; e < tot; e += NT) {
            const int rr = rs + static_cast<int>(e / k);
            const int c = static_cast<int>(e % k);
            const double2 v = rr == i ? make_double2(1.0, 0.0) : ldc(A + static_cast<int64_t>(rr) * n + i);
            double2* q = Q + static_cast<int64_t>(rr) * n + c;
            *q = csub(ldc(q), cmul(v, cmul(tau, g_sh.qw[c])));
        }
        __syncthreads();
    }
    T.sync();
}

// ---------------------------------------------------------------- small dense expm
// Within one wavefront the 64 lanes run in lockstep, so LDS hand-offs between them need only a
// wavefront-scope fence instead of a workgroup barrier.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ void argmax_take(double& v, int& idx, double ov, int oi) {
    if (ov > v || (ov == v && oi < idx)) {
        v = ov;
        idx = oi;
    }
}

// max over the wave with the lowest index winning ties (for finite values the result does not
// depend on the order of the pairings); the same butterfly as wave_sum
__device__ __forceinline__ void wave_argmax(double& v, int& idx) {
    argmax_take(v, idx, dpp_move<0xB1>(v), __builtin_amdgcn_update_dpp(0, idx, 0xB1, 0xF, 0xF, false));
    argmax_take(v, idx, dpp_move<0x4E>(v), __builtin_amdgcn_update_dpp(0, idx, 0x4E, 0xF, 0xF, false));
    argmax_take(v, idx, dpp_move<0x141>(v), __builtin_amdgcn_update_dpp(0, idx, 0x141, 0xF, 0xF, false));
    argmax_take(v, idx, dpp_move<0x140>(v), __builtin_amdgcn_update_dpp(0, idx, 0x140, 0xF, 0xF, false));
    argmax_take(v, idx, __shfl_xor(v, 16, 64), __shfl_xor(idx, 16, 64));
    argmax_take(v, idx, __shfl_xor(v, 32, 64), __shfl_xor(idx, 32, 64));
}

// one entry of C = A B per thread (n <= 16), summed over q in ascending order as in
// wv_matmul, with the operands of four q at a time loaded before their products
__device__ __forceinline__ void blk_matmul(const double2* A, const double2* B, double2* C, int n) {
    const int e = threadIdx.x;
    const bool on = e < n * n;
    double2 s = cz();
    if (on) {
        const int i = e / n, j = e % n;
        int q = 0;
        for (; q + 4 <= n; q += 4) {
            double2 a[4], bb[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                a[u] = A[i * n + q + u];
                bb[u] = B[(q + u) * n + j];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) s = cadd(s, cmul(a[u], bb[u]));
        }
        for (; q < n; ++q) s = cadd(s, cmul(A[i * n + q], B[q * n + j]));
    }
    __syncthreads();
    if (on) C[e] = s;
    __syncthreads();
}

// Pade(13) scaling and squaring (Higham 2005, the oracle's expm) by the whole workgroup: the
// matrix products take one entry per thread, the 1-norm and the pivoted LU solve run on wave
// 0 (lockstep lanes, wavefront fences).  Called by every thread of the workgroup.
__device__ void block_expm(const double2* S, double2* Eout, int n) {
    const double b[14] = {64764752532480000.0, 32382376266240000.0, 7771770303897600.0, 1187353796428800.0,
                          129060195264000.0,   10559470521600.0,   670442572800.0,     33522128640.0,
                          1323241920.0,        40840800.0,         960960.0,           16380.0,
                          182.0,               1.0};
    const double theta13 = 5.371920351148152;
    const int lane = threadIdx.x & 63, nn = n * n;
    const bool w0 = (threadIdx.x >> 6) == 0;
    double2 *A = g_sh.E[0], *A2 = g_sh.E[1], *A4 = g_sh.E[2], *A6 = g_sh.E[3], *T1 = g_sh.E[4], *T2 = g_sh.E[5], *U = g_sh.E[6],
            *V = g_sh.E[7];
    if (w0) {
        // 1-norm: lane j sums column j, then a wave max
        double colsum = 0.0;
        if (lane < n)
            for (int i = 0; i < n; ++i) colsum += cabsv(S[i * n + lane]);
        int dummy = lane;
        wave_argmax(colsum, dummy);
        const double norm1 = colsum;
        int sc = 0;
        if (norm1 > theta13) sc = static_cast<int>(ceil(log2(norm1 / theta13)));
        if (lane == 0) g_sh.expm_sc = sc;
    }
    __syncthreads();
    const int sc = g_sh.expm_sc;
    const double scale = ldexp(1.0, -sc);
    for (int e = threadIdx.x; e < nn; e += NT) A[e] = cscale(S[e], scale);
    __syncthreads();
    blk_matmul(A, A, A2, n);
    blk_matmul(A2, A2, A4, n);
    blk_matmul(A4, A2, A6, n);
    for (int e = threadIdx.x; e < nn; e += NT)
        T1[e] = cadd(cadd(cscale(A6[e], b[13]), cscale(A4[e], b[11])), cscale(A2[e], b[9]));
    __syncthreads();
    blk_matmul(A6, T1, T2, n);
    for (int e = threadIdx.x; e < nn; e += NT) {
        double2 v = cadd(T2[e], cadd(cadd(cscale(A6[e], b[7]), cscale(A4[e], b[5])), cscale(A2[e], b[3])));
        if (e / n == e % n) v.x += b[1];
        T2[e] = v;
    }
    __syncthreads();
    blk_matmul(A, T2, U, n);
    for (int e = threadIdx.x; e < nn; e += NT)
        T1[e] = cadd(cadd(cscale(A6[e], b[12]), cscale(A4[e], b[10])), cscale(A2[e], b[8]));
    __syncthreads();
    blk_matmul(A6, T1, V, n);
    for (int e = threadIdx.x; e < nn; e += NT) {
        double2 v = cadd(V[e], cadd(cadd(cscale(A6[e], b[6]), cscale(A4[e], b[4])), cscale(A2[e], b[2])));
        if (e / n == e % n) v.x += b[0];
        V[e] = v;
    }
    __syncthreads();
    double2 *P = A, *X = T1;
    for (int e = threadIdx.x; e < nn; e += NT) {
        const double2 u = U[e], v = V[e];
        P[e] = csub(v, u);
        X[e] = cadd(v, u);
    }
    __syncthreads();
    if (w0) {
        for (int kk = 0; kk < n; ++kk) {
            // partial pivoting: lanes kk..n-1 take |P(i, kk)|, first maximum wins
            double best = (lane >= kk && lane < n) ? cabsv(P[lane * n + kk]) : -1.0;
            int piv = lane;
            wave_argmax(best, piv);
            if (piv != kk && lane < n) {
                double2 a = P[kk * n + lane];
                P[kk * n + lane] = P[piv * n + lane];
                P[piv * n + lane] = a;
                a = X[kk * n + lane];
                X[kk * n + lane] = X[piv * n + lane];
                X[piv * n + lane] = a;
            }
            wave_sync();
            if (lane > kk && lane < n) g_sh.lu_f[lane] = cdiv(P[lane * n + kk], P[kk * n + kk]);
            wave_sync();
            for (int e = lane; e < nn; e += 64) {
                const int i = e / n, j = e % n;
                if (i > kk) {
                    const double2 f = g_sh.lu_f[i];
                    if (j >= kk) P[i * n + j] = csub(P[i * n + j], cmul(f, P[kk * n + j]));
                    X[i * n + j] = csub(X[i * n + j], cmul(f, X[kk * n + j]));
                }
            }
            wave_sync();
        }
        for (int kk = n - 1; kk >= 0; --kk) {
            if (lane < n) {
                double2 s = X[kk * n + lane];
                for (int i = kk + 1; i < n; ++i) s = csub(s, cmul(P[kk * n + i], X[i * n + lane]));
                X[kk * n + lane] = cdiv(s, P[kk * n + kk]);
            }
            wave_sync();
        }
    }
    __syncthreads();
    double2* cur = X;
    double2* nxt = T2;
    for (int s = 0; s < sc; ++s) {
        blk_matmul(cur, cur, nxt, n);
        double2* tmp = cur;
        cur = nxt;
        nxt = tmp;
    }
    for (int e = threadIdx.x; e < nn; e += NT) Eout[e] = cur[e];
    __syncthreads();
}

}  // namespace ttnd
