// Complex128 GEMM tiles on the FP64 matrix cores of gfx950.
//
// One wave computes a 16x16 complex tile with V_MFMA_F64_16X16X4F64: a complex
// multiply-accumulate over k is four real MFMAs (re*re - im*im into the real
// accumulator, re*im + im*re into the imaginary one); no 3M trick, so rounding matches a
// plain complex product sum.  Operand layout of the instruction (one double per lane):
//   A (16 x 4): lane l holds A[l % 16][l / 16];  B (4 x 16): lane l holds B[l / 16][l % 16];
//   D (16 x 16): lane l holds D[l / 16 + 4 * j][l % 16] in accumulator element j
//   (measured on gfx950 with a lane-id probe).
// Operands come straight from L2 (node tensors are too large for LDS) as 16-byte coherent
// buffer loads; the loads of up to eight k-steps are issued before their MFMAs.
#pragma once

#include <hip/hip_runtime.h>

namespace ttnd {

typedef double d4 __attribute__((ext_vector_type(4)));

// Coherent loads.  Workgroups of a team run on different CUs of one XCD and exchange
// data through the XCD's L2; the per-CU vector L1 is not kept coherent with it, and
// invalidating it costs about 6 us on gfx950.  Every load of data another workgroup may
// have written therefore uses device scope, which misses in the L1 and reads the L2.
__device__ __forceinline__ double ldc(const double* p) {
    // through the global address space: a global_load (a generic pointer would give a flat
    // load, which also waits on the LDS counter)
    return __hip_atomic_load((const __attribute__((address_space(1))) double*)p, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double2 ldc(const double2* p) {
    const double* q = reinterpret_cast<const double*>(p);
    return make_double2(ldc(q), ldc(q + 1));
}

// The same coherent loads, 16 bytes at a time, from a wave-uniform base: a buffer load with the
// device-scope cache policy (sc1, as the atomic loads above get) through a raw buffer
// descriptor.  The base must be the same in every lane of the wave (it is read from lane 0);
// offsets are element indices below 2^27 (2 GiB, the descriptor's extent; the hardware returns
// zero beyond it instead of faulting).
struct Coh {
    __amdgpu_buffer_rsrc_t r;
    __device__ __forceinline__ explicit Coh(const double2* base) {
        const uint64_t a = reinterpret_cast<uint64_t>(base);
        const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(a));
        const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(a >> 32));
        void* u = reinterpret_cast<void*>((static_cast<uint64_t>(hi) << 32) | lo);
        r = __builtin_amdgcn_make_buffer_rsrc(u, 0, 0x7fffffff, 0x00020000);
    }
    __device__ __forceinline__ double2 operator[](int64_t i) const {
        typedef int v4i __attribute__((ext_vector_type(4)));
        const v4i v = __builtin_amdgcn_raw_buffer_load_b128(r, static_cast<int>(i * 16), 0, 16);
        double2 d;
        __builtin_memcpy(&d, &v, 16);
        return d;
    }
};

struct ZAcc {
    d4 re, im;
};

__device__ __forceinline__ void zacc_zero(ZAcc& a) {
    a.re = d4{0.0, 0.0, 0.0, 0.0};
    a.im = d4{0.0, 0.0, 0.0, 0.0};
}

__device__ __forceinline__ void zmfma(ZAcc& acc, double2 a, double2 b) {
    acc.re = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.x, acc.re, 0, 0, 0);
    acc.re = __builtin_amdgcn_mfma_f64_16x16x4f64(-a.y, b.y, acc.re, 0, 0, 0);
    acc.im = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.y, acc.im, 0, 0, 0);
    acc.im = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, b.x, acc.im, 0, 0, 0);
}

// acc += op(A) B over k in [0, K): A(i, k) at a[i*sar + k*sak] (conjugated if CONJA),
// B(k, j) at b[k*sbk + j*sbc]; rows >= mv and columns >= nv read as zero.
struct LoadL2 {  // operands in global memory, possibly written by other workgroups
    __device__ double2 operator()(const double2* p) const { return ldc(p); }
};
struct LoadLocal {  // operands in this workgroup's LDS
    __device__ double2 operator()(const double2* p) const { return *p; }
};

template <bool CONJA, class LD = LoadL2>
__device__ __forceinline__ void ztile(ZAcc& acc, const double2* __restrict__ a, int64_t sar, int64_t sak, int mv,
                                      const double2* __restrict__ b, int64_t sbk, int64_t sbc, int nv, int64_t K,
                                      LD ld = LD()) {
    const int lane = threadIdx.x & 63, i = lane & 15, kq = lane >> 4;
    const bool arow = i < mv, bcol = i < nv;
    const double2 z = make_double2(0.0, 0.0);
    // global operands: one 16-byte coherent buffer load per element (the bases are
    // wave-uniform, every offset lies inside the instance's arena)
    constexpr bool coh = __is_same(LD, LoadL2);
    const Coh ca(coh ? a : nullptr), cb(coh ? b : nullptr);
    // up to 8 k-steps (k <= 32) of operands are loaded before the first MFMA
    for (int64_t k0 = 0; k0 < K; k0 += 32) {
        const int nks = static_cast<int>((K - k0 + 3) / 4 < 8 ? (K - k0 + 3) / 4 : 8);  // wave-uniform
        double2 av[8], bv[8];
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            if (s < nks) {
                const int64_t k = k0 + 4 * s + kq;
                const bool ok = k < K;
                if constexpr (coh) {
                    av[s] = (arow && ok) ? ca[i * sar + k * sak] : z;
                    bv[s] = (bcol && ok) ? cb[k * sbk + i * sbc] : z;
                } else {
                    av[s] = (arow && ok) ? ld(a + i * sar + k * sak) : z;
                    bv[s] = (bcol && ok) ? ld(b + k * sbk + i * sbc) : z;
                }
            }
        }
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            if (s < nks) {
                double2 x = av[s];
                if (CONJA) x.y = -x.y;
                zmfma(acc, x, bv[s]);
            }
        }
    }
}

// C(row, col) (=|+=) alpha * acc for the tile's valid rows/cols; C(row, col) at c[row*scr + col*scc]
__device__ __forceinline__ void ztile_store(const ZAcc& acc, double2* c, int64_t scr, int64_t scc, int mv, int nv,
                                            double2 alpha, bool accumulate) {
    const int lane = threadIdx.x & 63, col = lane & 15, rb = lane >> 4;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int row = rb + 4 * j;
        if (row < mv && col < nv) {
            const double re = acc.re[j], im = acc.im[j];
            double2 v = make_double2(alpha.x * re - alpha.y * im, alpha.x * im + alpha.y * re);
            double2* p = c + row * scr + col * scc;
            if (accumulate) {
                const double2 o = ldc(p);
                v.x += o.x;
                v.y += o.y;
            }
            *p = v;
        }
    }
}

}  // namespace ttnd
