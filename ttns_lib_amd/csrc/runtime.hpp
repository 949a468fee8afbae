// Host-side runtime shared by the C-ABI translation units (capi.cpp, node.cpp): device
// buffers, HIP error checks, the context (one per GPU, one HIP stream) and the exception
// guard that turns failures into status codes plus a message (ttn_last_error).
#pragma once

#include <hip/hip_runtime.h>

#include <exception>
#include <array>
#include <map>
#include <string>

#include "../../include/ttn_capi.h"
#include "device/desc.hpp"
#include "host/common.hpp"

namespace ttnrt {

using ttnh::ERR_DEVICE;
using ttnh::fail;

inline void hip_check(hipError_t e, const char* what) {
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

inline int64_t align16(int64_t x) { return (x + 15) & ~int64_t(15); }

}  // namespace ttnrt

struct ttn_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    // team launches: every workgroup of a grid of ncu * occupancy is resident at once and
    // the dispatcher spreads them evenly over nxcd XCDs (xcd_cap each), checked by a
    // placement census when the context is created
    bool teams_ok = false;
    int grid = 0, nxcd = 0, xcd_cap = 0;
    int64_t team_min = 2048;      // smaller operations run on a team's leader alone
    double team_gain = 0.01;      // a team grows only while a member cuts its predicted time by this much
    int64_t small_max = 512;      // propagations of at most this many entries run in one workgroup's LDS
    ttnrt::DevBuf<unsigned> team_words;  // XCD slot counters (MAX_XCD lines) + TeamWords per instance
    int64_t team_words_n = 0;     // instances the buffer holds
    ttnrt::DevBuf<ttnd::TeamPlan> plan;
    void* noalloc_scratch = nullptr;  // this device's copy of the allocation-free scratch (node.cpp)
    // tile autotuner of the node path's GEMMs (node.cpp): for every GEMM shape (M, N, K, batch,
    // K block, operand layouts, accumulation) the output tile and K split measured fastest,
    // with that time in ms
    bool autotune = true;
    std::map<std::array<int64_t, 8>, std::array<double, 3>> tune;
    ttnrt::DevBuf<double2> tune_out;  // output of the timing runs (the caller's C is untouched)
    // for every node H-action shape (d0, d1, r, terms): whether the mode-0 and mean-field
    // stages run fused (1) or staged (0), with the faster form's time in ms
    std::map<std::array<int64_t, 4>, std::array<double, 2>> fuse;
    std::map<std::array<int64_t, 5>, std::array<double, 2>> fuse_mf;  // the same per (child, shape)
    int heff_form = 0;  // 0: the autotuner's choice, 1: always staged, 2: fused where it fits
};

// Takes the device address of the allocation-free scratch on the context's (current) device.
void ttnrt_noalloc_setup(ttn_ctx* c);

namespace ttnrt {

template <class F>
int guard(ttn_ctx* ctx, F&& f) {
    try {
        f();
        return TTN_OK;
    } catch (const ttnh::Error& e) {
        if (ctx) ctx->err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        if (ctx) ctx->err = e.what();
        return TTN_ERR_INVALID;
    }
}

}  // namespace ttnrt
