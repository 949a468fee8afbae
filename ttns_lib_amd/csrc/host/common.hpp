// Common host-side types for the spin-boson TTNS TDVP path.
//
// Complex128 everywhere: the reference instantiates its whole engine with
// ttns::complex<double> (src/spin_boson.cpp:161) and so does this library.
#pragma once

#include <complex>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace ttnh {

using cplx = std::complex<double>;
using i64 = int64_t;

// Status codes shared by the C-ABI (include/ttn_capi.h) and the host code.
enum Status : int {
    OK = 0,
    ERR_NUMERIC = 1,   // NaN/Inf or a failed numerical kernel
    ERR_CAPACITY = 2,  // a requested size exceeds the allocated capacity
    ERR_INVALID = 3,   // malformed input (tree, SOP, state shapes ...)
    ERR_DEVICE = 4,    // HIP runtime failure / no device
};

// Exception carrying a status code; the C-ABI converts it into a return code
// plus a message retrievable with ttn_last_error().
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void fail(int code, const std::string& msg) { throw Error(code, msg); }

inline void check(bool cond, int code, const std::string& msg) {
    if (!cond) fail(code, msg);
}

}  // namespace ttnh
