#include "modes.hpp"

#include <cmath>

namespace ttnh {

namespace {
// leaf dimensions beyond this are not a tree-network leaf any more
constexpr i64 MAX_BLOCK_STATES = i64(1) << 20;

void finish(ModeBlock& b) {
    const size_t n = b.nper.size();
    b.stride.assign(n, 1);
    i64 s = 1;
    for (size_t k = n; k-- > 0;) {
        check(b.nper[k] >= 1, ERR_INVALID, "mode block: a mode keeps no Fock state (ecut below |w|)");
        b.stride[k] = s;
        check(s <= MAX_BLOCK_STATES / b.nper[k], ERR_INVALID, "mode block: more than 2^20 states");
        s *= b.nper[k];
    }
    b.nstates = s;
}
}  // namespace

i64 mode_cutoff(double w, double ecut, i64 nmax) {
    check(std::isfinite(w) && w != 0.0, ERR_INVALID, "mode block: frequencies must be finite and nonzero");
    check(std::isfinite(ecut) && ecut > 0.0 && nmax >= 1, ERR_INVALID, "mode block: need ecut > 0, nmax >= 1");
    const double q = ecut / std::fabs(w);
    return q >= static_cast<double>(nmax) ? nmax : static_cast<i64>(q);
}

ModeBlock mode_block(const std::vector<double>& w, const std::vector<double>& g, double ecut, i64 nmax) {
    check(!w.empty() && w.size() == g.size(), ERR_INVALID, "mode block: need equally many frequencies and couplings");
    ModeBlock b;
    b.w = w;
    b.g = g;
    for (double x : w) b.nper.push_back(mode_cutoff(x, ecut, nmax));
    finish(b);
    return b;
}

ModeBlock single_mode(double w, double g, i64 d) {
    check(d >= 1, ERR_INVALID, "mode block: boson dimension must be positive");
    ModeBlock b;
    b.w = {w};
    b.g = {g};
    b.nper = {d};
    finish(b);
    return b;
}

PrimOp block_h0(const ModeBlock& b) {
    PrimOp op;
    op.kind = OP_DIAG;
    op.dim = b.nstates;
    op.diag.resize(b.nstates);
    for (i64 s = 0; s < b.nstates; ++s) {
        double v = 0.0;
        for (size_t k = 0; k < b.w.size(); ++k) v += b.w[k] * (static_cast<double>(b.occ(s, k)) + 0.5 - 0.5);
        op.diag[s] = cplx(v, 0.0);
    }
    return op;
}

PrimOp block_hc(const ModeBlock& b) {
    PrimOp op;
    op.kind = OP_CSR;
    op.dim = b.nstates;
    op.rowptr.push_back(0);
    const size_t N = b.w.size();
    for (i64 s = 0; s < b.nstates; ++s) {
        for (size_t k = 0; k < N; ++k) {  // lowering, mode order
            const i64 n = b.occ(s, k);
            if (n > 0) {
                op.colind.push_back(s - b.stride[k]);
                op.vals.push_back(cplx(std::sqrt(static_cast<double>(n)) * b.g[k], 0.0));
            }
        }
        for (size_t k = N; k-- > 0;) {  // raising, reverse mode order
            const i64 n = b.occ(s, k);
            if (n + 1 < b.nper[k]) {
                op.colind.push_back(s + b.stride[k]);
                op.vals.push_back(cplx(std::sqrt(static_cast<double>(n) + 1.0) * b.g[k], 0.0));
            }
        }
        op.rowptr.push_back(static_cast<i64>(op.vals.size()));
    }
    return op;
}

std::vector<i64> partition_modes(const std::vector<double>& wk, double ecut, double wc, i64 nmax, i64 nmax4,
                                 i64 nmax2, i64 nmax3) {
    const size_t N = wk.size();
    check(N != 0, ERR_INVALID, "partition_modes: no modes");
    for (size_t i = 0; i + 1 < N; ++i)
        check(std::fabs(wk[i + 1]) >= std::fabs(wk[i]), ERR_INVALID,
              "partition_modes: frequencies must be sorted by magnitude");
    std::vector<i64> part;
    size_t count = 0, nremaining = N;
    while (nremaining != 0) {
        const bool high = wk[count] > wc;
        const i64 nmaxt = high ? nmax3 : nmax2;
        const i64 nmaxp = high ? nmax4 : nmax;
        const i64 nper = mode_cutoff(wk[count], ecut, nmaxt);
        size_t ntoadd = 1;
        for (;;) {
            // states of a block of ntoadd modes, counted with the first mode's cutoff
            // (compute_nstates, mode_data.hpp:228-233, 785-792), saturated above nmaxp
            i64 nt = 1;
            bool over = false;  // the product exceeds the bound (tracked apart: nmaxp may be INT64_MAX)
            for (size_t i = 0; i < ntoadd; ++i) {
                if (nper > 0 && nt > nmaxp / nper) {
                    over = true;
                    break;
                }
                nt *= nper;
            }
            if (over || nt > nmaxp) {
                if (ntoadd > 1) --ntoadd;
                break;
            }
            if (nremaining > ntoadd)
                ++ntoadd;
            else
                break;
        }
        part.push_back(static_cast<i64>(ntoadd));
        nremaining -= ntoadd;
        count += ntoadd;
    }
    return part;
}

}  // namespace ttnh
