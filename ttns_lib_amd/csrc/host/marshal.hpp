// Conversions between the plain C arrays of include/ttn_capi.h and the host structures.
#pragma once

#include "../../../include/ttn_capi.h"
#include "sop.hpp"
#include "spin_boson.hpp"
#include "topology.hpp"

namespace ttnh {

inline cplx cget(const double* p, i64 i) { return cplx(p[2 * i], p[2 * i + 1]); }

inline SOP sop_from_arrays(i64 nmodes, const i64* mode_dims, i64 nterms, const double* coeff, i64 nops,
                           const int32_t* op_mode, const int32_t* op_kind, const i64* op_coff,
                           const double* cdata, const i64* op_ioff, const i64* idata, const i64* op_roff,
                           const i64* r) {
    check(nmodes >= 1 && nterms >= 1 && nops >= 0, ERR_INVALID, "sop: empty operator");
    std::vector<i64> dims(mode_dims, mode_dims + nmodes);
    for (i64 d : dims) check(d >= 1, ERR_INVALID, "sop: mode dimensions must be positive");
    SOP s(static_cast<int>(nterms), dims);
    if (coeff)
        for (i64 i = 0; i < nterms; ++i) s.coeff[i] = cget(coeff, i);
    for (i64 k = 0; k < nops; ++k) {
        const int nu = op_mode[k];
        check(nu >= 0 && nu < nmodes, ERR_INVALID, "sop: operator mode out of range");
        const i64 d = dims[nu];
        PrimOp op;
        op.kind = op_kind[k];
        op.dim = d;
        const i64 c0 = op_coff[k], c1 = op_coff[k + 1];
        const i64 i0 = op_ioff ? op_ioff[k] : 0, i1 = op_ioff ? op_ioff[k + 1] : 0;
        switch (op.kind) {
            case OP_DENSE:
                check(c1 - c0 == d * d, ERR_INVALID, "sop: dense operator needs d*d values");
                for (i64 i = c0; i < c1; ++i) op.dense.push_back(cget(cdata, i));
                break;
            case OP_DIAG:
                check(c1 - c0 == d, ERR_INVALID, "sop: diagonal operator needs d values");
                for (i64 i = c0; i < c1; ++i) op.diag.push_back(cget(cdata, i));
                break;
            case OP_CSR: {
                const i64 nnz = c1 - c0;
                check(i1 - i0 == d + 1 + nnz, ERR_INVALID, "sop: CSR operator needs rowptr (d+1) and colind (nnz)");
                op.rowptr.assign(idata + i0, idata + i0 + d + 1);
                op.colind.assign(idata + i0 + d + 1, idata + i1);
                check(op.rowptr[0] == 0 && op.rowptr[d] == nnz, ERR_INVALID, "sop: CSR rowptr inconsistent");
                for (i64 i = 0; i < d; ++i)
                    check(op.rowptr[i] <= op.rowptr[i + 1], ERR_INVALID, "sop: CSR rowptr not monotone");
                for (i64 c : op.colind) check(c >= 0 && c < d, ERR_INVALID, "sop: CSR column out of range");
                for (i64 i = c0; i < c1; ++i) op.vals.push_back(cget(cdata, i));
                break;
            }
            case OP_IDENTITY:
                break;
            default:
                fail(ERR_INVALID, "sop: unknown operator kind");
        }
        std::vector<int> rr;
        for (i64 i = op_roff[k]; i < op_roff[k + 1]; ++i) rr.push_back(static_cast<int>(r[i]));
        s.bind(op, rr, nu);
    }
    return s;
}

inline SBParams sb_from_c(const ttn_sb_params* p) {
    check(p != nullptr && p->nmodes >= 1, ERR_INVALID, "spin-boson: need at least one bath mode");
    SBParams s;
    s.wk.assign(p->wk, p->wk + p->nmodes);
    s.gk.assign(p->gk, p->gk + p->nmodes);
    check(p->wk != nullptr && p->gk != nullptr, ERR_INVALID, "spin-boson: wk / gk is NULL");
    s.nspf = p->nspf;
    s.nspf_lower = p->nspf_lower;
    s.nspf_cap = p->nspf_cap;
    s.nspf_lower_cap = p->nspf_lower_cap;
    s.nhilb = p->nhilb;
    for (i64 i = 0; i < p->nhilb * p->nhilb; ++i) {
        s.hsys.push_back(cget(p->hsys, i));
        s.scoup.push_back(cget(p->scoup, i));
    }
    for (i64 i = 0; i < p->nhilb; ++i) s.psi0.push_back(cget(p->psi0, i));
    s.seed = p->seed;
    if (p->nblocks != 0) {
        check(p->nblocks > 0 && p->block_sizes != nullptr && p->block_nmax != nullptr, ERR_INVALID,
              "spin-boson: nblocks needs block_sizes and block_nmax");
        s.block_sizes.assign(p->block_sizes, p->block_sizes + p->nblocks);
        s.block_nmax.assign(p->block_nmax, p->block_nmax + p->nblocks);
        s.ecut = p->ecut;
    } else {
        check(p->boson_dim != nullptr, ERR_INVALID, "spin-boson: boson_dim is NULL");
        s.boson_dim.assign(p->boson_dim, p->boson_dim + p->nmodes);
    }
    return s;
}

inline void sb_partition_c(int64_t nmodes, const double* wk, double ecut, double wc, int64_t nmax, int64_t nmax4,
                           int64_t nmax2, int64_t nmax3, int64_t* nblocks, int64_t* partition) {
    check(nmodes >= 1 && wk != nullptr && nblocks != nullptr, ERR_INVALID, "partition_modes: bad arguments");
    const auto part = partition_modes(std::vector<double>(wk, wk + nmodes), ecut, wc, nmax, nmax4, nmax2, nmax3);
    *nblocks = static_cast<int64_t>(part.size());
    if (partition)
        for (size_t i = 0; i < part.size(); ++i) partition[i] = part[i];
}

inline void sb_mode_block_c(int64_t nmodes, const double* w, const double* g, double ecut, int64_t nmax,
                            int64_t* nstates, int64_t* nnz, int64_t* occupations, double* h0, int64_t* hc_rowptr,
                            int64_t* hc_colind, double* hc_vals) {
    check(nmodes >= 1 && w != nullptr && g != nullptr && nstates != nullptr && nnz != nullptr, ERR_INVALID,
          "mode block: bad arguments");
    const ModeBlock b = mode_block(std::vector<double>(w, w + nmodes), std::vector<double>(g, g + nmodes), ecut, nmax);
    const PrimOp hc = block_hc(b);
    *nstates = b.nstates;
    *nnz = hc.nnz();
    if (occupations)
        for (i64 s = 0; s < b.nstates; ++s)
            for (int64_t k = 0; k < nmodes; ++k) occupations[s * nmodes + k] = b.occ(s, static_cast<size_t>(k));
    if (h0) {
        const PrimOp d = block_h0(b);
        for (i64 s = 0; s < b.nstates; ++s) h0[s] = d.diag[s].real();
    }
    if (hc_rowptr)
        for (size_t i = 0; i < hc.rowptr.size(); ++i) hc_rowptr[i] = hc.rowptr[i];
    if (hc_colind)
        for (size_t i = 0; i < hc.colind.size(); ++i) hc_colind[i] = hc.colind[i];
    if (hc_vals)
        for (size_t i = 0; i < hc.vals.size(); ++i) hc_vals[i] = hc.vals[i].real();
}

}  // namespace ttnh
