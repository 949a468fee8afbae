// Test infrastructure: the ttn_* entry points that include/ttns_amd.hpp calls, forwarded
// to the oracle's orc_* C ABI, so the C++ drop-in driver also runs on the CPU with the
// oracle as its backend (the fake backend of the test strategy).  Never part of the product.
#include "../../oracle/orc_capi.h"
#include "../../include/ttn_capi.h"

#define O(T) reinterpret_cast<orc_##T*>
#define OC(T) reinterpret_cast<const orc_##T*>

extern "C" {
const char* ttn_last_error(const ttn_ctx* c) { return orc_last_error(OC(ctx)(c)); }
int ttn_ctx_create(int device, ttn_ctx** out) { return orc_ctx_create(device, reinterpret_cast<orc_ctx**>(out)); }
int ttn_ctx_destroy(ttn_ctx* c) { return orc_ctx_destroy(O(ctx)(c)); }
int ttn_tree_create(ttn_ctx* c, int64_t n, const int64_t* parent, const int64_t* hrank, const int64_t* cap,
                    const int64_t* leaf, ttn_tree** out) {
    return orc_tree_create(O(ctx)(c), n, parent, hrank, cap, leaf, reinterpret_cast<orc_tree**>(out));
}
int ttn_sb_partition_modes(ttn_ctx* c, int64_t n, const double* wk, double ecut, double wc, int64_t nmax, int64_t nmax4,
                           int64_t nmax2, int64_t nmax3, int64_t* nb, int64_t* part) {
    return orc_sb_partition_modes(O(ctx)(c), n, wk, ecut, wc, nmax, nmax4, nmax2, nmax3, nb, part);
}
int ttn_sb_mode_block(ttn_ctx* c, int64_t n, const double* w, const double* g, double ecut, int64_t nmax,
                      int64_t* ns, int64_t* nnz, int64_t* occ, double* h0, int64_t* rp, int64_t* ci, double* v) {
    return orc_sb_mode_block(O(ctx)(c), n, w, g, ecut, nmax, ns, nnz, occ, h0, rp, ci, v);
}
int ttn_tree_destroy(ttn_tree* t) { return orc_tree_destroy(O(tree)(t)); }
int ttn_sop_create(ttn_ctx* c, int64_t nmodes, const int64_t* dims, int64_t nterms, const double* coeff, int64_t nops,
                   const int32_t* mode, const int32_t* kind, const int64_t* coff, const double* cdata,
                   const int64_t* ioff, const int64_t* idata, const int64_t* roff, const int64_t* r, ttn_sop** out) {
    return orc_sop_create(O(ctx)(c), nmodes, dims, nterms, coeff, nops, mode, kind, coff, cdata, ioff, idata, roff, r,
                          reinterpret_cast<orc_sop**>(out));
}
int ttn_sop_destroy(ttn_sop* s) { return orc_sop_destroy(O(sop)(s)); }
int ttn_integrator_create(ttn_ctx* c, const ttn_tree* t, const ttn_sop* s, const ttn_integ_params* p,
                          ttn_integ** out) {
    return orc_integrator_create(O(ctx)(c), OC(tree)(t), OC(sop)(s), p, reinterpret_cast<orc_integ**>(out));
}
int ttn_integrator_destroy(ttn_integ* i) { return orc_integrator_destroy(O(integ)(i)); }
int ttn_state_set(ttn_integ* i, int64_t node, const double* d, int64_t r, int64_t c) {
    return orc_state_set(O(integ)(i), node, d, r, c);
}
int ttn_state_get(ttn_integ* i, int64_t node, double* d, int64_t r, int64_t c) {
    return orc_state_get(O(integ)(i), node, d, r, c);
}
int ttn_step(ttn_integ* i, double cr, double ci, int64_t n) { return orc_step(O(integ)(i), cr, ci, n); }
int ttn_time(const ttn_integ* i, double* t) { return orc_time(OC(integ)(i), t); }
int ttn_counters(const ttn_integ* i, int64_t* nh, int64_t* ns) { return orc_counters(OC(integ)(i), nh, ns); }
int ttn_expect(ttn_integ* i, int64_t mode, const double* op, double* out) {
    return orc_expect(O(integ)(i), mode, op, out);
}
int ttn_expansion_counters(const ttn_integ* i, int64_t* n1, int64_t* n2) {
    return orc_expansion_counters(OC(integ)(i), n1, n2);
}
int ttn_hranks(const ttn_integ* i, int64_t* hr) { return orc_hranks(OC(integ)(i), hr); }
int ttn_state_shape(const ttn_integ* i, int64_t node, int64_t* r, int64_t* c) {
    return orc_state_shape(OC(integ)(i), node, r, c);
}
int ttn_set_dt(ttn_integ* i, double dt) { return orc_set_dt(O(integ)(i), dt); }
int ttn_reset_nsteps(ttn_integ* i) { return orc_reset_nsteps(O(integ)(i)); }
int ttn_set_coefficients(ttn_integ* i, int64_t n, const double* c) { return orc_set_coefficients(O(integ)(i), n, c); }
int ttn_set_expansion_params(ttn_integ* i, const ttn_integ_params* p) {
    return orc_set_expansion_params(O(integ)(i), p);
}
// checkpoints belong to the device library only
int ttn_checkpoint_save(ttn_integ*, int64_t*, void*) { return TTN_ERR_INVALID; }
int ttn_checkpoint_load(ttn_integ*, const void*, int64_t) { return TTN_ERR_INVALID; }
int ttn_get_dt(const ttn_integ* i, double* dt) { return orc_get_dt(OC(integ)(i), dt); }
}
