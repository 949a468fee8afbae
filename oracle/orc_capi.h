/*
 * ORACLE — test infrastructure only.  CPU restatement of the TDVP step behind the same
 * call signatures as include/ttn_capi.h, with the prefix orc_ instead of ttn_.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 */
#ifndef ORC_CAPI_H
#define ORC_CAPI_H

#include "../include/ttn_capi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;
typedef struct orc_tree orc_tree;
typedef struct orc_sop orc_sop;
typedef struct orc_integ orc_integ;

int orc_ctx_create(int device, orc_ctx** out);
int orc_ctx_destroy(orc_ctx* ctx);
const char* orc_last_error(const orc_ctx* ctx);
int orc_tree_create(orc_ctx* ctx, int64_t n_nodes, const int64_t* parent, const int64_t* hrank,
                    const int64_t* cap_hrank, const int64_t* leaf_dim, orc_tree** out);
int orc_tree_destroy(orc_tree* tree);
int orc_tree_node(const orc_tree* tree, int64_t node, int64_t* nchild, int64_t* rows, int64_t* hrank,
                  int64_t* leaf_index);
int orc_tree_capacities(const orc_tree* tree, int64_t* cap_hranks);
int orc_tree_size(const orc_tree* tree, int64_t* n_nodes, int64_t* n_modes);
int orc_traversal(const orc_tree* tree, int64_t* len, int64_t* order);
int orc_sop_create(orc_ctx* ctx, int64_t nmodes, const int64_t* mode_dims, int64_t nterms, const double* coeff,
                   int64_t nops, const int32_t* op_mode, const int32_t* op_kind, const int64_t* op_coff,
                   const double* cdata, const int64_t* op_ioff, const int64_t* idata, const int64_t* op_roff,
                   const int64_t* r, orc_sop** out);
int orc_sop_destroy(orc_sop* sop);
int orc_sop_compiled_tables(const orc_tree* tree, const orc_sop* sop, int64_t* len, int64_t* buf);
int orc_integrator_create(orc_ctx* ctx, const orc_tree* tree, const orc_sop* sop, const ttn_integ_params* p,
                          orc_integ** out);
int orc_integrator_destroy(orc_integ* integ);
int orc_state_set(orc_integ* integ, int64_t node, const double* data, int64_t rows, int64_t cols);
int orc_state_get(orc_integ* integ, int64_t node, double* data, int64_t rows, int64_t cols);
int orc_step(orc_integ* integ, double coeff_re, double coeff_im, int64_t nsteps);
int orc_time(const orc_integ* integ, double* t);
int orc_counters(const orc_integ* integ, int64_t* nh_applications, int64_t* nsteps);
int orc_expansion_counters(const orc_integ* integ, int64_t* n_onesite, int64_t* n_twosite);
int orc_hranks(const orc_integ* integ, int64_t* hranks);
int orc_state_shape(const orc_integ* integ, int64_t node, int64_t* rows, int64_t* cols);
int orc_set_dt(orc_integ* integ, double dt);
int orc_get_dt(const orc_integ* integ, double* dt);
int orc_reset_nsteps(orc_integ* integ);
int orc_expansion_draws(const orc_integ* integ, int64_t* n);
int orc_set_coefficients(orc_integ* integ, int64_t nterms, const double* coeff);
int orc_set_expansion_params(orc_integ* integ, const ttn_integ_params* p);
int orc_set_expansion_log(orc_integ* integ, int64_t capacity);
int orc_expansion_log(const orc_integ* integ, int64_t* len, double* out);
int orc_expect(orc_integ* integ, int64_t mode, const double* op, double* out);
int orc_norm2(orc_integ* integ, double* out);
int orc_sb_problem(orc_ctx* ctx, const ttn_sb_params* p, orc_tree** tree, orc_sop** sop);
int orc_sb_initial_state(orc_integ* integ, const ttn_sb_params* p);
int orc_sb_partition_modes(orc_ctx* ctx, int64_t nmodes, const double* wk, double ecut, double wc, int64_t nmax,
                           int64_t nmax4, int64_t nmax2, int64_t nmax3, int64_t* nblocks, int64_t* partition);
int orc_sb_mode_block(orc_ctx* ctx, int64_t nmodes, const double* w, const double* g, double ecut, int64_t nmax,
                      int64_t* nstates, int64_t* nnz, int64_t* occupations, double* h0, int64_t* hc_rowptr,
                      int64_t* hc_colind, double* hc_vals);

/* gradient of Re<psi(T)|O|psi(T)> in the spin-boson couplings and frequencies by the costate
 * method (oracle/adjoint.cpp), same arguments as ttn_sb_adjoint_run */
int orc_sb_adjoint_run(orc_integ* integ, const ttn_sb_params* p, int64_t nsteps, int64_t every, int64_t nquad,
                       int64_t mode, const double* op, double* value, double* dg, double* dw);
/* leaf transition densities <bra|(|j><l|)_m|ket>, and with S inserted on leaf mode ins: per
 * leaf mode m, d_m^2 complex entries of each (len: complex entries of out) */
int orc_transition_densities(orc_integ* ket, orc_integ* bra, int64_t ins, const double* S, int64_t len, double* out);

/* building blocks for unit tests: apply the node / bond effective Hamiltonian with the
 * current environments, refresh environments, and the dense kernels */
int orc_heff_node(orc_integ* integ, int64_t node, const double* v, double* out);
int orc_heff_bond(orc_integ* integ, int64_t node, const double* v, double* out);
int orc_update_spf(orc_integ* integ, int64_t node);
int orc_update_mf(orc_integ* integ, int64_t node);
int orc_env_get(orc_integ* integ, int64_t node, int64_t term, int32_t which_mf, double* out);
int orc_env_set(orc_integ* integ, int64_t node, int64_t term, int32_t which_mf, const double* in);
int orc_node_expv(orc_integ* integ, int64_t node, double coeff_re, double coeff_im, double tau, int64_t kdim,
                  double tol, int64_t* napply);
int orc_expv(int64_t n, const double* Hdense, double* x, double c_re, double c_im, double tau, int64_t kdim,
             double tol, int64_t* napply);
int orc_expm(int64_t n, const double* M, double* E);
int orc_qr(int64_t m, int64_t n, const double* A, double* Q, double* R);
/* expansion helpers: normal draws consumed so far; singular values (descending) of a dense n x n
 * matrix; eigenpairs of a real symmetric tridiagonal matrix (eigenvectors as columns of Z) */
int orc_normals_used(const orc_integ* integ, int64_t* n);
int orc_singular_values(int64_t n, const double* M, double* sv);
int orc_tridiag_eigen(int64_t n, const double* a, const double* b, double* w, double* Z);

#ifdef __cplusplus
}
#endif

#endif /* ORC_CAPI_H */
