/*
 * ttn_capi.h - C ABI of the MI355X TDVP library for spin-boson tree tensor networks.
 *
 * A drop-in for the one-site projector-splitting TDVP time step of the reference
 * (llindoy/ttns_lib).  The library owns the tree layout, the per-node complex128
 * tensor storage in HBM, the compiled sum-of-products operator tables and the sweep
 * schedule; every time step runs on the GPU.  Plain C types only: no C++ types and no
 * exceptions cross this boundary.  Complex numbers are interleaved (re, im) doubles;
 * matrices are row-major.
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   ttn_tree_create      httensor<T>(topology, capacity)            include/ttns_lib/httensor.hpp:128-288
 *   ttn_sop_create       sop_operator<T>(nterms, mode_dims) + bind() include/ttns_lib/operators/sop_operator.hpp:252-392
 *   ttn_integrator_create  {fixed_step_,subspace_expansion_}projector_splitting_integrator(A, H, krylov_dim,
 *                        krylov_tol, nthreads) and initialise()     include/ttns_lib/integrators/projector_splitting_integrator.hpp:86-116,236-261
 *                        (operator_container::resize compiles the SOP:  tdvp_core/operator_container.hpp:43-96)
 *   ttn_state_set/get    httensor node matrix access (A[i]().as_matrix())  include/ttns_lib/ttn_nodes/httensor_node.hpp:20-285
 *   ttn_step             integrator operator()(A, H, coeff)           projector_splitting_integrator.hpp:136-169,281-326
 *   ttn_time             integrator t()                               projector_splitting_integrator.hpp:188,359
 *   ttn_counters         nh_applications() and the step count         projector_splitting_integrator.hpp:218,361,67
 *   ttn_expansion_counters  engine().Nonesite()/Ntwosite()          subspace_expansion_projector_splitting_engine.hpp:513-514
 *   ttn_set_dt, ttn_get_dt, ttn_reset_nsteps  set_dt()/dt(), reset_nsteps()  projector_splitting_integrator.hpp:33,184-187,352-353
 *   ttn_expect           mel()(op, mode, A)                           include/ttns_lib/observables/matrix_element.hpp:264-314
 *   ttn_traversal        tdvp_tree_traversal order                   include/ttns_lib/tdvp_core/tdvp_tree_traversal.hpp:37-87
 *   ttn_sb_problem, ttn_sb_initial_state
 *                        the spin-boson driver's topology, SOP binding and initial state
 *                                                                     src/spin_boson.cpp:447-649, src/mode_data.hpp:136-361
 * Error convention: every call returns TTN_OK or an error code; ttn_last_error() gives
 * the message of the last failure on a context (reference: CALL_AND_HANDLE / RAISE_*
 * exception chains).
 */
#ifndef TTN_CAPI_H
#define TTN_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TTN_OK 0
#define TTN_ERR_NUMERIC 1
#define TTN_ERR_CAPACITY 2
#define TTN_ERR_INVALID 3
#define TTN_ERR_DEVICE 4

#define TTN_OP_DENSE 0
#define TTN_OP_DIAG 1
#define TTN_OP_CSR 2
#define TTN_OP_IDENTITY 3

/* fixed_step_projector_splitting_integrator (projector_splitting_integrator.hpp:77-223) */
#define TTN_MODE_FIXED_STEP 0
/* subspace_expansion_projector_splitting_integrator (projector_splitting_integrator.hpp:225-397):
 * bonds below their capacity grow during the sweeps (subspace_expansion_projector_splitting_engine.hpp:
 * 207-510); node storage is laid out at capacity, so a bond grows inside a launch */
#define TTN_MODE_SUBSPACE_EXPANSION 1

typedef struct ttn_ctx ttn_ctx;
typedef struct ttn_tree ttn_tree;
typedef struct ttn_sop ttn_sop;
typedef struct ttn_integ ttn_integ;
typedef struct ttn_batch ttn_batch;

/* ---- context: one per GPU; owns a HIP stream ---- */
int ttn_ctx_create(int device, ttn_ctx** out);
int ttn_ctx_destroy(ttn_ctx* ctx);
const char* ttn_last_error(const ttn_ctx* ctx);
int ttn_version(int* major, int* minor);

/* ---- tree (nodes in pre-order) ----
 * parent[i]: parent id (-1 for node 0 only); hrank[i], cap_hrank[i]: bond dimension to
 * the parent and its capacity; leaf_dim[i] > 0 marks a leaf node carrying a primitive
 * mode of that dimension (modes are numbered in pre-order of the leaves). */
int ttn_tree_create(ttn_ctx* ctx, int64_t n_nodes, const int64_t* parent, const int64_t* hrank,
                    const int64_t* cap_hrank, const int64_t* leaf_dim, ttn_tree** out);
int ttn_tree_destroy(ttn_tree* tree);
/* node shape: number of children, rows (= product of mode dims), hrank, leaf mode index (-1 if branch) */
int ttn_tree_node(const ttn_tree* tree, int64_t node, int64_t* nchild, int64_t* rows, int64_t* hrank,
                  int64_t* leaf_index);
int ttn_tree_size(const ttn_tree* tree, int64_t* n_nodes, int64_t* n_modes);
/* bond capacity of every node, pre-order: the values of the capacity ntree the reference's
 * httensor(topology, capacity) takes (httensor.hpp:128-160) */
int ttn_tree_capacities(const ttn_tree* tree, int64_t* cap_hranks);
/* sweep order; order may be NULL to query the length */
int ttn_traversal(const ttn_tree* tree, int64_t* len, int64_t* order);

/* ---- sum-of-products operator ----
 * nops primitive operators; operator k acts on mode op_mode[k], has kind op_kind[k]
 * (TTN_OP_*), complex data cdata[2*op_coff[k] .. 2*op_coff[k+1]) (dense d*d row-major,
 * diagonal d, CSR values nnz), integer data idata[op_ioff[k] .. op_ioff[k+1]) (CSR only:
 * rowptr d+1 then colind nnz), and belongs to the terms r[op_roff[k] .. op_roff[k+1]).
 * Operators are bound in the order given (this fixes the per-mode operator index). */
int ttn_sop_create(ttn_ctx* ctx, int64_t nmodes, const int64_t* mode_dims, int64_t nterms,
                   const double* coeff, int64_t nops, const int32_t* op_mode, const int32_t* op_kind,
                   const int64_t* op_coff, const double* cdata, const int64_t* op_ioff, const int64_t* idata,
                   const int64_t* op_roff, const int64_t* r, ttn_sop** out);
int ttn_sop_destroy(ttn_sop* sop);
/* Compiled per-node term tables (operator_container's integer bookkeeping) serialised
 * as int64, buf may be NULL to query the length.  Layout, node by node in pre-order:
 *   ncommon_spf, ncommon_mf, nstandard, then per term:
 *     id_spf, id_mf, nspf, {len, (nu, cri) x len} x nspf, nmf, {parent, nsib, (nu, cri) x nsib} x nmf */
int ttn_sop_compiled_tables(const ttn_tree* tree, const ttn_sop* sop, int64_t* len, int64_t* buf);

/* ---- integrator ---- */
typedef struct {
    int64_t krylov_dim;          /* maximum Krylov subspace dimension (reference default 4) */
    double krylov_tol;           /* Krylov error tolerance (reference default 1e-12) */
    double dt;                   /* time step */
    int32_t mode;                /* TTN_MODE_* */
    double spawning_threshold;   /* subspace expansion knobs (projector_splitting_integrator.hpp:341-357) */
    double unoccupied_threshold;
    int64_t minimum_unoccupied;
    int64_t neigenvalues;
    int64_t team_size;           /* workgroups cooperating on this instance (0 = automatic) */
    double eigenvalue_tolerance; /* eigs tolerance of the downward expansion (0 = the engine's 1e-6) */
    int32_t only_apply_when_no_unoccupied;  /* expand only while fewer than minimum_unoccupied functions
                                               are unoccupied (projector_splitting_integrator.hpp:338-339) */
    uint32_t rng_seed;           /* std::mt19937 seed of the expansion's trial vectors; 0 = the engine's
                                    default-constructed generator (seed 5489,
                                    two_site_energy_variations.hpp:34) */
} ttn_integ_params;

int ttn_integrator_create(ttn_ctx* ctx, const ttn_tree* tree, const ttn_sop* sop, const ttn_integ_params* p,
                          ttn_integ** out);
int ttn_integrator_destroy(ttn_integ* integ);
/* node matrix, rows x cols must equal the node's current shape */
int ttn_state_set(ttn_integ* integ, int64_t node, const double* data, int64_t rows, int64_t cols);
int ttn_state_get(ttn_integ* integ, int64_t node, double* data, int64_t rows, int64_t cols);
/* nsteps integrator calls A <- TDVP(A, H, coeff) (coeff = -i for real time) */
int ttn_step(ttn_integ* integ, double coeff_re, double coeff_im, int64_t nsteps);
int ttn_time(const ttn_integ* integ, double* t);
/* cumulative node-propagation H applications (nh_applications(), projector_splitting_integrator.hpp:
 * 218, 361) and completed steps (the base class's m_nsteps, :67) */
int ttn_counters(const ttn_integ* integ, int64_t* nh_applications, int64_t* nsteps);
/* engine().Nonesite() / Ntwosite() (subspace_expansion_projector_splitting_engine.hpp:513-514, read by
 * src/spin_boson.cpp:91): random one-site and two-site expansion attempts so far (0 for the
 * fixed-step integrator) */
int ttn_expansion_counters(const ttn_integ* integ, int64_t* n_onesite, int64_t* n_twosite);
/* current bond dimension (hrank) of every node, pre-order (the root's is 1), as the driver prints
 * them (src/spin_boson.cpp:95-103); they grow under the expansion integrator */
int ttn_hranks(const ttn_integ* integ, int64_t* hranks);
/* current shape of a node matrix (httensor_node_data::shape, httensor_node.hpp:20-285) */
int ttn_state_shape(const ttn_integ* integ, int64_t node, int64_t* rows, int64_t* cols);
/* dt() / set_dt() (projector_splitting_integrator.hpp:184-187, 352-353): read by every following
 * ttn_step; t() stays dt * nsteps as in the reference */
int ttn_set_dt(ttn_integ* integ, double dt);
int ttn_get_dt(const ttn_integ* integ, double* dt);
/* reset_nsteps() (projector_splitting_integrator.hpp:33): the next ttn_step prepares the
 * evolution and runs the first-step ladder again */
int ttn_reset_nsteps(ttn_integ* integ);
/* new values of the SOP's coefficients (nterms complex, interleaved re/im; sop_operator::coeff(r),
 * sop_operator.hpp:471-473), which the reference's operator()(A, op, coeff) reads from `op` on
 * every call: the following ttn_step uses them */
int ttn_set_coefficients(ttn_integ* integ, int64_t nterms, const double* coeff);
/* the expansion knobs of a TTN_MODE_SUBSPACE_EXPANSION integrator, which the reference's engine
 * reads on every call (spawning_threshold(), unoccupied_threshold(), minimum_unoccupied(),
 * neigenvalues(), only_apply_when_no_unoccupied(), projector_splitting_integrator.hpp:338-357):
 * those fields of *p (and eigenvalue_tolerance) replace the current ones; the others are ignored */
int ttn_set_expansion_params(ttn_integ* integ, const ttn_integ_params* p);
/* Checkpoint / restart (the reference's cereal save/load of the integrator and the httensor,
 * projector_splitting_integrator.hpp:50-68, 198-216, 373-391, httensor.hpp:101-118): everything
 * a following step reads -- node tensors at their current shapes, bond dimensions, the gauge,
 * single-particle and mean-field matrices, step and H-application counters, Nonesite /
 * Ntwosite, dt and the expansion's random stream -- so that an integrator built on the same
 * tree, operator and parameters continues bit for bit after ttn_checkpoint_load.  Not kept: the
 * recorded trajectory and the profiling counters.  ttn_checkpoint_save with buf == NULL sets
 * *bytes to the size needed; otherwise *bytes is the buffer's size on entry and the size
 * written on return.  Loading clears a failure left by an earlier step. */
int ttn_checkpoint_save(ttn_integ* integ, int64_t* bytes, void* buf);
int ttn_checkpoint_load(ttn_integ* integ, const void* buf, int64_t bytes);
/* normal draws the expansion's trial vectors have taken from its std::mt19937 +
 * std::normal_distribution stream so far.  Instrumentation (no reference counterpart). */
int ttn_expansion_draws(const ttn_integ* integ, int64_t* n);
/* Expansion log (instrumentation, no reference counterpart): with a capacity > 0 every two-site
 * expansion attempt appends a record of 20 doubles: the child node, the direction (0 down, 1 up),
 * the number of Ritz values, the functions added and, for the 16 largest Ritz values S, the
 * spawning values sqrt(S) dt / 2 that the spawning threshold is applied to (descending).
 * ttn_expansion_log copies up to *len records (out may be NULL) and sets *len to the count. */
int ttn_set_expansion_log(ttn_integ* integ, int64_t capacity);
int ttn_expansion_log(const ttn_integ* integ, int64_t* len, double* out);
/* <psi|O|psi> for a dense operator O (d x d complex) on primitive mode `mode` */
int ttn_expect(ttn_integ* integ, int64_t mode, const double* op, double* out);
/* <psi|psi> */
int ttn_norm2(ttn_integ* integ, double* out);
/* Record <psi|O_j|psi> (j < nops <= 8) and <psi|psi> on the device after every completed step,
 * as the reference driver evaluates every mel(mops[j], 0, A) per step (src/spin_boson.cpp:81-105).
 * O_j: d_j x d_j complex on primitive mode modes[j], concatenated in ops.  Clears the records of
 * an earlier set. */
int ttn_set_observables(ttn_integ* integ, int64_t nops, const int64_t* modes, const double* ops);
/* the records, one per completed step of 2 nops + 1 doubles: re, im of every <O_j>, then <1>;
 * out may be NULL to query *len (steps) */
int ttn_observables(const ttn_integ* integ, int64_t* len, double* out);
/* one observable: ttn_set_observables with nops = 1 */
int ttn_set_observable(ttn_integ* integ, int64_t mode, const double* op);
/* recorded (Re<O_0>, <1>) pairs, one per completed step; out may be NULL to query *len */
int ttn_trajectory(const ttn_integ* integ, int64_t* len, double* out);

/* ---- spin-boson driver helpers (src/spin_boson.cpp:447-649) ---- */
typedef struct {
    int64_t nmodes;             /* bath modes */
    const double* wk;           /* frequencies [nmodes] */
    const double* gk;           /* couplings [nmodes] */
    const int64_t* boson_dim;   /* Fock truncation [nmodes], one mode per leaf (nblocks == 0) */
    int64_t nspf, nspf_lower, nspf_cap, nspf_lower_cap;
    int64_t nhilb;              /* system dimension */
    const double* hsys;         /* nhilb x nhilb complex */
    const double* scoup;        /* nhilb x nhilb complex */
    const double* psi0;         /* nhilb complex */
    uint64_t seed;              /* std::mt19937 seed of the initial state */
    /* Leaves holding several modes (src/spin_boson.cpp:418-441): nblocks > 0 splits the modes
     * into consecutive blocks of block_sizes[b] modes, one block per leaf, each mode keeping
     * min(block_nmax[b], floor(ecut / |w_k|)) Fock states (boson_dim is then unused).
     * nblocks == 0: one mode per leaf with boson_dim. */
    int64_t nblocks;
    const int64_t* block_sizes; /* [nblocks] */
    const int64_t* block_nmax;  /* [nblocks] */
    double ecut;
} ttn_sb_params;

int ttn_sb_problem(ttn_ctx* ctx, const ttn_sb_params* p, ttn_tree** tree, ttn_sop** sop);
int ttn_sb_initial_state(ttn_integ* integ, const ttn_sb_params* p);
/* partition_modes (src/mode_data.hpp:818-886, called at src/spin_boson.cpp:420 with
 * wc = 2 * high frequency bound, nmax = target dimension, nmax4 = 250, nmax2 = maximum
 * dimension, nmax3 = maximum dimension for high frequencies): sizes of consecutive mode
 * blocks; wk sorted by magnitude.  partition: room for nmodes entries, or NULL to query.
 * This and ttn_sb_mode_block are host-only and accept ctx == NULL (no error text is then kept). */
int ttn_sb_partition_modes(ttn_ctx* ctx, int64_t nmodes, const double* wk, double ecut, double wc, int64_t nmax,
                           int64_t nmax4, int64_t nmax2, int64_t nmax3, int64_t* nblocks, int64_t* partition);
/* One leaf's mode block (harmonic_mode_combination::construct_basis_topology, H0, Hc;
 * src/mode_data.hpp:136-361): *nstates and *nnz, then, for non-NULL outputs, the
 * occupations [nstates x nmodes], the H0 diagonal [nstates], and Hc in CSR
 * (rowptr [nstates + 1], colind [nnz], real values [nnz]). */
int ttn_sb_mode_block(ttn_ctx* ctx, int64_t nmodes, const double* w, const double* g, double ecut, int64_t nmax,
                      int64_t* nstates, int64_t* nnz, int64_t* occupations, double* h0, int64_t* hc_rowptr,
                      int64_t* hc_colind, double* hc_vals);

/* Gradient of J = Re<psi(T)|O|psi(T)> with respect to the bath couplings g_k and frequencies
 * w_k of a spin-boson integrator built from p by ttn_sb_problem (fixed-step, real time, not
 * yet stepped), T = nsteps * dt, O (d x d, complex) on leaf mode `mode`.  Costate method:
 * dJ/dtheta = 2 int_0^T Im <chi(t)|dH/dtheta|psi(t)> dt with chi(t) = exp(iH(T - t)) O_s psi(T),
 * O_s = (O + O^H)/2.  psi is advanced with a checkpoint every `every` calls and recomputed
 * from the checkpoints; chi is a second tree state propagated backwards by the same TDVP
 * step; each step's integral is an nquad-point Gauss-Legendre rule (1-16 nodes; 0: chosen
 * from dt, max |w_k| and |Hsys| so that exp(i w dt s) is integrated to 1e-13) whose
 * partial steps of psi and chi advance in one persistent launch, the matrix elements come
 * from the leaves' transition densities (one workgroup per node pair).  value: J; dg, dw:
 * nmodes doubles each.  On return the integrator holds psi(T).  The reference has no adjoint:
 * the forward is projector_splitting_integrator.hpp:136-169, and g_k, w_k enter the leaves'
 * H0 and Hc as src/mode_data.hpp:239-253, 309-361 builds them (Fock cutoffs held fixed). */
int ttn_sb_adjoint_run(ttn_integ* integ, const ttn_sb_params* p, int64_t nsteps, int64_t every, int64_t nquad,
                       int64_t mode, const double* op, double* value, double* dg, double* dw);

/* ---- batched persistent sweep: many instances advance in one launch ---- */
int ttn_batch_create(ttn_ctx* ctx, int64_t n, ttn_integ* const* integs, ttn_batch** out);
int ttn_batch_destroy(ttn_batch* batch);
int ttn_batch_step(ttn_batch* batch, double coeff_re, double coeff_im, int64_t nsteps);
/* Each batched step is one hipGraph launch (a fill kernel clearing the team words, then the
 * persistent sweep), captured on first use and replayed while the launch arguments and the
 * team plan are unchanged; enabled = 0 launches directly.  No reference counterpart. */
int ttn_batch_set_graph(ttn_batch* batch, int enabled);
int ttn_batch_graph_stats(const ttn_batch* batch, int64_t* captures, int64_t* replays);

/* ---- measurement ----
 * Algorithmic floating-point operations executed by this instance so far (8 real flops
 * per complex multiply-add): H-actions and Krylov orthogonalisation counted from the
 * measured per-node H-application counts, plus the environment updates, QR gauges and
 * gauge applications of every sweep.  No reference counterpart (instrumentation). */
int ttn_flops(const ttn_integ* integ, double* flops);
/* Device wall-clock ticks (100 MHz) spent per phase of the sweep, 16 slots: node H-action,
 * bond H-action, Gram-Schmidt, small exponentials, QR gauges, single-particle updates,
 * mean-field updates, gauge applications, node Krylov total, bond Krylov total,
 * observables, whole kernel, propagations run by the small-vector propagator; slot 13 counts
 * those propagations.  Instrumentation only. */
int ttn_profile(const ttn_integ* integ, int64_t* slots);

/* Per node (out[8 n .. 8 n + 7], n in pre-order): H applications of its node propagations,
 * their Krylov sizes summed over iterations, H applications of the bond propagations that
 * use its operators, their Krylov sizes summed, the device wall-clock ticks (100 MHz) of the
 * node and of the bond propagations, and the ticks of the H-actions inside those two (outside
 * the small-vector propagator).  Instrumentation only. */
int ttn_node_stats(const ttn_integ* integ, int64_t* out);

/* ---- workgroup teams (no reference counterpart: GPU execution tuning) ----
 * An instance is advanced by a team of workgroups on one XCD.  Operations on node
 * tensors of at least team_min complex entries are spread over the team; smaller ones run
 * on its first workgroup.  teams_enabled = 0 runs every instance on one workgroup.
 * Results do not depend on these settings beyond rounding. */
int ttn_ctx_set_tuning(ttn_ctx* ctx, int64_t team_min, int teams_enabled);
/* Krylov propagations of at most small_max entries (capped at 512) run on one workgroup with the
 * basis in registers and the operators in LDS; 0 runs every propagation on the general path.
 * Same algorithm, results equal to rounding.  No reference counterpart (a launch knob). */
int ttn_ctx_set_small(ttn_ctx* ctx, int64_t small_max);
/* team size chosen for every instance by the last ttn_batch_step (sizes[n]) */
int ttn_batch_teams(const ttn_batch* batch, int64_t* sizes);

/* ---- one large branch node on the whole GPU ----
 * A branch node with two children of dimensions d0, d1 and bond dimension r; its tensor A is
 * (d0 d1) x r, row-major, child 0 slowest.  Its effective Hamiltonian is given by the
 * environments a reference operator node holds, as nterms product terms
 *     H A = sum_t coeff_t (O0_t (x) O1_t) A MF_t^T
 * (O0_t: d0 x d0 single-particle matrix of child 0, O1_t: d1 x d1 of child 1, MF_t: r x r
 * mean field; arrays term-major, each matrix row-major, complex interleaved).  The tensor
 * stays resident in HBM between calls.
 *   ttn_node_heff     res = H A, reps times; coefficient_evolution_branch::operator()
 *                     (tdvp_core/projector_splitting_evolution_functions.hpp:166-258)
 *   ttn_node_mf       mf_t of child c for every term t: sum over all other indices of
 *                     conj(A) [(other child's O) A MF_t^T]; mean_field_operator_engine::evaluate
 *                     (tdvp_core/mean_field_operator.hpp:37-135); out: nterms x dc x dc
 *   ttn_node_expv     A <- exp(coeff tau H) A by the library's Krylov propagator; the
 *                     node update of projector_splitting_engine::update_node_tensor
 *                     (tdvp_core/projector_splitting_engine.hpp:117-156)
 *   ttn_node_last     device time (HIP events) and algorithmic flops of the last call */
typedef struct ttn_node ttn_node;
int ttn_node_create(ttn_ctx* ctx, int64_t d0, int64_t d1, int64_t r, int64_t nterms, const double* coeff,
                    const double* op0, const double* op1, const double* mf, ttn_node** out);
int ttn_node_destroy(ttn_node* node);
int ttn_node_set(ttn_node* node, const double* A);
int ttn_node_get(ttn_node* node, double* A);
int ttn_node_get_result(ttn_node* node, double* res);
int ttn_node_heff(ttn_node* node, int64_t reps);
int ttn_node_mf(ttn_node* node, int64_t child, int64_t reps, double* out);
int ttn_node_expv(ttn_node* node, double coeff_re, double coeff_im, double tau, int64_t krylov_dim, double tol,
                  int64_t* napply);
int ttn_node_last(const ttn_node* node, double* ms, double* flops);
/* Tile autotuner of the node-level GEMMs (on by default): the first time a context meets a GEMM
 * shape it times every output tile (32x32, 64x32, 32x64, 64x64, 128x64, 64x128 per workgroup)
 * with several K splits on the device and keeps the fastest; tiles never change results, K
 * splits change their rounding.  The allocation-free entry point uses the table but never
 * times.  ttn_autotune_table: *n entries (query with entries = ms = NULL), each M, N, K,
 * batch, K block, tile (0-5 in the order above), K split, and its time in ms. */
int ttn_ctx_set_autotune(ttn_ctx* ctx, int enabled);
int ttn_autotune_table(const ttn_ctx* ctx, int64_t* n, int64_t* entries, double* ms);
/* Form of the last two stages of the node-level H-action and mean-field update (no
 * counterpart in the reference; the reference contracts term by term, coefficient_evolution_branch,
 * projector_splitting_evolution_functions.hpp:120-258, mean_field_operator.hpp:37-184).
 * 0 (default): with the autotuner on, the first call on a shape times both forms and keeps the
 * faster; 1: staged, batched GEMMs with the per-term intermediates in memory; 2: fused where the
 * child dimension and r are at most 64, one kernel that keeps each slab of a term's intermediate
 * in LDS and sums the products in registers.  The forms differ only in the order of rounding. */
int ttn_ctx_set_heff_form(ttn_ctx* ctx, int form);

/* No counterpart in the reference.  For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the
 * ttn_node_heff H-action, res = sum_t coeff_t (O0_t (x) O1_t) A MF_t^T, with every
 * intermediate in a fixed device scratch array of 64 MiB that each GPU holds until the process
 * ends.  Same arguments as ttn_node_create + ttn_node_heff, but every array is a DEVICE pointer
 * the caller owns (coeff: nterms complex; op0: nterms d0 x d0; op1: nterms d1 x d1; mf: nterms
 * r x r; A and res: (d0 d1) x r), and the work is enqueued on `stream` (a hipStream_t, NULL for
 * the default stream) without synchronising.  Terms are processed in chunks that fit the
 * scratch; a node whose mean-field factors and one term's intermediates exceed it returns
 * TTN_ERR_CAPACITY.  Allocates nothing.  All calls on one GPU share its scratch: calls that
 * overlap on one GPU (other streams, other threads) race and must be serialised by the caller. */
int ttn_node_heff_noalloc(ttn_ctx* ctx, int64_t d0, int64_t d1, int64_t r, int64_t nterms, const double* coeff,
                          const double* op0, const double* op1, const double* mf, const double* A, double* res,
                          void* stream);

/* ---- timing helpers (device-side event timing of the last ttn_step / ttn_batch_step) ---- */
int ttn_last_step_ms(const ttn_integ* integ, double* ms);
int ttn_batch_last_step_ms(const ttn_batch* batch, double* ms);

#ifdef __cplusplus
}
#endif

#endif /* TTN_CAPI_H */
