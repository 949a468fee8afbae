// Device-side descriptors of one TTNS instance: node shapes and arena offsets, the
// compiled SOP term tables flattened into arrays, and the sweep order.  Built on the host
// (capi.cpp) and read by the persistent sweep kernel (sweep.hip).  Shared by host and
// device code, so plain structs only.
//
// All complex128 data of an instance lives in one HBM arena of double2 values; every
// "offset" below is an index into that arena.  Node tensors are stored row-major
// (prod dims) x hrank, child 0 slowest (reference ttn_nodes/httensor_node.hpp:20-285).
#pragma once

#include <cstdint>

namespace ttnd {

constexpr int MAXC = 4;        // max children of a branch node on the device path
constexpr int MAXK = 16;       // max Krylov dimension on the device path
constexpr int NWORK = 6;       // generic work buffers, each max node size
constexpr int MAX_NODES = 4096;  // nodes per tree (one byte of visit count each in LDS)

// device wall-clock phase slots (100 MHz ticks, accumulated by thread 0)
// entries per node of the per-node statistics: node-update H applications, their Krylov sizes
// summed, bond-update H applications, their sizes, then the wall-clock ticks of the node and of
// the bond propagations, and of the H-actions inside them (general path only; the small-vector
// propagator's H-actions are not timed apart)
constexpr int NODE_STATS = 8;
enum ProfSlot { PROF_HEFF = 0, PROF_HBOND, PROF_ORTH, PROF_EXPM, PROF_QR, PROF_SPF, PROF_MF, PROF_GAUGE,
                PROF_EXPV_NODE, PROF_EXPV_BOND, PROF_OBS, PROF_TOTAL,
                PROF_SMALL,    // ticks in propagations run by the small-vector propagator
                PROF_NSMALL,   // number of those propagations (a count, not ticks)
                PROF_NSLOTS = 16 };

// Static part of a node.  Its shape is not stored here: bond dimensions grow under the
// expansion integrator, so the sweep reads them from the instance's live bond-dimension
// table (DInst::hr, mirrored in LDS) and derives the mode dimensions from the children's.
// Storage (A, R and the environment matrices) is laid out at capacity.
struct DNode {
    int parent, child_id, nchild, leaf;  // leaf = primitive mode index, -1 for branch nodes
    int nmodes, leafdim, cap_hrank, pad;  // leafdim: primitive dimension of a leaf node
    int children[MAXC];
    int64_t A, R;                         // arena offsets: node tensor (capacity), gauge matrix (rmax^2)
    int term0, nterms;                    // slice of DTerm
};

// Live shape of a node: bond dimension, mode dimensions (the children's bond dimensions, or
// the primitive dimension of a leaf), their product and the matrix size dimen x hrank.
struct Shape {
    int hrank, nmodes;
    int dims[MAXC];
    int dimen, size;
};

// operator_term (reference ttn_nodes/operator_node.hpp:78-229); spf and mf matrices are
// hrank x hrank with the node's live hrank as their row stride, in slots sized for capacity
struct DTerm {
    int id_spf, id_mf;
    double cre, cim;       // coeff()
    double ire, iim;       // identity-spf coefficient: nspf > 1 ? sum(accum) : coeff
    int sub0, nsub;        // spf sub-terms (DSub)
    int mf0, nmf;          // mean-field contributions (DMf)
    int64_t spf, mf;       // arena offsets of the hrank x hrank matrices, -1 for identity
    double a0re, a0im;     // accum_coeff(0): the accumulation coefficient a single sub-term or
                           // mean-field contribution carries in the two-site energy
};

// one spf sub-term: accumulation coefficient and the chain of mode operators
struct DSub {
    double are, aim;       // accum[i]
    int k0, nk;            // slice of DK
};

// one operator of a chain.  Branch nodes: nu = child slot, mat = child's spf matrix (the
// child's live hrank square, kind = DENSE; dim is not used).  Leaf nodes: nu = primitive mode,
// kind/dim/mat/idx/nnz describe the primitive operator (dense d x d, diagonal d, or CSR values
// + (rowptr, colind) in the integer table).
struct DK {
    int nu, kind, dim, nnz;
    int64_t mat, idx;
};

// one mean-field contribution (operator_node.hpp:23-74 mf_index)
struct DMf {
    double are, aim;       // accum[it]
    int64_t pmf;           // parent's mf matrix offset, -1 for identity
    int k0, nk;            // sibling spf chain (DK, branch form, applied to the parent tensor)
};

// Failure record written by the sweep kernel: reason | bond << 4 | node << 5, where bond
// is 1 for the bond (R) propagation behind `node` and 0 for the node propagation.
enum FailReason {
    FAIL_NONE = 0,
    FAIL_INPUT_NONFINITE = 1,   // norm of the vector to propagate is not finite
    FAIL_KRYLOV_NONFINITE = 2,  // a Krylov basis norm is not finite
    FAIL_KRYLOV_STEP = 3,       // no sub-step size met the Krylov tolerance
    FAIL_TEAM_TIMEOUT = 4,      // a team barrier did not complete in time
    FAIL_TEAM_FORM = 5,         // the team's workgroups did not assemble (state untouched)
    FAIL_RNG_EXHAUSTED = 6,     // the expansion needed more normal draws than were uploaded
    FAIL_EXPANSION = 7,         // an expansion step found no orthogonal trial vector or a non-finite value
};

// Expansion state of an instance in global memory (int64 words)
enum ExpWord { EXP_RNG_POS = 0, EXP_ONESITE, EXP_TWOSITE, EXP_GREW, EXP_LOG_N, EXP_NWORDS = 8 };
// one record of the expansion log: node, direction (0 down, 1 up), Ritz values kept, functions
// added, then the Ritz values (descending, MAXK slots)
constexpr int EXP_LOG_REC = 4 + MAXK;

constexpr int MAX_OBS = 8;  // observables recorded on the device after every step

struct RunArgs;

struct DInst {
    double2* base;                  // arena
    const DNode* nodes;
    const DTerm* terms;
    const DSub* subs;
    const DK* ks;
    const DMf* mfs;
    const int* tour;
    const int64_t* ints;           // CSR rowptr/colind
    int nnodes, tour_len;
    int kdim, pad0;
    double tol;
    int64_t maxsize;               // largest node size (work-buffer stride)
    int64_t work[NWORK];           // work buffer offsets
    int64_t krylov;                // Krylov basis (kdim + 1) * maxsize
    int64_t obs_ops[MAX_OBS];      // observable operators (d x d each) offsets
    int obs_leaves[MAX_OBS];       // node id of each observable's leaf
    int nobs;                      // observables recorded after every step (0: none)
    int pad1;
    double* traj;                  // per step: re, im of every <O_j>, then <1> (2 nobs + 1 doubles),
                                   // or null
    int64_t* counters;             // [0] H applications (node updates), [1] steps done
    int64_t* node_stats;           // NODE_STATS per node: node-update H applications, their Krylov sizes
                                   // summed, bond-update applications (by operator node), sizes
    int* status;                   // [0]: failure record (0 = ok, see FailReason), [1]: step index
    int64_t* prof;                 // per-phase device wall-clock ticks (PROF_* slots), or null
    int64_t team_part;             // team partial sums (2 x TEAM_MAX x PART_K doubles)
    int64_t team_tiles;            // team partial products (TEAM_MAX x TILE_MAX)
    int* hr;                       // live bond dimension of every node (written by expansions)
    // subspace expansion (subspace_expansion_projector_splitting_engine.hpp:207-510); expand = 0
    // for the fixed-step integrator
    int expand, neig, min_unocc, only_unocc;
    double spawn, unocc_thr, eig_tol;
    const double* rng;             // normal draws of std::mt19937 + std::normal_distribution, in order
    int64_t rng_n;                 // draws uploaded
    int64_t* exp;                  // ExpWord words
    int64_t e_op1, e_op2;          // arena: two-site factors, nt2 slots of maxsize each
    int64_t e_vec;                 // arena: trial vector, restart vector and F temporaries (4 x maxsize)
    int nt2, pad2;                 // largest number of two-site terms of a node
    double* exp_log;               // expansion log (EXP_LOG_REC doubles per two-site attempt) or null
    int64_t exp_log_cap;           // records it holds
    // this instance's own call arguments (coefficient, dt, prepare and ladder), or null for the
    // launch's: lets one launch advance instances by different lengths and directions in time
    const RunArgs* own;
};

// ---------------------------------------------------------------- teams
constexpr int TEAM_MAX = 32;              // members per team (the CUs of one XCD)
constexpr int PART_K = 160;               // doubles per member and buffer of team partial sums
constexpr int TILE_MAX = 64 * 64;         // partial product (hrank x hrank) per member
constexpr uint64_t TEAM_TIMEOUT = 200000000ull;   // 2 s of 100 MHz ticks at one barrier

// Team words of one team in global memory (zeroed before every launch): barrier arrival
// count and completed generations on separate 128-byte lines, the team-wide failure
// record, and the roll-call outcome (0 undecided, 1 formed, 2 aborted).
struct TeamWords {
    unsigned count;
    unsigned pad0[31];
    unsigned gen;
    unsigned pad1[31];
    int fail;
    int outcome;
    unsigned pad2[30];
};

// Placement of teams on the XCDs for one launch: a workgroup reads the XCD it runs on,
// takes the next free slot of that XCD, and becomes member `rank` of the team of
// instance `inst`.  Workgroups beyond an XCD's slot count exit at once.
constexpr int MAX_XCD = 16;
constexpr int XCD_SLOTS = 64;
struct TeamSlot {
    int inst, rank, size, pad;
};
struct TeamPlan {
    int nxcd, pad;
    int count[MAX_XCD];
    TeamSlot slot[MAX_XCD][XCD_SLOTS];
};

struct RunArgs {
    double2 coeff;                 // -i for real time
    int prepare;                   // prepare_evolution sweep pairs to run first (0: none)
    int prepare_spf;               // run its bottom-up single-particle pass first
    int nladder;                   // ladder sub-steps (first step)
    double ladder_dt[8];
    int nregular;                  // regular steps after the ladder
    double dt;
    int64_t step0;                 // index of the first step of this call (trajectory slot)
    int64_t team_min;              // operations on fewer elements run on a team's leader alone
    int64_t small_max;             // propagations of at most this many entries run in LDS (0: never)
};

}  // namespace ttnd
