/* xtx.h -- C ABI of the MI355X (gfx950) engine for the xtax L1 congruence
 * annealer.
 *
 * The engine replaces the reference's CPU annealer for the L1 objective,
 * `canneal::run_annealer<L1Objective>(...)` (src/congruence_anneal.hpp:421-423)
 * at its xtax call sites (src/main.cpp:716-718 A mode, src/main.cpp:844-846
 * L mode). Where the reference passes a policy object by value and returns the
 * best one by value, this boundary takes caller-owned host buffers (row-major
 * n x n int64) and writes the best A and X back into caller-owned buffers.
 *
 * Conventions: every function returns XTX_OK (0) or a negative XTX_E* code;
 * no exception crosses the boundary; the message of the last failure on a
 * context is available from xtx_last_error(). The library writes no files.
 * Calls on one context are not thread-safe except xtx_request_stop(), which
 * may be called from a signal handler.
 */
#ifndef XTX_H
#define XTX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XTX_OK 0
#define XTX_E_ARG -1        /* invalid argument (shape, parameter range) */
#define XTX_E_BOUND -2      /* an input entry exceeds the 2^48 magnitude bound */
#define XTX_E_DEVICE -3     /* HIP runtime failure */
#define XTX_E_NOMEM -4      /* device or host allocation failed */
#define XTX_E_STATE -5      /* call not valid in the current state */
#define XTX_E_FULL -6       /* serve queue full */
#define XTX_E_VERIFY -7     /* a result failed exact host verification */

#define XTX_MAGNITUDE_LIMIT (((int64_t)1) << 48) /* mat_io.hpp:35 */
#define XTX_SHEAR_CAP 64                          /* mat_io.hpp:38 */
#define XTX_MAX_N 256                             /* largest order per replica (xdual:
                                                     128) */
#define XTX_MAX_TARGET_SAMPLES 4096  /* tournaments above 16 samples run the lane
                                        kernel's wide mode (xdual: 16) */

typedef struct xtx_ctx xtx_ctx;

/* Engine parameters. The first block mirrors canneal::EngineParams
 * (src/congruence_anneal.hpp:122-153) field for field with the same defaults
 * (see xtx_params_default); the second block is engine-specific. */
typedef struct {
    double t_init;            /* <= 0: auto, max(1, 2*score/n^2) (main.cpp:383-387) */
    double t_min;             /* 1e-3: cooling floor and ladder bottom */
    double cooling;           /* 0.999 */
    double reheat;            /* 1.0 */
    double reseed_factor;     /* 1.25 */
    double greedy_fraction;   /* 0.5 */
    double target_fraction;   /* 0.5 */
    double target_accept;     /* 0.44 */
    double worker_diversity;  /* 0.4 */
    double sweep_fraction;    /* 0.0 */
    double max_seconds;       /* 0: run until solved (or max_moves / max_epochs) */
    double save_interval;     /* 2.0 (used by the CLI for best_*.csv throttling) */
    int32_t moves_per_cool;   /* 200 */
    int32_t stuck_threshold;  /* 20000 */
    int32_t target_samples;   /* 8, at most XTX_MAX_TARGET_SAMPLES */
    int32_t tempering;        /* 1 (needs replicas >= 2, as threads >= 2) */
    int32_t adaptive_cooling; /* 0 */
    int32_t exchange_interval;/* 2000: moves per replica between exchange sweeps */
    uint64_t seed;            /* Philox key; 0 is a valid key (no random_device) */
    /* engine-specific */
    int32_t replicas;         /* R: tempering rungs / independent workers (-t) */
    int32_t move_log;         /* 1: record every accepted move of every replica */
    int64_t move_log_cap;     /* entries kept per replica when move_log = 1 */
    uint64_t max_moves;       /* per-replica proposal cap; 0 = none */
    int32_t max_epochs;       /* 0 = none */
    int32_t verify;           /* 1: exact host verification of every result */
    int32_t waves_per_block;  /* waves per workgroup for orders with no tuned
                                 configuration; 0 = 1 */
    int32_t autotune;         /* 1: xtx_batch_solve tunes every order of its
                                 batch that has no configuration yet (on the
                                 batch itself, before the run) */
    int32_t lookahead;        /* proposals of one replica evaluated at once
                                 (1, 2, 4, 8, 16 or 32; the L1 engine allows
                                 up to 32 to order 32, 8 to 64, 4 above; the
                                 xdual engine 16 to order 16, 8 to 32, 4 to
                                 64): the replica takes a whole wave (2, 4) or
                                 a workgroup of four or eight waves and
                                 resolves them in order, the same trajectory
                                 at up to that many moves per step. 0 = auto:
                                 plain and batch calls 16 (or the order's
                                 lookahead configuration from xtx_launch_set
                                 / xtx_autotune); a serve admission 16 while
                                 at most 32 replicas per compute unit are
                                 resident with it, else 1 (the order's
                                 one-group configuration or the default
                                 layout). 1: one group per replica. */
    int32_t persistent;       /* xtx_batch_solve: 1 (default) runs the whole
                                 batch as one persistent launch when every
                                 instance runs the lane kernel by its rule
                                 (orders to 48, 16 proposals per step to order
                                 32 and 32 above, at most 256 replicas): every
                                 wave takes the next ready unit (one wave's
                                 replicas of one instance for one epoch) and
                                 the instance's last unit of an epoch runs
                                 its epoch end, so instances advance with no
                                 device-wide join; 0: captured graphs of
                                 epochs. The same results either way. */
    /* Plain forms (xtx_solve_plain, xtx_dual_solve): called on the calling
     * thread at an epoch boundary whenever the instance best improves -- the
     * reference's "new best" report (congruence_anneal.hpp:309-352). seconds
     * since the call began; score (L1) or dscore (xdual; score = 0); the
     * replica holding the best. NULL: no reports. */
    void (*on_best)(void* user, double seconds, int64_t score, double dscore, int32_t replica);
    void* on_best_user;
} xtx_params;

typedef struct {
    int32_t compute_units;
    int32_t lds_bytes_per_cu;
    int32_t lds_bytes_per_block_max;
    int32_t wave_size;
    int32_t max_n;            /* XTX_MAX_N */
    int32_t max_n_lds;        /* largest n whose replica state fits one block */
    int64_t hbm_bytes;
    char arch[64];
} xtx_limits;

typedef struct {
    int64_t moves;            /* proposals executed, all replicas */
    int64_t accepts;          /* committed moves (incl. sweep commits) */
    int64_t greedy_calls;     /* best-shear computations */
    int64_t infeasible;       /* proposals rejected by the magnitude bound */
    int64_t sweeps;           /* stall reduction sweeps run */
    int64_t exchanges_tried;
    int64_t exchanges_accepted;
    int32_t solved;           /* a replica reached a diagonal */
    int32_t solve_replica;
    int64_t solve_move;       /* that replica's move index at the diagonal */
    int64_t best_score;       /* score of the returned A */
    int32_t best_replica;
    int32_t epochs;
    int32_t stop_reason;      /* 0 solved, 1 max_seconds, 2 stop request,
                                 3 max_moves/max_epochs, 4 trivial (n < 2) */
    int32_t verify_status;    /* 0 ok; bits as in xtx_verify_report.status; -1 not run */
    double seconds;           /* wall time of the call */
    double kernel_ms;         /* summed anneal-kernel time (HIP events) */
} xtx_stats;

/* One matrix for the batch and serve forms. A0 is the input matrix; X0 (null =
 * identity) is a start transform, so the engine starts from X0^T A0 X0 exactly
 * as `xtax -A A.csv -X X.csv` does (main.cpp:709-712). Results are written to
 * A_out / X_out (n x n, caller-owned; either may be null). */
typedef struct {
    const int64_t* A0;
    const int64_t* X0;
    int32_t n;
    int32_t replicas;         /* 0 = params.replicas */
    uint64_t seed;            /* Philox key of this instance */
    const int32_t* active;    /* null = all indices */
    int32_t n_active;
    int32_t reserved;
    int64_t* A_out;
    int64_t* X_out;
} xtx_instance;

typedef struct {
    int32_t solved;
    int32_t verify_status;    /* 0 ok, -1 not run, else xtx_verify_report bits */
    int64_t solve_move;
    int64_t best_score;
    int64_t moves;            /* proposals executed for this instance */
    double submit_s;          /* times relative to the batch start / serve open */
    double admit_s;
    double retire_s;
    uint64_t tag;
} xtx_result;

typedef struct {
    int64_t moves;
    int64_t accepts;
    int32_t instances;
    int32_t solved;
    int32_t verified;
    int32_t epochs;
    double seconds;
    double kernel_ms;
    int32_t buckets;          /* kernel launches per epoch (orders whose
                                 lookahead configurations match share one) */
    int32_t graph_epochs;     /* epochs per captured graph launch */
    double upload_seconds;    /* admission: host inputs to device-resident
                                 state; seconds - upload_seconds is the run
                                 with inputs resident in HBM */
    int32_t lane_instances;   /* instances run by the lane kernel (lanes = 1) */
    int32_t persistent;       /* 1: the batch ran as one persistent launch */
} xtx_batch_stats;

typedef struct {
    int32_t status;           /* 1: X^T A0 X != A, 2: A not diagonal,
                                 4: |det X| != 1, 8: undecided within budget */
    int32_t primes_used;
    int32_t det_sign;
    int32_t exact_int128;     /* 1 if the int128 direct product was used */
} xtx_verify_report;

/* One accepted move (16 bytes): the proposal index of the replica, pivot i,
 * target j, shear s; flags bit 0 marks a stall reduction-sweep commit, bit 1 a
 * reseed from the instance-best snapshot (cooling mode; i = j = -1, s = 0). */
typedef struct {
    int64_t move;
    int16_t i, j, s;
    uint16_t flags;
} xtx_log_entry;

/* ---- context ---- */
int xtx_open(int device, xtx_ctx** out);
int xtx_close(xtx_ctx* ctx);
const char* xtx_last_error(const xtx_ctx* ctx); /* ctx may be null: open errors */
int xtx_query_limits(xtx_ctx* ctx, xtx_limits* out);
void xtx_params_default(xtx_params* p);
/* Async-signal-safe: asks running solves to stop at the next poll point
 * (the reference's g_stop, stop_signal.hpp:40-64). */
void xtx_request_stop(xtx_ctx* ctx);
void xtx_clear_stop(xtx_ctx* ctx);

/* ---- launch configuration per order bucket ----
 * No counterpart in the reference (its CPU threads have no launch shape).
 * Instances are bucketed by padded order (n rounded up to a multiple of 8);
 * each bucket runs with one configuration: lanes per replica, columns per
 * lane, LDS layout of A, waves per workgroup and lookahead. Orders whose
 * lookahead configurations match run in one launch. The configuration changes
 * speed only: every configuration computes the same moves and results.
 * lanes = 1 is the lane kernel: one lane per proposal, the whole row pair in
 * its registers, 16 or 32 proposals per replica and step (cols = order,
 * triangular, lookahead 16 or 32, orders 8 to 48; orders 8 and 16, 24 and 32,
 * 40 and 48 share a launch); plain and batch calls with automatic lookahead
 * choose it when their replicas fill the device (16 per CU), at 16 proposals
 * up to order 32 and 32 above. */
typedef struct {
    int32_t order;            /* padded order: a multiple of 8 in [8, 128]; 192 or
                                 256: the wide lane kernel (orders above 128) */
    int32_t lanes;            /* 16, 32 or 64 lanes per replica; 1 = lane kernel */
    int32_t cols;             /* 1 or 2 columns per lane; lanes * cols >= order
                                 (the order itself for the lane kernel) */
    int32_t triangular;       /* 1: A as its upper triangle in LDS, 0: square */
    int32_t waves_per_block;  /* 1, 2 or 4; 0 = the params' waves_per_block */
    int32_t tuned;            /* out: 1 if set (by hand or autotuner), 0 = rule */
    int32_t lds_bytes;        /* out: LDS per replica */
    int32_t lookahead;        /* lane groups per replica (xtx_params.lookahead):
                                 0 or 1 = one; 2, 4 (one wave), 8, 16, 32 (a
                                 workgroup; waves_per_block must be 0); for
                                 the lane kernel, proposals per step: 16 or 32 */
    int32_t x_in_lds;         /* 1: each replica's X^T in LDS for the epoch
                                 (n x n int64 beside A; 64 lanes, 2 cols,
                                 triangular, lookahead 4, orders 72 to 112:
                                 A and X^T in one workgroup's LDS), 0: in HBM */
    int32_t reserved;         /* 0 */
    double moves_per_sec;     /* out of xtx_autotune: measured rate */
} xtx_launch;

/* The configuration set for order n (tuned = 1), else the one a plain or
 * batch call with automatic lookahead uses (tuned = 0). A set configuration
 * applies where its kind is chosen: a lookahead one under automatic
 * lookahead, a one-group one at lookahead 1 or a loaded server; a lane-kernel
 * one the autotuner chose, only to calls whose replicas fill the device. */
int xtx_launch_get(xtx_ctx* ctx, int32_t n, xtx_launch* out);
/* Sets the configuration of cfg->order (lanes = 0 restores the rule). Affects
 * instances admitted afterwards. XTX_E_ARG if no kernel has that shape or its
 * LDS does not fit a workgroup. */
int xtx_launch_set(xtx_ctx* ctx, const xtx_launch* cfg);
/* Autotuner: for every order present in sample[0..count), times each valid
 * configuration at the lookahead depth plain and batch calls use (16 where
 * the order allows it, else its largest) on those instances (repeated until
 * they fill the device) for `moves` proposals per replica and keeps the
 * fastest for the context.
 * Writes up to max_results measurements (every candidate, the chosen ones
 * with tuned = 1) to results; *n_results is the number written. */
int xtx_autotune(xtx_ctx* ctx, const xtx_instance* sample, int32_t count,
                 const xtx_params* params, int32_t moves, xtx_launch* results,
                 int32_t max_results, int32_t* n_results);

/* ---- plain form: one instance per launch ----
 * Replaces run_annealer(L1Objective(X0^T A0 X0, X0), params, active)
 * (congruence_anneal.hpp:421-569 via main.cpp:712-718). Writes the best (or
 * the solving) state to A_out / X_out. active (nullable) restricts pivots and
 * targets to those indices (the reference's `active`, :437-443). */
int xtx_solve_plain(xtx_ctx* ctx, const int64_t* A0, const int64_t* X0, int32_t n,
                    const xtx_params* params, const int32_t* active, int32_t n_active,
                    int64_t* A_out, int64_t* X_out, xtx_stats* stats);

/* The plain form without allocation. No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Same
 * arguments, results and errors as xtx_solve_plain, whose kernels it runs;
 * every device buffer of the solve (instance record, replica table, each
 * replica's A, X^T, best snapshots, scalars, ladder, active list and move log)
 * is carved out of one fixed 32 MiB device array per GPU, held until the
 * process ends, and the call makes no hipMalloc, hipFree, hipHostMalloc or
 * other allocating HIP call. Sizes whose state does not fit (about
 * 4 R n^2 8 bytes, plus R move_log_cap 16 with a move log) are XTX_E_ARG with
 * a message naming the bytes needed. All calls on one GPU share its array:
 * concurrent calls on one GPU (from any contexts) race and must not overlap;
 * xtx_last_replicas and the move log read it until the next call. */
int xtx_solve_plain_noalloc(xtx_ctx* ctx, const int64_t* A0, const int64_t* X0, int32_t n,
                            const xtx_params* params, const int32_t* active, int32_t n_active,
                            int64_t* A_out, int64_t* X_out, xtx_stats* stats);

/* The instance best as of the last epoch boundary of the running plain solve
 * (L1: A and X; xdual: P, Q and X; each pointer nullable). Valid only from
 * inside an on_best report (the reference's throttled best_*.csv writes,
 * congruence_anneal.hpp:327-340); XTX_E_STATE elsewhere. */
int xtx_best_so_far(xtx_ctx* ctx, int64_t* A_out, double* Q_out, int64_t* X_out);

/* Per-replica state of the last plain solve (for replay and trajectory checks):
 * A, Xt (R*n*n each, nullable) and 8 scalars per replica
 * {score, offnz, moves, best_score, best_move, rung, accepts, local_best}. */
int xtx_last_replicas(xtx_ctx* ctx, int64_t* A, int64_t* Xt, int64_t* scalars);

/* Move log of replica r of the last plain solve; the pointer stays valid until
 * xtx_movelog_free or the next solve. *truncated is set when the cap was hit. */
int xtx_movelog_get(xtx_ctx* ctx, int32_t replica, const xtx_log_entry** entries,
                    int64_t* count, int32_t* truncated);
int xtx_movelog_free(xtx_ctx* ctx);

/* ---- batched form: many instances of uneven n in one captured launch ---- */
int xtx_batch_solve(xtx_ctx* ctx, const xtx_instance* instances, int32_t count,
                    const xtx_params* params, xtx_result* results, xtx_batch_stats* stats);

/* ---- served form: continuous batching ---- */
typedef struct {
    int64_t lds_budget_bytes;   /* 0 = whole device */
    int64_t hbm_budget_bytes;   /* 0 = 8 GiB */
    int32_t max_resident;       /* 0 = unlimited (budget-bound) */
    int32_t queue_capacity;     /* 0 = 65536 */
} xtx_serve_config;

int xtx_serve_open(xtx_ctx* ctx, const xtx_params* params, const xtx_serve_config* cfg);
/* Queues one request; it is admitted at the next epoch boundary at which the
 * LDS/HBM budget allows. The instance's buffers must stay valid until its
 * result has been polled. */
int xtx_serve_submit(xtx_ctx* ctx, const xtx_instance* inst, uint64_t tag);
/* Runs one epoch over the resident requests, retiring solved ones and
 * admitting queued ones. *resident / *queued report the state after it. */
int xtx_serve_step(xtx_ctx* ctx, int32_t* resident, int32_t* queued);
/* Retired results since the last poll (at most max). */
int xtx_serve_poll(xtx_ctx* ctx, xtx_result* out, int32_t max, int32_t* count);
/* Force-retires a resident or queued request (its best so far is written). */
int xtx_serve_retire(xtx_ctx* ctx, uint64_t tag);
int xtx_serve_close(xtx_ctx* ctx);

/* ---- exact host verification ----
 * X^T A0 X == A (int128 when the bound allows, else multi-modular CRT),
 * A diagonal, |det X| == 1 (det mod primes within the Hadamard bound). */
int xtx_verify(const int64_t* A0, const int64_t* X, const int64_t* A, int32_t n,
               xtx_verify_report* report);

/* ---- host-side parts of xtax around the annealer (no device needed) ---- */
#define XTX_REORDER_RCM 1       /* --rcm */
#define XTX_REORDER_CENTROID 2  /* --centroid */
/* Publish-time symmetric reordering, in place: A <- A[p][p], X <- X[:, p]
 * (X may be NULL), Reverse Cuthill-McKee first, then the iterated row centre
 * of mass, as flags ask. Score- and nonzero-preserving. perm (nullable)
 * receives the composite permutation (new index a holds old index perm[a]).
 * Replaces L1Objective::reorder_for_publish (main.cpp:407-411) with
 * rcm_order (main.cpp:249-286), centroid_order (:292-324) and permute_state
 * (:330-350). */
int xtx_reorder(int64_t* A, int64_t* X, int32_t n, int32_t flags, int32_t* perm);
/* det(A) = +-1 by the reference's test: det mod 1000000007 and mod 998244353
 * both 1 or both -1 (is_unimodular, mat_io.hpp:635-644). */
int xtx_is_unimodular(const int64_t* A, int32_t n, int32_t* result);
/* Strict-deflation pivot clearing (deflate_index, main.cpp:435-455): with
 * M_ii = +-1, shear every active j != i by s = -M_ji / M_ii so that row and
 * column i of M become zero off the diagonal, updating M = X^T A0 X and
 * Xt = X^T. If any entry would exceed 2^48, M and Xt are restored and
 * *cleared = 0; otherwise *cleared = 1. */
int xtx_deflate_pivot(int64_t* M, int64_t* Xt, int32_t n, int32_t i, const int32_t* active,
                      int32_t n_active, int32_t* cleared);

/* ---- xdual: the primal/dual objective on the same engine ----
 * Replaces run_xdual (src/xdual_main.cpp:378-424), i.e.
 * run_annealer(DualObjective(P, P^-1, X0, c), params, rows): minimize
 * |offdiag P|^2 + c |offdiag Q|^2 over P = X^T A0 X (exact int64, the same
 * shears and 2^48 bound) with the double dual Q = P^-1 updated by the inverse
 * transpose of every shear. The dual is inverted once on the host (invert_to,
 * mat_io.hpp:297-337); c = lambda * |offdiag P0|^2 / |offdiag Q0|^2 (long
 * double, :388-393). Engine parameters are the xtx_params of the L1 engine
 * (waves_per_block, autotune and the launch configurations do not apply). */
typedef struct {
    double lambda;            /* 1.0: equal initial primal/dual weight (--lambda) */
    double dual_tol;          /* 1e-6: residual that triggers an early re-inversion */
    double lambda_ramp;       /* 0: seconds to ramp c from 0 to full (--lambda-ramp);
                                 applied per epoch from the host clock */
    int32_t dual_refresh;     /* 200000: moves between re-inversions (0 = never) */
    int32_t dual_check;       /* 25000: moves between sampled residual checks (0 = never) */
} xtx_dual_params;

typedef struct {
    int64_t moves;
    int64_t accepts;
    int64_t greedy_calls;
    int64_t infeasible;
    int64_t sweeps;
    int64_t exchanges_tried;
    int64_t exchanges_accepted;
    int64_t refreshes;        /* dual re-inversions (all replicas) */
    int64_t checks;           /* sampled residual checks */
    int64_t recomputes;       /* running-score recomputations before a new best */
    int32_t solved;           /* the primal P reached a diagonal */
    int32_t solve_replica;
    int64_t solve_move;
    double best_score;        /* |offdiag P|^2 + c |offdiag Q|^2 of the returned
                                 state at the full weight c */
    double start_score;       /* the same of the start state */
    double start_primal;      /* |offdiag P|^2 of the start state */
    double start_dual;        /* |offdiag Q|^2 of the start state */
    double c;                 /* normalized dual weight */
    double dual_residual;     /* max |P Q - I| of the returned state */
    int32_t best_replica;
    int32_t epochs;
    int32_t stop_reason;      /* as xtx_stats.stop_reason */
    int32_t verify_status;    /* X^T A0 X == P and |det X| = 1 (and P diagonal
                                 when solved); 0 ok, -1 not run */
    double seconds;
    double kernel_ms;
} xtx_dual_stats;

void xtx_dual_params_default(xtx_dual_params* d);
/* A0 (n x n, symmetric, |entries| <= 2^48), X0 (nullable = identity): the
 * annealer starts from P = X0^T A0 X0. Writes the best (or solving) P, Q and
 * X to the caller's buffers (each nullable). XTX_E_ARG if P is singular. */
int xtx_dual_solve(xtx_ctx* ctx, const int64_t* A0, const int64_t* X0, int32_t n,
                   const xtx_params* params, const xtx_dual_params* dparams,
                   const int32_t* active, int32_t n_active, int64_t* P_out, double* Q_out,
                   int64_t* X_out, xtx_dual_stats* stats);
/* Per-replica state of the last xtx_dual_solve: P, Xt (R*n*n int64), Q
 * (R*n*n double), 4 doubles {score, best_score, local_best, T} and 8 int64
 * {offnz, moves, best_move, rung, accepts, refreshes, checks, moves since
 * improvement} per replica (each pointer nullable). Its move log is read with
 * xtx_movelog_get. */
int xtx_dual_last_replicas(xtx_ctx* ctx, int64_t* P, double* Q, int64_t* Xt, double* dscal,
                           int64_t* scal);
/* The primal/dual move primitives as the device computes them, for queries
 * (matrix index, i, j, s) on n_mats matrices P (int64) and Q (double): per
 * query out_i = {feasible, d_nonzero, best_shear(i, j)}, out_d = delta
 * (d_off_p + c d_off_q), the proposed P row j and Q row i (n each), and the
 * score |offdiag P|^2 + c |offdiag Q|^2 of the matrix (each pointer nullable
 * except out_i). */
int xtx_dual_probe(xtx_ctx* ctx, const int64_t* P, const double* Q, int32_t n, int32_t n_mats,
                   double c, const int32_t* queries, int32_t n_queries, int64_t* out_i,
                   double* out_d, int64_t* rows_p, double* rows_q, double* scores);

/* ---- device primitives, for parity tests of the kernels' building blocks ----
 * Each query q is (matrix index, i, j, s); the device evaluates it with the
 * same device functions the annealer uses. out per query (6 values):
 *   {feasible, delta, d_nonzero, best_shear(i,j), xt_feasible,
 *    delta of Add(i, j, best_shear) as the annealer evaluates it after a
 *    greedy call (0 when best_shear is 0)} */
int xtx_probe_moves(xtx_ctx* ctx, const int64_t* A, const int64_t* Xt, int32_t n,
                    int32_t n_mats, const int32_t* queries, int32_t n_queries,
                    int64_t* out);
/* Philox words and exp values as the device computes them. */
int xtx_probe_rng(xtx_ctx* ctx, const uint32_t* ctr_key, int32_t count, uint32_t* out,
                  const double* x, int32_t nx, double* exp_out);

#ifdef __cplusplus
}
#endif
#endif /* XTX_H */
