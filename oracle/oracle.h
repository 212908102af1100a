/* oracle.h -- CPU oracle for the xtax L1 congruence annealer.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product (xtax_amd/, include/) links,
 * loads or calls this library. It is the checker the parity tests, smoke() and
 * bench.py's cpu_baseline leg compare the HIP engine against.
 *
 * Contents, each restating the reference (flinty-mathematics/xtax) in code of
 * our own:
 *   - exact int64 move primitives of the L1 objective    (src/main.cpp:46-242)
 *   - xdual's primal/dual move primitives and dual     (src/xdual_main.cpp:67-364,
 *     inverse, in the engine's double specification      src/mat_io.hpp:297-337)
 *   - Matrix helpers: score, nonzero count, congruence,  (src/mat_io.hpp:41-48,
 *     Gram, X^T L                                         219-235, 340-372, 545-580)
 *   - Philox4x32-10 (Salmon et al., SC'11, "Parallel random numbers: as easy as
 *     1, 2, 3"), pinned by the Random123 known-answer vectors in tests/
 *   - a deterministic exp() (Cody-Waite reduction + Taylor, all fma) that the
 *     engine's Metropolis test is specified with
 *   - an exact verifier: X^T A0 X == A checked modulo enough 61-bit primes to
 *     cover the integer bound (CRT), det X mod the same primes
 *   - move-log replay through the restated primitives
 *   - a trajectory model of the engine's epoch-synchronous algorithm (the
 *     reference's anneal_worker loop, congruence_anneal.hpp:177-416, with the
 *     reference's mt19937 replaced by counter-based Philox and its 5 ms monitor
 *     exchange replaced by an exchange every `exchange_interval` moves), whose
 *     loop body also runs on the reference's own draws (std::mt19937 and
 *     libstdc++'s distributions) and publish rule: in that mode it reproduces
 *     the reference's single-threaded runs and its worker loop exactly, which
 *     pins the loop body the engine is tested against
 *
 * Each function is pinned by tests/test_oracle_golden.py (golden vectors from
 * the reference's own code, closed forms, known answers).
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAGNITUDE_LIMIT (((int64_t)1) << 48) /* mat_io.hpp:35 */
#define ORC_SHEAR_CAP 64                          /* mat_io.hpp:38 */

/* ---- primitives (restated from the reference) ---- */
int64_t orc_rounded_div(int64_t a, int64_t b);
int64_t orc_score(const int64_t* A, int32_t n);
int64_t orc_offdiag_nonzero(const int64_t* A, int32_t n);
int32_t orc_try_add(const int64_t* A, const int64_t* Xt, int32_t n, int32_t i, int32_t j,
                    int64_t s, int64_t* new_a_row, int64_t* new_xt_row /* nullable */,
                    int64_t* delta, int64_t* d_nonzero);
int32_t orc_add_transform_row(const int64_t* Xt, int32_t n, int32_t i, int32_t j, int64_t s,
                              int64_t* new_xt_row);
void orc_commit_add(int64_t* A, int64_t* Xt, int32_t n, int32_t j, const int64_t* new_a_row,
                    const int64_t* new_xt_row, int64_t* row_mass);
void orc_compute_row_mass(const int64_t* A, int32_t n, int64_t* row_mass);
int64_t orc_score_delta_l1(const int64_t* A, int32_t n, int32_t i, int32_t j, int64_t s);
int64_t orc_best_shear_l1(const int64_t* A, int32_t n, int32_t i, int32_t j);
/* congruence_of: out = X^T A X (int64, unchecked like the reference) */
void orc_congruence_of(const int64_t* A, const int64_t* X, int32_t n, int64_t* out);
void orc_gram_of(const int64_t* L, int32_t m, int32_t d, int64_t* out);
void orc_xt_times(const int64_t* X, const int64_t* L, int32_t m, int32_t d, int64_t* out);

/* ---- xdual primal/dual objective (oracle_dual.cpp), in the engine's double
 * specification: pairwise column sums, exact primal squares rounded once ---- */
double orc_pairsum(const double* x, int32_t n);
int32_t orc_dual_eval(const int64_t* P, const double* Q, int32_t n, int32_t i, int32_t j,
                      int64_t s, int64_t* new_p_row, double* new_q_row, double* d_off_p,
                      double* d_off_q, int64_t* d_nonzero);
int64_t orc_dual_best_shear(const int64_t* P, const double* Q, int32_t n, double c, int32_t i,
                            int32_t j);
int32_t orc_invert(const int64_t* P, int32_t n, double* Q);
double orc_frob2_off_p(const int64_t* P, int32_t n);
double orc_frob2_off_q(const double* Q, int32_t n);
double orc_dual_score(const int64_t* P, const double* Q, int32_t n, double c);
double orc_dual_residual(const int64_t* P, const double* Q, int32_t n, int32_t k);

/* ---- engine building blocks ---- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double orc_xexp(double x);

/* ---- exact verification ----
 * status bits: 1 = X^T A0 X != A, 2 = A not diagonal, 4 = |det X| != 1,
 *              8 = |det X| = 1 not decidable within the prime budget.
 */
typedef struct {
    int32_t status;
    int32_t primes_used;
    int32_t det_sign; /* +1 / -1 when |det X| = 1 */
    int32_t reserved;
} orc_verify_report;
int32_t orc_verify(const int64_t* A0, const int64_t* X, const int64_t* A, int32_t n,
                   orc_verify_report* rep);

/* ---- move-log replay ----
 * Entry layout (16 bytes, shared with the engine's log format):
 *   int64 move; int16 i; int16 j; int16 s; uint16 flags (bit0 = sweep commit,
 *   bit1 = reseed from the instance-best snapshot, i = j = -1)
 * Applies the moves in order to (A, Xt) with the restated primitives; returns
 * the index of the first entry that is infeasible or a reseed marker (or count
 * on success).
 */
typedef struct {
    int64_t move;
    int16_t i, j, s;
    uint16_t flags;
} orc_log_entry;
int64_t orc_replay(int64_t* A, int64_t* Xt, int32_t n, const orc_log_entry* log, int64_t count);

/* ---- trajectory model of the engine ---- */
typedef struct {
    double t_init;          /* <= 0: auto (suggest_t_init of the start state) */
    double t_min;
    double cooling;
    double reheat;
    double reseed_factor;
    double greedy_fraction;
    double target_fraction;
    double target_accept;
    double worker_diversity;
    double sweep_fraction;
    int32_t moves_per_cool;
    int32_t stuck_threshold;
    int32_t target_samples;
    int32_t tempering;
    int32_t adaptive_cooling;
    int32_t exchange_interval;
    int32_t replicas;
    int32_t reserved;
    uint64_t seed;
    uint64_t max_moves;     /* per replica; 0 = unbounded */
} orc_params;

typedef struct {
    int32_t solved;          /* 1 if some replica reached a diagonal */
    int32_t solve_replica;
    int64_t solve_move;
    int64_t best_score;      /* instance best (argmin of score, move, replica) */
    int32_t best_replica;
    int32_t epochs;
    int64_t total_moves;     /* proposals executed by all replicas */
    int64_t accepts;
    int64_t exchanges;
} orc_model_result;

/* Runs the model for at most max_epochs epochs (stops at the first epoch in
 * which the instance is solved). Per-replica outputs (each may be null):
 *   A_out, Xt_out   R*n*n  final state of each replica
 *   best_A, best_Xt R*n*n  best state of each replica
 *   scal_out        R*8    {score, offnz, moves, best_score, best_move, rung,
 *                           accepts, local_best}
 * log_out/log_count: per-replica accepted-move logs, log_cap entries each. */
int32_t orc_model_run(const int64_t* A_start, const int64_t* Xt_start, int32_t n,
                      const int32_t* active, int32_t n_active, const orc_params* p,
                      int32_t max_epochs, int64_t* A_out, int64_t* Xt_out,
                      int64_t* best_A, int64_t* best_Xt, int64_t* scal_out,
                      orc_log_entry* log_out, int64_t* log_count, int64_t log_cap,
                      orc_model_result* res);

/* ---- trajectory model of the engine on xdual's objective ---- */
typedef struct {
    double c;               /* dual weight, already normalized */
    double dual_tol;        /* residual that triggers an early re-inversion */
    int32_t dual_refresh;   /* moves between re-inversions (0 = never) */
    int32_t dual_check;     /* moves between residual checks (0 = never) */
} orc_dual_params;

typedef struct {
    int32_t solved;
    int32_t solve_replica;
    int64_t solve_move;
    double best_score;       /* of the returned state */
    int32_t best_replica;    /* replica whose state is returned */
    int32_t epochs;
    int64_t total_moves;
    int64_t accepts;
    int64_t exchanges;
    int64_t refreshes;
    int64_t checks;
    int64_t recomputes;
} orc_dual_result;

/* The model on (P, Q, X^T) for at most max_epochs epochs. Per replica (each
 * output may be null): P_out/Xt_out R*n*n int64, Q_out R*n*n double,
 * dscal_out R*4 {score, best_score, local_best, T}, scal_out R*8 {offnz,
 * moves, best_move, rung, accepts, refreshes, checks, msi}; best_P/best_Q/
 * best_Xt the returned state (the solving replica's, else the instance best). */
int32_t orc_dual_model_run(const int64_t* P_start, const double* Q_start,
                           const int64_t* Xt_start, int32_t n, const int32_t* active,
                           int32_t n_active, const orc_params* p, const orc_dual_params* dp,
                           int32_t max_epochs, int64_t* P_out, double* Q_out, int64_t* Xt_out,
                           double* dscal_out, int64_t* scal_out, int64_t* best_P,
                           double* best_Q, int64_t* best_Xt, orc_log_entry* log_out,
                           int64_t* log_count, int64_t log_cap, orc_dual_result* res);

/* ---- reference-draw mode of the trajectory model ----
 * One worker of the reference's annealer (anneal_worker,
 * congruence_anneal.hpp:177-416) run through the model's own loop body, with
 * the reference's draws in place of Philox: std::mt19937 seeded with the
 * worker seed, libstdc++'s uniform_int_distribution / uniform_real_distribution,
 * the redraw loops, the raw-draw sign bit and std::exp(-d / T); and with the
 * reference's publish rule (the global best is replaced by a strictly better
 * state after an accepted downhill move; the run stops when that state is
 * diagonal). base.replicas is the pool size (threads: it sets the diversity
 * offsets and, below 2, turns tempering off), thread_id the worker, rung_T its
 * fixed tempering rung (no exchange). base.max_moves caps the run (0: until
 * solved). Outputs: the worker's final state, the global best, the accepted
 * moves (flag bit 0 sweep commit, bit 1 reseed). */
typedef struct {
    orc_params base;
    uint32_t worker_seed;
    int32_t thread_id;
    double rung_T;
} orc_ref_params;

typedef struct {
    int32_t solved;
    int32_t reserved;
    int64_t moves;       /* proposals, the reference's "done moves=" */
    int64_t accepts;
    int64_t score;       /* final state */
    int64_t gb_score;    /* global best */
} orc_ref_result;

typedef struct {
    int32_t solved;
    int32_t reserved;
    int64_t moves;        /* the reference's "done moves=" */
    int64_t accepts;
    int64_t refreshes;
    int64_t checks;
    int64_t recomputes;
    double score;         /* final running score */
    double gb_score;      /* global best score */
    double c;             /* the scaled dual weight */
} orc_dual_ref_result;

/* seed_rng() ^ t * 0x9e3779b9 for worker t (congruence_anneal.hpp:484-488) */
uint32_t orc_ref_worker_seed(uint64_t seed, int32_t t);
int32_t orc_ref_worker_run(const int64_t* A_start, const int64_t* Xt_start, int32_t n,
                           const int32_t* active, int32_t n_active, const orc_ref_params* p,
                           int64_t* A_out, int64_t* Xt_out, int64_t* gb_A, int64_t* gb_Xt,
                           orc_log_entry* log_out, int64_t* log_count, int64_t log_cap,
                           orc_ref_result* res);

/* The dual model's loop body in the reference's mode: one worker of xdual's
 * annealer (its long double arithmetic, mt19937 draws from p->worker_seed,
 * publish rule and per-move maintenance; the diversity offsets of worker
 * p->thread_id of p->base.replicas; tempering on the fixed rung p->rung_T).
 * dp->c is unused: the weight is lambda scaled by the start's primal/dual
 * norm ratio. Returns 1 if the run stopped at a diagonal, 0 at the move cap,
 * -1 if P is singular. P/Q/Xt_out: the worker's final state; gb_P/gb_Q/gb_Xt:
 * the global best (the reference's best_P, best_Q, best_X^T). */
int32_t orc_dual_ref_run(const int64_t* P_start, const int64_t* Xt_start, int32_t n,
                         const int32_t* active, int32_t n_active, const orc_ref_params* p,
                         const orc_dual_params* dp, double lambda, int64_t* P_out, double* Q_out,
                         int64_t* Xt_out, int64_t* gb_P, double* gb_Q, int64_t* gb_Xt,
                         orc_log_entry* log_out, int64_t* log_count, int64_t log_cap,
                         orc_dual_ref_result* res);

/* The model's replica-exchange sweep after epoch `epoch` (rung[r] = rung of
 * replica r, updated in place; score[r] its current score); returns the
 * number of swaps. */
int64_t orc_exchange_sweep(const double* temps, const double* score, int32_t* rung, int32_t R,
                           uint64_t seed, uint64_t epoch);
/* Test hook: plants one of a few plausible mistakes in the models (0 = none,
 * 1 strict auto-accept, 2 reversed hot-row tournament, 3 diversity offset
 * sign, 4 exchange argument sign, 5 Metropolis exponent sign; 1, 2 and 5 in
 * the dual model's loop body too; 6 the dual model's Q update committed to
 * row and column j instead of i), so tests can show that the checks pinning
 * the models reject each. */
void orc_set_mutation(int32_t which);
int32_t orc_get_mutation(void);
/* The model's Metropolis decision for a proposal of score change d at
 * temperature T, given the move's two uniform words: 1 = accept. */
int32_t orc_metropolis(int64_t d, double T, uint32_t u_hi, uint32_t u_lo);

/* Ladder and per-replica fractions exactly as the model derives them (exposed
 * so tests can check the engine's host setup). */
void orc_ladder(double top, double t_min, int32_t R, double* temps);
void orc_fractions(const orc_params* p, int32_t r, double* greedy, double* target);
double orc_suggest_t_init(const int64_t* A, int32_t n);

#ifdef __cplusplus
}
#endif
