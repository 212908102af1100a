// engine_types.h -- records shared by the host engine and the gfx950 kernels.
//
// Layout in HBM (one instance = one matrix being diagonalized, R replicas):
//   A, Xt      R * n * n int64   current state of every replica (row-major)
//   bA, bXt    R * n * n int64   best state of every replica (lazily flushed)
//   gA, gXt    n * n int64       instance-best snapshot (reseed source)
//   rs         R RepState        per-replica scalars
//   temps      R double          tempering ladder, rung -> temperature
//   act        N uint8           active indices (the reference's `active` list)
//   log        R * log_cap       accepted-move log per replica (optional)
// During an epoch a replica's A lives in LDS (row stride P, odd) and its row
// masses in registers (A returns to HBM at the epoch end); X^T stays in HBM.
// An instance of the primal/dual objective adds Q, bQ (R * n * n double), gQ
// (n * n) and a Gauss-Jordan work area (R * n * 2n double).
#pragma once
#include <stdint.h>

namespace xtx {

constexpr int64_t kLimit = (int64_t)1 << 48;  // mat_io.hpp:35 (strict >)
constexpr int kShearCap = 64;                 // mat_io.hpp:38
constexpr int kMaxN = 256;       // largest order (the active list holds bytes)
constexpr int kMaxGroupN = 128;  // largest order with A in LDS (all but the wide lane kernel)
constexpr int kMaxSamples = 16;     // tournament samples of every kernel but the wide mode
constexpr int kMaxSamplesWide = 4096;  // the wide mode's (its pivot key holds 12 bits)
constexpr int kRepBits = 20;                  // solve key = move << 20 | replica

// Weighted-median histogram of the greedy shear (one per replica, LDS): bin
// t + 64 holds the weight of breakpoints with ceiling t, t in [-64, 64]. A
// group of G lanes scans it with med_per_lane(G) consecutive bins per lane;
// med_slots(G) bins are allocated (whole lanes).
constexpr int kMedBins = 2 * kShearCap + 1;
constexpr int med_per_lane(int G) { return (kMedBins + G - 1) / G; }
constexpr int med_slots(int G) {
    return (kMedBins + med_per_lane(G) - 1) / med_per_lane(G) * med_per_lane(G);
}
constexpr unsigned long long kNoKey = ~0ull;

enum InstStatus : int32_t {
    ST_IDLE = 0,
    ST_RUNNING = 1,
    ST_SOLVED = 2,
    ST_EXHAUSTED = 3,  // every replica hit max_moves
    ST_STOPPED = 4,    // host stop / time budget / forced retire
};

enum RepFlags : int32_t {
    RF_CUR_BEST = 1,    // current state is this replica's best (not flushed)
    RF_STOPPED = 2,
    RF_SOLVED = 4,
    RF_BEST_STALE = 8,  // bA/bXt do not hold best_score's state (cannot win)
    RF_CAPPED = 16,
};

struct LogEnt {
    int64_t move;
    int16_t i, j, s;
    uint16_t flags;
};

struct RepState {
    int64_t score;
    int64_t offnz;
    uint64_t moves;
    int64_t local_best;
    int64_t msi;           // moves since improvement
    double T;              // cooling-mode temperature
    double cooling;        // cooling-mode factor (adaptive)
    int32_t cool_counter;
    int32_t win_p;
    int32_t win_a;
    int32_t rung;          // tempering rung
    int64_t best_score;
    int64_t best_move;
    int32_t flags;
    int32_t pad0;
    double gf;             // per-replica greedy fraction (diversity applied)
    double tf;             // per-replica target fraction
    int64_t accepts;
    int64_t greedy;
    int64_t infeasible;
    int64_t sweeps;
    int64_t log_count;
    int64_t solve_move;
    // primal/dual objective (Inst::dual): the score is a double
    double dscore;
    double dlocal_best;
    double dbest_score;
    double dc;             // dual weight the running score was computed with
    uint64_t m_maint;      // move count at the last maintenance
    int64_t refreshes;     // dual re-inversions
    int64_t checks;        // sampled residual checks
    int64_t recomputes;    // running-score recomputations before a new best
};

struct InstParams {
    double t_init;         // effective (auto already resolved)
    double t_min;
    double cooling;
    double reheat;
    double reseed_factor;
    double target_accept;
    double sweep_fraction;
    int32_t moves_per_cool;
    int32_t stuck_threshold;
    int32_t target_samples;
    int32_t tempering;
    int32_t adaptive;
    int32_t epoch_moves;
    uint64_t max_moves;
    int32_t max_epochs;    // 0 = none: the instance is exhausted after this many
    int32_t pad0;
};

struct Inst {
    int32_t n;
    int32_t N;             // active count
    int32_t R;
    int32_t status;
    uint32_t key0, key1;   // Philox key
    uint32_t epoch;
    int32_t log_on;
    int64_t* A;
    int64_t* Xt;
    int64_t* bA;
    int64_t* bXt;
    int64_t* gA;
    int64_t* gXt;
    RepState* rs;
    double* temps;
    uint8_t* act;
    LogEnt* log;
    int64_t log_cap;
    InstParams P;
    unsigned long long solve_key;
    int64_t gb_score;      // instance best score at the last epoch end
    int64_t gb_offnz;
    int32_t gb_rep;
    int32_t solve_rep;
    int64_t solve_move;
    int64_t exch_tried;
    int64_t exch_acc;
    // primal/dual objective (xdual's DualObjective): P lives in A/bA/gA, the
    // double dual Q = P^-1 in Q/bQ/gQ (R * n * n, n * n), work holds one
    // n x 2n Gauss-Jordan matrix per replica
    int32_t dual;
    int32_t dual_refresh;  // moves between re-inversions (0 = never)
    int32_t dual_check;    // moves between residual checks (0 = never)
    int32_t pad1;
    double dual_tol;
    double c;              // dual weight (normalized)
    double gb_dscore;      // instance best score at the last epoch end
    double* Q;
    double* bQ;
    double* gQ;
    double* work;
    // host-mapped flag the epoch end raises when the instance leaves
    // ST_RUNNING (null: none), so a host can collect it while later epochs run
    int32_t* done_flag;
};

// Order-preserving map of a double onto int64 (for a non-NaN x and y,
// x < y <=> okey(x) < okey(y); -0 and +0 map to the same key).
#ifdef __HIPCC__
__host__ __device__
#endif
inline int64_t okey(double x) {
    x = x + 0.0;
    const int64_t b = __builtin_bit_cast(int64_t, x);
    return b ^ ((b >> 63) & 0x7FFFFFFFFFFFFFFFll);
}

struct RepRef {
    int32_t inst;
    int32_t rep;
};

}  // namespace xtx
