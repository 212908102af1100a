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

// Whole 128-byte cache lines per record: the persistent launch runs different
// instances on different XCDs, whose L2s are not coherent with each other.
struct alignas(128) Inst {
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

// The batch's persistent launch (anneal_persist.hip). A work unit is one
// wave's replicas of one instance (64 / Q replicas from r0, Q the lane
// kernel's proposals per step at order nmax) for one epoch; the units of an
// instance are consecutive from ubase[inst].
struct PUnit {
    int32_t inst;
    int16_t r0;
    int16_t nmax;
};

// Queues and counters of a persistent launch, in device memory. Queues per
// XCD (an accelerator die with its own L2), order class (orders 8-16, 24-32,
// 40-48) and epoch slot (the epoch a unit runs, modulo kEpochSlots): an
// instance's units only ever run on the XCD whose queues hold them, so what
// one unit leaves for the next passes through that XCD's L2 (stores complete,
// L1 invalidated) with no L2 write-back or invalidation. Queue qi = (x
// kClasses + k) kEpochSlots + b holds units in q[qbase[qi] ..] (a whole line)
// in the order they became ready (-1: not yet written; else the unit id with
// the epoch it runs); [head, tail) are the units not yet taken. A wave of
// hardware XCD id h serves queue set xq[h] and takes the unit of the lowest
// epoch ready there (the instance furthest behind, whose remaining epochs
// decide when the batch ends), from its home class (home[x 128 + CU-pair
// key]: the two compute units sharing an instruction cache keep to one
// class's code) while that class is at most one epoch later. The epoch end of
// an instance appends its next epoch's units to its queue. pend[inst]: units
// of the instance's current epoch not yet done; live[x]: instances on XCD x
// that can still append units. An instance whose epoch reaches epoch_limit
// appends nothing (the host launches again).
constexpr int kMaxXcd = 16;
constexpr int kClasses = 3;
constexpr int kEpochSlots = 8;
constexpr int kQueues = kMaxXcd * kClasses * kEpochSlots;
constexpr int kCuKeys = 128;
// int32 words per 128-byte cache line: counters that different XCDs update
// (the queues' counters, each instance's pending count) are a line apart
constexpr int kLineWords = 32;
#ifdef __HIPCC__
__host__ __device__
#endif
inline int order_class(int nmax) { return nmax <= 16 ? 0 : nmax <= 32 ? 1 : 2; }
struct PersistCtl {
    int32_t* q;
    const uint32_t* qbase;  // [qi]
    const uint32_t* qcap;   // [qi]
    const PUnit* units;
    const int32_t* ubase;
    const int32_t* unum;
    const int8_t* home;  // [x * kCuKeys + key]
    int32_t* pend;       // [inst * kLineWords]
    uint32_t* head;      // [qi * kLineWords]
    uint32_t* tail;      // [qi * kLineWords]
    int32_t* live;       // [x * kLineWords]
    int32_t* err;        // 1: an append did not fit its queue
    // host-mapped outcomes: an instance that ends has its outcome written at
    // out + out_off[inst] before its done flag is raised (null: none)
    unsigned char* out;
    const int64_t* out_off;
    int32_t xq[kMaxXcd];  // hardware XCD id -> queue set, -1: none
    int32_t nq;
    uint32_t epoch_limit;
    uint32_t epoch_base;  // the lowest epoch of the launch's instances
};

}  // namespace xtx
