// ref_harness_dual.cpp -- exposes the reference's own xdual primal/dual
// primitives (src/xdual_main.cpp) and its dual inverse (invert_to,
// src/mat_io.hpp) through a C ABI so tests/golden/gen_golden.py can record
// golden vectors from them.
//
// Built by `make -C oracle ref` into oracle/_ref/ (git-ignored) directly from
// the reference sources under REF; nothing of the reference is copied here. The
// reference's main() is renamed so this file can be a shared library.
#define main xdual_reference_main
#include "xdual_main.cpp"
#undef main

#include <cmath>
#include <cstring>

namespace {
Matrix to_matrix(const int64_t* p, int n) {
    Matrix M((size_t)n);
    std::memcpy(M.data.data(), p, sizeof(int64_t) * (size_t)n * (size_t)n);
    return M;
}
Matrixd to_matrixd(const double* p, int n) {
    Matrixd M((size_t)n);
    std::memcpy(M.data.data(), p, sizeof(double) * (size_t)n * (size_t)n);
    return M;
}
}  // namespace

extern "C" {

// eval_add (primal, Add(i, j, s)) and eval_add_d (dual, Add(j, i, -s)), as
// DualObjective::evaluate combines them; the long double deltas as doubles.
int refd_eval(const int64_t* P, const double* Q, int n, int i, int j, int64_t s,
              int64_t* new_p_row, double* new_q_row, double* d_off_p, double* d_off_q,
              int64_t* dnz) {
    const Matrix MP = to_matrix(P, n);
    const Matrixd MQ = to_matrixd(Q, n);
    std::vector<int64_t> prow((size_t)n, 0);
    std::vector<double> qrow((size_t)n, 0.0);
    long double dp = 0.0L, dq = 0.0L;
    int64_t nz = 0;
    if (!eval_add(MP, i, j, s, prow, dp, nz)) return 0;
    eval_add_d(MQ, j, i, -s, qrow, dq);
    std::memcpy(new_p_row, prow.data(), sizeof(int64_t) * (size_t)n);
    std::memcpy(new_q_row, qrow.data(), sizeof(double) * (size_t)n);
    *d_off_p = (double)dp;
    *d_off_q = (double)dq;
    *dnz = nz;
    return 1;
}

int64_t refd_best_shear(const int64_t* P, const double* Q, int n, double c, int i, int j) {
    Matrix X((size_t)n);
    X.fill_identity();
    const DualObjective obj(to_matrix(P, n), to_matrixd(Q, n), X, (long double)c);
    return obj.best_shear(i, j);
}

int refd_invert(const int64_t* P, int n, double* Q) {
    Matrixd MQ;
    if (!invert_to(to_matrix(P, n), MQ)) return 0;
    std::memcpy(Q, MQ.data.data(), sizeof(double) * (size_t)n * (size_t)n);
    return 1;
}

double refd_frob2_off_p(const int64_t* P, int n) { return (double)to_matrix(P, n).offdiag_frob2(); }

double refd_frob2_off_q(const double* Q, int n) { return (double)to_matrixd(Q, n).offdiag_frob2(); }

}  // extern "C"

// ---- one worker of the reference's annealer on xdual's objective, recorded ----
//
// canneal::anneal_worker (congruence_anneal.hpp:177-416) itself, run on the
// calling thread with a recording DualObjective, as ref_harness.cpp records
// it on the L1 objective: commits, reseeds and the per-move maintenance calls
// are logged (the attempt's evaluate is the last evaluate before a move's
// maintenance call, so every event is assigned to its move: accepted moves
// flags 0, sweep commits 1, reseeds 2), the maintenance call raises the done
// flag at a move cap, then runs the objective's own maintenance.

namespace {

struct DEv {
    char kind;  // 'E' evaluate, 'C' commit, 'R' reseed
    int16_t i, j, s;
};

struct DRec {
    int64_t move;
    int16_t i, j, s;
    uint16_t flags;
};

std::vector<DEv> d_pending;
std::vector<DRec> d_trace;
uint64_t d_moves = 0;
uint64_t d_cap = 0;
std::atomic<bool>* d_done = nullptr;
const void* d_cur = nullptr;
Matrix d_finalP, d_finalXt;
Matrixd d_finalQ;
long double d_final_score = 0.0L;

void d_flush(uint64_t before_move, uint64_t split_move, bool split) {
    size_t last_e = d_pending.size();
    if (split)
        for (size_t k = d_pending.size(); k-- > 0;)
            if (d_pending[k].kind == 'E') {
                last_e = k;
                break;
            }
    for (size_t k = 0; k < d_pending.size(); ++k) {
        const bool attempt = split && k >= last_e;
        const uint64_t m = attempt ? split_move : before_move;
        const DEv& e = d_pending[k];
        if (e.kind == 'C')
            d_trace.push_back({(int64_t)m, e.i, e.j, e.s, (uint16_t)(attempt ? 0 : 1)});
        else if (e.kind == 'R')
            d_trace.push_back({(int64_t)m, -1, -1, 0, 2});
    }
    d_pending.clear();
}

struct RecDual : DualObjective {
    using DualObjective::DualObjective;
    RecDual() = default;
    RecDual(const RecDual&) = default;
    RecDual(RecDual&&) = default;
    RecDual& operator=(RecDual&&) = default;
    RecDual& operator=(const RecDual& o) {
        if (this == d_cur) d_pending.push_back({'R', -1, -1, 0});
        DualObjective::operator=(o);
        return *this;
    }
    ~RecDual() {
        if (this != d_cur) return;
        const bool capped = d_cap && d_moves >= d_cap;
        d_flush(d_moves, d_moves + 1, !capped);
        d_finalP = P;
        d_finalQ = Q;
        d_finalXt = Xt;
        d_final_score = score_;
        d_cur = nullptr;
    }
    bool evaluate(int i, int j, int64_t s, score_t& d, int64_t& dnz) {
        if (this == d_cur) d_pending.push_back({'E', (int16_t)i, (int16_t)j, (int16_t)s});
        return DualObjective::evaluate(i, j, s, d, dnz);
    }
    bool commit(int i, int j, int64_t s, score_t d, int64_t dnz) {
        const bool ok = DualObjective::commit(i, j, s, d, dnz);
        if (ok && this == d_cur) d_pending.push_back({'C', (int16_t)i, (int16_t)j, (int16_t)s});
        return ok;
    }
    // the worker refreshes its own state's cache first (:194)
    void refresh_cache() {
        if (d_cur == nullptr) d_cur = this;
        DualObjective::refresh_cache();
    }
    void periodic_maintenance(uint64_t moves) {
        d_flush(moves - 1, moves, true);
        d_moves = moves;
        if (d_cap && moves >= d_cap) d_done->store(true, std::memory_order_relaxed);
        DualObjective::periodic_maintenance(moves);
    }
};

void put(const Matrix& M, int64_t* p) {
    std::memcpy(p, M.data.data(), sizeof(int64_t) * M.n * M.n);
}
void putd(const Matrixd& M, double* p) {
    std::memcpy(p, M.data.data(), sizeof(double) * M.n * M.n);
}

}  // namespace

extern "C" {

// dp: t_init, t_min, cooling, reheat, reseed_factor, greedy_fraction,
//     target_fraction, target_accept, worker_diversity, sweep_fraction, rung_T,
//     lambda, dual_tol
// ip: moves_per_cool, stuck_threshold, target_samples, tempering,
//     adaptive_cooling, threads, thread_id, worker seed, move cap (> 0),
//     dual_refresh, dual_check
// Start: P_start (the working Gram) and X (the start transform); the dual is
// the reference's inverse and the weight lambda scaled by the start's norms
// (run_xdual, xdual_main.cpp:378-404). Outputs: the worker's final P, Q, Xt;
// the global best's P, Q, Xt; the trace (move, i, j, s | flags << 32);
// scal = {moves, global best solved, trace entries}; dscal = {final score,
// global best score, c}.
int refd_worker_trace(const int64_t* P_start, const int64_t* X, int n, const int* active,
                      int n_active, const double* dp, const int64_t* ip, int64_t* P_out,
                      double* Q_out, int64_t* Xt_out, int64_t* gbP, double* gbQ, int64_t* gbXt,
                      int64_t* trace, int64_t trace_cap, int64_t* scal, double* dscal) {
    canneal::EngineParams ep;
    ep.t_init = dp[0];
    ep.t_min = dp[1];
    ep.cooling = dp[2];
    ep.reheat = dp[3];
    ep.reseed_factor = dp[4];
    ep.greedy_fraction = dp[5];
    ep.target_fraction = dp[6];
    ep.target_accept = dp[7];
    ep.worker_diversity = dp[8];
    ep.sweep_fraction = dp[9];
    const double rung_T = dp[10];
    ep.moves_per_cool = (int)ip[0];
    ep.stuck_threshold = (int)ip[1];
    ep.target_samples = (int)ip[2];
    ep.tempering = ip[3] != 0;
    ep.adaptive_cooling = ip[4] != 0;
    ep.threads = (int)ip[5];
    const int thread_id = (int)ip[6];
    const uint32_t seed = (uint32_t)ip[7];
    if (ip[8] <= 0) return -1;
    ep.quiet = true;
    if (ep.threads < 2) ep.tempering = false;  // as run_annealer (:427)

    const Matrix P0 = to_matrix(P_start, n);
    Matrix XS = to_matrix(X, n);
    Matrixd Q0;
    if (!invert_to(P0, Q0)) return -2;
    const long double s_p0 = P0.offdiag_frob2(), s_q0 = Q0.offdiag_frob2();
    long double c = (long double)dp[11];
    if (s_p0 > 0.0L && s_q0 > 0.0L && std::isfinite((double)s_q0)) c *= s_p0 / s_q0;

    canneal::Shared<RecDual> sh;
    sh.global_best = RecDual(P0, Q0, XS, c);
    sh.global_best.dual_tol = dp[12];
    sh.global_best.dual_refresh = (int)ip[9];
    sh.global_best.dual_check = (int)ip[10];
    sh.best_score_relaxed.store((double)sh.global_best.score(), std::memory_order_relaxed);
    if (ep.tempering) {
        sh.temps.assign((size_t)ep.threads, rung_T);
        sh.rung_of_worker = std::vector<std::atomic<int>>((size_t)ep.threads);
        sh.worker_score = std::vector<std::atomic<double>>((size_t)ep.threads);
        for (int k = 0; k < ep.threads; ++k) {
            sh.rung_of_worker[(size_t)k].store(k, std::memory_order_relaxed);
            sh.worker_score[(size_t)k].store((double)sh.global_best.score(),
                                             std::memory_order_relaxed);
        }
    }
    std::vector<int> act;
    if (active && n_active > 0) {
        act.assign(active, active + n_active);
    } else {
        act.resize((size_t)n);
        for (int k = 0; k < n; ++k) act[(size_t)k] = k;
    }
    d_pending.clear();
    d_trace.clear();
    d_moves = 0;
    d_cap = (uint64_t)ip[8];
    d_done = &sh.done_flag;
    d_cur = nullptr;
    canneal::anneal_worker<RecDual>(ep, seed, sh, thread_id, act, std::chrono::steady_clock::now());
    put(d_finalP, P_out);
    putd(d_finalQ, Q_out);
    put(d_finalXt, Xt_out);
    put(sh.global_best.P, gbP);
    putd(sh.global_best.Q, gbQ);
    put(sh.global_best.Xt, gbXt);
    const int64_t cnt = std::min<int64_t>((int64_t)d_trace.size(), trace_cap);
    for (int64_t k = 0; k < cnt; ++k) {
        const DRec& t = d_trace[(size_t)k];
        trace[4 * k + 0] = t.move;
        trace[4 * k + 1] = t.i;
        trace[4 * k + 2] = t.j;
        trace[4 * k + 3] = (int64_t)(uint16_t)t.s | ((int64_t)t.flags << 32);
    }
    scal[0] = (int64_t)sh.total_moves.load();
    scal[1] = sh.global_best.solved() ? 1 : 0;
    scal[2] = (int64_t)d_trace.size();
    dscal[0] = (double)d_final_score;
    dscal[1] = (double)sh.global_best.score();
    dscal[2] = (double)c;
    return 0;
}

}  // extern "C"
