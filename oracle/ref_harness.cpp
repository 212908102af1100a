// ref_harness.cpp -- exposes the reference's own xtax move primitives through a
// C ABI so tests/golden/gen_golden.py can record golden vectors from them.
//
// Built by `make -C oracle ref` into oracle/_ref/ (git-ignored) directly from
// the reference sources under REF (default /root/reference); nothing of the
// reference is copied here. The reference's main() is renamed so this file can
// be a shared library.
#define main xtax_reference_main
#include "main.cpp"
#undef main

#include <cstring>

namespace {
Matrix to_matrix(const int64_t* p, int n) {
    Matrix M((size_t)n);
    std::memcpy(M.data.data(), p, sizeof(int64_t) * (size_t)n * (size_t)n);
    return M;
}
void from_matrix(const Matrix& M, int64_t* p) {
    std::memcpy(p, M.data.data(), sizeof(int64_t) * M.data.size());
}
}  // namespace

extern "C" {

int64_t ref_rounded_div(int64_t a, int64_t b) { return rounded_div(a, b); }

int64_t ref_score(const int64_t* A, int n) { return to_matrix(A, n).score(); }

int64_t ref_offdiag_nonzero(const int64_t* A, int n) {
    return to_matrix(A, n).count_offdiag_nonzero();
}

int ref_try_add(const int64_t* A, const int64_t* Xt, int n, int i, int j, int64_t s,
                int64_t* new_a_row, int64_t* new_xt_row, int64_t* delta, int64_t* dnz) {
    const Matrix MA = to_matrix(A, n);
    const Matrix MX = to_matrix(Xt, n);
    std::vector<int64_t> row((size_t)n, 0), xrow((size_t)n, 0);
    const bool ok = try_add(MA, MX, i, j, s, row, new_xt_row ? &xrow : nullptr, *delta, *dnz);
    std::memcpy(new_a_row, row.data(), sizeof(int64_t) * (size_t)n);
    if (new_xt_row) std::memcpy(new_xt_row, xrow.data(), sizeof(int64_t) * (size_t)n);
    return ok ? 1 : 0;
}

int ref_add_transform_row(const int64_t* Xt, int n, int i, int j, int64_t s, int64_t* out) {
    const Matrix MX = to_matrix(Xt, n);
    std::vector<int64_t> xrow((size_t)n, 0);
    const bool ok = add_transform_row(MX, i, j, s, xrow);
    std::memcpy(out, xrow.data(), sizeof(int64_t) * (size_t)n);
    return ok ? 1 : 0;
}

void ref_commit_add(int64_t* A, int64_t* Xt, int n, int j, const int64_t* new_a_row,
                    const int64_t* new_xt_row, int64_t* row_mass) {
    Matrix MA = to_matrix(A, n);
    Matrix MX = to_matrix(Xt, n);
    std::vector<int64_t> row(new_a_row, new_a_row + n), xrow(new_xt_row, new_xt_row + n);
    std::vector<int64_t> mass(row_mass, row_mass + n);
    commit_add(MA, MX, j, row, xrow, mass);
    from_matrix(MA, A);
    from_matrix(MX, Xt);
    std::memcpy(row_mass, mass.data(), sizeof(int64_t) * (size_t)n);
}

void ref_compute_row_mass(const int64_t* A, int n, int64_t* row_mass) {
    std::vector<int64_t> mass((size_t)n, 0);
    compute_row_mass(to_matrix(A, n), mass);
    std::memcpy(row_mass, mass.data(), sizeof(int64_t) * (size_t)n);
}

int64_t ref_score_delta_l1(const int64_t* A, int n, int i, int j, int64_t s) {
    return score_delta_l1(to_matrix(A, n), i, j, s);
}

int64_t ref_best_shear_l1(const int64_t* A, int n, int i, int j) {
    return best_shear_l1(to_matrix(A, n), i, j);
}

void ref_congruence_of(const int64_t* A, const int64_t* X, int n, int64_t* out) {
    from_matrix(congruence_of(to_matrix(A, n), to_matrix(X, n)), out);
}

void ref_gram_of(const int64_t* L, int m, int d, int64_t* out) {
    Lattice lat;
    lat.m = m;
    lat.d = d;
    lat.data.assign(L, L + (size_t)m * d);
    from_matrix(gram_of(lat), out);
}

void ref_xt_times(const int64_t* X, const int64_t* L, int m, int d, int64_t* out) {
    Lattice lat;
    lat.m = m;
    lat.d = d;
    lat.data.assign(L, L + (size_t)m * d);
    const Lattice R = xt_times(to_matrix(X, m), lat);
    std::memcpy(out, R.data.data(), sizeof(int64_t) * R.data.size());
}

// ---- publish-time reordering and deflation (main.cpp:244-350, :435-455;
// mat_io.hpp:597-644) ----

void ref_rcm_order(const int64_t* A, int n, int* perm) {
    const std::vector<int> p = rcm_order(to_matrix(A, n));
    std::memcpy(perm, p.data(), sizeof(int) * p.size());
}

void ref_centroid_order(const int64_t* A, int n, int* perm) {
    const std::vector<int> p = centroid_order(to_matrix(A, n));
    std::memcpy(perm, p.data(), sizeof(int) * p.size());
}

// L1Objective::reorder_for_publish(true) on (A, Xt) with the given flags.
void ref_reorder_publish(int64_t* A, int64_t* Xt, int n, int band, int centroid) {
    Matrix MA = to_matrix(A, n), MX = to_matrix(Xt, n);
    std::vector<int64_t> mass((size_t)n, 0);
    compute_row_mass(MA, mass);
    if (band) permute_state(MA, MX, rcm_order(MA), mass);
    if (centroid) permute_state(MA, MX, centroid_order(MA), mass);
    from_matrix(MA, A);
    from_matrix(MX, Xt);
}

int ref_is_unimodular(const int64_t* A, int n) { return is_unimodular(to_matrix(A, n)) ? 1 : 0; }

int64_t ref_det_mod_prime(const int64_t* A, int n, int64_t p) {
    return det_mod_prime(to_matrix(A, n), p);
}

int ref_deflate_index(int64_t* M, int64_t* Xt, int n, int i, const int* active, int n_active) {
    Matrix MM = to_matrix(M, n), MX = to_matrix(Xt, n);
    std::vector<int> act(active, active + n_active);
    std::vector<int64_t> row((size_t)n), xrow((size_t)n), mass((size_t)n, 0);
    const bool ok = deflate_index(MM, MX, i, act, row, xrow, mass);
    from_matrix(MM, M);
    from_matrix(MX, Xt);
    return ok ? 1 : 0;
}

}  // extern "C"

// ---- one worker of the reference's annealer, recorded ----
//
// canneal::anneal_worker (congruence_anneal.hpp:177-416) itself, run on the
// calling thread with a recording objective: the reference's L1Objective
// whose commits, reseeds and per-move maintenance calls are logged, and whose
// maintenance call raises the done flag at a move cap. Within one iteration
// of the worker loop the attempt's evaluate is the last evaluate before the
// maintenance call (stall sweeps and reseeds come after it), which assigns
// every logged event to its move: accepted moves (flags 0), sweep commits
// (flags 1) and reseeds (flags 2, i = j = -1), in the engine's log format.

namespace {

struct Ev {
    char kind;  // 'E' evaluate, 'C' commit, 'R' reseed
    int16_t i, j, s;
};

struct TraceRec {
    int64_t move;
    int16_t i, j, s;
    uint16_t flags;
};

std::vector<Ev> g_pending;
std::vector<TraceRec> g_trace;
uint64_t g_moves = 0;
uint64_t g_cap = 0;
std::atomic<bool>* g_done = nullptr;
const void* g_cur = nullptr;
Matrix g_finalA, g_finalXt;
int64_t g_final_score = 0;

void flush_events(std::vector<Ev>& ev, uint64_t before_move, uint64_t split_move, bool split) {
    // events before the last evaluate belong to before_move's stall; from the
    // last evaluate on to split_move's attempt (when split)
    size_t last_e = ev.size();
    if (split)
        for (size_t k = ev.size(); k-- > 0;)
            if (ev[k].kind == 'E') {
                last_e = k;
                break;
            }
    for (size_t k = 0; k < ev.size(); ++k) {
        const bool attempt = split && k >= last_e;
        const uint64_t m = attempt ? split_move : before_move;
        if (ev[k].kind == 'C')
            g_trace.push_back({(int64_t)m, ev[k].i, ev[k].j, ev[k].s, (uint16_t)(attempt ? 0 : 1)});
        else if (ev[k].kind == 'R')
            g_trace.push_back({(int64_t)m, -1, -1, 0, 2});
    }
    ev.clear();
}

struct RecObj : L1Objective {
    using L1Objective::L1Objective;
    RecObj() = default;
    RecObj(const RecObj&) = default;
    RecObj(RecObj&&) = default;
    RecObj& operator=(RecObj&&) = default;
    RecObj& operator=(const RecObj& o) {
        if (this == g_cur) g_pending.push_back({'R', -1, -1, 0});
        L1Objective::operator=(o);
        return *this;
    }
    ~RecObj() {
        if (this != g_cur) return;
        const bool capped = g_cap && g_moves >= g_cap;
        flush_events(g_pending, g_moves, g_moves + 1, !capped);
        g_finalA = A;
        g_finalXt = Xt;
        g_final_score = score_;
        g_cur = nullptr;
    }
    bool evaluate(int i, int j, int64_t s, score_t& d, int64_t& dnz) {
        if (this == g_cur) g_pending.push_back({'E', (int16_t)i, (int16_t)j, (int16_t)s});
        return L1Objective::evaluate(i, j, s, d, dnz);
    }
    bool commit(int i, int j, int64_t s, score_t d, int64_t dnz) {
        const bool ok = L1Objective::commit(i, j, s, d, dnz);
        if (ok && this == g_cur) g_pending.push_back({'C', (int16_t)i, (int16_t)j, (int16_t)s});
        return ok;
    }
    // the worker calls refresh_cache on its own state first (:194), before
    // any move: that identifies the worker's object
    void refresh_cache() {
        if (g_cur == nullptr) g_cur = this;
        L1Objective::refresh_cache();
    }
    void periodic_maintenance(uint64_t moves) {
        flush_events(g_pending, moves - 1, moves, true);
        g_moves = moves;
        if (g_cap && moves >= g_cap) g_done->store(true, std::memory_order_relaxed);
    }
};

}  // namespace

extern "C" {

// dp: t_init, t_min, cooling, reheat, reseed_factor, greedy_fraction,
//     target_fraction, target_accept, worker_diversity, sweep_fraction, rung_T
// ip: moves_per_cool, stuck_threshold, target_samples, tempering,
//     adaptive_cooling, threads, thread_id, worker seed, move cap (> 0)
// Start state: A_start (the working matrix) and X (the start transform).
// Outputs: the worker's final A, Xt; the global best's A, Xt; the trace
// (4 int64 per entry: move, i, j, s | flags << 32); scal = {moves, final
// score, global best score, global best solved, trace entries}.
int ref_worker_trace(const int64_t* A_start, const int64_t* X, int n, const int* active,
                     int n_active, const double* dp, const int64_t* ip, int64_t* A_out,
                     int64_t* Xt_out, int64_t* gbA, int64_t* gbXt, int64_t* trace,
                     int64_t trace_cap, int64_t* scal) {
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

    canneal::Shared<RecObj> sh;
    sh.global_best = RecObj(to_matrix(A_start, n), to_matrix(X, n), false, false);
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
    g_pending.clear();
    g_trace.clear();
    g_moves = 0;
    g_cap = (uint64_t)ip[8];
    g_done = &sh.done_flag;
    g_cur = nullptr;
    canneal::anneal_worker<RecObj>(ep, seed, sh, thread_id, act, std::chrono::steady_clock::now());
    from_matrix(g_finalA, A_out);
    from_matrix(g_finalXt, Xt_out);
    from_matrix(sh.global_best.A, gbA);
    from_matrix(sh.global_best.Xt, gbXt);
    const int64_t cnt = std::min<int64_t>((int64_t)g_trace.size(), trace_cap);
    for (int64_t k = 0; k < cnt; ++k) {
        const TraceRec& t = g_trace[(size_t)k];
        trace[4 * k + 0] = t.move;
        trace[4 * k + 1] = t.i;
        trace[4 * k + 2] = t.j;
        trace[4 * k + 3] = (int64_t)(uint16_t)t.s | ((int64_t)t.flags << 32);
    }
    scal[0] = (int64_t)sh.total_moves.load();
    scal[1] = g_final_score;
    scal[2] = sh.global_best.score();
    scal[3] = sh.global_best.solved() ? 1 : 0;
    scal[4] = (int64_t)g_trace.size();
    return 0;
}

}  // extern "C"
