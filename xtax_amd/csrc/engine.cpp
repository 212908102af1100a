// engine.cpp -- C++20 host side of the gfx950 xtax annealer (the C ABI of
// include/xtx.h).
//
// The reference runs one std::thread per replica plus a monitor that swaps
// tempering rungs every 5 ms (congruence_anneal.hpp:483-558). Here every
// replica of every instance is a lane group on the GPU and the host only
// sequences epochs: an epoch is `exchange_interval` moves of every replica
// (k_anneal, one launch per LDS layout bucket) followed by k_epoch_end (solve
// retirement, exchange sweep, instance best). The host polls instance status
// between epochs (plain form), between captured multi-epoch graphs (batched
// form), or per epoch while admitting new work (served form).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <tuple>
#include <vector>

#include "engine_types.h"
#include "host_internal.h"
#include "xtx.h"

namespace xtx {
hipError_t launch_anneal(hipStream_t st, int G, int C, bool tri, int K, bool xl, int units_per_block,
                         const RepRef* table, int count, Inst* insts, int nmax, int rep_bytes,
                         const volatile int* stop);
bool anneal_variant_exists(int G, int C, bool tri, int K, bool xl = false);
hipError_t launch_epoch_end(hipStream_t st, Inst* insts, int count, int max_r,
                            const int* list = nullptr);
hipError_t launch_probe(hipStream_t st, int G, int C, const int64_t* A, const int64_t* Xt, int n,
                        const int32_t* q, int nq, int64_t* out);
hipError_t launch_probe_rng(hipStream_t st, const uint32_t* ck, int count, uint32_t* out,
                            const double* x, int nx, double* ex);
hipError_t launch_replicate(hipStream_t st, int64_t* dst, const int64_t* src, int64_t nn, int R);
hipError_t launch_gather_outcomes(hipStream_t st, const Inst* insts, const int* slots,
                                  const int64_t* offs, int count, unsigned char* out);
hipError_t init_kernel_attrs(int max_lds_per_block);
hipError_t init_dual_kernel_attrs(int max_lds_per_block);
// the lane kernel (anneal_lane.hip)
hipError_t launch_anneal_lane(hipStream_t st, int nmax, int Q, int units_per_block,
                              const RepRef* table, int count, int split, Inst* insts,
                              int rep_bytes, const volatile int* stop);
hipError_t init_lane_kernel_attrs(int max_lds_per_block);
// the batch's persistent launch (anneal_persist.hip)
hipError_t launch_lane_persist(hipStream_t st, const PersistCtl& c, Inst* insts, int blocks,
                               const volatile int* stop);
hipError_t init_persist_kernel_attrs(int max_lds_per_block);
int lane_persist_blocks_per_cu();
hipError_t launch_xcc_probe(hipStream_t st, int32_t* out, int blocks);
// the no-alloc solve's device scratch (scratch.hip)
hipError_t scratch_init(int device);
void* scratch_address(int device);
size_t scratch_bytes();
}  // namespace xtx

using namespace xtx;
using namespace xtx::host;

namespace xtx {
namespace host {

int g_lds_block = 64 * 1024;

int64_t score_of(const int64_t* A, int n) {
    int64_t full = 0, diag = 0;
    for (int64_t q = 0; q < (int64_t)n * n; ++q) full += iabs(A[q]);
    for (int i = 0; i < n; ++i) diag += iabs(A[(int64_t)i * n + i]);
    return 2 * full - diag;
}

int64_t offnz_of(const int64_t* A, int n) {
    int64_t c = 0;
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) c += A[(int64_t)i * n + j] != 0;
    return c;
}

void check_params(const xtx_params& p) {
    auto bad = [](const char* m) { throw Err{XTX_E_ARG, m}; };
    if (p.replicas < 1) bad("replicas must be >= 1");
    if (p.replicas > (1 << kRepBits) - 1) bad("replicas too large");
    if (p.target_samples < 1 || p.target_samples > kMaxSamplesWide)
        bad("target_samples must be in [1, 4096]");
    if (p.exchange_interval < 1) bad("exchange_interval must be >= 1");
    if (p.moves_per_cool < 1) bad("moves_per_cool must be >= 1");
    if (!(p.cooling > 0.0)) bad("cooling must be > 0");
    if (p.move_log && p.move_log_cap < 1) bad("move_log_cap must be >= 1 when move_log is set");
    if (p.tempering && p.replicas > 8000) bad("tempering supports at most 8000 replicas");
    if (p.lookahead < 0 || p.lookahead > 32 || (p.lookahead & (p.lookahead - 1)) != 0)
        bad("lookahead must be 0, 1, 2, 4, 8, 16 or 32");
}

// Builds the start state exactly: A_start = X0^T A0 X0 in __int128, rejecting
// any input or start entry beyond the 2^48 bound (the reference does not check
// its inputs, mat_io.hpp:432-458; here such inputs are an explicit error).
void prepare(HostInst& h, const int64_t* A0, const int64_t* X0, int n, int R, uint64_t seed,
             const int32_t* active, int n_active, const xtx_params& p) {
    if (!A0 || n < 1) throw Err{XTX_E_ARG, "A0 must be non-null and n >= 1"};
    if (n > kMaxN) throw Err{XTX_E_ARG, "n exceeds the maximum order"};
    h.n = n;
    h.R = R;
    h.seed = seed;
    const size_t nn = (size_t)n * n;
    h.A0.assign(A0, A0 + nn);
    for (size_t q = 0; q < nn; ++q)
        if (iabs(A0[q]) > kLimit) throw Err{XTX_E_BOUND, "an entry of A exceeds 2^48"};
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < i; ++j)
            if (A0[(size_t)i * n + j] != A0[(size_t)j * n + i])
                throw Err{XTX_E_ARG, "A must be symmetric"};
    h.Xt0.assign(nn, 0);
    if (X0) {
        for (size_t q = 0; q < nn; ++q)
            if (iabs(X0[q]) > kLimit) throw Err{XTX_E_BOUND, "an entry of X exceeds 2^48"};
        for (int r = 0; r < n; ++r)
            for (int c = 0; c < n; ++c) h.Xt0[(size_t)c * n + r] = X0[(size_t)r * n + c];
        std::vector<__int128> AX(nn, 0);
        for (int r = 0; r < n; ++r)
            for (int k = 0; k < n; ++k) {
                const __int128 a = A0[(size_t)r * n + k];
                if (!a) continue;
                for (int c = 0; c < n; ++c) AX[(size_t)r * n + c] += a * X0[(size_t)k * n + c];
            }
        h.Astart.assign(nn, 0);
        for (int r = 0; r < n; ++r)
            for (int c = 0; c < n; ++c) {
                __int128 s = 0;
                for (int k = 0; k < n; ++k) {
                    s += (__int128)X0[(size_t)k * n + r] * AX[(size_t)k * n + c];
                    if (s > ((__int128)1 << 100) || s < -((__int128)1 << 100))
                        throw Err{XTX_E_BOUND, "X^T A X exceeds the 2^48 bound"};
                }
                if (s > kLimit || s < -kLimit)
                    throw Err{XTX_E_BOUND, "X^T A X exceeds the 2^48 bound"};
                h.Astart[(size_t)r * n + c] = (int64_t)s;
            }
    } else {
        for (int i = 0; i < n; ++i) h.Xt0[(size_t)i * n + i] = 1;
        h.Astart = h.A0;
    }
    if (active && n_active > 0) {
        std::vector<char> seen((size_t)n, 0);
        for (int q = 0; q < n_active; ++q) {
            if (active[q] < 0 || active[q] >= n) throw Err{XTX_E_ARG, "active index out of range"};
            if (seen[(size_t)active[q]]) throw Err{XTX_E_ARG, "duplicate active index"};
            seen[(size_t)active[q]] = 1;
            h.act.push_back(active[q]);
        }
    } else {
        for (int q = 0; q < n; ++q) h.act.push_back(q);
    }
    h.N = (int)h.act.size();
    h.score0 = score_of(h.Astart.data(), n);
    h.offnz0 = offnz_of(h.Astart.data(), n);
    // suggest_t_init (main.cpp:383-387)
    const double avg = (double)h.score0 / (double)((int64_t)n * n);
    h.t_init = p.t_init > 0.0 ? p.t_init : std::max(1.0, 2.0 * avg);
    const bool tempering = p.tempering && R >= 2;
    if (tempering) h.temps = ladder(h.t_init, p.t_min, R);
    h.P.t_init = h.t_init;
    h.P.t_min = p.t_min;
    h.P.cooling = p.cooling;
    h.P.reheat = p.reheat;
    h.P.reseed_factor = p.reseed_factor;
    h.P.target_accept = p.target_accept;
    h.P.sweep_fraction = p.sweep_fraction;
    h.P.moves_per_cool = p.moves_per_cool;
    h.P.stuck_threshold = p.stuck_threshold;
    h.P.target_samples = p.target_samples;
    h.P.tempering = tempering ? 1 : 0;
    h.P.adaptive = p.adaptive_cooling;
    h.P.epoch_moves = p.exchange_interval;
    h.P.max_moves = p.max_moves;
    h.P.max_epochs = p.max_epochs;
    h.log_on = p.move_log != 0;
    h.log_cap = h.log_on ? p.move_log_cap : 0;
    h.lay = p.target_samples > kMaxSamples ? wide_layout(order_of(n)) : layout_for(n);
    if (h.lay.rep_bytes > g_lds_block)
        throw Err{XTX_E_ARG, "n too large for an LDS-resident replica"};
}

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

std::vector<double> ladder(double top, double t_min, int R) {
    std::vector<double> temps((size_t)R);
    const double bot = std::min(t_min, top);
    const double ratio = std::pow(top / std::max(bot, 1e-12), 1.0 / (double)(R - 1));
    double t = bot;
    for (int k = 0; k < R; ++k) {
        temps[(size_t)k] = t;
        t *= ratio;
    }
    return temps;
}

void fractions(const xtx_params& p, int r, int R, double* greedy, double* target) {
    // replica index in place of the thread id
    double g = p.greedy_fraction, t = p.target_fraction;
    if (p.worker_diversity > 0.0 && R > 1) {
        const double off = p.worker_diversity * ((double)r / (double)(R - 1) - 0.5);
        g = std::clamp(g + off, 0.0, 1.0);
        t = std::clamp(t + off, 0.0, 1.0);
    }
    *greedy = g;
    *target = t;
}

}  // namespace host
}  // namespace xtx

namespace {

size_t device_bytes(const HostInst& h) {
    const size_t nn = (size_t)h.n * h.n;
    size_t b = 0;
    b += 4 * align256((size_t)h.R * nn * 8);            // A, Xt, bA, bXt
    b += 2 * align256(nn * 8);                          // gA, gXt
    b += align256((size_t)h.R * sizeof(RepState));      // rs
    b += align256((size_t)h.R * 8);                     // temps
    b += align256((size_t)h.n);                         // act
    b += align256((size_t)h.R * h.log_cap * sizeof(LogEnt));
    return b;
}

// Lays out the instance's device buffers in dmem and uploads its start state
// (staging vectors live in h until the stream is synchronized).
struct Staging {
    std::vector<RepState> rs;
    std::vector<uint8_t> act;
};

void upload(HostInst& h, char* dmem, hipStream_t st, Staging& stg) {
    h.dmem = dmem;
    const size_t nn = (size_t)h.n * h.n;
    char* p = dmem;
    auto take = [&](size_t bytes) {
        char* q = p;
        p += align256(bytes);
        return q;
    };
    Inst& I = h.rec;
    I = Inst{};
    I.n = h.n;
    I.N = h.N;
    I.R = h.R;
    I.status = ST_RUNNING;
    I.key0 = (uint32_t)h.seed;
    I.key1 = (uint32_t)(h.seed >> 32);
    I.epoch = 0;
    I.log_on = h.log_on;
    I.A = (int64_t*)take((size_t)h.R * nn * 8);
    I.Xt = (int64_t*)take((size_t)h.R * nn * 8);
    I.bA = (int64_t*)take((size_t)h.R * nn * 8);
    I.bXt = (int64_t*)take((size_t)h.R * nn * 8);
    I.gA = (int64_t*)take(nn * 8);
    I.gXt = (int64_t*)take(nn * 8);
    I.rs = (RepState*)take((size_t)h.R * sizeof(RepState));
    I.temps = (double*)take((size_t)h.R * 8);
    I.act = (uint8_t*)take((size_t)h.n);
    I.log = h.log_on ? (LogEnt*)take((size_t)h.R * h.log_cap * sizeof(LogEnt)) : nullptr;
    I.log_cap = h.log_cap;
    I.P = h.P;
    I.solve_key = kNoKey;
    I.gb_score = h.score0;
    I.gb_offnz = h.offnz0;
    I.gb_rep = 0;
    I.solve_rep = -1;
    I.solve_move = -1;

    stg.rs.assign((size_t)h.R, RepState{});
    const bool div = h.R > 1;
    for (int r = 0; r < h.R; ++r) {
        RepState& s = stg.rs[(size_t)r];
        s.score = h.score0;
        s.offnz = h.offnz0;
        s.local_best = h.score0;
        s.T = h.t_init;
        s.cooling = h.P.cooling;
        s.rung = r;
        s.best_score = h.score0;
        s.best_move = 0;
        s.flags = RF_CUR_BEST;
        s.gf = 0;
        s.tf = 0;
        (void)div;
    }
    stg.act.assign(h.act.begin(), h.act.end());
    HIPCHK(hipMemcpyAsync(I.gA, h.Astart.data(), nn * 8, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(I.gXt, h.Xt0.data(), nn * 8, hipMemcpyHostToDevice, st));
    HIPCHK(launch_replicate(st, I.A, I.gA, (int64_t)nn, h.R));
    HIPCHK(launch_replicate(st, I.Xt, I.gXt, (int64_t)nn, h.R));
    HIPCHK(hipMemcpyAsync(I.rs, stg.rs.data(), (size_t)h.R * sizeof(RepState),
                          hipMemcpyHostToDevice, st));
    if (!h.temps.empty())
        HIPCHK(hipMemcpyAsync(I.temps, h.temps.data(), (size_t)h.R * 8, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(I.act, stg.act.data(), stg.act.size(), hipMemcpyHostToDevice, st));
}

void set_fractions(HostInst& h, Staging& stg, const xtx_params& p) {
    // per-worker diversity (congruence_anneal.hpp:200-207), replica index in
    // place of the thread id
    for (int r = 0; r < h.R; ++r) fractions(p, r, h.R, &stg.rs[(size_t)r].gf, &stg.rs[(size_t)r].tf);
}

// Result of one instance, read back from the device.
struct Outcome {
    bool solved = false;
    int rep = 0;
    int64_t solve_move = -1;
    int64_t best_score = 0;
    std::vector<int64_t> A, Xt;
    int64_t moves = 0, accepts = 0, greedy = 0, infeasible = 0, sweeps = 0;
    std::vector<RepState> rs;
};

// An instance's outcome in three stages, so that many instances can share one
// wait between them: the replica records' copy; the result replica (the
// solving one, else the instance best) and its A and X^T copied; the score.
void outcome_records(const HostInst& h, const Inst& rec, hipStream_t st, Outcome& out) {
    out.rs.resize((size_t)h.R);
    HIPCHK(hipMemcpyAsync(out.rs.data(), rec.rs, (size_t)h.R * sizeof(RepState),
                          hipMemcpyDeviceToHost, st));
}

// the replica records' counters summed into the outcome
void outcome_tally(Outcome& out) {
    out.moves = out.accepts = out.greedy = out.infeasible = out.sweeps = 0;
    for (const auto& s : out.rs) {
        out.moves += (int64_t)s.moves;
        out.accepts += s.accepts;
        out.greedy += s.greedy;
        out.infeasible += s.infeasible;
        out.sweeps += s.sweeps;
    }
}

void outcome_state(const HostInst& h, const Inst& rec, hipStream_t st, Outcome& out) {
    const size_t nn = (size_t)h.n * h.n;
    outcome_tally(out);
    out.A.resize(nn);
    out.Xt.resize(nn);
    const int64_t* srcA;
    const int64_t* srcX;
    if (rec.status == ST_SOLVED) {
        out.solved = true;
        out.rep = rec.solve_rep;
        out.solve_move = rec.solve_move;
        srcA = rec.A + (size_t)out.rep * nn;
        srcX = rec.Xt + (size_t)out.rep * nn;
    } else {
        // instance best: argmin (best_score, best_move, replica)
        int b = 0;
        for (int r = 1; r < h.R; ++r) {
            const auto& x = out.rs[(size_t)r];
            const auto& y = out.rs[(size_t)b];
            if (x.best_score < y.best_score ||
                (x.best_score == y.best_score && x.best_move < y.best_move))
                b = r;
        }
        const auto& x = out.rs[(size_t)b];
        if (!(x.flags & RF_CUR_BEST) && (x.flags & RF_BEST_STALE))
            throw Err{XTX_E_STATE, "internal: best replica has a stale snapshot"};
        out.rep = b;
        const bool cur = (x.flags & RF_CUR_BEST) != 0;
        srcA = (cur ? rec.A : rec.bA) + (size_t)b * nn;
        srcX = (cur ? rec.Xt : rec.bXt) + (size_t)b * nn;
    }
    HIPCHK(hipMemcpyAsync(out.A.data(), srcA, nn * 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(out.Xt.data(), srcX, nn * 8, hipMemcpyDeviceToHost, st));
}

void fetch_outcome(const HostInst& h, const Inst& rec, hipStream_t st, Outcome& out) {
    outcome_records(h, rec, st, out);
    HIPCHK(hipStreamSynchronize(st));
    outcome_state(h, rec, st, out);
    HIPCHK(hipStreamSynchronize(st));
    out.best_score = score_of(out.A.data(), h.n);
}

// fn(i) for i < count on up to 16 host threads (results are independent)
template <class F>
void parallel_for(size_t count, F fn) {
    const size_t nt = std::min<size_t>(count, std::min(16u, std::max(1u, std::thread::hardware_concurrency())));
    if (nt <= 1) {
        for (size_t i = 0; i < count; ++i) fn(i);
        return;
    }
    std::vector<std::thread> pool;
    std::atomic<size_t> next{0};
    for (size_t t = 0; t < nt; ++t)
        pool.emplace_back([&]() {
            for (size_t i = next++; i < count; i = next++) fn(i);
        });
    for (auto& th : pool) th.join();
}

void write_outputs(const HostInst& h, const Outcome& o, int64_t* A_out, int64_t* X_out) {
    const int n = h.n;
    if (A_out) std::memcpy(A_out, o.A.data(), o.A.size() * 8);
    if (X_out)
        for (int r = 0; r < n; ++r)
            for (int c = 0; c < n; ++c) X_out[(size_t)r * n + c] = o.Xt[(size_t)c * n + r];
}

int verify_outcome(const HostInst& h, const Outcome& o) {
    std::vector<int64_t> X((size_t)h.n * h.n);
    write_outputs(h, o, nullptr, X.data());
    xtx_verify_report rep{};
    xtx_verify(h.A0.data(), X.data(), o.A.data(), h.n, &rep);
    // a diagonal is required only of solved results
    return o.solved ? rep.status : (rep.status & ~2);
}

// Pads a replica table to whole waves: the kernel reads the instance of a wave
// from its first group, so a wave never mixes instances (rep -1 = idle group).
// U: lanes per replica (G K).
void pad_to_wave(std::vector<RepRef>& refs, int U, int inst) {
    const size_t per_wave = (size_t)std::max(1, 64 / U);
    while (refs.size() % per_wave) refs.push_back(RepRef{inst, -1});
}

// Per-bucket replica tables for a set of resident instances.
struct Bucket {
    Layout lay;
    std::vector<RepRef> refs;
    RepRef* d_refs = nullptr;
    int cap = 0;
    int units_per_block = 1;  // replicas (table entries) per block
    // lane kernel: entries of the launch pair's higher order come first, this
    // many (whole blocks); -1: every entry is of the layout's order
    int split = -1;
    // the instances (slot indices) whose replicas the table holds: their
    // epoch ends follow this bucket's launch, so buckets advance independently
    std::vector<int> insts;
    int* d_insts = nullptr;
    int insts_cap = 0;
};

// The lane kernel's split for a table whose entries are all of order L.nmax.
int whole_split(const Layout& L, int count, int units_per_block) {
    if (L.nmax != lane_pair_hi(L.nmax)) return 0;
    return (count + units_per_block - 1) / units_per_block * units_per_block;
}

// Orders a lane-kernel bucket's table: the pair's higher order first, padded
// to whole blocks with idle entries, then the lower order. The kernel reads
// which order a block runs from `split`.
void split_lane_bucket(Bucket& b, const std::vector<std::unique_ptr<HostInst>>& slots) {
    if (b.lay.G != 1 || b.refs.empty()) return;
    const int hi = lane_pair_hi(b.lay.nmax);
    std::vector<RepRef> top, rest;
    for (const RepRef& e : b.refs) (slots[(size_t)e.inst]->lay.nmax == hi ? top : rest).push_back(e);
    if (!top.empty() && !rest.empty()) {
        const int inst = top.back().inst;
        while (top.size() % (size_t)b.units_per_block) top.push_back(RepRef{inst, -1});
    }
    b.split = top.empty() ? 0 : (int)top.size();
    if (!top.empty() && rest.empty()) b.split = whole_split(b.lay, (int)top.size(), b.units_per_block);
    top.insert(top.end(), rest.begin(), rest.end());
    b.refs.swap(top);
}

}  // namespace

static std::string g_open_error;

namespace {

// The layout instances of order n run in. `requested`: the params' lookahead
// (0 = automatic); `rule`: the form's lookahead rule (16 for plain and batch
// calls; by load for a serve admission). A requested lookahead gives its
// layout (1: the order's configuration from xtx_launch_set / xtx_autotune if
// it has one lane group, else the default layout). Automatic: the order's
// configuration if it has lookahead (rule above 1) or one group (rule 1),
// else the rule's layout. `lane`: the form has replicas enough to fill the
// device with the lane kernel (one lane per proposal, four or two replicas
// per wave), which then runs the orders it covers under automatic lookahead,
// unless the order has a lookahead configuration of its own. A lane-kernel
// configuration the autotuner chose (timed with the device full) applies only
// to such forms; one set by hand applies wherever its kind is chosen.
Layout choose_layout(const xtx_ctx* ctx, int n, int requested, int rule, bool lane = false) {
    const int nmax = order_of(n);
    if (nmax > kMaxGroupN) return wide_layout(nmax);
    const auto it = ctx->tuned.find(nmax);
    const bool has = it != ctx->tuned.end();
    if (requested > 1) return lookahead_layout(nmax, requested);
    const int K = requested == 1 ? 1 : rule;
    if (has && (it->second.K > 1) == (K > 1) && (lane || !it->second.tuner || it->second.G > 1))
        return it->second;
    if (requested == 0 && lane && lane_layout_ok(nmax)) return lane_layout(nmax);
    return K > 1 ? lookahead_layout(nmax, K) : layout_for(n);
}

// Replicas that fill the device with the lane kernel: 16 per compute unit
// (four waves of four replicas; the lookahead kernels serve fewer replicas
// better).
bool lane_load(const xtx_ctx* ctx, int64_t replicas) {
    return replicas >= (int64_t)16 * ctx->prop.multiProcessorCount;
}

void apply_layout(const xtx_ctx* ctx, HostInst& h, int requested, int rule, bool lane = false) {
    // tournaments of more than 16 samples: only the wide mode draws them
    h.lay = h.P.target_samples > kMaxSamples ? wide_layout(order_of(h.n))
                                             : choose_layout(ctx, h.n, requested, rule, lane);
    if (h.lay.rep_bytes > g_lds_block)
        throw Err{XTX_E_ARG, "n too large for an LDS-resident replica"};
}

int lookahead_rule(const xtx_ctx* ctx, int64_t replicas) {
    return xtx::host::lookahead_rule(ctx->prop.multiProcessorCount, replicas);
}

int units_per_block_for(const Layout& L, int waves) {
    if (L.G * L.K > 64) return 1;  // a replica over several waves: one per block
    const int per_wave = 64 / (L.G * L.K);
    int w = L.waves > 0 ? L.waves : (waves > 0 ? waves : 1);
    // keep a block within the LDS of one CU
    while (w > 1 && (int64_t)w * per_wave * L.rep_bytes > g_lds_block) --w;
    return w * per_wave;
}

// One launch of layout L's kernel over a replica table (split: the lane
// kernel's, -1 when every entry is of order L.nmax).
hipError_t launch_layout(hipStream_t st, const Layout& L, int units_per_block, const RepRef* table,
                         int count, Inst* insts, const volatile int* stop, int split = -1) {
    if (L.G == 1)
        return launch_anneal_lane(st, L.nmax, L.K, units_per_block, table, count,
                                  split < 0 ? whole_split(L, count, units_per_block) : split,
                                  insts, L.rep_bytes, stop);
    return launch_anneal(st, L.G, L.C, L.tri, L.K, L.xl, units_per_block, table, count, insts, L.nmax,
                         L.rep_bytes, stop);
}

// Launches k epochs on ctx->stream. Every bucket is a chain of its own: its
// kernel, then the epoch end (exchange, instance best, retirement) of its own
// instances, k times; the buckets' instances never meet, so no bucket waits
// for another between epochs. Chains go to the main and side streams (the
// largest work first, so the device fills while the short ones finish); two
// chains on one stream interleave epoch by epoch. Captured, the fork and join
// become graph edges.
void launch_epochs_chained(xtx_ctx* ctx, std::vector<Bucket>& buckets, Inst* d_insts, int max_r,
                           int k) {
    std::vector<const Bucket*> order;
    for (auto& b : buckets)
        if (!b.refs.empty()) order.push_back(&b);
    std::stable_sort(order.begin(), order.end(), [](const Bucket* a, const Bucket* b) {
        return (int64_t)a->refs.size() * a->lay.nmax > (int64_t)b->refs.size() * b->lay.nmax;
    });
    const int ns = std::min<int>(kSideStreams, std::max<int>(0, (int)order.size() - 1));
    if (ns > 0) {
        HIPCHK(hipEventRecord(ctx->fork_ev, ctx->stream));
        for (int q = 0; q < ns; ++q) HIPCHK(hipStreamWaitEvent(ctx->side[q], ctx->fork_ev, 0));
    }
    for (int e = 0; e < k; ++e)
        for (size_t q = 0; q < order.size(); ++q) {
            const Bucket& b = *order[q];
            const int lane = (int)(q % (size_t)(ns + 1));
            hipStream_t st = lane == 0 ? ctx->stream : ctx->side[lane - 1];
            HIPCHK(launch_layout(st, b.lay, b.units_per_block, b.d_refs, (int)b.refs.size(),
                                 d_insts, ctx->stop_dev, b.split));
            HIPCHK(launch_epoch_end(st, d_insts, (int)b.insts.size(), max_r, b.d_insts));
        }
    for (int q = 0; q < ns; ++q) {
        HIPCHK(hipEventRecord(ctx->join_ev[q], ctx->side[q]));
        HIPCHK(hipStreamWaitEvent(ctx->stream, ctx->join_ev[q], 0));
    }
}

}  // namespace

// ---------------------------------------------------------------- basics
extern "C" {

void xtx_params_default(xtx_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->t_init = 0.0;
    p->t_min = 1e-3;
    p->cooling = 0.999;
    p->reheat = 1.0;
    p->reseed_factor = 1.25;
    p->greedy_fraction = 0.5;
    p->target_fraction = 0.5;
    p->target_accept = 0.44;
    p->worker_diversity = 0.4;
    p->sweep_fraction = 0.0;
    p->max_seconds = 0.0;
    p->save_interval = 2.0;
    p->moves_per_cool = 200;
    p->stuck_threshold = 20000;
    p->target_samples = 8;
    p->tempering = 1;
    p->adaptive_cooling = 0;
    p->exchange_interval = 2000;
    p->seed = 0;
    p->replicas = 64;
    p->move_log = 0;
    p->move_log_cap = 0;
    p->max_moves = 0;
    p->max_epochs = 0;
    p->verify = 1;
    p->waves_per_block = 0;
    p->lookahead = 0;
    p->persistent = 1;
}

int xtx_open(int device, xtx_ctx** out) {
    if (!out) return XTX_E_ARG;
    *out = nullptr;
    auto ctx = std::make_unique<xtx_ctx>();
    try {
        int count = 0;
        HIPCHK(hipGetDeviceCount(&count));
        if (device < 0 || device >= count) throw Err{XTX_E_ARG, "no such device"};
        ctx->device = device;
        HIPCHK(hipSetDevice(device));
        HIPCHK(hipGetDeviceProperties(&ctx->prop, device));
        if (std::strncmp(ctx->prop.gcnArchName, "gfx950", 6) != 0)
            throw Err{XTX_E_DEVICE, std::string("built for gfx950, device is ") +
                                        ctx->prop.gcnArchName};
        g_lds_block = (int)ctx->prop.sharedMemPerBlock;
        HIPCHK(init_kernel_attrs(g_lds_block));
        HIPCHK(init_dual_kernel_attrs(g_lds_block));
        HIPCHK(init_lane_kernel_attrs(g_lds_block));
        HIPCHK(init_persist_kernel_attrs(g_lds_block));
        HIPCHK(scratch_init(device));
        HIPCHK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        HIPCHK(hipEventCreate(&ctx->ev0));
        HIPCHK(hipEventCreate(&ctx->ev1));
        HIPCHK(hipEventCreateWithFlags(&ctx->fork_ev, hipEventDisableTiming));
        for (int q = 0; q < kSideStreams; ++q) {
            HIPCHK(hipStreamCreateWithFlags(&ctx->side[q], hipStreamNonBlocking));
            HIPCHK(hipEventCreateWithFlags(&ctx->join_ev[q], hipEventDisableTiming));
        }
        HIPCHK(hipHostMalloc((void**)&ctx->stop_host, sizeof(int), hipHostMallocMapped));
        *ctx->stop_host = 0;
        HIPCHK(hipHostGetDevicePointer((void**)&ctx->stop_dev, ctx->stop_host, 0));
    } catch (const Err& e) {
        g_open_error = e.msg;
        return e.code;
    }
    *out = ctx.release();
    return XTX_OK;
}

int xtx_serve_close(xtx_ctx* ctx);

int xtx_close(xtx_ctx* ctx) {
    if (!ctx) return XTX_E_ARG;
    (void)hipSetDevice(ctx->device);
    xtx_serve_close(ctx);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->d_plain) (void)hipFree(ctx->d_plain);
    ctx->plain_arena.release();
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->fork_ev) (void)hipEventDestroy(ctx->fork_ev);
    for (int q = 0; q < kSideStreams; ++q) {
        if (ctx->side[q]) (void)hipStreamSynchronize(ctx->side[q]);
        if (ctx->side[q]) (void)hipStreamDestroy(ctx->side[q]);
        if (ctx->join_ev[q]) (void)hipEventDestroy(ctx->join_ev[q]);
    }
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    if (ctx->stop_host) (void)hipHostFree(ctx->stop_host);
    if (ctx->zc_host) (void)hipHostFree(ctx->zc_host);
    delete ctx;
    return XTX_OK;
}

const char* xtx_last_error(const xtx_ctx* ctx) {
    return ctx ? ctx->err.c_str() : g_open_error.c_str();
}

void xtx_request_stop(xtx_ctx* ctx) {
    // a lock-free atomic OR: async-signal-safe
    if (ctx && ctx->stop_host) __atomic_fetch_or(ctx->stop_host, kStopRequest, __ATOMIC_SEQ_CST);
}

void xtx_clear_stop(xtx_ctx* ctx) {
    if (ctx && ctx->stop_host) __atomic_fetch_and(ctx->stop_host, ~kStopRequest, __ATOMIC_SEQ_CST);
}

int xtx_query_limits(xtx_ctx* ctx, xtx_limits* out) {
    if (!ctx || !out) return XTX_E_ARG;
    std::memset(out, 0, sizeof(*out));
    out->compute_units = ctx->prop.multiProcessorCount;
    out->lds_bytes_per_cu = kLdsPerCu;
    out->lds_bytes_per_block_max = (int)ctx->prop.sharedMemPerBlock;
    out->wave_size = ctx->prop.warpSize;
    out->max_n = kMaxN;
    int mx = 0;
    for (int n = 2; n <= kMaxGroupN; ++n)
        if (layout_for(n).rep_bytes <= g_lds_block) mx = n;
    out->max_n_lds = mx;
    out->hbm_bytes = (int64_t)ctx->prop.totalGlobalMem;
    std::snprintf(out->arch, sizeof(out->arch), "%s", ctx->prop.gcnArchName);
    return XTX_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- plain form
namespace {

// Device memory a plain solve may use instead of allocating: the fixed
// scratch array of the no-alloc entry point (its device address and size).
struct ScratchRegion {
    char* base;
    size_t bytes;
};

// One instance per launch (xtx_solve_plain). With `scr` null the solve
// allocates its device buffers (the instance arena, the record and the
// replica table) and a pinned host record; otherwise every device buffer is
// carved out of scr (checked against its size before any device work) and the
// record is read back into host memory of the call's own frame, so the solve
// calls no allocating HIP routine.
int solve_plain(xtx_ctx* ctx, const int64_t* A0, const int64_t* X0, int32_t n,
                const xtx_params* params, const int32_t* active, int32_t n_active,
                int64_t* A_out, int64_t* X_out, xtx_stats* stats, const ScratchRegion* scr) {
    if (!ctx || !params) return XTX_E_ARG;
    const auto t0 = clk::now();
    try {
        HIPCHK(hipSetDevice(ctx->device));
        const xtx_params& p = *params;
        check_params(p);
        ctx->plain_log_valid = false;
        HostInst& h = ctx->plain;
        h = HostInst{};
        prepare(h, A0, X0, n, p.replicas, p.seed, active, n_active, p);
        apply_layout(ctx, h, p.lookahead, kLookahead);
        xtx_stats st{};
        st.verify_status = -1;
        if (n < 2 || h.N < 2 || h.offnz0 == 0) {
            // nothing to anneal (congruence_anneal.hpp:435, :444); an already
            // diagonal start is reported solved
            Outcome o;
            o.A = h.Astart;
            o.Xt = h.Xt0;
            o.solved = h.offnz0 == 0;
            o.best_score = h.score0;
            write_outputs(h, o, A_out, X_out);
            st.solved = o.solved;
            st.solve_move = 0;
            st.best_score = h.score0;
            st.stop_reason = 4;
            st.verify_status = p.verify ? verify_outcome(h, o) : -1;
            st.seconds = std::chrono::duration<double>(clk::now() - t0).count();
            if (stats) *stats = st;
            ctx->plain_rs.clear();
            return XTX_OK;
        }
        std::vector<Bucket> buckets(1);
        buckets[0].lay = h.lay;
        buckets[0].units_per_block = units_per_block_for(h.lay, p.waves_per_block);
        for (int r = 0; r < h.R; ++r) buckets[0].refs.push_back(RepRef{0, r});
        pad_to_wave(buckets[0].refs, h.lay.G * h.lay.K, 0);
        const size_t inst_bytes = device_bytes(h);
        const size_t refs_bytes = buckets[0].refs.size() * sizeof(RepRef);
        char* dm = nullptr;
        Inst* d_rec = nullptr;
        Inst frame_rec{};
        Inst* pin = nullptr;
        // allocated buffers are released on every exit, including an error
        // thrown below
        struct Release {
            bool owned;
            RepRef*& refs;
            Inst*& pin;
            ~Release() {
                if (!owned) return;
                if (refs) (void)hipFree(refs);
                if (pin) (void)hipHostFree(pin);
            }
        } release{scr == nullptr, buckets[0].d_refs, pin};
        if (scr) {
            const size_t need = align256(sizeof(Inst)) + align256(refs_bytes) + inst_bytes;
            if (need > scr->bytes)
                throw Err{XTX_E_ARG, "the no-alloc solve needs " + std::to_string(need) +
                                         " bytes of device scratch for this order, replica"
                                         " count and move log; it has " +
                                         std::to_string(scr->bytes)};
            d_rec = reinterpret_cast<Inst*>(scr->base);
            buckets[0].d_refs = reinterpret_cast<RepRef*>(scr->base + align256(sizeof(Inst)));
            dm = scr->base + align256(sizeof(Inst)) + align256(refs_bytes);
            pin = &frame_rec;
        } else {
            ctx->plain_arena.reserve(inst_bytes);
            dm = ctx->plain_arena.alloc(inst_bytes);
            if (!dm) throw Err{XTX_E_NOMEM, "arena"};
            if (!ctx->d_plain) HIPCHK(hipMalloc(&ctx->d_plain, sizeof(Inst)));
            d_rec = ctx->d_plain;
            HIPCHK(hipMalloc(&buckets[0].d_refs, refs_bytes));
            HIPCHK(hipHostMalloc((void**)&pin, sizeof(Inst), hipHostMallocDefault));
        }
        Staging stg;
        upload(h, dm, ctx->stream, stg);
        set_fractions(h, stg, p);
        HIPCHK(hipMemcpyAsync(h.rec.rs, stg.rs.data(), (size_t)h.R * sizeof(RepState),
                              hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipMemcpyAsync(d_rec, &h.rec, sizeof(Inst), hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipMemcpyAsync(buckets[0].d_refs, buckets[0].refs.data(), refs_bytes,
                              hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));

        double kernel_ms = 0;
        int epochs = 0;
        int reason = 3;
        int64_t reported = h.score0;  // the start state is never reported
        ctx->best_so_far = [ctx, &h, &pin](int64_t* A, double*, int64_t* X) {
            Outcome o;
            fetch_outcome(h, *pin, ctx->stream, o);
            write_outputs(h, o, A, X);
        };
        struct Clear {
            xtx_ctx* c;
            ~Clear() { c->best_so_far = nullptr; }
        } clear_snapshot{ctx};
        {
            Deadline dl(ctx->stop_host, p.max_seconds);
            for (;;) {
                HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
                for (auto& b : buckets)
                    HIPCHK(launch_layout(ctx->stream, b.lay, b.units_per_block, b.d_refs,
                                         (int)b.refs.size(), d_rec, ctx->stop_dev));
                HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
                HIPCHK(launch_epoch_end(ctx->stream, d_rec, 1, h.R));
                HIPCHK(hipMemcpyAsync(pin, d_rec, sizeof(Inst), hipMemcpyDeviceToHost,
                                      ctx->stream));
                HIPCHK(hipStreamSynchronize(ctx->stream));
                float ms = 0;
                HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
                kernel_ms += ms;
                ++epochs;
                if (p.on_best && pin->status != ST_SOLVED && pin->gb_score < reported) {
                    reported = pin->gb_score;
                    p.on_best(p.on_best_user,
                              std::chrono::duration<double>(clk::now() - t0).count(), reported,
                              (double)reported, pin->gb_rep);
                }
                if (pin->status == ST_SOLVED) {
                    reason = 0;
                    break;
                }
                if (pin->status == ST_EXHAUSTED) {
                    reason = 3;
                    break;
                }
                if (*(volatile int*)ctx->stop_host) {
                    reason = stop_reason_of(ctx->stop_host);
                    break;
                }
                if (p.max_epochs > 0 && epochs >= p.max_epochs) {
                    reason = 3;
                    break;
                }
            }
        }
        h.rec = *pin;
        Outcome o;
        fetch_outcome(h, h.rec, ctx->stream, o);
        ctx->plain_rs = o.rs;
        write_outputs(h, o, A_out, X_out);
        st.moves = o.moves;
        st.accepts = o.accepts;
        st.greedy_calls = o.greedy;
        st.infeasible = o.infeasible;
        st.sweeps = o.sweeps;
        st.exchanges_tried = h.rec.exch_tried;
        st.exchanges_accepted = h.rec.exch_acc;
        st.solved = o.solved;
        st.solve_replica = o.solved ? o.rep : -1;
        st.solve_move = o.solve_move;
        st.best_score = o.best_score;
        st.best_replica = o.rep;
        st.epochs = epochs;
        st.stop_reason = reason;
        st.kernel_ms = kernel_ms;
        st.verify_status = p.verify ? verify_outcome(h, o) : -1;
        st.seconds = std::chrono::duration<double>(clk::now() - t0).count();
        if (stats) *stats = st;
        if (h.log_on) {
            ctx->plain_log.resize((size_t)h.R * h.log_cap);
            HIPCHK(hipMemcpy(ctx->plain_log.data(), h.rec.log, ctx->plain_log.size() * sizeof(LogEnt),
                             hipMemcpyDeviceToHost));
            ctx->plain_log_count.resize((size_t)h.R);
            for (int r = 0; r < h.R; ++r) ctx->plain_log_count[(size_t)r] = o.rs[(size_t)r].log_count;
            ctx->plain_log_valid = true;
        }
        if (p.verify && st.verify_status != 0) {
            ctx->err = "result failed exact verification";
            return XTX_E_VERIFY;
        }
        return XTX_OK;
    } catch (const Err& e) {
        return fail(ctx, e);
    }
}

}  // namespace

extern "C" {

int xtx_solve_plain(xtx_ctx* ctx, const int64_t* A0, const int64_t* X0, int32_t n,
                    const xtx_params* params, const int32_t* active, int32_t n_active,
                    int64_t* A_out, int64_t* X_out, xtx_stats* stats) {
    return solve_plain(ctx, A0, X0, n, params, active, n_active, A_out, X_out, stats, nullptr);
}

int xtx_solve_plain_noalloc(xtx_ctx* ctx, const int64_t* A0, const int64_t* X0, int32_t n,
                            const xtx_params* params, const int32_t* active, int32_t n_active,
                            int64_t* A_out, int64_t* X_out, xtx_stats* stats) {
    if (!ctx || !params) return XTX_E_ARG;
    void* base = scratch_address(ctx->device);
    if (!base) {
        ctx->err = "no device scratch address for this device (xtx_open resolves it)";
        return XTX_E_STATE;
    }
    const ScratchRegion scr{static_cast<char*>(base), scratch_bytes()};
    return solve_plain(ctx, A0, X0, n, params, active, n_active, A_out, X_out, stats, &scr);
}

int xtx_best_so_far(xtx_ctx* ctx, int64_t* A_out, double* Q_out, int64_t* X_out) {
    if (!ctx) return XTX_E_ARG;
    try {
        if (!ctx->best_so_far)
            throw Err{XTX_E_STATE, "xtx_best_so_far is valid only inside an on_best report"};
        ctx->best_so_far(A_out, Q_out, X_out);
        return XTX_OK;
    } catch (const Err& e) {
        return fail(ctx, e);
    }
}

int xtx_last_replicas(xtx_ctx* ctx, int64_t* A, int64_t* Xt, int64_t* scal) {
    if (!ctx) return XTX_E_ARG;
    try {
        const HostInst& h = ctx->plain;
        if (ctx->plain_rs.empty() || h.dmem == nullptr) throw Err{XTX_E_STATE, "no plain solve"};
        const size_t nn = (size_t)h.n * h.n;
        if (A)
            HIPCHK(hipMemcpy(A, h.rec.A, (size_t)h.R * nn * 8, hipMemcpyDeviceToHost));
        if (Xt)
            HIPCHK(hipMemcpy(Xt, h.rec.Xt, (size_t)h.R * nn * 8, hipMemcpyDeviceToHost));
        if (scal)
            for (int r = 0; r < h.R; ++r) {
                const RepState& s = ctx->plain_rs[(size_t)r];
                int64_t* o = scal + (size_t)r * 8;
                o[0] = s.score;
                o[1] = s.offnz;
                o[2] = (int64_t)s.moves;
                o[3] = s.best_score;
                o[4] = s.best_move;
                o[5] = s.rung;
                o[6] = s.accepts;
                o[7] = s.local_best;
            }
        return XTX_OK;
    } catch (const Err& e) {
        return fail(ctx, e);
    }
}

int xtx_movelog_get(xtx_ctx* ctx, int32_t replica, const xtx_log_entry** entries, int64_t* count,
                    int32_t* truncated) {
    if (!ctx || !entries || !count) return XTX_E_ARG;
    if (!ctx->plain_log_valid) {
        ctx->err = "no move log (set move_log before the solve)";
        return XTX_E_STATE;
    }
    const HostInst& h = ctx->plain;
    if (replica < 0 || replica >= h.R) return XTX_E_ARG;
    const int64_t c = ctx->plain_log_count[(size_t)replica];
    *entries = reinterpret_cast<const xtx_log_entry*>(ctx->plain_log.data() +
                                                      (size_t)replica * h.log_cap);
    *count = std::min(c, h.log_cap);
    if (truncated) *truncated = c > h.log_cap;
    return XTX_OK;
}

int xtx_movelog_free(xtx_ctx* ctx) {
    if (!ctx) return XTX_E_ARG;
    ctx->plain_log.clear();
    ctx->plain_log.shrink_to_fit();
    ctx->plain_log_count.clear();
    ctx->plain_log_valid = false;
    return XTX_OK;
}

// ---------------------------------------------------------------- probes
int xtx_probe_moves(xtx_ctx* ctx, const int64_t* A, const int64_t* Xt, int32_t n, int32_t n_mats,
                    const int32_t* queries, int32_t n_queries, int64_t* out) {
    if (!ctx || !A || !Xt || !queries || !out || n < 2 || n > kMaxGroupN || n_mats < 1)
        return XTX_E_ARG;
    try {
        for (int q = 0; q < n_queries; ++q) {
            const int32_t* x = queries + 4 * q;
            if (x[0] < 0 || x[0] >= n_mats || x[1] < 0 || x[1] >= n || x[2] < 0 || x[2] >= n ||
                x[1] == x[2] || x[3] < -kShearCap || x[3] > kShearCap)
                throw Err{XTX_E_ARG, "bad probe query"};
        }
        HIPCHK(hipSetDevice(ctx->device));
        const size_t nn = (size_t)n * n * n_mats;
        int64_t *dA = nullptr, *dX = nullptr, *dout = nullptr;
        int32_t* dq = nullptr;
        HIPCHK(hipMalloc(&dA, nn * 8));
        HIPCHK(hipMalloc(&dX, nn * 8));
        HIPCHK(hipMalloc(&dq, (size_t)n_queries * 16));
        HIPCHK(hipMalloc(&dout, (size_t)n_queries * 48));
        HIPCHK(hipMemcpy(dA, A, nn * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dX, Xt, nn * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dq, queries, (size_t)n_queries * 16, hipMemcpyHostToDevice));
        const Layout L = layout_for(n);
        HIPCHK(launch_probe(ctx->stream, L.G, L.C, dA, dX, n, dq, n_queries, dout));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        HIPCHK(hipMemcpy(out, dout, (size_t)n_queries * 48, hipMemcpyDeviceToHost));
        (void)hipFree(dA);
        (void)hipFree(dX);
        (void)hipFree(dq);
        (void)hipFree(dout);
        return XTX_OK;
    } catch (const Err& e) {
        return fail(ctx, e);
    }
}

int xtx_probe_rng(xtx_ctx* ctx, const uint32_t* ctr_key, int32_t count, uint32_t* out,
                  const double* x, int32_t nx, double* exp_out) {
    if (!ctx || count < 0 || nx < 0) return XTX_E_ARG;
    try {
        HIPCHK(hipSetDevice(ctx->device));
        uint32_t *dck = nullptr, *dout = nullptr;
        double *dx = nullptr, *dex = nullptr;
        HIPCHK(hipMalloc(&dck, (size_t)std::max(count, 1) * 24));
        HIPCHK(hipMalloc(&dout, (size_t)std::max(count, 1) * 16));
        HIPCHK(hipMalloc(&dx, (size_t)std::max(nx, 1) * 8));
        HIPCHK(hipMalloc(&dex, (size_t)std::max(nx, 1) * 8));
        if (count) HIPCHK(hipMemcpy(dck, ctr_key, (size_t)count * 24, hipMemcpyHostToDevice));
        if (nx) HIPCHK(hipMemcpy(dx, x, (size_t)nx * 8, hipMemcpyHostToDevice));
        HIPCHK(launch_probe_rng(ctx->stream, dck, count, dout, dx, nx, dex));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        if (count) HIPCHK(hipMemcpy(out, dout, (size_t)count * 16, hipMemcpyDeviceToHost));
        if (nx) HIPCHK(hipMemcpy(exp_out, dex, (size_t)nx * 8, hipMemcpyDeviceToHost));
        (void)hipFree(dck);
        (void)hipFree(dout);
        (void)hipFree(dx);
        (void)hipFree(dex);
        return XTX_OK;
    } catch (const Err& e) {
        return fail(ctx, e);
    }
}

}  // extern "C"

// =====================================================================
// Batched and served forms share one scheduler: a set of resident
// instances, grouped into LDS layout buckets, advanced epoch by epoch.
// =====================================================================
struct Server {
    xtx_ctx* ctx = nullptr;
    xtx_params p{};
    xtx_serve_config cfg{};
    Arena arena;
    std::vector<std::unique_ptr<HostInst>> slots;  // resident, index = slot id
    std::vector<int> free_slots;
    std::deque<std::unique_ptr<HostInst>> queue;
    std::vector<xtx_result> done;
    std::vector<Inst> h_insts;
    Inst* d_insts = nullptr;
    int cap_insts = 0;
    std::map<BucketKey, Bucket> buckets;
    bool dirty = true;
    int64_t lds_in_use = 0;
    int64_t resident_reps = 0;
    bool auto_look = false;  // serve: the lookahead rule at each admission
    int max_r = 1;
    clk::time_point t0 = clk::now();
    double kernel_ms = 0;
    int epochs = 0;
    int64_t moves_retired = 0, accepts_retired = 0;
    int retired_since_rebuild = 0;
    // graph of several epochs, rebuilt when residency changes
    hipGraphExec_t gexec = nullptr;
    int graph_epochs = 0;
    // the overlapped batch loop: a pinned snapshot of the instance records
    // copied after the epochs in flight, and a stream for retirement copies
    // while they run
    Inst* pin_insts = nullptr;
    int pin_cap = 0;
    // gathered outcomes (k_gather_outcomes): device staging, its pinned host
    // copy, and the slot and offset lists
    unsigned char* d_stage = nullptr;
    unsigned char* h_stage = nullptr;
    size_t stage_cap = 0;
    int* d_gslots = nullptr;
    int64_t* d_goffs = nullptr;
    int gather_cap = 0;
    // per slot, host-mapped: raised by the device when the instance finishes
    int32_t* done_host = nullptr;
    int32_t* done_dev = nullptr;
    int done_cap = 0;
    int inflight_ns = 0, inflight_k = 0;
    float last_ms = 0;
    hipStream_t fetch_stream = nullptr;
    // the persistent launch (anneal_persist.hip): units, per-instance index
    // (first unit | units | pending), queue slots and counters (head, tail,
    // live, err)
    PUnit* d_units = nullptr;
    size_t units_cap = 0;
    int32_t* d_pidx = nullptr;
    size_t pidx_cap = 0;
    int32_t* d_pend = nullptr;
    size_t pend_cap = 0;
    int32_t* d_q = nullptr;
    size_t q_cap = 0;
    uint32_t* d_pcnt = nullptr;
    size_t pcnt_cap = 0;
    bool inflight_persist = false;
    // zero-copy retirement of the persistent launch: each instance's outcome
    // area in host-mapped memory (offsets per slot), which slots have been
    // taken from it, and outcomes read but not yet retired
    int64_t* d_zoff = nullptr;
    size_t zoff_cap = 0;
    std::vector<int64_t> zc_off;
    std::vector<char> zc_taken;
    std::map<size_t, Outcome> staged;
    bool zero_copy = false;

    ~Server() {
        // nothing in flight may still use what is freed here (a call that
        // throws can leave a launch running)
        if (ctx && ctx->stream) (void)hipStreamSynchronize(ctx->stream);
        if (fetch_stream) (void)hipStreamSynchronize(fetch_stream);
        if (gexec) (void)hipGraphExecDestroy(gexec);
        for (auto& kv : buckets) {
            if (kv.second.d_refs) (void)hipFree(kv.second.d_refs);
            if (kv.second.d_insts) (void)hipFree(kv.second.d_insts);
        }
        if (d_insts) (void)hipFree(d_insts);
        if (pin_insts) (void)hipHostFree(pin_insts);
        if (done_host) (void)hipHostFree(done_host);
        if (d_stage) (void)hipFree(d_stage);
        if (h_stage) (void)hipHostFree(h_stage);
        if (d_gslots) (void)hipFree(d_gslots);
        if (d_goffs) (void)hipFree(d_goffs);
        if (fetch_stream) (void)hipStreamDestroy(fetch_stream);
        if (d_units) (void)hipFree(d_units);
        if (d_pidx) (void)hipFree(d_pidx);
        if (d_pend) (void)hipFree(d_pend);
        if (d_q) (void)hipFree(d_q);
        if (d_pcnt) (void)hipFree(d_pcnt);
        if (d_zoff) (void)hipFree(d_zoff);
    }

    double now() const { return std::chrono::duration<double>(clk::now() - t0).count(); }

    // LDS the resident set occupies, charged by launch: every replica of a
    // launch holds the LDS share of that launch's largest order (orders that
    // share a lookahead kernel run in one launch sized for the largest).
    // launch_reps[key][order] = resident replicas of that order in that launch.
    std::map<BucketKey, std::map<int, int64_t>> launch_reps;

    static int64_t launch_lds(const BucketKey& k, const std::map<int, int64_t>& orders) {
        int64_t reps = 0;
        for (const auto& kv : orders) reps += kv.second;
        if (orders.empty()) return 0;
        const int nmax = orders.rbegin()->first;
        return reps * rep_bytes_of(nmax, std::get<1>(k), std::get<3>(k), std::get<5>(k),
                                   std::get<6>(k));
    }

    // LDS in use after adding (sign +1) or removing (-1) h's replicas
    int64_t lds_after(const HostInst& h, int sign) const {
        const BucketKey key = launch_key(h.lay);
        int64_t total = 0;
        bool seen = false;
        for (const auto& kv : launch_reps) {
            if (kv.first != key) {
                total += launch_lds(kv.first, kv.second);
                continue;
            }
            seen = true;
            auto m = kv.second;
            m[h.lay.nmax] += sign * (int64_t)h.R;
            if (m[h.lay.nmax] <= 0) m.erase(h.lay.nmax);
            total += launch_lds(kv.first, m);
        }
        if (!seen && sign > 0) {
            std::map<int, int64_t> m{{h.lay.nmax, (int64_t)h.R}};
            total += launch_lds(key, m);
        }
        return total;
    }

    void charge(const HostInst& h, int sign) {
        const BucketKey key = launch_key(h.lay);
        auto& m = launch_reps[key];
        m[h.lay.nmax] += sign * (int64_t)h.R;
        if (m[h.lay.nmax] <= 0) m.erase(h.lay.nmax);
        if (m.empty()) launch_reps.erase(key);
        lds_in_use = 0;
        for (const auto& kv : launch_reps) lds_in_use += launch_lds(kv.first, kv.second);
    }

    int64_t lds_budget() const {
        if (cfg.lds_budget_bytes > 0) return cfg.lds_budget_bytes;
        // every resident replica gets LDS once per pass of the device; allow
        // a few passes of oversubscription so the dispatcher always has work
        return (int64_t)ctx->prop.multiProcessorCount * kLdsPerCu * 4;
    }

    // Admits one instance if the budgets allow. With `keep`, its staging
    // buffers are parked there and the caller synchronizes the stream once
    // after a run of admissions; without, this call synchronizes.
    bool try_admit(std::unique_ptr<HostInst>& h, std::vector<Staging>* keep = nullptr) {
        if (cfg.max_resident > 0 && (int)(slots.size() - free_slots.size()) >= cfg.max_resident)
            return false;
        if (auto_look && p.lookahead == 0)
            apply_layout(ctx, *h, 0, lookahead_rule(ctx, resident_reps + h->R),
                         lane_load(ctx, resident_reps + h->R));
        if (lds_in_use > 0 && lds_after(*h, +1) > lds_budget()) return false;
        char* dm = arena.alloc(h->bytes);
        if (!dm) return false;
        Staging stg;
        upload(*h, dm, ctx->stream, stg);
        set_fractions(*h, stg, p);
        HIPCHK(hipMemcpyAsync(h->rec.rs, stg.rs.data(), (size_t)h->R * sizeof(RepState),
                              hipMemcpyHostToDevice, ctx->stream));
        if (keep)
            keep->push_back(std::move(stg));
        else
            HIPCHK(hipStreamSynchronize(ctx->stream));
        h->admit_s = now();
        charge(*h, +1);
        resident_reps += h->R;
        int slot;
        if (!free_slots.empty()) {
            slot = free_slots.back();
            free_slots.pop_back();
            slots[(size_t)slot] = std::move(h);
        } else {
            slot = (int)slots.size();
            slots.push_back(std::move(h));
        }
        (void)slot;
        dirty = true;
        return true;
    }

    void admit_queued() {
        while (!queue.empty()) {
            if (!try_admit(queue.front())) break;
            queue.pop_front();
        }
    }

    // Rebuild the device instance array and bucket tables after admissions or
    // retirements.
    void rebuild() {
        if (!dirty) return;
        const int ns = (int)slots.size();
        if (ns > cap_insts) {
            if (d_insts) (void)hipFree(d_insts);
            cap_insts = std::max(ns, 2 * cap_insts);
            HIPCHK(hipMalloc(&d_insts, (size_t)cap_insts * sizeof(Inst)));
        }
        h_insts.assign((size_t)ns, Inst{});
        if (ns > done_cap) {
            if (done_host) (void)hipHostFree(done_host);
            done_host = nullptr;
            done_cap = std::max(ns, 2 * done_cap);
            HIPCHK(hipHostMalloc((void**)&done_host, (size_t)done_cap * sizeof(int32_t),
                                 hipHostMallocMapped));
            HIPCHK(hipHostGetDevicePointer((void**)&done_dev, done_host, 0));
            std::memset(done_host, 0, (size_t)done_cap * sizeof(int32_t));
        }
        for (auto& kv : buckets) {
            kv.second.refs.clear();
            kv.second.insts.clear();
        }
        max_r = 1;
        for (int s = 0; s < ns; ++s) {
            if (!slots[(size_t)s]) {
                h_insts[(size_t)s].status = ST_IDLE;
                continue;
            }
            HostInst& h = *slots[(size_t)s];
            h_insts[(size_t)s] = h.rec;
            h_insts[(size_t)s].done_flag = done_dev + s;
            if (h.rec.status == ST_RUNNING) done_host[s] = 0;
            max_r = std::max(max_r, h.R);
            Bucket& b = buckets[launch_key(h.lay)];
            if (b.refs.empty() || h.lay.nmax > b.lay.nmax) {
                b.lay = h.lay;
                b.units_per_block = units_per_block_for(h.lay, p.waves_per_block);
            }
            for (int r = 0; r < h.R; ++r) b.refs.push_back(RepRef{s, r});
            pad_to_wave(b.refs, h.lay.G * h.lay.K, s);
            b.insts.push_back(s);
        }
        for (auto& kv : buckets) {
            Bucket& b = kv.second;
            split_lane_bucket(b, slots);
            if ((int)b.insts.size() > b.insts_cap) {
                if (b.d_insts) (void)hipFree(b.d_insts);
                b.insts_cap = std::max((int)b.insts.size(), 2 * b.insts_cap);
                HIPCHK(hipMalloc(&b.d_insts, (size_t)b.insts_cap * sizeof(int)));
            }
            if (!b.insts.empty())
                HIPCHK(hipMemcpyAsync(b.d_insts, b.insts.data(), b.insts.size() * sizeof(int),
                                      hipMemcpyHostToDevice, ctx->stream));
            if ((int)b.refs.size() > b.cap) {
                if (b.d_refs) (void)hipFree(b.d_refs);
                b.cap = std::max((int)b.refs.size(), 2 * b.cap);
                HIPCHK(hipMalloc(&b.d_refs, (size_t)b.cap * sizeof(RepRef)));
            }
            if (!b.refs.empty())
                HIPCHK(hipMemcpyAsync(b.d_refs, b.refs.data(), b.refs.size() * sizeof(RepRef),
                                      hipMemcpyHostToDevice, ctx->stream));
        }
        if (ns)
            HIPCHK(hipMemcpyAsync(d_insts, h_insts.data(), (size_t)ns * sizeof(Inst),
                                  hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        if (gexec) {
            (void)hipGraphExecDestroy(gexec);
            gexec = nullptr;
        }
        dirty = false;
        retired_since_rebuild = 0;
    }

    std::vector<Bucket> bucket_list() {
        std::vector<Bucket> v;
        for (auto& kv : buckets)
            if (!kv.second.refs.empty()) v.push_back(kv.second);
        return v;
    }

    // Runs k epochs: captured once into a graph per residency, replayed.
    // Launches k epochs as one captured graph (recaptured when k or the
    // launch tables change) and the copy of the instance records into pinned
    // memory, without waiting; finish_epochs waits.
    void launch_epochs(int k) {
        const int ns = (int)slots.size();
        if (ns == 0) return;
        if (pin_cap < ns) {
            if (pin_insts) (void)hipHostFree(pin_insts);
            pin_insts = nullptr;
            pin_cap = std::max(ns, 2 * pin_cap);
            HIPCHK(hipHostMalloc((void**)&pin_insts, (size_t)pin_cap * sizeof(Inst),
                                 hipHostMallocDefault));
        }
        if (!gexec || graph_epochs != k) {
            if (gexec) (void)hipGraphExecDestroy(gexec);
            gexec = nullptr;
            std::vector<Bucket> bl = bucket_list();
            hipGraph_t graph;
            HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            launch_epochs_chained(ctx, bl, d_insts, max_r, k);
            HIPCHK(hipStreamEndCapture(ctx->stream, &graph));
            HIPCHK(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
            (void)hipGraphDestroy(graph);
            graph_epochs = k;
        }
        HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
        HIPCHK(hipGraphLaunch(gexec, ctx->stream));
        HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
        HIPCHK(hipMemcpyAsync(pin_insts, d_insts, (size_t)ns * sizeof(Inst),
                              hipMemcpyDeviceToHost, ctx->stream));
        inflight_ns = ns;
        inflight_k = k;
    }

    // Whether every running instance can run in the persistent launch: the
    // lane kernel by its rule (orders to 48, lane_q proposals per step), the
    // L1 objective, at most kPersistMaxR replicas.
    static constexpr int kPersistMaxR = 256;
    bool persistent_ok() const {
        bool any = false;
        for (size_t s = 0; s < slots.size() && s < h_insts.size(); ++s) {
            if (!slots[s] || h_insts[s].status != ST_RUNNING) continue;
            const HostInst& h = *slots[s];
            const Layout& L = h.lay;
            if (L.G != 1 || L.nmax > kLaneMaxOrder || L.K != lane_q(L.nmax) || L.xl ||
                h.R > kPersistMaxR || h.rec.dual)
                return false;
            any = true;
        }
        return any;
    }

    template <typename T>
    static void ensure(T*& p, size_t& cap, size_t count) {
        if (count <= cap && p) return;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = std::max(count, 2 * cap);
        HIPCHK(hipMalloc((void**)&p, cap * sizeof(T)));
    }

    // The device's XCDs (dies) and compute-unit pairs, as k_xcc_probe reports
    // them (XCD id << 8 | CU-pair key): one-wave workgroups dispatched across
    // the device, each reporting where it ran. Once per context.
    const std::vector<int>& cu_pairs() {
        if (!ctx->xcc_ids.empty()) return ctx->xcc_ids;
        const int blocks = std::max(256, 16 * ctx->prop.multiProcessorCount);
        int32_t* d = nullptr;
        HIPCHK(hipMalloc((void**)&d, (size_t)blocks * sizeof(int32_t)));
        std::vector<int32_t> h((size_t)blocks, -1);
        hipError_t err = launch_xcc_probe(ctx->stream, d, blocks);
        if (err == hipSuccess)
            err = hipMemcpyAsync(h.data(), d, (size_t)blocks * sizeof(int32_t),
                                 hipMemcpyDeviceToHost, ctx->stream);
        if (err == hipSuccess) err = hipStreamSynchronize(ctx->stream);
        (void)hipFree(d);
        HIPCHK(err);
        std::vector<int> ids;
        for (int32_t v : h)
            if (v >= 0) ids.push_back(((v >> 8) & (kMaxXcd - 1)) << 8 | (v & (kCuKeys - 1)));
        std::sort(ids.begin(), ids.end());
        ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
        if (ids.empty()) throw Err{XTX_E_DEVICE, "XCD probe: no ids"};
        ctx->xcc_ids = ids;
        return ctx->xcc_ids;
    }

    // Launches up to k more epochs of every running instance as one
    // persistent kernel (k_lane_persist), without waiting; finish_epochs
    // waits. Every instance is placed on one XCD (the most replicas first,
    // each onto the die with the fewest so far: a replica's epoch costs about
    // the same at every order) and its units go to that die's queue of its
    // order class and epoch slot, the largest orders first. Each die's
    // compute-unit pairs are given home classes in proportion to the die's
    // replicas of each class (a unit's time follows its replicas: measured
    // 2.0 ms for the four of an order-24 unit, 1.1 ms for the two of an
    // order-48 one). Each queue is sized for every unit its instances' k
    // epochs can append. The grid leaves two workgroups of the device free,
    // so the copies that retire finished instances (blit kernels) run while
    // it does.
    uint32_t persist_limit = 0;
    void launch_persistent(int k) {
        const int ns = (int)slots.size();
        if (ns == 0) return;
        if (pin_cap < ns) {
            if (pin_insts) (void)hipHostFree(pin_insts);
            pin_insts = nullptr;
            pin_cap = std::max(ns, 2 * pin_cap);
            HIPCHK(hipHostMalloc((void**)&pin_insts, (size_t)pin_cap * sizeof(Inst),
                                 hipHostMallocDefault));
        }
        const std::vector<int>& pairs = cu_pairs();
        std::vector<int> ids;  // XCD ids
        for (int v : pairs)
            if (ids.empty() || ids.back() != (v >> 8)) ids.push_back(v >> 8);
        const int nq = (int)ids.size();
        uint32_t e_hi = 0;
        std::vector<int> run;
        for (int s = 0; s < ns; ++s)
            if (slots[(size_t)s] && h_insts[(size_t)s].status == ST_RUNNING) {
                run.push_back(s);
                e_hi = std::max(e_hi, h_insts[(size_t)s].epoch);
            }
        const uint32_t limit = e_hi + (uint32_t)k;
        uint32_t e_lo = e_hi;
        for (int s : run) e_lo = std::min(e_lo, h_insts[(size_t)s].epoch);
        // placement
        std::stable_sort(run.begin(), run.end(), [&](int a, int b) {
            return slots[(size_t)a]->R > slots[(size_t)b]->R;
        });
        std::vector<int64_t> load((size_t)nq, 0);
        std::vector<std::vector<int>> on((size_t)nq);
        for (int s : run) {
            const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            load[(size_t)x] += slots[(size_t)s]->R;
            on[(size_t)x].push_back(s);
        }
        std::vector<PUnit> units;
        std::vector<int32_t> idx((size_t)2 * ns, 0);
        std::vector<int32_t> pend((size_t)ns * kLineWords, 0);
        for (int s : run) {
            const HostInst& h = *slots[(size_t)s];
            const int per = 64 / lane_q(h.lay.nmax);
            const int nu = (h.R + per - 1) / per;
            idx[(size_t)s] = (int32_t)units.size();
            idx[(size_t)ns + s] = nu;
            pend[(size_t)s * kLineWords] = nu;
            for (int u = 0; u < nu; ++u)
                units.push_back(PUnit{s, (int16_t)(u * per), (int16_t)h.lay.nmax});
        }
        if (units.size() >= ((size_t)1 << 24)) throw Err{XTX_E_NOMEM, "persistent launch: too many units"};
        const int bpc = lane_persist_blocks_per_cu();
        if (bpc < 1) throw Err{XTX_E_DEVICE, "persistent kernel: no occupancy"};
        const int blocks = std::max(1, bpc * ctx->prop.multiProcessorCount - 2);
        PersistCtl c{};
        std::vector<int32_t> q0;
        std::vector<uint32_t> qbase((size_t)kQueues, 0u), qcap((size_t)kQueues, 0u);
        const size_t nlines = (size_t)2 * kQueues + kMaxXcd + 1;
        std::vector<uint32_t> cnt(nlines * kLineWords, 0u);
        uint32_t* tail_h = cnt.data() + (size_t)kQueues * kLineWords;
        uint32_t* live_h = cnt.data() + (size_t)2 * kQueues * kLineWords;
        std::vector<int8_t> home((size_t)kMaxXcd * kCuKeys, (int8_t)(kClasses - 1));
        size_t qtotal = 0;
        for (int x = 0; x < nq; ++x) {
            auto& v = on[(size_t)x];
            std::stable_sort(v.begin(), v.end(), [&](int a, int b) {
                return slots[(size_t)a]->lay.nmax > slots[(size_t)b]->lay.nmax;
            });
            int64_t creps[kClasses] = {};  // a unit's time follows its replicas
            for (int cl = 0; cl < kClasses; ++cl)
                for (int b = 0; b < kEpochSlots; ++b) {
                    // epochs e0 .. limit-1 of each instance; those in slot b
                    size_t tasks = 0;
                    const size_t first = q0.size();
                    for (int s : v) {
                        if (order_class(slots[(size_t)s]->lay.nmax) != cl) continue;
                        const uint32_t e0 = h_insts[(size_t)s].epoch;
                        const int nu = idx[(size_t)ns + s];
                        for (uint32_t ep = e0; ep < limit; ++ep)
                            if ((int)(std::min<uint32_t>(ep - e_lo, 125) & (kEpochSlots - 1)) == b)
                                tasks += (size_t)nu;
                        const int rel = std::min<int>((int)(e0 - e_lo), 125);
                        if ((rel & (kEpochSlots - 1)) == b)
                            for (int u = 0; u < nu; ++u) q0.push_back(rel << 24 | (idx[(size_t)s] + u));
                        if (b == 0) creps[cl] += slots[(size_t)s]->R;
                    }
                    const size_t cap = (tasks + 64 + kLineWords - 1) / kLineWords * kLineWords;
                    const int qi = (x * kClasses + cl) * kEpochSlots + b;
                    qbase[(size_t)qi] = (uint32_t)qtotal;
                    qcap[(size_t)qi] = (uint32_t)cap;
                    tail_h[(size_t)qi * kLineWords] = (uint32_t)(q0.size() - first);
                    // the queue's initial units, at its base
                    std::vector<int32_t> mine(q0.begin() + (ptrdiff_t)first, q0.end());
                    q0.resize(first);
                    q0.resize(qtotal, -1);
                    q0.insert(q0.end(), mine.begin(), mine.end());
                    qtotal += cap;
                    if (qtotal >= (size_t)INT32_MAX)
                        throw Err{XTX_E_NOMEM, "persistent queue too large"};
                }
            live_h[(size_t)x * kLineWords] = (uint32_t)v.size();
            // home classes of the die's CU pairs, in proportion to its replicas
            std::vector<int> keys;
            for (int pv : pairs)
                if ((pv >> 8) == ids[(size_t)x]) keys.push_back(pv & (kCuKeys - 1));
            const int64_t tot = creps[0] + creps[1] + creps[2];
            int at = 0;
            for (int cl = 0; cl < kClasses && tot > 0; ++cl) {
                int want = (int)((creps[cl] * (int64_t)keys.size() + tot / 2) / tot);
                if (creps[cl] > 0 && want == 0) want = 1;
                for (int i = 0; i < want && at < (int)keys.size(); ++i, ++at)
                    home[(size_t)x * kCuKeys + keys[(size_t)at]] = (int8_t)cl;
            }
        }
        // every slot's outcome area in host-mapped memory
        zc_off.assign((size_t)ns, -1);
        size_t zbytes = 0;
        for (int s = 0; s < ns; ++s) {
            if (!slots[(size_t)s]) continue;
            const HostInst& h = *slots[(size_t)s];
            zc_off[(size_t)s] = (int64_t)zbytes;
            zbytes += (32 + (size_t)h.R * sizeof(RepState) + 16 * (size_t)h.n * h.n + 127) / 128 * 128;
        }
        if (zbytes > ctx->zc_cap) {
            if (ctx->zc_host) (void)hipHostFree(ctx->zc_host);
            ctx->zc_host = ctx->zc_dev = nullptr;
            ctx->zc_cap = 0;
            HIPCHK(hipHostMalloc((void**)&ctx->zc_host, zbytes, hipHostMallocMapped));
            HIPCHK(hipHostGetDevicePointer((void**)&ctx->zc_dev, ctx->zc_host, 0));
            ctx->zc_cap = zbytes;
        }
        zc_taken.assign((size_t)ns, 0);
        ensure(d_zoff, zoff_cap, zc_off.size());
        zero_copy = true;
        ensure(d_units, units_cap, units.size());
        ensure(d_pidx, pidx_cap, idx.size() + (home.size() + 3) / 4 + 2 * (size_t)kQueues);
        ensure(d_pend, pend_cap, pend.size());
        ensure(d_q, q_cap, qtotal);
        ensure(d_pcnt, pcnt_cap, cnt.size());
        uint32_t* d_qbase = (uint32_t*)(d_pidx + idx.size());
        uint32_t* d_qcap = d_qbase + kQueues;
        int8_t* d_home = (int8_t*)(d_qcap + kQueues);
        HIPCHK(hipMemsetAsync(d_q, 0xff, qtotal * sizeof(int32_t), ctx->stream));
        HIPCHK(hipMemcpyAsync(d_q, q0.data(), q0.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                              ctx->stream));
        HIPCHK(hipMemcpyAsync(d_units, units.data(), units.size() * sizeof(PUnit),
                              hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipMemcpyAsync(d_pidx, idx.data(), idx.size() * sizeof(int32_t),
                              hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipMemcpyAsync(d_home, home.data(), home.size(), hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipMemcpyAsync(d_qbase, qbase.data(), qbase.size() * sizeof(uint32_t),
                              hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipMemcpyAsync(d_qcap, qcap.data(), qcap.size() * sizeof(uint32_t),
                              hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipMemcpyAsync(d_pend, pend.data(), pend.size() * sizeof(int32_t),
                              hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipMemcpyAsync(d_zoff, zc_off.data(), zc_off.size() * sizeof(int64_t),
                              hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipMemcpyAsync(d_pcnt, cnt.data(), cnt.size() * sizeof(uint32_t),
                              hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        c.q = d_q;
        c.qbase = d_qbase;
        c.qcap = d_qcap;
        c.units = d_units;
        c.ubase = d_pidx;
        c.unum = d_pidx + ns;
        c.home = d_home;
        c.pend = d_pend;
        c.head = d_pcnt;
        c.tail = d_pcnt + (size_t)kQueues * kLineWords;
        c.live = (int32_t*)(d_pcnt + (size_t)2 * kQueues * kLineWords);
        c.err = (int32_t*)(d_pcnt + ((size_t)2 * kQueues + kMaxXcd) * kLineWords);
        c.out = ctx->zc_dev;
        c.out_off = d_zoff;
        for (int h = 0; h < kMaxXcd; ++h) c.xq[h] = -1;
        for (int x = 0; x < nq; ++x) c.xq[ids[(size_t)x]] = x;
        c.nq = nq;
        c.epoch_limit = limit;
        c.epoch_base = e_lo;
        persist_limit = limit;
        HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
        HIPCHK(launch_lane_persist(ctx->stream, c, d_insts, blocks, ctx->stop_dev));
        HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
        HIPCHK(hipMemcpyAsync(pin_insts, d_insts, (size_t)ns * sizeof(Inst),
                              hipMemcpyDeviceToHost, ctx->stream));
        inflight_ns = ns;
        inflight_k = k;
        inflight_persist = true;
    }

    // Collects, while epochs are in flight, the instances the device has
    // flagged as finished: each one's record, then its outcome on the fetch
    // stream, verified on host threads (retire). Their state is final, and
    // nothing the running epochs do touches it.
    void collect_done() {
        if (!done_host) return;
        bool any = false;
        const int ns = std::min((int)slots.size(), std::min(done_cap, (int)h_insts.size()));
        if (zero_copy) {
            // the persistent launch wrote each ended instance's outcome to
            // host-mapped memory before raising its flag: read it there
            for (int s = 0; s < ns && s < (int)zc_taken.size(); ++s) {
                if (!slots[(size_t)s] || zc_taken[(size_t)s] || zc_off[(size_t)s] < 0) continue;
                if (__atomic_load_n(&done_host[s], __ATOMIC_ACQUIRE) == 0) continue;
                const HostInst& h = *slots[(size_t)s];
                const unsigned char* src = ctx->zc_host + zc_off[(size_t)s];
                const int32_t* head = (const int32_t*)src;
                if (head[1]) throw Err{XTX_E_STATE, "internal: best replica has a stale snapshot"};
                Outcome out;
                out.rep = head[0];
                const int64_t sm = ((const int64_t*)src)[2];
                Inst& rec = h_insts[(size_t)s];
                rec.status = head[2];
                if (rec.status == ST_SOLVED) {
                    out.solved = true;
                    out.solve_move = sm;
                    rec.solve_move = sm;
                    rec.solve_rep = head[0];
                }
                out.rs.resize((size_t)h.R);
                std::memcpy(out.rs.data(), src + 32, (size_t)h.R * sizeof(RepState));
                outcome_tally(out);
                const size_t nn = (size_t)h.n * h.n;
                const int64_t* a = (const int64_t*)(src + 32 + (size_t)h.R * sizeof(RepState));
                out.A.assign(a, a + nn);
                out.Xt.assign(a + nn, a + 2 * nn);
                staged[(size_t)s] = std::move(out);
                zc_taken[(size_t)s] = 1;
                any = true;
            }
            if (any) retire(false, 0, false, fetch_stream);
            return;
        }
        for (int s = 0; s < ns; ++s) {
            if (!slots[(size_t)s] || h_insts[(size_t)s].status != ST_RUNNING) continue;
            if (__atomic_load_n(&done_host[s], __ATOMIC_ACQUIRE) == 0) continue;
            HIPCHK(hipMemcpyAsync(&h_insts[(size_t)s], d_insts + s, sizeof(Inst),
                                  hipMemcpyDeviceToHost, fetch_stream));
            any = true;
        }
        if (!any) return;
        HIPCHK(hipStreamSynchronize(fetch_stream));
        retire(false, 0, false, fetch_stream);
    }

    void finish_epochs() {
        if (inflight_ns == 0) return;
        HIPCHK(hipStreamSynchronize(ctx->stream));
        const int ns = std::min(inflight_ns, (int)std::min(slots.size(), h_insts.size()));
        for (int s = 0; s < ns; ++s) {
            h_insts[(size_t)s] = pin_insts[s];
            if (slots[(size_t)s]) slots[(size_t)s]->rec = pin_insts[s];
        }
        HIPCHK(hipEventElapsedTime(&last_ms, ctx->ev0, ctx->ev1));
        kernel_ms += last_ms;
        epochs += inflight_k;
        inflight_ns = 0;
        if (inflight_persist) {
            inflight_persist = false;
            int32_t err = 0;
            HIPCHK(hipMemcpy(&err, d_pcnt + ((size_t)2 * kQueues + kMaxXcd) * kLineWords,
                             sizeof(err), hipMemcpyDeviceToHost));
            if (err) throw Err{XTX_E_DEVICE, "persistent launch: queue overflow"};
            // every instance ran to its end or to the launch's epoch limit
            // (unless stopped): a queue whose die ran no wave would leave its
            // instances behind
            if (__atomic_load_n(ctx->stop_host, __ATOMIC_SEQ_CST) == 0)
                for (int s = 0; s < ns; ++s)
                    if (slots[(size_t)s] && h_insts[(size_t)s].status == ST_RUNNING &&
                        h_insts[(size_t)s].epoch < persist_limit)
                        throw Err{XTX_E_DEVICE, "persistent launch ended before its instances"};
        }
    }

    // graph: replay the k epochs as one captured graph (recaptured when k or
    // the launch tables change), also for k = 1
    void run_epochs(int k, bool graph = false) {
        const int ns = (int)slots.size();
        if (ns == 0) return;
        std::vector<Bucket> bl = bucket_list();
        if (k > 1 || graph) {
            if (!gexec || graph_epochs != k) {
                if (gexec) (void)hipGraphExecDestroy(gexec);
                hipGraph_t graph;
                HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
                launch_epochs_chained(ctx, bl, d_insts, max_r, k);
                HIPCHK(hipStreamEndCapture(ctx->stream, &graph));
                HIPCHK(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
                (void)hipGraphDestroy(graph);
                graph_epochs = k;
            }
            HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
            HIPCHK(hipGraphLaunch(gexec, ctx->stream));
            HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
        } else {
            HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
            launch_epochs_chained(ctx, bl, d_insts, max_r, 1);
            HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
        }
        HIPCHK(hipMemcpyAsync(h_insts.data(), d_insts, (size_t)ns * sizeof(Inst),
                              hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        for (int s = 0; s < ns; ++s)
            if (slots[(size_t)s]) slots[(size_t)s]->rec = h_insts[(size_t)s];
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        kernel_ms += ms;
        epochs += k;
    }

    // Retires every resident instance whose status left RUNNING (or all, when
    // force); writes outputs and queues results.
    // st: the stream for the outcome copies (default: the engine's). The
    // copies of all retiring instances share two waits; outputs and exact
    // verification run on host threads.
    // Outcomes of many instances (as outcome_records + outcome_state give
    // them) through one gather kernel and one copy: four small copies per
    // instance cost about 50 ms when a batch's 800 capped instances retire
    // together.
    static constexpr size_t kGatherMin = 16;
    void gather(const std::vector<size_t>& fin, hipStream_t st, std::vector<Outcome>& outs) {
        const int m = (int)fin.size();
        std::vector<int> sl((size_t)m);
        std::vector<int64_t> off((size_t)m);
        size_t total = 0;
        for (int i = 0; i < m; ++i) {
            const HostInst& h = *slots[fin[(size_t)i]];
            sl[(size_t)i] = (int)fin[(size_t)i];
            off[(size_t)i] = (int64_t)total;
            total += 16 + (size_t)h.R * sizeof(RepState) + 2 * (size_t)h.n * h.n * 8;
        }
        if (total > stage_cap) {
            if (d_stage) (void)hipFree(d_stage);
            if (h_stage) (void)hipHostFree(h_stage);
            d_stage = h_stage = nullptr;
            stage_cap = std::max(total, 2 * stage_cap);
            HIPCHK(hipMalloc(&d_stage, stage_cap));
            HIPCHK(hipHostMalloc((void**)&h_stage, stage_cap, hipHostMallocDefault));
        }
        if (m > gather_cap) {
            if (d_gslots) (void)hipFree(d_gslots);
            if (d_goffs) (void)hipFree(d_goffs);
            gather_cap = std::max(m, 2 * gather_cap);
            HIPCHK(hipMalloc(&d_gslots, (size_t)gather_cap * sizeof(int)));
            HIPCHK(hipMalloc(&d_goffs, (size_t)gather_cap * sizeof(int64_t)));
        }
        HIPCHK(hipMemcpyAsync(d_gslots, sl.data(), (size_t)m * sizeof(int), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d_goffs, off.data(), (size_t)m * sizeof(int64_t),
                              hipMemcpyHostToDevice, st));
        HIPCHK(launch_gather_outcomes(st, d_insts, d_gslots, d_goffs, m, d_stage));
        HIPCHK(hipMemcpyAsync(h_stage, d_stage, total, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        for (int i = 0; i < m; ++i) {
            const HostInst& h = *slots[fin[(size_t)i]];
            const Inst& rec = h_insts[fin[(size_t)i]];
            const unsigned char* src = h_stage + off[(size_t)i];
            const size_t nn = (size_t)h.n * h.n;
            Outcome& out = outs[(size_t)i];
            const int32_t* head = (const int32_t*)src;
            if (head[1]) throw Err{XTX_E_STATE, "internal: best replica has a stale snapshot"};
            out.rs.resize((size_t)h.R);
            std::memcpy(out.rs.data(), src + 16, (size_t)h.R * sizeof(RepState));
            outcome_tally(out);
            out.rep = head[0];
            if (rec.status == ST_SOLVED) {
                out.solved = true;
                out.solve_move = rec.solve_move;
            }
            const int64_t* a = (const int64_t*)(src + 16 + (size_t)h.R * sizeof(RepState));
            out.A.assign(a, a + nn);
            out.Xt.assign(a + nn, a + 2 * nn);
        }
    }

    void retire(bool force, uint64_t only_tag = 0, bool by_tag = false, hipStream_t st = nullptr) {
        if (!st) st = ctx->stream;
        std::vector<size_t> fin;
        for (size_t s = 0; s < slots.size(); ++s) {
            if (!slots[s]) continue;
            HostInst& h = *slots[s];
            if (by_tag && h.tag != only_tag) continue;
            if (h_insts[s].status == ST_RUNNING && !force && !by_tag) continue;
            h.rec = h_insts[s];
            fin.push_back(s);
        }
        std::vector<Outcome> outs(fin.size());
        // outcomes already read from host-mapped memory, then the others
        // copied from the device
        std::vector<size_t> dev;
        std::vector<size_t> dev_at;
        for (size_t i = 0; i < fin.size(); ++i) {
            auto it = staged.find(fin[i]);
            if (it != staged.end()) {
                outs[i] = std::move(it->second);
                staged.erase(it);
            } else {
                dev.push_back(fin[i]);
                dev_at.push_back(i);
            }
        }
        std::vector<Outcome> douts(dev.size());
        if (dev.size() >= kGatherMin) {
            gather(dev, st, douts);
        } else {
            for (size_t i = 0; i < dev.size(); ++i)
                outcome_records(*slots[dev[i]], h_insts[dev[i]], st, douts[i]);
            if (!dev.empty()) HIPCHK(hipStreamSynchronize(st));
            for (size_t i = 0; i < dev.size(); ++i)
                outcome_state(*slots[dev[i]], h_insts[dev[i]], st, douts[i]);
            if (!dev.empty()) HIPCHK(hipStreamSynchronize(st));
        }
        for (size_t i = 0; i < dev.size(); ++i) outs[dev_at[i]] = std::move(douts[i]);
        std::vector<int> vstat(fin.size(), -1);
        parallel_for(fin.size(), [&](size_t i) {
            const HostInst& h = *slots[fin[i]];
            outs[i].best_score = score_of(outs[i].A.data(), h.n);
            write_outputs(h, outs[i], h.A_out, h.X_out);
            if (p.verify) vstat[i] = verify_outcome(h, outs[i]);
        });
        for (size_t i = 0; i < fin.size(); ++i) {
            const size_t s = fin[i];
            HostInst& h = *slots[s];
            const Outcome& o = outs[i];
            xtx_result res{};
            res.solved = o.solved;
            res.verify_status = vstat[i];
            res.solve_move = o.solve_move;
            res.best_score = o.best_score;
            res.moves = o.moves;
            res.submit_s = h.submit_s;
            res.admit_s = h.admit_s;
            res.retire_s = now();
            res.tag = h.tag;
            done.push_back(res);
            moves_retired += o.moves;
            accepts_retired += o.accepts;
            charge(h, -1);
            resident_reps -= h.R;
            arena.free(h.dmem, h.bytes);
            slots[s].reset();
            free_slots.push_back((int)s);
            dirty = true;
            ++retired_since_rebuild;
        }
        // trim trailing free slots so the instance array stays compact
        while (!slots.empty() && !slots.back()) {
            slots.pop_back();
            free_slots.erase(std::remove(free_slots.begin(), free_slots.end(), (int)slots.size()),
                             free_slots.end());
            dirty = true;
        }
    }

    int resident() const { return (int)(slots.size() - free_slots.size()); }
};

namespace {

std::unique_ptr<HostInst> make_inst(const xtx_ctx* ctx, const xtx_instance& in,
                                    const xtx_params& p) {
    auto h = std::make_unique<HostInst>();
    const int R = in.replicas > 0 ? in.replicas : p.replicas;
    xtx_params q = p;
    q.replicas = R;
    check_params(q);
    prepare(*h, in.A0, in.X0, in.n, R, in.seed, in.active, in.n_active, q);
    apply_layout(ctx, *h, p.lookahead, 1);  // batch and serve apply their rule
    if (h->n < 2 || h->N < 2)
        throw Err{XTX_E_ARG, "instance needs n >= 2 and at least two active indices"};
    h->bytes = device_bytes(*h);
    h->A_out = in.A_out;
    h->X_out = in.X_out;
    return h;
}

}  // namespace

// ---------------------------------------------------------------- launch shapes
namespace {

void fill_launch(const Layout& L, bool tuned, xtx_launch* out) {
    *out = xtx_launch{};
    out->order = L.nmax;
    out->lanes = L.G;
    out->cols = L.C;
    out->triangular = L.tri ? 1 : 0;
    out->waves_per_block = L.waves;
    out->tuned = tuned ? 1 : 0;
    out->lds_bytes = L.rep_bytes;
    out->lookahead = L.K;
    out->x_in_lds = L.xl ? 1 : 0;
}

// A configuration the kernels can run for this order, or an error message.
const char* launch_invalid(int nmax, int G, int C, bool tri, int waves, int K, bool xl) {
    if (nmax < 8 || nmax > kMaxN || nmax % 8) return "order must be a multiple of 8 in [8, 256]";
    if (nmax > kMaxGroupN && (G != 1 || (nmax != 192 && nmax != 256)))
        return "orders above 128 run the wide lane kernel only (lanes 1, cols = order, order "
               "192 or 256, lookahead 32)";
    if (G * C < nmax) return "lanes * cols must cover the order";
    if (G == 1) {
        // the lane kernel: one lane per proposal, the whole row per lane
        if (C != nmax || !tri || !lane_variant_exists(nmax, K))
            return "no lane kernel with this order/lookahead (lanes 1, cols = order, "
                   "triangular, lookahead 16 or 32, orders to 48)";
        if (waves < 0 || waves > 4 || waves == 3) return "waves_per_block must be 0, 1, 2 or 4";
        if ((int64_t)std::max(1, waves) * (64 / K) * rep_bytes_of(nmax, G, tri, K) > g_lds_block)
            return "replica LDS exceeds a workgroup";
        return nullptr;
    }
    if (!anneal_variant_exists(G, C, tri, K, xl))
        return xl ? "no kernel with X^T in LDS at these lanes/cols/layout/lookahead"
                  : "no kernel with these lanes/cols/layout/lookahead";
    if (waves < 0 || waves > 4 || waves == 3) return "waves_per_block must be 0, 1, 2 or 4";
    if (K > 1 && waves != 0) return "a lookahead replica takes its own wave or workgroup (waves 0)";
    const int rb = rep_bytes_of(nmax, G, tri, K, xl);
    if ((int64_t)std::max(1, 64 / (G * K)) * rb > g_lds_block) return "replica LDS exceeds a workgroup";
    return nullptr;
}

// Every valid configuration of an order: G with C = ceil(nmax / G) <= 2,
// both layouts, 1, 2 or 4 waves per workgroup (those the LDS allows); and the
// lookahead configurations (K = 2 .. 32 groups per replica, triangular A).
std::vector<Layout> candidates(int nmax) {
    std::vector<Layout> v;
    for (int G : {16, 32, 64}) {
        const int C = (nmax + G - 1) / G;
        for (bool tri : {true, false})
            for (int w : {1, 2, 4}) {
                if (launch_invalid(nmax, G, C, tri, w, 1, false)) continue;
                const Layout L = make_layout(nmax, G, C, tri, w);
                if ((int64_t)w * (64 / G) * L.rep_bytes > g_lds_block) continue;
                v.push_back(L);
            }
    }
    for (int K : {2, 4, 8, 16, 32})
        for (int G : {16, 32, 64}) {
            const int C = (nmax + G - 1) / G;
            for (bool tri : {true, false})
                for (bool xl : {false, true})
                    if (!launch_invalid(nmax, G, C, tri, 0, K, xl))
                        v.push_back(make_layout(nmax, G, C, tri, 0, K, xl));
        }
    for (int Q : {16, 32})
        for (int w : {1, 2, 4})
            if (nmax <= kLaneMaxOrder && !launch_invalid(nmax, 1, nmax, true, w, Q, false))
                v.push_back(make_layout(nmax, 1, nmax, true, w, Q));
    return v;
}

// One timed launch of `moves` proposals per replica over copies of the given
// instances in layout L; returns the kernel milliseconds.
double time_layout(xtx_ctx* ctx, const std::vector<const HostInst*>& proto, const Layout& L,
                   const xtx_params& p) {
    std::vector<HostInst> hs;
    hs.reserve(proto.size());
    size_t total = 0;
    for (const HostInst* h : proto) {
        hs.push_back(*h);
        hs.back().lay = L;
        total += h->bytes;
    }
    Arena arena;
    arena.reserve(total + 4096);
    std::vector<Staging> keep(hs.size());
    std::vector<Inst> recs(hs.size());
    Bucket b;
    b.lay = L;
    b.units_per_block = units_per_block_for(L, p.waves_per_block);
    for (size_t q = 0; q < hs.size(); ++q) {
        HostInst& h = hs[q];
        upload(h, arena.alloc(h.bytes), ctx->stream, keep[q]);
        set_fractions(h, keep[q], p);
        HIPCHK(hipMemcpyAsync(h.rec.rs, keep[q].rs.data(), (size_t)h.R * sizeof(RepState),
                              hipMemcpyHostToDevice, ctx->stream));
        recs[q] = h.rec;
        for (int r = 0; r < h.R; ++r) b.refs.push_back(RepRef{(int)q, r});
        pad_to_wave(b.refs, L.G * L.K, (int)q);
    }
    Inst* d_insts = nullptr;
    HIPCHK(hipMalloc(&d_insts, recs.size() * sizeof(Inst)));
    HIPCHK(hipMalloc(&b.d_refs, b.refs.size() * sizeof(RepRef)));
    double ms_out = 0;
    try {
        HIPCHK(hipMemcpyAsync(d_insts, recs.data(), recs.size() * sizeof(Inst),
                              hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipMemcpyAsync(b.d_refs, b.refs.data(), b.refs.size() * sizeof(RepRef),
                              hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
        HIPCHK(launch_layout(ctx->stream, L, b.units_per_block, b.d_refs, (int)b.refs.size(),
                             d_insts, ctx->stop_dev));
        HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        ms_out = ms;
    } catch (...) {
        (void)hipFree(d_insts);
        (void)hipFree(b.d_refs);
        throw;
    }
    (void)hipFree(d_insts);
    (void)hipFree(b.d_refs);
    return ms_out;
}

// Tunes every order of `sample` not yet in ctx->tuned (all of them when
// `all`), appending measurements to `out`.
void autotune_orders(xtx_ctx* ctx, const xtx_instance* sample, int count,
                     const xtx_params& params, int moves, bool all,
                     std::vector<xtx_launch>* out) {
    // tournaments of more than 16 samples run the wide mode whatever is set
    if (params.target_samples > kMaxSamples) return;
    xtx_params p = params;
    p.exchange_interval = moves;
    p.max_epochs = 1;
    p.max_moves = 0;
    p.move_log = 0;
    p.verify = 0;
    // Each sample instance is timed with 256 replicas (independent chains on
    // the instance's ladder, the same move mix), repeated until the copies
    // fill the device twice over in the narrowest shape (16 resident waves
    // per CU, 4 replicas per wave) or take 8 GiB.
    constexpr int kTuneReplicas = 256;
    std::map<int, std::vector<std::unique_ptr<HostInst>>> by_order;
    for (int q = 0; q < count; ++q) {
        const int nmax = order_of(sample[q].n);
        if (!all && ctx->tuned.count(nmax)) continue;
        xtx_instance in = sample[q];
        in.replicas = kTuneReplicas;
        by_order[nmax].push_back(make_inst(ctx, in, p));
    }
    const int64_t target_replicas = (int64_t)ctx->prop.multiProcessorCount * 16 * 4 * 2;
    // the lane kernel's timings per order and shape (proposals, waves)
    std::map<int, std::map<std::pair<int, int>, double>> lane_ms;
    for (auto& kv : by_order) {
        const int nmax = kv.first;
        std::vector<const HostInst*> proto;
        int64_t reps = 0;
        size_t bytes = 0;
        for (size_t q = 0; reps < target_replicas && bytes < ((size_t)8 << 30);
             q = (q + 1) % kv.second.size()) {
            proto.push_back(kv.second[q].get());
            reps += kv.second[q]->R;
            bytes += kv.second[q]->bytes;
        }
        double best_rate = -1, rule_rate = -1;
        Layout best;
        // what the batch runs by rule at full load (waves 0: one per block)
        const Layout rule = lane_layout_ok(nmax) ? lane_layout(nmax) : lookahead_layout(nmax, kLookahead);
        Layout rule_t = rule;
        // The batch and plain forms run the rule's lookahead depth; the tuner
        // chooses the lane shape at that depth only. Depths do not rank here
        // as they do in a run: one epoch from the start matrices commits
        // often, and lookahead makes most of its proposals again after a
        // commit, while later epochs commit rarely. The lane kernel (one
        // lane per proposal, 16 or 32 per step) competes at every order it covers.
        const int rule_K = lookahead_layout(nmax, kLookahead).K;
        for (const Layout& L : candidates(nmax)) {
            if (L.K != rule_K && L.G != 1) continue;
            const double ms = time_layout(ctx, proto, L, p);
            const double rate = ms > 0 ? (double)reps * moves / (ms * 1e-3) : 0.0;
            if (L.G == 1 && ms > 0) lane_ms[nmax][{L.K, L.waves}] = ms;
            if (L.G == rule.G && L.C == rule.C && L.tri == rule.tri && L.K == rule.K &&
                L.xl == rule.xl && std::max(1, L.waves) == std::max(1, rule.waves)) {
                rule_rate = rate;
                rule_t = L;
            }
            xtx_launch r;
            fill_launch(L, false, &r);
            r.moves_per_sec = rate;
            out->push_back(r);
            if (rate > best_rate) {
                best_rate = rate;
                best = L;
            }
        }
        // The tuner times one epoch of each order alone from the start
        // matrices, not the batch's mix of orders in steady state: it keeps
        // the rule unless another shape is more than 10% faster there
        // (measured: 16 proposals per step at orders 40/48 timed faster here
        // and ran the 1024-instance batch 27% slower).
        if (rule_rate > 0 && best_rate < 1.1 * rule_rate) best = rule_t;
        if (best_rate > 0) {
            best.tuner = true;
            ctx->tuned[nmax] = best;
        }
    }
    // The two orders of a lane-kernel pair share a launch only with one shape:
    // where both chose the lane kernel with different shapes, both take the
    // shape of the least summed time.
    for (auto& kv : lane_ms) {
        const int o = kv.first, hi = lane_pair_hi(o), lo = hi - 8;
        if (o != hi || !lane_ms.count(lo)) continue;
        const auto a = ctx->tuned.find(lo), b = ctx->tuned.find(hi);
        if (a == ctx->tuned.end() || b == ctx->tuned.end() || a->second.G != 1 || b->second.G != 1)
            continue;
        if (a->second.K == b->second.K && a->second.waves == b->second.waves) continue;
        double best_sum = -1;
        std::pair<int, int> shape{};
        for (const auto& t : kv.second) {
            const auto u = lane_ms[lo].find(t.first);
            if (u == lane_ms[lo].end()) continue;
            const double sum = t.second + u->second;
            if (best_sum < 0 || sum < best_sum) {
                best_sum = sum;
                shape = t.first;
            }
        }
        if (best_sum < 0) continue;
        // the rule's shape unless the pair's summed time is 10% lower
        const std::pair<int, int> rs{lane_q(hi), 1};
        const auto rh = kv.second.find(rs), rl = lane_ms[lo].find(rs);
        if (rh != kv.second.end() && rl != lane_ms[lo].end() &&
            best_sum > 0.9 * (rh->second + rl->second))
            shape = rs;
        for (int order : {lo, hi}) {
            Layout L = make_layout(order, 1, order, true, shape.second, shape.first);
            L.tuner = true;
            ctx->tuned[order] = L;
        }
    }
    for (auto& r : *out) {
        const auto it = ctx->tuned.find(r.order);
        if (it == ctx->tuned.end() || !it->second.tuner) continue;
        const Layout& t = it->second;
        if (r.lanes == t.G && r.cols == t.C && r.triangular == (t.tri ? 1 : 0) &&
            r.waves_per_block == t.waves && r.lookahead == t.K && r.x_in_lds == (t.xl ? 1 : 0))
            r.tuned = 1;
    }
}

}  // namespace

extern "C" {

int xtx_launch_get(xtx_ctx* ctx, int32_t n, xtx_launch* out) {
    if (!ctx || !out || n < 2 || n > kMaxN) return XTX_E_ARG;
    // the order's configuration if one is set, else what a plain or batch
    // call with automatic lookahead runs
    const auto it = ctx->tuned.find(order_of(n));
    const bool tuned = it != ctx->tuned.end();
    fill_launch(tuned ? it->second : choose_layout(ctx, n, 0, kLookahead), tuned, out);
    return XTX_OK;
}

int xtx_launch_set(xtx_ctx* ctx, const xtx_launch* cfg) {
    if (!ctx || !cfg) return XTX_E_ARG;
    if (cfg->lanes == 0) {
        ctx->tuned.erase(cfg->order);
        return XTX_OK;
    }
    const int K = cfg->lookahead > 1 ? cfg->lookahead : 1;
    const bool xl = cfg->x_in_lds != 0;
    if (const char* why = launch_invalid(cfg->order, cfg->lanes, cfg->cols, cfg->triangular != 0,
                                         cfg->waves_per_block, K, xl))
        return fail(ctx, Err{XTX_E_ARG, why});
    ctx->tuned[cfg->order] = make_layout(cfg->order, cfg->lanes, cfg->cols,
                                         cfg->triangular != 0, cfg->waves_per_block, K, xl);
    return XTX_OK;
}

int xtx_autotune(xtx_ctx* ctx, const xtx_instance* sample, int32_t count,
                 const xtx_params* params, int32_t moves, xtx_launch* results,
                 int32_t max_results, int32_t* n_results) {
    if (!ctx || !sample || count < 1 || !params || moves < 1) return XTX_E_ARG;
    try {
        HIPCHK(hipSetDevice(ctx->device));
        check_params(*params);
        std::vector<xtx_launch> out;
        autotune_orders(ctx, sample, count, *params, moves, true, &out);
        const int w = results ? std::min<int>(max_results, (int)out.size()) : 0;
        for (int q = 0; q < w; ++q) results[q] = out[(size_t)q];
        if (n_results) *n_results = w;
        return XTX_OK;
    } catch (const Err& e) {
        return fail(ctx, e);
    }
}

}  // extern "C"

extern "C" {

int xtx_batch_solve(xtx_ctx* ctx, const xtx_instance* instances, int32_t count,
                    const xtx_params* params, xtx_result* results, xtx_batch_stats* stats) {
    if (!ctx || !instances || count < 1 || !params) return XTX_E_ARG;
    try {
        HIPCHK(hipSetDevice(ctx->device));
        if (params->autotune) {
            std::vector<xtx_launch> tuned;
            autotune_orders(ctx, instances, count, *params,
                            std::min(params->exchange_interval, 512), false, &tuned);
        }
        Server S;
        S.ctx = ctx;
        S.p = *params;
        S.cfg.lds_budget_bytes = INT64_MAX / 4;  // batch: everything resident
        std::vector<std::unique_ptr<HostInst>> hs;
        size_t total = 0;
        for (int q = 0; q < count; ++q) {
            hs.push_back(make_inst(ctx, instances[q], *params));
            hs.back()->tag = (uint64_t)q;
            total += hs.back()->bytes;
        }
        int64_t batch_reps = 0;
        for (auto& h : hs) batch_reps += h->R;
        const bool lane = lane_load(ctx, batch_reps);
        int lane_instances = 0;
        for (auto& h : hs) {
            apply_layout(ctx, *h, params->lookahead, kLookahead, lane);
            lane_instances += h->lay.G == 1;
        }
        S.arena.reserve(total + 4096);
        S.t0 = clk::now();
        std::vector<Staging> keep;
        keep.reserve(hs.size());
        for (auto& h : hs) {
            h->submit_s = 0;
            if (!S.try_admit(h, &keep)) throw Err{XTX_E_NOMEM, "batch does not fit the device"};
        }
        HIPCHK(hipStreamSynchronize(ctx->stream));
        keep.clear();
        S.rebuild();
        const double upload_s = S.now();
        // Epochs per poll: about 100 ms of work between polls (ten epochs at
        // the full batch), at most 16. Between polls every launch bucket runs
        // its epochs as a chain of its own (launch_epochs_chained), so the
        // buckets meet only at the poll: measured on the 1024-instance batch,
        // 8 ms per poll (one epoch) 314 matrices/s at p50/p99 time-to-diagonal
        // 0.073/0.57 s, 40 ms 348 at 0.07/0.50 s, 100 ms 390 at 0.10/0.47 s. The
        // next epochs are launched before a poll's retirements (finished
        // instances' waves exit at once), so the device stays busy while the
        // host copies out and verifies results; the launch tables are rebuilt,
        // with nothing in flight, once a quarter of the residents have gone
        // (each rebuild costs a graph capture).
        HIPCHK(hipStreamCreate(&S.fetch_stream));
        // Every epoch still to run goes into one captured graph (at most 64;
        // 16 at a time without an epoch cap), each launch bucket a chain of
        // its own (launch_epochs_chained), so the buckets never wait for one
        // another. While it runs, the host collects every instance the
        // device flags as finished (collect_done) and verifies it on host
        // threads, so a result's latency does not depend on the graph's
        // length. Measured on the 1024-instance batch with the buckets'
        // chains: a poll after every epoch (8 ms of work) 314 matrices/s at
        // p50/p99 time-to-diagonal 0.073/0.57 s, every ten epochs 390 at
        // 0.10/0.47 s. At a graph's end the launch tables are rebuilt once a
        // quarter of the residents have gone.
        Deadline dl(ctx->stop_host, params->max_seconds);
        // One persistent launch (k_lane_persist) when every instance runs the
        // lane kernel by its rule: waves take ready units of any instance and
        // each instance's epoch end runs in the wave that finishes its epoch,
        // so instances never wait for one another (64 epochs per launch
        // without an epoch cap).
        const bool persist = params->persistent != 0 && S.persistent_ok();
        auto graph_k = [&]() {
            if (params->max_epochs > 0) return std::clamp(params->max_epochs - S.epochs, 1, 64);
            return persist ? 64 : 16;
        };
        auto launch = [&](int kk) {
            if (persist)
                S.launch_persistent(kk);
            else
                S.launch_epochs(kk);
        };
        int k = graph_k();
        launch(k);
        for (;;) {
            while (hipStreamQuery(ctx->stream) == hipErrorNotReady) {
                S.collect_done();
                std::this_thread::sleep_for(std::chrono::microseconds(500));
            }
            S.finish_epochs();
            S.collect_done();
            bool running = false;
            for (size_t s = 0; s < S.slots.size(); ++s)
                if (S.slots[s] && S.h_insts[s].status == ST_RUNNING) running = true;
            const bool stop = *(volatile int*)ctx->stop_host != 0;
            const bool capped = params->max_epochs > 0 && S.epochs >= params->max_epochs;
            S.retire(false);
            if (stop) break;
            if (!running || capped || S.resident() == 0) break;
            if (4 * S.retired_since_rebuild > S.resident()) S.rebuild();
            k = graph_k();
            launch(k);
        }
        S.finish_epochs();
        S.retire(true);
        int solved = 0, verified = 0;
        for (const auto& r : S.done) {
            if (results) results[r.tag] = r;
            solved += r.solved;
            verified += r.solved && r.verify_status == 0;
        }
        if (stats) {
            xtx_batch_stats bs{};
            bs.moves = S.moves_retired;
            bs.accepts = S.accepts_retired;
            bs.instances = count;
            bs.solved = solved;
            bs.verified = verified;
            bs.epochs = S.epochs;
            bs.seconds = S.now();
            bs.kernel_ms = S.kernel_ms;
            bs.buckets = (int)S.buckets.size();
            bs.graph_epochs = k;
            bs.upload_seconds = upload_s;
            bs.lane_instances = lane_instances;
            bs.persistent = persist ? 1 : 0;
            *stats = bs;
        }
        return XTX_OK;
    } catch (const Err& e) {
        return fail(ctx, e);
    }
}

int xtx_serve_open(xtx_ctx* ctx, const xtx_params* params, const xtx_serve_config* cfg) {
    if (!ctx || !params) return XTX_E_ARG;
    try {
        if (ctx->server) throw Err{XTX_E_STATE, "serve already open"};
        check_params(*params);
        HIPCHK(hipSetDevice(ctx->device));
        auto S = std::make_unique<Server>();
        S->ctx = ctx;
        S->p = *params;
        S->auto_look = true;
        if (cfg) S->cfg = *cfg;
        const int64_t hbm = S->cfg.hbm_budget_bytes > 0 ? S->cfg.hbm_budget_bytes
                                                         : (int64_t)8 << 30;
        S->arena.reserve((size_t)hbm);
        S->t0 = clk::now();
        ctx->server = std::move(S);
        return XTX_OK;
    } catch (const Err& e) {
        return fail(ctx, e);
    }
}

int xtx_serve_submit(xtx_ctx* ctx, const xtx_instance* inst, uint64_t tag) {
    if (!ctx || !inst) return XTX_E_ARG;
    if (!ctx->server) {
        ctx->err = "serve not open";
        return XTX_E_STATE;
    }
    try {
        Server& S = *ctx->server;
        const int qcap = S.cfg.queue_capacity > 0 ? S.cfg.queue_capacity : 65536;
        if ((int)S.queue.size() >= qcap) throw Err{XTX_E_FULL, "serve queue full"};
        auto h = make_inst(ctx, *inst, S.p);
        if ((int64_t)h->bytes > (int64_t)S.arena.size())
            throw Err{XTX_E_NOMEM, "request larger than the HBM budget"};
        h->tag = tag;
        h->submit_s = S.now();
        S.queue.push_back(std::move(h));
        return XTX_OK;
    } catch (const Err& e) {
        return fail(ctx, e);
    }
}

int xtx_serve_step(xtx_ctx* ctx, int32_t* resident, int32_t* queued) {
    if (!ctx) return XTX_E_ARG;
    if (!ctx->server) {
        ctx->err = "serve not open";
        return XTX_E_STATE;
    }
    try {
        HIPCHK(hipSetDevice(ctx->device));
        Server& S = *ctx->server;
        S.admit_queued();
        S.rebuild();
        if (S.resident() > 0) {
            S.run_epochs(1);
            S.retire(false);
        }
        S.admit_queued();
        if (resident) *resident = S.resident();
        if (queued) *queued = (int32_t)S.queue.size();
        return XTX_OK;
    } catch (const Err& e) {
        return fail(ctx, e);
    }
}

int xtx_serve_poll(xtx_ctx* ctx, xtx_result* out, int32_t max, int32_t* count) {
    if (!ctx || !count || (max > 0 && !out)) return XTX_E_ARG;
    if (!ctx->server) {
        ctx->err = "serve not open";
        return XTX_E_STATE;
    }
    Server& S = *ctx->server;
    const int k = std::min<int>(max, (int)S.done.size());
    for (int t = 0; t < k; ++t) out[t] = S.done[(size_t)t];
    S.done.erase(S.done.begin(), S.done.begin() + k);
    *count = k;
    return XTX_OK;
}

int xtx_serve_retire(xtx_ctx* ctx, uint64_t tag) {
    if (!ctx) return XTX_E_ARG;
    if (!ctx->server) {
        ctx->err = "serve not open";
        return XTX_E_STATE;
    }
    try {
        Server& S = *ctx->server;
        for (auto it = S.queue.begin(); it != S.queue.end(); ++it) {
            if ((*it)->tag == tag) {
                xtx_result res{};
                res.tag = tag;
                res.verify_status = -1;
                res.submit_s = (*it)->submit_s;
                res.retire_s = S.now();
                S.done.push_back(res);
                S.queue.erase(it);
                return XTX_OK;
            }
        }
        const size_t before = S.done.size();
        // Requests admitted since the last epoch (a step admits after its
        // retirements) have records only on the host: rebuild writes every
        // slot's record to the device instance array (growing it as needed),
        // then the array is read back so each slot's record is its own.
        HIPCHK(hipSetDevice(ctx->device));
        S.rebuild();
        const int ns = (int)S.slots.size();
        S.h_insts.resize((size_t)ns);
        if (ns)
            HIPCHK(hipMemcpy(S.h_insts.data(), S.d_insts, (size_t)ns * sizeof(Inst),
                             hipMemcpyDeviceToHost));
        S.retire(false, tag, true);
        if (S.done.size() == before) throw Err{XTX_E_ARG, "unknown tag"};
        return XTX_OK;
    } catch (const Err& e) {
        return fail(ctx, e);
    }
}

int xtx_serve_close(xtx_ctx* ctx) {
    if (!ctx) return XTX_E_ARG;
    ctx->server.reset();
    return XTX_OK;
}

}  // extern "C"
