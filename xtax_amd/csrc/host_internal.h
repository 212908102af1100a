// host_internal.h -- host-side pieces shared by the engine's translation
// units (engine.cpp: plain, batched and served forms of the L1 objective;
// dual.cpp: the primal/dual objective). Not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "engine_types.h"
#include "xtx.h"

namespace xtx {
namespace host {

using clk = std::chrono::steady_clock;

struct Err {
    int code;
    std::string msg;
};

#define HIPCHK(expr)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            throw Err{XTX_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)}; \
    } while (0)

inline int64_t iabs(int64_t v) { return v < 0 ? -v : v; }

// ---------------------------------------------------------------- layout
// Launch configuration of an order bucket: padded order nmax (multiple of 8),
// lane-group width G, columns per lane C (G C >= nmax), A as its upper
// triangle or square in LDS, waves per workgroup (0: the params' choice) and
// the LDS bytes of one replica. Every configuration computes the same
// trajectory (the random stream and the arithmetic do not depend on it); it is
// a performance choice, made by a rule or by xtx_autotune.
struct Layout {
    int G = 0, C = 0, nmax = 0, rep_bytes = 0;
    bool tri = true;
    int waves = 0;
    int K = 1;  // lookahead: lane groups per replica (G K == 64 when K > 1)
    bool xl = false;  // X^T in LDS for the epoch (else in HBM)
    // chosen by the autotuner, which times the device full of replicas: a
    // lane-kernel choice then applies only to calls that fill the device
    bool tuner = false;
};

inline int order_of(int n) { return std::max(8, (n + 7) / 8 * 8); }

}  // namespace host
// LDS of one replica of the lane kernel (anneal_lane.hip) at order nmax
int lane_rep_bytes(int nmax);
bool lane_variant_exists(int nmax, int Q);
// the higher order of the launch pair order nmax runs in (0: no lane kernel)
int lane_pair_hi(int nmax);
namespace host {

// The lane kernel (G = 1: one lane per proposal, C = nmax columns per lane,
// K = Q proposals per replica and step): A's upper triangle at the order, the
// row masses and the active list.
// Otherwise: A, then per lane group 5 staged Philox blocks and a median
// histogram, and the active list; a replica over several waves adds two
// buffers of K 64-byte proposals and a poll word (16-byte aligned). X^T stays
// in HBM.
inline int rep_bytes_of(int nmax, int G, bool tri, int K = 1, bool xl = false) {
    if (G == 1) return lane_rep_bytes(nmax);
    const int64_t a = tri ? (int64_t)nmax * (nmax + 1) / 2 : (int64_t)nmax * (nmax + 1);
    int64_t bytes = a * 8 + K * (5 * 16 + 4 * med_slots(G)) + nmax;
    if (G * K > 64) bytes = (bytes + 15) / 16 * 16 + 2 * K * 64 + 16;
    bytes = (bytes + 15) / 16 * 16;
    if (xl) bytes += (int64_t)nmax * nmax * 8;  // X^T last
    return (int)bytes;
}

inline Layout make_layout(int nmax, int G, int C, bool tri, int waves, int K = 1,
                          bool xl = false) {
    Layout L;
    L.nmax = nmax;
    L.G = G;
    L.C = C;
    L.tri = tri;
    L.waves = waves;
    L.K = K;
    L.xl = xl;
    L.rep_bytes = rep_bytes_of(nmax, G, tri, K, xl);
    return L;
}

// The lane kernel's wide mode: A exact in HBM, 32 proposals per replica and
// step, each evaluated by the replica's 32 lanes (2, 4, 6 or 8 columns per
// lane to order 64, 128, 192, 256), tournaments of any size. Orders above 128
// (A does not fit LDS) and calls whose tournaments take more than 16 samples
// run it. The layout's order is the kernel's column bound.
inline Layout wide_layout(int nmax) {
    const int nw = nmax <= 64 ? 64 : nmax <= 128 ? 128 : nmax <= 192 ? 192 : 256;
    return make_layout(nw, 1, nw, true, 0, 32);
}

// the rule used when a bucket has not been tuned
inline Layout layout_for(int n) {
    const int nmax = order_of(n);
    if (nmax > kMaxGroupN) return wide_layout(nmax);
    if (nmax <= 16) return make_layout(nmax, 16, 1, true, 0);
    if (nmax <= 32) return make_layout(nmax, 32, 1, true, 0);
    if (nmax <= 64) return make_layout(nmax, 64, 1, true, 0);
    return make_layout(nmax, 64, 2, true, 0);
}

// The lookahead layout of order nmax: K groups of G lanes per replica (one
// wave for K G = 64, a workgroup of K G / 64 waves above), K the largest of
// 32, 16, 8, 4 and 2 that is at most `want` and has a kernel for the order (C <= 2;
// above order 64 four whole-wave groups); K = 1 (the order's default layout)
// otherwise.
inline Layout lookahead_layout(int nmax, int want) {
    if (want >= 32 && nmax <= 16) return make_layout(nmax, 16, 1, true, 0, 32);
    if (want >= 32 && nmax <= 32) return make_layout(nmax, 16, 2, true, 0, 32);
    if (want >= 16 && nmax <= 16) return make_layout(nmax, 16, 1, true, 0, 16);
    if (want >= 16 && nmax <= 32) return make_layout(nmax, 16, 2, true, 0, 16);
    if (want >= 8 && nmax <= 64) return make_layout(nmax, 32, 2, true, 0, 8);
    if (want >= 4 && nmax <= 16) return make_layout(nmax, 16, 1, true, 0, 4);
    if (want >= 4 && nmax <= 32) return make_layout(nmax, 16, 2, true, 0, 4);
    if (want >= 2 && nmax <= 32) return make_layout(nmax, 32, 1, true, 0, 2);
    if (want >= 2 && nmax <= 64) return make_layout(nmax, 32, 2, true, 0, 2);
    if (want >= 4 && nmax <= 128) return make_layout(nmax, 64, 2, true, 0, 4);
    return layout_for(nmax);
}

// The lane kernel's layout at order nmax (orders to 48): 16 proposals of a
// replica per step (four replicas per wave) up to order 32, 32 proposals (two
// replicas per wave) at orders 40 and 48. Measured on the 1024-instance batch
// (orders paired in three launches, 16-bit LDS words): kernel time 615 ms,
// against 685 ms at 16 proposals for every order and 649 ms at 32: at the
// high orders two replicas per wave let twice as many waves be resident,
// which outweighs the proposals made again after a commit; at the low orders
// residency is not the bound.
constexpr int kLaneMaxOrder = 48;
inline int lane_q(int nmax) { return nmax >= 40 ? 32 : 16; }
inline Layout lane_layout(int nmax) { return make_layout(nmax, 1, nmax, true, 0, lane_q(nmax)); }
inline bool lane_layout_ok(int nmax) { return nmax <= kLaneMaxOrder; }


// The lookahead rule. Plain and batch calls: 16 (measured: 256 replicas of
// order 24 run 5.4x the moves/s of one group per replica, the full batch of
// 1024 instances x 32 replicas 4.5% faster). Serve admissions, by the
// replicas resident with the request on a device of `cus` compute units: 16
// up to 32 per compute unit, else 1 (at 200 requests/s, about 4000 resident,
// lookahead halves the latency; overloaded at 400/s a lookahead replica's
// larger LDS share admits fewer requests and one group per replica serves
// more).
constexpr int kLookahead = 16;
inline int lookahead_rule(int cus, int64_t replicas) {
    return replicas <= 32 * (int64_t)cus ? kLookahead : 1;
}

using BucketKey = std::tuple<int, int, int, bool, int, int, bool>;
// The launch a resident instance joins. A kernel runs any order up to the
// launch's padded order (A's rows in LDS are laid out by the instance's own
// n; only the offsets of the replica's other LDS arrays follow the launch's
// order). Lookahead layouts give each replica its own workgroup, so orders
// that share a kernel share one launch sized for the largest: fewer launches
// than streams, all of an epoch's kernels concurrent. One-group layouts keep
// one launch per order (a larger order's LDS would cost them occupancy).
// A server's admission budget charges every replica of a launch the LDS share
// of the launch's largest order.
inline BucketKey launch_key(const Layout& L) {
    // the lane kernel runs orders in pairs (8 and 16, 24 and 32, 40 and 48):
    // one launch per pair
    if (L.G == 1) return {-lane_pair_hi(L.nmax), 1, 0, L.tri, L.waves, L.K, false};
    return {(L.K > 1 && L.G > 1) ? 0 : L.nmax, L.G, L.C, L.tri, L.waves, L.K, L.xl};
}

constexpr int kLdsPerCu = 160 * 1024;
// Largest LDS one block may use on the open device (set by xtx_open).
extern int g_lds_block;

// ---------------------------------------------------------------- arena
// First-fit free list over one device allocation (256-byte granules).
class Arena {
  public:
    ~Arena() { release(); }
    void reserve(size_t bytes) {
        release();
        bytes = (bytes + 255) / 256 * 256;
        HIPCHK(hipMalloc(&base_, bytes));
        size_ = bytes;
        free_.clear();
        free_[0] = bytes;
    }
    void release() {
        if (base_) (void)hipFree(base_);
        base_ = nullptr;
        size_ = 0;
        free_.clear();
    }
    char* alloc(size_t bytes) {
        bytes = (bytes + 255) / 256 * 256;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                const size_t off = it->first, len = it->second;
                free_.erase(it);
                if (len > bytes) free_[off + bytes] = len - bytes;
                return base_ + off;
            }
        }
        return nullptr;
    }
    void free(char* p, size_t bytes) {
        bytes = (bytes + 255) / 256 * 256;
        size_t off = (size_t)(p - base_);
        auto next = free_.lower_bound(off);
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                off = prev->first;
                bytes += prev->second;
                free_.erase(prev);
            }
        }
        next = free_.lower_bound(off);
        if (next != free_.end() && off + bytes == next->first) {
            bytes += next->second;
            free_.erase(next);
        }
        free_[off] = bytes;
    }
    size_t size() const { return size_; }
    bool empty() const { return base_ == nullptr; }

  private:
    char* base_ = nullptr;
    size_t size_ = 0;
    std::map<size_t, size_t> free_;
};

// ---------------------------------------------------------------- instance
// Host-side description of one instance: validated inputs, start state,
// ladder, and where its device buffers live.
struct HostInst {
    int n = 0, N = 0, R = 0;
    uint64_t seed = 0;
    std::vector<int64_t> A0, Astart, Xt0;  // original input, X0^T A0 X0, X0^T
    std::vector<int32_t> act;
    int64_t score0 = 0, offnz0 = 0;
    double t_init = 0;
    std::vector<double> temps;
    InstParams P{};
    bool log_on = false;
    int64_t log_cap = 0;
    Layout lay{};
    size_t bytes = 0;
    char* dmem = nullptr;
    Inst rec{};
    int64_t* A_out = nullptr;
    int64_t* X_out = nullptr;
    uint64_t tag = 0;
    double submit_s = 0, admit_s = 0;
};

int64_t score_of(const int64_t* A, int n);
int64_t offnz_of(const int64_t* A, int n);
void check_params(const xtx_params& p);
// Validates the inputs and builds the start state A_start = X0^T A0 X0
// exactly, the active list, the L1 score and temperature setup, and the
// instance parameters.
void prepare(HostInst& h, const int64_t* A0, const int64_t* X0, int n, int R, uint64_t seed,
             const int32_t* active, int n_active, const xtx_params& p);
// geometric ladder from min(t_min, top) to top (congruence_anneal.hpp:454-468)
std::vector<double> ladder(double top, double t_min, int R);
// per-replica greedy/target fractions (congruence_anneal.hpp:200-207)
void fractions(const xtx_params& p, int r, int R, double* greedy, double* target);
size_t align256(size_t b);

// The stop word the kernels poll (nonzero = stop at the next poll point) holds
// two bits: a caller's stop request (xtx_request_stop, the reference's g_stop,
// stop_signal.hpp:40-64) and a solve's own max_seconds deadline. Each owner
// sets and clears only its bit, so a deadline that fires while a solve winds
// down cannot stop the next solve, and a deadline never clears a request.
constexpr int kStopRequest = 1;
constexpr int kStopDeadline = 2;

// Watchdog: raises the deadline bit at a wall-clock deadline (the reference's
// monitor-thread max_seconds check, congruence_anneal.hpp:512-519) and clears
// it again when the solve ends (the destructor joins the watchdog first).
class Deadline {
  public:
    Deadline(int* flag, double seconds) : flag_(flag) {
        if (!(seconds > 0.0)) return;
        th_ = std::thread([this, seconds] {
            std::unique_lock<std::mutex> lk(m_);
            if (!cv_.wait_for(lk, std::chrono::duration<double>(seconds), [this] { return done_; })) {
                fired_ = true;
                __atomic_fetch_or(flag_, kStopDeadline, __ATOMIC_SEQ_CST);
            }
        });
    }
    ~Deadline() {
        if (th_.joinable()) {
            {
                std::lock_guard<std::mutex> lk(m_);
                done_ = true;
            }
            cv_.notify_all();
            th_.join();
        }
        if (fired_) __atomic_fetch_and(flag_, ~kStopDeadline, __ATOMIC_SEQ_CST);
    }
    bool fired() const { return fired_; }

  private:
    int* flag_;
    std::thread th_;
    std::mutex m_;
    std::condition_variable cv_;
    bool done_ = false;
    std::atomic<bool> fired_{false};
};

// Why a solve loop saw a nonzero stop word: 2 = a caller's request (it wins
// when both bits are set), 1 = the solve's deadline.
inline int stop_reason_of(const int* flag) {
    return (__atomic_load_n(flag, __ATOMIC_SEQ_CST) & kStopRequest) ? 2 : 1;
}

// Host copy of the last primal/dual solve (dual.cpp).
struct DualLast {
    bool valid = false;
    double c = 0.0;
    std::vector<double> Q0;
    std::vector<RepState> rs;
};

}  // namespace host
}  // namespace xtx

struct Server;


constexpr int kSideStreams = 3;

struct xtx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // buckets of one epoch run concurrently: forked from `stream` onto these
    // (default priority) and joined back before the epoch end
    hipStream_t side[kSideStreams] = {};
    hipEvent_t fork_ev = nullptr;
    hipEvent_t join_ev[kSideStreams] = {};
    int* stop_host = nullptr;  // pinned, mapped
    int* stop_dev = nullptr;
    std::string err;
    hipDeviceProp_t prop{};
    // last plain solve
    xtx::host::Arena plain_arena;
    xtx::host::HostInst plain;
    xtx::Inst* d_plain = nullptr;
    std::vector<xtx::RepState> plain_rs;
    std::vector<xtx::LogEnt> plain_log;
    std::vector<int64_t> plain_log_count;
    bool plain_log_valid = false;
    std::unique_ptr<Server> server;
    // set by a running plain solve (L1 or xdual) for xtx_best_so_far: writes
    // the instance best as of the last epoch boundary (A or P, Q, X)
    std::function<void(int64_t*, double*, int64_t*)> best_so_far;
    // launch configuration per padded order, set by xtx_launch_set or
    // xtx_autotune; other orders use layout_for's rule
    std::map<int, xtx::host::Layout> tuned;
    // last solve of the primal/dual objective (its instance is `plain`)
    xtx::host::DualLast dual;
    // the device's compute-unit pairs as XCD (die) id << 8 | CU-pair key,
    // probed once for the batch's persistent launch
    std::vector<int> xcc_ids;
    // the persistent launch's host-mapped outcome buffer, kept for the next
    // batch (pinning tens of MB per call would cost milliseconds each time)
    unsigned char* zc_host = nullptr;
    unsigned char* zc_dev = nullptr;
    size_t zc_cap = 0;
};

namespace xtx {
namespace host {

inline int fail(xtx_ctx* ctx, const Err& e) {
    if (ctx) ctx->err = e.msg;
    return e.code;
}

}  // namespace host
}  // namespace xtx
