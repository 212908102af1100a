// lane_eval.h -- one Add proposal evaluated by a single lane, serially over the
// columns: the exact best integer shear (best_shear_l1, main.cpp:181-242) and
// the evaluation of Add(i, j, s) (try_add without the transform row,
// main.cpp:46-91), for the kernel that gives every lane its own proposal
// (anneal_lane.hip). Plain integer code, also compiled for the host, where a
// check against the oracle runs on the CPU (tests/lane_eval_check.cpp).
//
// Rows i (pivot) and j (target) of A enter as, for every column k != j,
//   p_k = |a_ik|,  q_k = a_jk if a_ik >= 0 else -a_jk,
// and 0 for k == j and beyond the order (those columns then add nothing to any
// sum below). Then |a_jk + s a_ik| = |q_k + s p_k|, and with the breakpoint
// b_k = -a_jk / a_ik of a column with a_ik != 0:
//   ceil(b_k) <= t  <=>  p_k t + q_k >= 0,     floor(b_k) < t  <=>  p_k t + q_k > 0,
// so the weighted lower median's ceiling and floor come from integer counts,
// with no division: ceil(m) = min{t : 2 W(ceil b <= t) >= W} by binary search
// over t in [-64, 65] (65 standing for "beyond the cap"), and floor(m) =
// ceil(m) - 1 iff 2 W(b < ceil m) >= W. With |a| <= 2^48 these are the
// reference's floor and ceil of its double median: every breakpoint is at
// least 1/|a_ik| from an integer it is not equal to, and the double rounding
// of the reference moves it by less.
//
// The small path holds p and q of a row pair whose entries are all below 2^16
// in int32 registers (every new entry is below 2^23 and every row sum below
// 2^30: no entry can reach the 2^48 bound); the int64 path reads the rows
// again for every pass and applies the bound.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define XTX_HD __host__ __device__ __forceinline__
#else
#define XTX_HD inline
#endif

// Keeps the instruction scheduler from overlapping two phases of a lane's
// evaluation (on the device; nothing on the host): overlapping the median
// search with the candidates' sums holds both phases' values at once and
// costs more registers than it gains in latency.
#if defined(__HIP_DEVICE_COMPILE__)
#define XTX_PHASE() __builtin_amdgcn_sched_barrier(0)
#else
#define XTX_PHASE() ((void)0)
#endif

namespace xtx {
namespace lane {
// a b for |a|, |b| < 2^23: one full-rate 24-bit multiply on the device (a
// 32-bit multiply there is a quarter-rate 64-bit multiply-add)
XTX_HD int32_t mul24(int32_t a, int32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __mul24(a, b);
#else
    return a * b;
#endif
}
}  // namespace lane
}  // namespace xtx

namespace xtx {
namespace lane {

constexpr int64_t kLim = (int64_t)1 << 48;  // mat_io.hpp:35 (strict >)
constexpr int kCap = 64;                    // mat_io.hpp:38
constexpr int32_t kSmallEntry = 1 << 16;
constexpr int64_t kPenalty = (int64_t)1 << 62;  // score_delta_l1's overflow value

XTX_HD int64_t abs64(int64_t v) { return v < 0 ? -v : v; }
XTX_HD int32_t abs32(int32_t v) { return v < 0 ? -v : v; }
XTX_HD bool over(int64_t v) { return v > kLim || v < -kLim; }
XTX_HD int64_t clamp_cap(int64_t v) { return v > kCap ? kCap : (v < -kCap ? -kCap : v); }

// Nearest integer a/b, ties away from zero (mat_io.hpp:41-48).
XTX_HD int64_t rounded_div(int64_t a, int64_t b) {
    int64_t q = a / b;
    const int64_t r = a - q * b;
    if (2 * abs64(r) >= abs64(b)) q += ((a < 0) == (b < 0)) ? 1 : -1;
    return q;
}
XTX_HD int32_t rounded_div32(int32_t a, int32_t b) {
    int32_t q = a / b;
    const int32_t r = a - q * b;
    if (2 * abs32(r) >= abs32(b)) q += ((a < 0) == (b < 0)) ? 1 : -1;
    return q;
}

// Candidate shears in the reference's order (main.cpp:214-231): floor and ceil
// of the median, then q, q+1, q-1 with q = -round(a_ji/a_ii) when a_ii != 0,
// each clamped to +/-64; a candidate equal to an earlier one is skipped.
struct Cands {
    int64_t v[5];
    bool ok[5];
};
XTX_HD void make_cands(int64_t cf, int64_t cc, bool has_q, int64_t q, Cands& cd) {
    cd.v[0] = clamp_cap(cf);
    cd.v[1] = clamp_cap(cc);
    cd.v[2] = clamp_cap(q);
    cd.v[3] = clamp_cap(q + 1);
    cd.v[4] = clamp_cap(q - 1);
#pragma unroll
    for (int t = 0; t < 5; ++t) {
        bool v = t < 2 || has_q;
#pragma unroll
        for (int u = 0; u < t; ++u) v &= cd.v[u] != cd.v[t];
        cd.ok[t] = v;
    }
}

// Floor and ceiling of the weighted lower median (clamped to the cap) from
// le(t) = W(ceil b <= t) and lt(t) = W(b < t); both 0 when W = 0.
template <class LE, class LT>
XTX_HD void median_bounds(int64_t tot, LE le, LT lt, int64_t& cf, int64_t& cc) {
    cf = cc = 0;
    if (tot == 0) return;
    int lo = -kCap, hi = kCap + 1;  // P(hi) holds by convention
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;  // floor
        if (2 * le(mid) >= tot)
            hi = mid;
        else
            lo = mid + 1;
    }
    if (lo <= -kCap) {
        cf = cc = -kCap;
    } else if (lo > kCap) {
        cf = cc = kCap;
    } else {
        cc = lo;
        cf = (2 * lt(lo) >= tot) ? lo - 1 : lo;
    }
}

// ---- small path: p, q in registers -------------------------------------

// The best shear of the pair (0 when no candidate lowers the score) and its
// off-diagonal change sum_k |q + s p| - |q| in *off_out. sumq = sum_k |q_k|.
template <int NMAX>
XTX_HD int64_t shear_small(const int32_t (&p)[NMAX], const int32_t (&q)[NMAX], int32_t sumq,
                           int32_t a_ii, int32_t a_ji, int32_t a_jj, int32_t* off_out) {
    int32_t tot = 0;
#pragma unroll
    for (int k = 0; k < NMAX; ++k) tot += p[k];
    auto le = [&](int t) -> int64_t {
        int32_t acc = 0;
#pragma unroll
        for (int k = 0; k < NMAX; ++k) acc += (mul24(p[k], t) + q[k] >= 0) ? p[k] : 0;
        return acc;
    };
    auto lt = [&](int t) -> int64_t {
        int32_t acc = 0;
#pragma unroll
        for (int k = 0; k < NMAX; ++k) acc += (mul24(p[k], t) + q[k] > 0) ? p[k] : 0;
        return acc;
    };
    int64_t cf, cc;
    median_bounds(tot, le, lt, cf, cc);
    XTX_PHASE();
    const bool has_q = a_ii != 0;
    const int64_t qq = has_q ? -(int64_t)rounded_div32(a_ji, a_ii) : 0;
    Cands cd;
    make_cands(cf, cc, has_q, qq, cd);
    int64_t best_s = 0, best_d = 0;
    int32_t best_off = 0;
#pragma unroll
    for (int t = 0; t < 5; ++t) {
        XTX_PHASE();
        const int64_t s = cd.v[t];
        if (!cd.ok[t] || s == 0) continue;
        const int32_t s32 = (int32_t)s;
        int32_t sum = 0;
#pragma unroll
        for (int k = 0; k < NMAX; ++k) sum += abs32(q[k] + mul24(s32, p[k]));
        const int32_t off = sum - sumq;
        const int32_t jj = a_jj + mul24(s32, 2 * a_ji + mul24(s32, a_ii));
        const int64_t d = 4 * (int64_t)off + (int64_t)(abs32(jj) - abs32(a_jj));
        if (d < best_d) {
            best_d = d;
            best_s = s;
            best_off = off;
        }
    }
    XTX_PHASE();
    *off_out = best_off;
    return best_s;
}

// Evaluation of Add(i, j, s) on the small path (nothing can leave the bound).
// With `know`, the off-diagonal change of s is already known.
template <int NMAX>
XTX_HD void eval_small(const int32_t (&p)[NMAX], const int32_t (&q)[NMAX], int32_t sumq,
                       int32_t s32, int32_t a_ii, int32_t a_ji, int32_t a_jj, bool know,
                       int32_t known_off, int64_t* delta, int32_t* dnz, int32_t* off_out) {
    int32_t sum = 0, nz = 0;
#pragma unroll
    for (int k = 0; k < NMAX; ++k) {
        const int32_t nv = q[k] + mul24(s32, p[k]);
        if (!know) sum += abs32(nv);
        nz += (q[k] == 0) ? (nv != 0) : -(int32_t)(nv == 0);
    }
    const int32_t off = know ? known_off : sum - sumq;
    const int32_t jj = a_jj + mul24(s32, 2 * a_ji + mul24(s32, a_ii));
    *delta = 4 * (int64_t)off + (int64_t)(abs32(jj) - abs32(a_jj));
    *dnz = nz;
    *off_out = off;
}

// ---- int64 path: rows read through an accessor -----------------------------
// row(k, ai, aj) gives a_ik and a_jk for a column k < n.

template <class Row>
XTX_HD int64_t shear_big(Row row, int n, int j, int64_t a_ii, int64_t a_ji, int64_t a_jj,
                         int64_t* off_out) {
    int64_t tot = 0;
    for (int k = 0; k < n; ++k) {
        if (k == j) continue;
        int64_t ai, aj;
        row(k, ai, aj);
        tot += abs64(ai);
    }
    auto le = [&](int t) -> int64_t {
        int64_t acc = 0;
        for (int k = 0; k < n; ++k) {
            if (k == j) continue;
            int64_t ai, aj;
            row(k, ai, aj);
            const int64_t v = aj + (int64_t)t * ai;
            acc += (ai > 0 ? v >= 0 : (ai < 0 && v <= 0)) ? abs64(ai) : 0;
        }
        return acc;
    };
    auto lt = [&](int t) -> int64_t {
        int64_t acc = 0;
        for (int k = 0; k < n; ++k) {
            if (k == j) continue;
            int64_t ai, aj;
            row(k, ai, aj);
            const int64_t v = aj + (int64_t)t * ai;
            acc += (ai > 0 ? v > 0 : (ai < 0 && v < 0)) ? abs64(ai) : 0;
        }
        return acc;
    };
    int64_t cf, cc;
    median_bounds(tot, le, lt, cf, cc);
    const bool has_q = a_ii != 0;
    const int64_t qq = has_q ? -rounded_div(a_ji, a_ii) : 0;
    Cands cd;
    make_cands(cf, cc, has_q, qq, cd);
    int64_t best_s = 0, best_d = 0, best_off = 0;
#pragma unroll
    for (int t = 0; t < 5; ++t) {
        const int64_t s = cd.v[t];
        if (!cd.ok[t] || s == 0) continue;
        int64_t off = 0;
        bool bad = false;
        for (int k = 0; k < n; ++k) {
            if (k == j) continue;
            int64_t ai, aj;
            row(k, ai, aj);
            const int64_t nv = aj + s * ai;
            bad |= over(nv);
            off += abs64(nv) - abs64(aj);
        }
        const int64_t jj = a_jj + 2 * s * a_ji + s * s * a_ii;
        const int64_t d = (bad || over(jj)) ? kPenalty : 4 * off + (abs64(jj) - abs64(a_jj));
        if (d < best_d) {
            best_d = d;
            best_s = s;
            best_off = off;
        }
    }
    *off_out = best_off;
    return best_s;
}

// try_add's outcome for Add(i, j, s) on the int64 path: false when some new
// entry (or the new diagonal) leaves the bound.
template <class Row>
XTX_HD bool eval_big(Row row, int n, int j, int64_t s, int64_t a_ii, int64_t a_ji, int64_t a_jj,
                     int64_t* delta, int32_t* dnz, int64_t* off_out) {
    int64_t off = 0;
    int32_t nz = 0;
    bool bad = false;
    for (int k = 0; k < n; ++k) {
        if (k == j) continue;
        int64_t ai, aj;
        row(k, ai, aj);
        const int64_t nv = aj + s * ai;
        bad |= over(nv);
        off += abs64(nv) - abs64(aj);
        nz += (aj == 0) ? (nv != 0) : -(int32_t)(nv == 0);
    }
    const int64_t jj = a_jj + 2 * s * a_ji + s * s * a_ii;
    bad |= over(jj);
    *delta = 4 * off + (abs64(jj) - abs64(a_jj));
    *dnz = nz;
    *off_out = off;
    return !bad;
}

}  // namespace lane
}  // namespace xtx
