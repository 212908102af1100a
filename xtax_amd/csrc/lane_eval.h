// lane_eval.h -- one Add proposal evaluated by a single lane, serially over the
// columns: the exact best integer shear (best_shear_l1, main.cpp:181-242) and
// the score change of Add(i, j, s) (try_add without the transform row,
// main.cpp:46-91), for the kernel that gives every lane its own proposal
// (anneal_lane.hip), on rows whose entries are all below 2^16 in magnitude
// (the kernel evaluates other rows as a lane group, group_eval.h). Plain
// integer code, also compiled for the host, where a check against the oracle
// runs on the CPU (tests/lane_eval_check.cpp).
//
// Rows i (pivot) and j (target) enter as, for every column k of the order,
//   p_k = |a_ik|,  qb_k = q_k + 2^31 (mod 2^32),  q_k = a_jk if a_ik >= 0 else -a_jk,
// column j included and 0 past n (p = 0, qb = 2^31). Column j does not take
// part in a shear's sums; its terms are removed from every sum through pj, qbj
// (its p, qb), which costs a few operations per sum instead of one per column.
// With the breakpoint b_k = -a_jk / a_ik of a column with a_ik != 0,
//   ceil(b_k) <= t  <=>  p_k t + q_k >= 0,     floor(b_k) < t  <=>  p_k t + q_k > 0,
// so the weighted lower median's ceiling and floor come from integer counts,
// with no division: ceil(m) = min{t : 2 W(ceil b <= t) >= W} by a search over
// t in [-64, 65] (65 standing for "beyond the cap": 0, then -1 or 1, then
// bisection of the rest), and floor(m) =
// ceil(m) - 1 iff 2 W(b < ceil m) >= W. These are the reference's floor and
// ceil of its double median: every breakpoint is at least 1/|a_ik| from an
// integer it is not equal to, and the double rounding of the reference moves
// it by less.
//
// Every quantity fits 32 bits: |q + s p| < 65 * 2^16 < 2^23, sums over at most
// 64 columns below 2^29, and no new entry can reach the 2^48 bound. With the
// bias 2^31 on q, p t + q >= 0 is the top bit of p t + qb, and |q + s p| is the
// unsigned distance of qb and 2^31 - s p: one 24-bit multiply-add and one
// shift, or one multiply-add and one sum of absolute differences, per column.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define XTX_HD __host__ __device__ __forceinline__
#else
#define XTX_HD inline
#endif

namespace xtx {
namespace lane {

constexpr int kCap = 64;  // mat_io.hpp:38
constexpr int32_t kSmallEntry = 1 << 16;
constexpr uint32_t kBias = 0x80000000u;

// a b + c (mod 2^32) for |a|, |b| < 2^23: one full-rate 24-bit multiply-add on
// the device (a 32-bit multiply there is a quarter-rate 64-bit one)
XTX_HD uint32_t mad24(int32_t a, int32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
    return (uint32_t)__mul24(a, b) + c;
#else
    return (uint32_t)(a * b) + c;
#endif
}
XTX_HD int32_t mul24(int32_t a, int32_t b) { return (int32_t)mad24(a, b, 0u); }
// |a - b| + c, unsigned (v_sad_u32 on the device)
XTX_HD uint32_t sad(uint32_t a, uint32_t b, uint32_t c) { return (a > b ? a - b : b - a) + c; }
// -1 when the top bit of v is set, else 0
XTX_HD int32_t top_mask(uint32_t v) { return (int32_t)v >> 31; }

XTX_HD int32_t abs32(int32_t v) { return v < 0 ? -v : v; }
XTX_HD int32_t clamp_cap(int32_t v) { return v > kCap ? kCap : (v < -kCap ? -kCap : v); }

// Nearest integer a/b, ties away from zero (mat_io.hpp:41-48).
XTX_HD int32_t rounded_div32(int32_t a, int32_t b) {
    int32_t q = a / b;
    const int32_t r = a - q * b;
    if (2 * abs32(r) >= abs32(b)) q += ((a < 0) == (b < 0)) ? 1 : -1;
    return q;
}

// p, qb of one column from a_ik, a_jk: with g = sign(a_ik) (+1 for 0),
// p = g a_ik and qb = g a_jk + 2^31, two 24-bit multiply(-add)s
XTX_HD void to_pq(int32_t ai, int32_t aj, int32_t& p, uint32_t& qb) {
    const int32_t g = (ai >> 31) | 1;
    p = mul24(ai, g);
    qb = mad24(aj, g, kBias);
}

// Candidate shears in the reference's order (main.cpp:214-231): floor and ceil
// of the median, then q, q+1, q-1 with q = -round(a_ji/a_ii) when a_ii != 0,
// each clamped to +/-64; a candidate equal to an earlier one is skipped.
struct Cands {
    int32_t v[5];
    bool ok[5];
};
XTX_HD void make_cands(int32_t cf, int32_t cc, bool has_q, int32_t q, Cands& cd) {
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

// The row pair of one proposal and its column-j terms.
template <int NMAX>
struct Pair {
    int32_t p[NMAX];
    uint32_t qb[NMAX];
    int32_t pj;
    uint32_t qbj;
    int32_t a_ii, a_ji, a_jj;
};

// sum_{k != j} |q_k + s p_k|
template <int NMAX>
XTX_HD int32_t abs_sum(const Pair<NMAX>& r, int32_t s) {
    const int32_t ns = -s;
    uint32_t acc = 0;
#pragma unroll
    for (int k = 0; k < NMAX; ++k) acc = sad(r.qb[k], mad24(r.p[k], ns, kBias), acc);
    return (int32_t)(acc - sad(r.qbj, mad24(r.pj, ns, kBias), 0u));
}

// sum_{k != j} p_k and sum_{k != j} |q_k| (the latter is abs_sum at s = 0)
template <int NMAX>
XTX_HD void totals(const Pair<NMAX>& r, int32_t& tot, int32_t& sumq) {
    int32_t t = 0;
    uint32_t a = 0;
#pragma unroll
    for (int k = 0; k < NMAX; ++k) {
        t += r.p[k];
        a = sad(r.qb[k], kBias, a);
    }
    tot = t - r.pj;
    sumq = (int32_t)(a - sad(r.qbj, kBias, 0u));
}

// W(ceil b <= t) = sum_{k != j} p_k [p_k t + q_k >= 0]
template <int NMAX>
XTX_HD int32_t weight_le(const Pair<NMAX>& r, int32_t t) {
    uint32_t acc = 0;  // minus the sum
#pragma unroll
    for (int k = 0; k < NMAX; ++k) acc = mad24(top_mask(mad24(r.p[k], t, r.qb[k])), r.p[k], acc);
    return (int32_t)((uint32_t)mul24(top_mask(mad24(r.pj, t, r.qbj)), r.pj) - acc);
}

// W(b < t) = sum_{k != j} p_k [p_k t + q_k > 0]
template <int NMAX>
XTX_HD int32_t weight_lt(const Pair<NMAX>& r, int32_t t) {
    int32_t acc = 0;
#pragma unroll
    for (int k = 0; k < NMAX; ++k) acc += mad24(r.p[k], t, r.qb[k]) > kBias ? r.p[k] : 0;
    return acc - (mad24(r.pj, t, r.qbj) > kBias ? r.pj : 0);
}

// score change of Add(i, j, s) given the off-diagonal change
// off = sum_{k != j} |q_k + s p_k| - |q_k| (score_delta_l1, main.cpp:150-168)
template <int NMAX>
XTX_HD int32_t delta_of(const Pair<NMAX>& r, int32_t s, int32_t off) {
    const int32_t jj = r.a_jj + mul24(s, 2 * r.a_ji + mul24(s, r.a_ii));
    return 4 * off + (abs32(jj) - abs32(r.a_jj));
}

// The best shear of the pair (0 when no candidate lowers the score), with its
// off-diagonal change in *off_out. tot, sumq from totals().
template <int NMAX>
XTX_HD int32_t best_shear(const Pair<NMAX>& r, int32_t tot, int32_t sumq, int32_t* off_out) {
    int32_t cf = 0, cc = 0;
    if (tot != 0) {
        // ceil(m) = min{t in [-64, 65] : 2 W(ceil b <= t) >= W}, P(65) by
        // convention. On the annealer's proposals it is 0 or 1 for 99% of
        // pairs, so t = 0 and then t = -1 or 1 are tested first (two passes
        // over the columns settle those), and what is left is bisected (six
        // more passes, as many as plain bisection takes in all).
        int lo = -kCap, hi = kCap + 1;
        if (2 * weight_le(r, 0) >= tot)
            hi = 0;
        else
            lo = 1;
        const int t1 = hi == 0 ? -1 : 1;
        if (2 * weight_le(r, t1) >= tot)
            hi = t1;
        else
            lo = t1 + 1;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;  // floor
            if (2 * weight_le(r, mid) >= tot)
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
            cf = (2 * weight_lt(r, lo) >= tot) ? lo - 1 : lo;
        }
    }
    const bool has_q = r.a_ii != 0;
    const int32_t qq = has_q ? -rounded_div32(r.a_ji, r.a_ii) : 0;
    Cands cd;
    make_cands(cf, cc, has_q, qq, cd);
    int32_t best_s = 0, best_d = 0, best_off = 0;
#pragma unroll
    for (int t = 0; t < 5; ++t) {
        const int32_t s = cd.v[t];
        if (!cd.ok[t] || s == 0) continue;
        const int32_t off = abs_sum(r, s) - sumq;
        const int32_t d = delta_of(r, s, off);
        if (d < best_d) {
            best_d = d;
            best_s = s;
            best_off = off;
        }
    }
    *off_out = best_off;
    return best_s;
}

}  // namespace lane
}  // namespace xtx
