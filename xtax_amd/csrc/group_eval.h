// group_eval.h -- exact int64 evaluation of an Add proposal by a lane group
// (G lanes, lane g owning columns g + c G, c < C; sums as group reductions):
// try_add without the transform row and best_shear_l1 with its candidate
// order (main.cpp:46-91, 181-242). Used by the lane-group annealer for every
// proposal with an entry of 2^16 or more, and by the lane kernel for its
// stall sweeps and for proposals whose rows hold such entries.
#pragma once
#include <hip/hip_runtime.h>

#include "device_common.h"
#include "engine_types.h"

namespace xtx {

// Result of evaluating Add(i, j, s) (main.cpp:46-91) over one group.
template <int C>
struct MoveEval {
    int64_t nv[C];    // proposed A[j][k] for the lane's columns (k == j: new A_jj)
    int64_t off;      // sum_{k != j} |new| - |old|
    int64_t delta;
    int32_t dnz;
    bool bad;         // some entry would exceed the bound
};

// try_add without the transform row (main.cpp:46-91).
template <int G, int C>
__device__ __forceinline__ void evaluate(const int64_t (&ai)[C], const int64_t (&aj)[C], int g,
                                         int n, int j, int64_t s, int64_t a_ii, int64_t a_ji,
                                         int64_t a_jj, MoveEval<C>& e, bool know = false,
                                         int64_t known_off = 0) {
    int64_t off = 0;
    int32_t dnz = 0;
    bool bad = false;
    const int64_t new_jj = a_jj + 2 * s * a_ji + s * s * a_ii;
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const int k = g + c * G;
        e.nv[c] = new_jj;
        if (k < n && k != j) {
            const int64_t old = aj[c];
            const int64_t nv = old + s * ai[c];
            bad |= over(nv);
            off += iabs64(nv) - iabs64(old);
            dnz += (old == 0) ? (nv != 0) : -(int32_t)(nv == 0);
            e.nv[c] = nv;
        }
    }
    off = know ? known_off : Grp<G>::sum(off);
    dnz = Grp<G>::sum32(dnz);
    bad = Grp<G>::any(bad) || over(new_jj);
    e.off = off;
    e.dnz = dnz;
    e.bad = bad;
    e.delta = 4 * off + (iabs64(new_jj) - iabs64(a_jj));
}

// Candidate shears in the reference's order (main.cpp:214-231): floor and ceil
// of the median, then q, q+1, q-1 with q = -round(a_ji/a_ii) when a_ii != 0,
// each clamped to +/-64; a candidate equal to an earlier one is skipped.
struct Cands {
    int64_t v[5];
    bool ok[5];
};

__device__ __forceinline__ void make_cands(int64_t cf, int64_t cc, bool has_q, int64_t q,
                                           Cands& cd) {
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

// Ceiling and floor of a breakpoint b = -aj/ai (ai != 0, |ai|, |aj| <= 2^48),
// clamped to [-65, 65]. For |b| <= 66 the correctly rounded double quotient is
// within 66 * 2^-53 of b and equals b when b is an integer, so its ceiling is
// ceil(b) or ceil(b) - 1; one exact comparison (b <= t <=> -aj <= t ai for
// ai > 0) settles which, and b is an integer iff t ai == -aj.
__device__ __forceinline__ void breakpoint64(int64_t ai, int64_t aj, int32_t& ce, int32_t& fl) {
    double qc = __builtin_ceil(-(double)aj / (double)ai);
    qc = qc < -66.0 ? -66.0 : (qc > 66.0 ? 66.0 : qc);
    int32_t t = (int32_t)qc;
    const int64_t ta = (int64_t)t * ai, na = -aj;
    const bool le = ai > 0 ? na <= ta : na >= ta;
    const bool exact = le && ta == na;
    t += le ? 0 : 1;
    const int32_t f = exact ? t : t - 1;
    ce = t < -65 ? -65 : (t > 65 ? 65 : t);
    fl = f < -65 ? -65 : (f > 65 ? 65 : f);
}

// Exact best integer shear (main.cpp:181-242). The |a_ik|-weighted lower median
// m of the breakpoints b_k = -a_jk/a_ik is not formed; its floor and ceiling are
// found exactly from each breakpoint's ceiling and floor (clamped to [-65, 65],
// which leaves every weight sum W(b <= t), t in [-64, 64], unchanged):
// ceil(m) = min{t : 2 W(b <= t) >= W} (binary search over t in [-64, 64],
// clamped) and floor(m) = ceil(m) - 1 iff 2 W(b < ceil m) >= W. Since
// |a| <= 2^48 these equal the reference's double floor/ceil of the median
// (every breakpoint is at least 1/|a_ik| from an integer while double rounding
// moves it by at most 2^-5/|a_ik|).
// Candidates, clamping, dedup order and the strict-improvement rule follow the
// reference; returns 0 when no candidate lowers the score.
// best_shear_inl is the body, for callers that need it inline (no call, no
// stack arrays); best_shear is the same as a function of its own.
template <int G, int C>
__device__ __forceinline__ int64_t best_shear_inl(const int64_t (&ai)[C], const int64_t (&aj)[C],
                                                  int g, int n, int j, int64_t a_ii, int64_t a_ji,
                                                  int64_t a_jj, int64_t* off_out) {
    int64_t w[C];
    int32_t ce[C], fl[C];
    int64_t wl = 0;
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const int k = g + c * G;
        const bool bp = k < n && k != j && ai[c] != 0;
        w[c] = bp ? iabs64(ai[c]) : 0;
        wl += w[c];
        ce[c] = 0;
        fl[c] = 0;
        if (bp) breakpoint64(ai[c], aj[c], ce[c], fl[c]);
    }
    const int64_t tot = Grp<G>::sum(wl);
    int64_t cf = 0, cc = 0;
    if (tot != 0) {
        int lo = -kShearCap, hi = kShearCap + 1;  // P(hi) holds by convention
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            int64_t acc = 0;
#pragma unroll
            for (int c = 0; c < C; ++c) acc += ce[c] <= mid ? w[c] : 0;
            acc = Grp<G>::sum(acc);
            if (2 * acc >= tot)
                hi = mid;
            else
                lo = mid + 1;
        }
        if (lo <= -kShearCap) {
            cf = cc = -kShearCap;
        } else if (lo > kShearCap) {
            cf = cc = kShearCap;
        } else {
            cc = lo;
            int64_t acc = 0;
#pragma unroll
            for (int c = 0; c < C; ++c) acc += fl[c] < lo ? w[c] : 0;
            acc = Grp<G>::sum(acc);
            cf = (2 * acc >= tot) ? lo - 1 : lo;
        }
    }
    const bool has_q = a_ii != 0;
    const int64_t q = has_q ? -rounded_div(a_ji, a_ii) : 0;
    Cands cd;
    make_cands(cf, cc, has_q, q, cd);
    // Score every candidate (score_delta_l1, main.cpp:150-168).
    int64_t best_s = 0, best_d = 0, best_off = 0;
#pragma unroll
    for (int t = 0; t < 5; ++t) {
        const int64_t s = cd.v[t];
        if (cd.ok[t] && s != 0) {
            int64_t part = 0;
            bool bad = false;
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const int k = g + c * G;
                if (k < n && k != j) {
                    const int64_t nv = aj[c] + s * ai[c];
                    bad |= over(nv);
                    part += iabs64(nv) - iabs64(aj[c]);
                }
            }
            const int64_t off = Grp<G>::sum(part);
            const bool b = Grp<G>::any(bad);
            const int64_t new_jj = a_jj + 2 * s * a_ji + s * s * a_ii;
            const int64_t d = (b || over(new_jj)) ? ((int64_t)1 << 62)
                                                  : 4 * off + (iabs64(new_jj) - iabs64(a_jj));
            if (d < best_d) {
                best_d = d;
                best_s = s;
                best_off = off;
            }
        }
    }
    *off_out = best_off;
    return best_s;
}

template <int G, int C>
__device__ int64_t best_shear(const int64_t (&ai)[C], const int64_t (&aj)[C], int g, int n, int j,
                              int64_t a_ii, int64_t a_ji, int64_t a_jj, int64_t* off_out) {
    return best_shear_inl<G, C>(ai, aj, g, n, j, a_ii, a_ji, a_jj, off_out);
}

}  // namespace xtx
