// oracle_dual.cpp -- CPU oracle for the xdual primal/dual objective.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h). Restates, in code of our own, the
// move primitives of xdual's DualObjective (src/xdual_main.cpp:67-364) in the
// engine's double specification, the reference's double Gauss-Jordan inverse
// (invert_to, src/mat_io.hpp:297-337) exactly, and a trajectory model of the
// engine running that objective, whose loop body also runs in the reference's
// mode (its long double arithmetic, draws, publish rule and per-move
// maintenance: orc_dual_ref_run) to be pinned against the reference's own
// xdual runs. The engine's specification:
//   * a sum over the columns k of a row folds column k + 64 onto column k
//     (zero-padded to 128 columns), then adds the 64 results as a binary tree
//     of adjacent pairs;
//   * a primal term nv^2 - old^2 is (nv - old) * (nv + old) with both factors
//     exact doubles (|.| <= 2^49 for a feasible shear), i.e. the exact integer
//     rounded once; a dual term is nv*nv - old*old in double;
//   * a squared off-diagonal norm sums each column sequentially over the rows
//     (r = 0, 1, ...), then the column sums as above;
//   * products and sums are single IEEE operations (no fused multiply-add).
// The reference accumulates in long double; tests/test_oracle_dual.py pins
// these functions against its outputs to a relative tolerance and the inverse
// bit for bit.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <array>
#include <random>
#include <vector>

#include "oracle.h"
#include "ref_draw.h"

namespace {

constexpr int kPad = 128;

// x[0..n) zero-padded to 128 terms: x[k] + x[k + 64], then the adjacent-pair
// binary tree over those 64
double pairsum(const double* x, int n) {
    double buf[kPad / 2];
    for (int k = 0; k < kPad / 2; ++k)
        buf[k] = (k < n ? x[k] : 0.0) + (k + kPad / 2 < n ? x[k + kPad / 2] : 0.0);
    for (int w = kPad / 4; w >= 1; w /= 2)
        for (int k = 0; k < w; ++k) buf[k] = buf[2 * k] + buf[2 * k + 1];
    return buf[0];
}

bool over(__int128 v) { return v > ORC_MAGNITUDE_LIMIT || v < -ORC_MAGNITUDE_LIMIT; }

}  // namespace

extern "C" {

double orc_pairsum(const double* x, int32_t n) { return pairsum(x, n); }

// eval_add on P and eval_add_d on Q (xdual_main.cpp:67-158, DualObjective::
// evaluate :225-234): the primal shear Add(i, j, s) on P and the dual shear
// Add(j, i, -s) on Q. Returns 0 (infeasible) when a P entry would exceed 2^48.
int32_t orc_dual_eval(const int64_t* P, const double* Q, int32_t n, int32_t i, int32_t j,
                      int64_t s, int64_t* new_p_row, double* new_q_row, double* d_off_p,
                      double* d_off_q, int64_t* d_nonzero) {
    std::vector<double> tp((size_t)n, 0.0), tq((size_t)n, 0.0);
    int64_t dnz = 0;
    for (int k = 0; k < n; ++k) {
        if (k == j) continue;
        const int64_t old = P[(size_t)j * n + k];
        const __int128 nv = (__int128)old + (__int128)s * P[(size_t)i * n + k];
        if (over(nv)) return 0;
        new_p_row[k] = (int64_t)nv;
        // (nv - old)(nv + old): exact factors, one rounding
        tp[(size_t)k] = (double)(int64_t)(nv - old) * (double)(int64_t)(nv + old);
        if (old == 0) {
            if (nv != 0) ++dnz;
        } else if (nv == 0) {
            --dnz;
        }
    }
    const int64_t a_pp = P[(size_t)i * n + i], a_tp = P[(size_t)j * n + i];
    const __int128 new_tt =
        (__int128)P[(size_t)j * n + j] + 2 * (__int128)s * a_tp + (__int128)s * s * a_pp;
    if (over(new_tt)) return 0;
    new_p_row[j] = (int64_t)new_tt;
    // dual: target i, pivot j, shear -s
    const double sf = (double)(-s);
    for (int k = 0; k < n; ++k) {
        if (k == i) continue;
        const double old = Q[(size_t)i * n + k];
        const double prod = sf * Q[(size_t)j * n + k];
        const double nv = old + prod;
        new_q_row[k] = nv;
        const double a = nv * nv, b = old * old;
        tq[(size_t)k] = a - b;
    }
    {
        const double q_pp = Q[(size_t)j * n + j], q_tp = Q[(size_t)i * n + j];
        const double t1 = 2.0 * sf;
        const double t2 = t1 * q_tp;
        const double t3 = sf * sf;
        const double t4 = t3 * q_pp;
        const double t5 = Q[(size_t)i * n + i] + t2;
        new_q_row[i] = t5 + t4;
    }
    *d_off_p = 2.0 * pairsum(tp.data(), n);
    *d_off_q = 2.0 * pairsum(tq.data(), n);
    *d_nonzero = dnz;
    return 1;
}

// DualObjective::best_shear (xdual_main.cpp:254-296): the real minimizer of
// the quadratic change f(s) = 2 (quad s^2 + 2 lin s), its floor and floor + 1
// clamped to +-64, the one with the lower f if it is below f(0) = 0.
int64_t orc_dual_best_shear(const int64_t* P, const double* Q, int32_t n, double c, int32_t i,
                            int32_t j) {
    std::vector<double> ta((size_t)n, 0.0), tb((size_t)n, 0.0), tc((size_t)n, 0.0),
        td((size_t)n, 0.0);
    for (int k = 0; k < n; ++k) {
        if (k != j) {
            const double pik = (double)P[(size_t)i * n + k];
            ta[(size_t)k] = pik * pik;
            tb[(size_t)k] = (double)P[(size_t)j * n + k] * pik;
        }
        if (k != i) {
            const double qjk = Q[(size_t)j * n + k];
            tc[(size_t)k] = qjk * qjk;
            td[(size_t)k] = Q[(size_t)i * n + k] * qjk;
        }
    }
    const double a2 = pairsum(ta.data(), n), b2 = pairsum(tb.data(), n);
    const double c2 = pairsum(tc.data(), n), d2 = pairsum(td.data(), n);
    const double cc2 = c * c2;
    const double quad = a2 + cc2;
    if (!(quad > 0.0)) return 0;
    const double cd2 = c * d2;
    const double lin = b2 - cd2;
    double s_star = -lin / quad;
    if (s_star != s_star) return 0;  // no minimizer from a non-finite state
    s_star = s_star > 4e18 ? 4e18 : (s_star < -4e18 ? -4e18 : s_star);
    int64_t lo = (int64_t)std::floor(s_star);
    int64_t hi = lo + 1;
    lo = lo > ORC_SHEAR_CAP ? ORC_SHEAR_CAP : (lo < -ORC_SHEAR_CAP ? -ORC_SHEAR_CAP : lo);
    hi = hi > ORC_SHEAR_CAP ? ORC_SHEAR_CAP : (hi < -ORC_SHEAR_CAP ? -ORC_SHEAR_CAP : hi);
    int64_t best_s = 0;
    double best_f = 0.0;
    const int64_t cand[2] = {lo, hi};
    for (int t = 0; t < 2; ++t) {
        const int64_t s = cand[t];
        if (s == 0) continue;
        const double sd = (double)s;
        const double u = quad * sd;
        const double v = u * sd;
        const double w2 = 2.0 * lin;
        const double x = w2 * sd;
        const double f = 2.0 * (v + x);
        if (f < best_f) {
            best_f = f;
            best_s = s;
        }
    }
    return best_s;
}

// invert_to (mat_io.hpp:297-337): Gauss-Jordan with partial pivoting on
// [P | I] in double, then the symmetrized right half. Returns 0 if singular.
int32_t orc_invert(const int64_t* P, int32_t n, double* Q) {
    const int w = 2 * n;
    std::vector<double> a((size_t)n * w, 0.0);
    auto A = [&](int r, int c) -> double& { return a[(size_t)r * w + c]; };
    for (int r = 0; r < n; ++r) {
        for (int c = 0; c < n; ++c) A(r, c) = (double)P[(size_t)r * n + c];
        A(r, n + r) = 1.0;
    }
    for (int col = 0; col < n; ++col) {
        int piv = col;
        double big = std::fabs(A(col, col));
        for (int r = col + 1; r < n; ++r)
            if (std::fabs(A(r, col)) > big) {
                big = std::fabs(A(r, col));
                piv = r;
            }
        if (big == 0.0) return 0;
        if (piv != col)
            for (int c = 0; c < w; ++c) std::swap(A(piv, c), A(col, c));
        const double inv = 1.0 / A(col, col);
        for (int c = 0; c < w; ++c) A(col, c) *= inv;
        for (int r = 0; r < n; ++r) {
            if (r == col) continue;
            const double f = A(r, col);
            if (f == 0.0) continue;
            for (int c = 0; c < w; ++c) A(r, c) -= f * A(col, c);
        }
    }
    for (int r = 0; r < n; ++r)
        for (int c = 0; c < n; ++c) Q[(size_t)r * n + c] = A(r, n + c);
    for (int r = 0; r < n; ++r)
        for (int c = r + 1; c < n; ++c) {
            const double m = 0.5 * (Q[(size_t)r * n + c] + Q[(size_t)c * n + r]);
            Q[(size_t)r * n + c] = m;
            Q[(size_t)c * n + r] = m;
        }
    return 1;
}

// Off-diagonal squared Frobenius norms: per column k the sequential sum over
// rows r != k of the squares (exact entries, one rounding each), then the
// column sums pairwise as above.
double orc_frob2_off_p(const int64_t* P, int32_t n) {
    std::vector<double> col((size_t)n, 0.0);
    for (int k = 0; k < n; ++k)
        for (int r = 0; r < n; ++r)
            if (r != k) {
                const double v = (double)P[(size_t)r * n + k];
                col[(size_t)k] += v * v;
            }
    return pairsum(col.data(), n);
}

double orc_frob2_off_q(const double* Q, int32_t n) {
    std::vector<double> col((size_t)n, 0.0);
    for (int k = 0; k < n; ++k)
        for (int r = 0; r < n; ++r)
            if (r != k) {
                const double v = Q[(size_t)r * n + k];
                col[(size_t)k] += v * v;
            }
    return pairsum(col.data(), n);
}

// combined_score (xdual_main.cpp:160-162): |offdiag P|^2 + c |offdiag Q|^2
double orc_dual_score(const int64_t* P, const double* Q, int32_t n, double c) {
    const double fp = orc_frob2_off_p(P, n), fq = orc_frob2_off_q(Q, n);
    const double cq = c * fq;
    return fp + cq;
}

// Sampled residual of the dual (DualObjective::periodic_maintenance,
// xdual_main.cpp:316-333): max over rows r of |(P Q)_rk - delta_rk|, each
// product row the sequential sum over t of P_rt * Q_tk.
double orc_dual_residual(const int64_t* P, const double* Q, int32_t n, int32_t k) {
    double worst = 0.0;
    for (int r = 0; r < n; ++r) {
        double acc = 0.0;
        for (int t = 0; t < n; ++t) {
            const double pr = (double)P[(size_t)r * n + t] * Q[(size_t)t * n + k];
            acc += pr;
        }
        if (r == k) acc -= 1.0;
        const double a = std::fabs(acc);
        if (a > worst) worst = a;
    }
    return worst;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Trajectory model of the engine running the primal/dual objective
// ---------------------------------------------------------------------------
//
// The L1 model's per-move algorithm (oracle.cpp) with DualObjective in the
// objective slot (congruence_anneal.hpp:276-413 over xdual_main.cpp:169-364):
// the same Philox words, pair draws (row masses and pivot magnitudes of P) and
// stall handling; the shear from the dual best_shear; the delta
// d = d_off_p + c * d_off_q; Metropolis on d <= 0 or u < exp(d * (-1/T)).
// On top of the L1 model:
//   * a committed downhill move whose running score is below both its
//     replica's best and the instance best at the epoch start recomputes the
//     running score from scratch (the reference's recompute before publishing
//     a global best, congruence_anneal.hpp:310-315), and is a new replica best
//     if the recomputed score is below the replica's best;
//   * at every epoch start, for each running replica, the maintenance of the
//     moves made since the last one (periodic_maintenance,
//     xdual_main.cpp:305-334): a re-inversion when the move count crossed a
//     multiple of dual_refresh, else a residual check of column
//     (that multiple of dual_check) mod n when it crossed one, re-inverting
//     when the residual exceeds dual_tol; a re-inversion, or a changed dual
//     weight, recomputes the running score.
namespace {

inline uint32_t dmulhi32(uint32_t w, uint32_t n) { return (uint32_t)(((uint64_t)w * n) >> 32); }
inline double du01(uint32_t w) { return (double)w * 0x1p-32; }
inline double du53(uint32_t hi, uint32_t lo) {
    return (double)(((uint64_t)(hi >> 5) << 26) | (lo >> 6)) * 0x1p-53;
}

// The engine's draws for one move of a replica: the Philox words of its
// slots, in the L1 model's slot map (oracle.cpp).
struct DualPhiloxDraw {
    std::vector<std::array<uint32_t, 4>> W;
    uint32_t N = 0;
    int row = 0, piv = 0;  // tournament candidates drawn so far
    bool coin(double f) const { return du01(W[0][2]) < f; }
    uint32_t index() { return dmulhi32(W[(size_t)row++][0], N); }
    uint32_t pivot_index() { return dmulhi32(W[(size_t)piv++][1], N); }
    uint32_t fallback(uint32_t tpos) const {
        return (tpos + 1 + dmulhi32(W[3][3], N - 1)) % N;
    }
    void pair(uint32_t& ip, uint32_t& jp) const {
        ip = dmulhi32(W[2][3], N);
        jp = (ip + 1 + dmulhi32(W[3][2], N - 1)) % N;
    }
    bool greedy_coin(double f) const { return du01(W[0][3]) < f; }
    int64_t sign() const { return (W[1][2] & 1u) ? 1 : -1; }
    bool metropolis(double d, double T) const {
        return du53(W[1][3], W[2][2]) < orc_xexp(d * (-1.0 / T));
    }
    bool sweep_coin(double f) const { return du01(W[4][2]) < f; }
    uint32_t sweep_index() const { return dmulhi32(W[4][3], N); }
};

// The engine's arithmetic: the double specification at the top of this file.
struct EngineArith {
    using S = double;
    static constexpr bool kRef = false;
    static bool eval(const int64_t* P, const double* Q, int n, int i, int j, int64_t s, S c,
                     int64_t* prow, double* qrow, S* d, int64_t* dnz) {
        double d_p = 0, d_q = 0;
        if (!orc_dual_eval(P, Q, n, i, j, s, prow, qrow, &d_p, &d_q, dnz)) return false;
        const double cq = c * d_q;
        *d = d_p + cq;
        return true;
    }
    static int64_t best_shear(const int64_t* P, const double* Q, int n, S c, int i, int j) {
        return orc_dual_best_shear(P, Q, n, c, i, j);
    }
    static S score(const int64_t* P, const double* Q, int n, S c) {
        return orc_dual_score(P, Q, n, c);
    }
    static double residual(const int64_t* P, const double* Q, int n, int k) {
        return orc_dual_residual(P, Q, n, k);
    }
};

// The reference's arithmetic: long double sums in the order of its loops
// (eval_add and eval_add_d, xdual_main.cpp:67-158; best_shear :254-296;
// combined_score :160-162 over offdiag_frob2, mat_io.hpp:240-251 and
// :280-291; the sampled residual :316-333). P's entries are exact, Q's are
// doubles, the score and the dual weight long doubles.
struct RefArith {
    using S = long double;
    static constexpr bool kRef = true;
    static bool eval(const int64_t* P, const double* Q, int n, int i, int j, int64_t s, S c,
                     int64_t* prow, double* qrow, S* d, int64_t* dnz) {
        long double dp = 0.0L, dq = 0.0L;
        int64_t z = 0;
        for (int k = 0; k < n; ++k) {
            if (k == j) continue;
            const int64_t old = P[(size_t)j * n + k];
            const __int128 nv = (__int128)old + (__int128)s * P[(size_t)i * n + k];
            if (over(nv)) return false;
            prow[k] = (int64_t)nv;
            dp += (long double)prow[k] * (long double)prow[k] - (long double)old * (long double)old;
            if (old == 0) {
                if (nv != 0) ++z;
            } else if (nv == 0) {
                --z;
            }
        }
        const __int128 new_tt = (__int128)P[(size_t)j * n + j] +
                                2 * (__int128)s * P[(size_t)j * n + i] +
                                (__int128)s * s * P[(size_t)i * n + i];
        if (over(new_tt)) return false;
        prow[j] = (int64_t)new_tt;
        // dual: target i, pivot j, shear -s
        const double sf = (double)(-s);
        for (int k = 0; k < n; ++k) {
            if (k == i) continue;
            const double old = Q[(size_t)i * n + k];
            const double nv = old + sf * Q[(size_t)j * n + k];
            qrow[k] = nv;
            dq += (long double)nv * (long double)nv - (long double)old * (long double)old;
        }
        qrow[i] = Q[(size_t)i * n + i] + 2.0 * sf * Q[(size_t)i * n + j] +
                  sf * sf * Q[(size_t)j * n + j];
        *d = 2.0L * dp + c * (2.0L * dq);
        *dnz = z;
        return true;
    }
    static int64_t best_shear(const int64_t* P, const double* Q, int n, S c, int i, int j) {
        long double a2 = 0.0L, b2 = 0.0L, c2 = 0.0L, d2 = 0.0L;
        for (int k = 0; k < n; ++k) {
            if (k == j) continue;
            const long double pik = (long double)P[(size_t)i * n + k];
            a2 += pik * pik;
            b2 += (long double)P[(size_t)j * n + k] * pik;
        }
        for (int k = 0; k < n; ++k) {
            if (k == i) continue;
            const long double qjk = (long double)Q[(size_t)j * n + k];
            c2 += qjk * qjk;
            d2 += (long double)Q[(size_t)i * n + k] * qjk;
        }
        const long double quad = a2 + c * c2;
        if (!(quad > 0.0L)) return 0;
        const long double lin = b2 - c * d2;
        const long double s_star = -lin / quad;
        // the double's floor converted as x86-64 converts it (out of range or
        // NaN: the lowest int64)
        const double fl = std::floor((double)s_star);
        int64_t lo = (fl >= -0x1p63 && fl < 0x1p63) ? (int64_t)fl : INT64_MIN;
        int64_t hi = lo + 1;
        lo = std::clamp(lo, -(int64_t)ORC_SHEAR_CAP, (int64_t)ORC_SHEAR_CAP);
        hi = std::clamp(hi, -(int64_t)ORC_SHEAR_CAP, (int64_t)ORC_SHEAR_CAP);
        int64_t best_s = 0;
        long double best_f = 0.0L;
        for (const int64_t sv : {lo, hi}) {
            if (sv == 0) continue;
            const long double x = (long double)sv;
            const long double f = 2.0L * (quad * x * x + 2.0L * lin * x);
            if (f < best_f) {
                best_f = f;
                best_s = sv;
            }
        }
        return best_s;
    }
    static long double frob2_p(const int64_t* P, int n) {
        long double t = 0.0L;
        for (int r = 0; r < n; ++r)
            for (int k = 0; k < n; ++k)
                if (r != k) {
                    const long double v = (long double)P[(size_t)r * n + k];
                    t += v * v;
                }
        return t;
    }
    static long double frob2_q(const double* Q, int n) {
        long double t = 0.0L;
        for (int r = 0; r < n; ++r)
            for (int k = 0; k < n; ++k)
                if (r != k) {
                    const long double v = (long double)Q[(size_t)r * n + k];
                    t += v * v;
                }
        return t;
    }
    static S score(const int64_t* P, const double* Q, int n, S c) {
        return frob2_p(P, n) + c * frob2_q(Q, n);
    }
    static double residual(const int64_t* P, const double* Q, int n, int k) {
        long double worst = 0.0L;
        for (int r = 0; r < n; ++r) {
            long double acc = 0.0L;
            for (int t = 0; t < n; ++t)
                acc += (long double)P[(size_t)r * n + t] * (long double)Q[(size_t)t * n + k];
            if (r == k) acc -= 1.0L;
            const long double a = acc < 0.0L ? -acc : acc;
            if (a > worst) worst = a;
        }
        return (double)worst;
    }
};

template <class S>
struct DualRep {
    std::vector<int64_t> P, Xt, rm, bestP, bestXt;
    std::vector<double> Q, bestQ;
    S score = 0, local_best = 0, best_score = 0;
    double T = 0, cooling = 0, dc = 0;
    int64_t offnz = 0, msi = 0, best_move = 0, accepts = 0, solve_move = 0;
    int64_t refreshes = 0, checks = 0, recomputes = 0;
    uint64_t moves = 0, m_maint = 0;
    int32_t cool_counter = 0, win_p = 0, win_a = 0, rung = 0;
    bool stopped = false, solved = false;
    double gf = 0, tf = 0;
    std::vector<orc_log_entry> log;
};

// The loop body over an arithmetic (EngineArith, RefArith) and a draw source
// (DualPhiloxDraw, RefDraw). The engine's mode keeps per-replica bests, its
// recompute rule and epoch-start maintenance (above); the reference's mode
// (Ar::kRef, one worker) its publish rule: an accepted downhill move whose
// running score is below the global best recomputes the score and, if still
// below, becomes the global best, which stops the run when it is diagonal
// (congruence_anneal.hpp:309-352); and its maintenance after every move
// (:361).
template <class Ar>
struct DualModelT {
    using S = typename Ar::S;
    using Rep = DualRep<S>;
    static constexpr bool kRef = Ar::kRef;

    int n = 0, N = 0;
    std::vector<int> act;
    orc_params p{};
    orc_dual_params dp{};
    S c = 0;  // dual weight
    bool tempering = false;
    double rung_T = 0;  // reference mode, tempering: the worker's (fixed) rung temperature
    double t_init = 0;
    std::vector<double> temps;
    uint32_t key[2] = {0, 0};
    std::vector<Rep> reps;
    // engine: the instance best at the epoch start; reference: its global best
    S gb_score = 0;
    int64_t gb_offnz = 0;
    std::vector<int64_t> gbP, gbXt;
    std::vector<double> gbQ;
    int64_t log_cap = 0;

    void block(uint64_t m, uint32_t r, uint32_t cs, uint32_t w[4]) const {
        const uint64_t q = (m - 1) >> 2;
        const uint32_t t = (uint32_t)((m - 1) & 3);
        const uint32_t ctr[4] = {(uint32_t)q, (uint32_t)(q >> 32), r,
                                 64 * (cs >> 4) + 16 * t + (cs & 15)};
        orc_philox4x32_10(ctr, key, w);
    }
    void log_commit(Rep& R, uint64_t m, int i, int j, int64_t s, uint16_t flags) {
        if ((int64_t)R.log.size() < log_cap)
            R.log.push_back(orc_log_entry{(int64_t)m, (int16_t)i, (int16_t)j, (int16_t)s, flags});
    }
    S rescore(const Rep& R) const { return Ar::score(R.P.data(), R.Q.data(), n, c); }

    // DualObjective::commit (xdual_main.cpp:236-252): the transform row is
    // bound-checked first; P row/column j with the row masses, Q row/column
    // i, X^T row j.
    bool commit(Rep& R, int i, int j, int64_t s, const std::vector<int64_t>& prow,
                const std::vector<double>& qrow) {
        std::vector<int64_t> xrow((size_t)n);
        if (!orc_add_transform_row(R.Xt.data(), n, i, j, s, xrow.data())) return false;
        int64_t off = 0;
        for (int k = 0; k < n; ++k) {
            if (k == j) continue;
            const int64_t o = R.P[(size_t)j * n + k], v = prow[(size_t)k];
            R.rm[(size_t)k] += std::llabs(v) - std::llabs(o);
            off += std::llabs(v) - std::llabs(o);
        }
        R.rm[(size_t)j] += off;
        const int qi = orc_get_mutation() == 6 ? j : i;
        for (int k = 0; k < n; ++k) {
            R.P[(size_t)j * n + k] = prow[(size_t)k];
            R.P[(size_t)k * n + j] = prow[(size_t)k];
            R.Q[(size_t)qi * n + k] = qrow[(size_t)k];
            R.Q[(size_t)k * n + qi] = qrow[(size_t)k];
            R.Xt[(size_t)j * n + k] = xrow[(size_t)k];
        }
        return true;
    }
    void snapshot(Rep& R) {
        R.bestP = R.P;
        R.bestQ = R.Q;
        R.bestXt = R.Xt;
    }
    void track_best(Rep& R, uint64_t m) {
        if (kRef) return;  // the reference keeps no per-worker best state
        if (R.score < R.best_score) {
            R.best_score = R.score;
            R.best_move = (int64_t)m;
            snapshot(R);
        }
    }

    // Greedy reduction sweep from a pivot (congruence_anneal.hpp:246-260).
    bool sweep(Rep& R, int pivot, uint64_t m) {
        const S before = R.score;
        // The reference's sweep ends as soon as the run is done (:250): in the
        // reference's mode the move cap stands for that stop, raised before the
        // capped move's stall handling.
        if (kRef && p.max_moves && R.moves >= p.max_moves) return false;
        std::vector<int64_t> prow((size_t)n);
        std::vector<double> qrow((size_t)n);
        for (int q = 0; q < N; ++q) {
            const int j = act[(size_t)q];
            if (j == pivot) continue;
            const int64_t s = Ar::best_shear(R.P.data(), R.Q.data(), n, c, pivot, j);
            if (s == 0) continue;
            S d = 0;
            int64_t dnz = 0;
            if (!Ar::eval(R.P.data(), R.Q.data(), n, pivot, j, s, c, prow.data(), qrow.data(), &d,
                          &dnz))
                continue;
            if ((double)d > 0.0) continue;
            if (!commit(R, pivot, j, s, prow, qrow)) continue;
            R.score += d;
            R.offnz += dnz;
            ++R.accepts;
            log_commit(R, m, pivot, j, s, 1);
        }
        return R.score < before;
    }

    // One move with the engine's draws: Philox slots of (replica, move).
    void step(Rep& R, uint32_t r) {
        const uint64_t m = ++R.moves;
        DualPhiloxDraw D;
        D.N = (uint32_t)N;
        const int G = std::max(p.target_samples, 5);
        D.W.resize((size_t)G);
        for (int g = 0; g < G; ++g) block(m, r, (uint32_t)g, D.W[(size_t)g].data());
        step_with(R, D, m);
    }

    // One iteration of the worker loop (congruence_anneal.hpp:276-413).
    template <class D>
    void step_with(Rep& R, D& draw, uint64_t m) {
        const int mut = orc_get_mutation();
        if (tempering) R.T = kRef ? rung_T : temps[(size_t)R.rung];

        // pair: hot-row and pivot tournaments, else uniform (:211-228, :283-288)
        int i, j;
        if (draw.coin(R.tf)) {
            uint32_t tpos = draw.index();
            for (int q = 1; q < p.target_samples; ++q) {
                const uint32_t cand = draw.index();
                const int64_t wc = R.rm[(size_t)act[cand]], wt = R.rm[(size_t)act[tpos]];
                if (mut == 2 ? wc < wt : wc > wt) tpos = cand;
            }
            const int target = act[tpos];
            int ibest = -1;
            int64_t vbest = -1;
            for (int q = 0; q < p.target_samples; ++q) {
                const int cand = act[draw.pivot_index()];
                if (cand == target) continue;
                const int64_t v = std::llabs(R.P[(size_t)target * n + cand]);
                if (v > vbest) {
                    vbest = v;
                    ibest = cand;
                }
            }
            if (ibest < 0) ibest = act[draw.fallback(tpos)];
            i = ibest;
            j = target;
        } else {
            uint32_t ip, jp;
            draw.pair(ip, jp);
            i = act[ip];
            j = act[jp];
        }
        // shear: the best with probability greedy, +-1 otherwise or when the
        // best is 0 (:290-296)
        int64_t s = 0;
        if (draw.greedy_coin(R.gf)) s = Ar::best_shear(R.P.data(), R.Q.data(), n, c, i, j);
        if (s == 0) s = draw.sign();

        // attempt_add (:232-240): evaluate, Metropolis on the double delta, commit
        int res = -1;
        {
            std::vector<int64_t> prow((size_t)n);
            std::vector<double> qrow((size_t)n);
            S d = 0;
            int64_t dnz = 0;
            if (Ar::eval(R.P.data(), R.Q.data(), n, i, j, s, c, prow.data(), qrow.data(), &d,
                         &dnz)) {
                const double dd = (double)d;
                const bool ok = (mut == 1 ? dd < 0.0 : dd <= 0.0) ||
                                draw.metropolis(mut == 5 ? -dd : dd, R.T);
                if (ok && commit(R, i, j, s, prow, qrow)) {
                    if (kRef) {
                        R.score += d;
                        R.offnz += dnz;
                    } else {
                        const S ns = R.score + d;
                        const bool rec = d < 0 && ns < R.best_score && ns < gb_score;
                        bool newbest = ns < R.best_score;
                        R.score = ns;
                        R.offnz += dnz;
                        if (rec) {
                            R.score = rescore(R);
                            ++R.recomputes;
                            newbest = R.score < R.best_score;
                        }
                        if (newbest) {
                            R.best_score = R.score;
                            R.best_move = (int64_t)m;
                            snapshot(R);
                        }
                    }
                    ++R.accepts;
                    log_commit(R, m, i, j, s, 0);
                    res = dd < 0.0 ? 1 : 0;
                }
            }
        }
        ++R.win_p;
        if (res >= 0) ++R.win_a;
        if (R.score < R.local_best) {
            R.local_best = R.score;
            R.msi = 0;
        } else {
            ++R.msi;
        }
        if (kRef) {
            if (res == 1 && (double)R.score < (double)gb_score) {
                R.score = rescore(R);
                ++R.recomputes;
                if ((double)R.score < (double)gb_score) {
                    gb_score = R.score;
                    gb_offnz = R.offnz;
                    gbP = R.P;
                    gbQ = R.Q;
                    gbXt = R.Xt;
                    if (R.offnz == 0) {
                        R.solved = R.stopped = true;
                        R.solve_move = (int64_t)m;
                        return;
                    }
                }
            }
            maintain_move(R, m);
        } else if (R.offnz == 0) {
            R.solved = R.stopped = true;
            R.solve_move = (int64_t)m;
            return;
        }
        if (tempering) {
            if (R.msi > p.stuck_threshold) {
                if (p.sweep_fraction > 0.0 && draw.sweep_coin(p.sweep_fraction)) {
                    sweep(R, act[draw.sweep_index()], m);
                    track_best(R, m);
                }
                R.local_best = R.score;
                R.msi = 0;
            }
        } else {
            if (R.msi > p.stuck_threshold) {
                if (p.sweep_fraction > 0.0 && draw.sweep_coin(p.sweep_fraction)) {
                    if (sweep(R, act[draw.sweep_index()], m)) {
                        R.local_best = R.score;
                        R.msi = 0;
                    }
                    track_best(R, m);
                }
                R.T = p.reheat * t_init;
                if ((double)R.score > (double)gb_score * p.reseed_factor) {
                    log_commit(R, m, -1, -1, 0, 2);
                    R.P = gbP;
                    R.Q = gbQ;
                    R.Xt = gbXt;
                    orc_compute_row_mass(R.P.data(), n, R.rm.data());
                    R.score = gb_score;
                    R.offnz = gb_offnz;
                    track_best(R, m);
                }
                R.local_best = R.score;
                R.msi = 0;
                R.cool_counter = 0;
            } else if (++R.cool_counter >= p.moves_per_cool) {
                R.cool_counter = 0;
                if (p.adaptive_cooling && R.win_p > 0) {
                    const double acc = (double)R.win_a / (double)R.win_p;
                    if (acc > p.target_accept)
                        R.cooling *= 0.999;
                    else
                        R.cooling /= 0.999;
                    R.cooling = std::min(0.99999, std::max(0.90, R.cooling));
                }
                R.win_p = 0;
                R.win_a = 0;
                R.T *= R.cooling;
                if (R.T < p.t_min) R.T = p.t_min;
            }
        }
        if (!kRef && R.offnz == 0) {
            R.solved = R.stopped = true;
            R.solve_move = (int64_t)m;
            return;
        }
        if (p.max_moves && R.moves >= p.max_moves) R.stopped = true;
    }

    // The reference's maintenance after move m (DualObjective::
    // periodic_maintenance, xdual_main.cpp:305-334, the weight ramp off).
    void maintain_move(Rep& R, uint64_t m) {
        bool refresh = dp.dual_refresh > 0 && m % (uint64_t)dp.dual_refresh == 0;
        if (!refresh && dp.dual_check > 0 && m % (uint64_t)dp.dual_check == 0) {
            ++R.checks;
            refresh = Ar::residual(R.P.data(), R.Q.data(), n, (int)(m % (uint64_t)n)) > dp.dual_tol;
        }
        if (refresh) {
            std::vector<double> Qn((size_t)n * n);
            if (orc_invert(R.P.data(), n, Qn.data())) {
                R.Q.swap(Qn);
                ++R.refreshes;
                R.score = rescore(R);
            }
        }
    }

    // The engine's maintenance at an epoch start, for the moves since the last.
    void maintain(Rep& R) {
        if (R.stopped) return;
        const uint64_t m0 = R.m_maint, m1 = R.moves;
        R.m_maint = m1;
        bool refresh = false;
        if (dp.dual_refresh > 0 && m1 / (uint64_t)dp.dual_refresh > m0 / (uint64_t)dp.dual_refresh) {
            refresh = true;
        } else if (dp.dual_check > 0 &&
                   m1 / (uint64_t)dp.dual_check > m0 / (uint64_t)dp.dual_check) {
            const uint64_t at = m1 / (uint64_t)dp.dual_check * (uint64_t)dp.dual_check;
            const int k = (int)(at % (uint64_t)n);
            ++R.checks;
            refresh = Ar::residual(R.P.data(), R.Q.data(), n, k) > dp.dual_tol;
        }
        bool rescore_now = R.dc != (double)c;
        if (refresh) {
            std::vector<double> Qn((size_t)n * n);
            if (orc_invert(R.P.data(), n, Qn.data())) {
                R.Q.swap(Qn);
                ++R.refreshes;
                rescore_now = true;
            }
        }
        if (rescore_now) {
            R.score = rescore(R);
            R.dc = (double)c;
        }
    }

    int best_rep() const {
        int b = 0;
        for (int r = 1; r < (int)reps.size(); ++r) {
            const auto& x = reps[(size_t)r];
            const auto& y = reps[(size_t)b];
            if (x.best_score < y.best_score ||
                (x.best_score == y.best_score && x.best_move < y.best_move))
                b = r;
        }
        return b;
    }
};

using EngineModel = DualModelT<EngineArith>;

}  // namespace

extern "C" int32_t orc_dual_model_run(const int64_t* P_start, const double* Q_start,
                                      const int64_t* Xt_start, int32_t n, const int32_t* active,
                                      int32_t n_active, const orc_params* pp,
                                      const orc_dual_params* dpp, int32_t max_epochs,
                                      int64_t* P_out, double* Q_out, int64_t* Xt_out,
                                      double* dscal_out, int64_t* scal_out,
                                      int64_t* best_P, double* best_Q, int64_t* best_Xt,
                                      orc_log_entry* log_out, int64_t* log_count,
                                      int64_t log_cap, orc_dual_result* res) {
    EngineModel M;
    M.n = n;
    M.p = *pp;
    M.dp = *dpp;
    M.c = dpp->c;
    M.log_cap = log_cap;
    if (active && n_active > 0) {
        M.act.assign(active, active + n_active);
    } else {
        M.act.resize((size_t)n);
        for (int k = 0; k < n; ++k) M.act[(size_t)k] = k;
    }
    M.N = (int)M.act.size();
    const int R = std::max(1, pp->replicas);
    M.tempering = pp->tempering && R >= 2;
    M.key[0] = (uint32_t)pp->seed;
    M.key[1] = (uint32_t)(pp->seed >> 32);
    const size_t nn = (size_t)n * n;
    const std::vector<int64_t> P0(P_start, P_start + nn), X0(Xt_start, Xt_start + nn);
    const std::vector<double> Q0(Q_start, Q_start + nn);
    const double score0 = orc_dual_score(P0.data(), Q0.data(), n, dpp->c);
    const int64_t offnz0 = orc_offdiag_nonzero(P0.data(), n);
    // suggest_t_init (xdual_main.cpp:212-219)
    const double avg = score0 / (double)((int64_t)n * n);
    M.t_init = pp->t_init > 0.0 ? pp->t_init : std::max(1.0, 2.0 * avg);
    if (M.tempering) {
        M.temps.resize((size_t)R);
        orc_ladder(M.t_init, pp->t_min, R, M.temps.data());
    }
    M.reps.resize((size_t)R);
    for (int r = 0; r < R; ++r) {
        EngineModel::Rep& x = M.reps[(size_t)r];
        x.P = P0;
        x.Q = Q0;
        x.Xt = X0;
        x.rm.resize((size_t)n);
        orc_compute_row_mass(x.P.data(), n, x.rm.data());
        x.score = x.local_best = x.best_score = score0;
        x.dc = dpp->c;
        x.offnz = offnz0;
        x.T = M.t_init;
        x.cooling = pp->cooling;
        x.rung = r;
        M.snapshot(x);
        orc_fractions(pp, r, &x.gf, &x.tf);
    }
    M.gb_score = score0;
    M.gb_offnz = offnz0;
    M.gbP = P0;
    M.gbQ = Q0;
    M.gbXt = X0;

    orc_dual_result out{};
    out.solve_replica = -1;
    const bool trivial = n < 2 || M.N < 2;
    const bool start_diag = offnz0 == 0;
    if (start_diag) {
        out.solved = 1;
        out.solve_replica = 0;
    }
    int epochs = 0;
    int64_t exchanges = 0;
    const uint64_t E = (uint64_t)std::max(1, pp->exchange_interval);
    for (int e = 0; !trivial && !start_diag && e < max_epochs; ++e) {
        for (int r = 0; r < R; ++r) M.maintain(M.reps[(size_t)r]);
        bool any_running = false;
        for (int r = 0; r < R; ++r) {
            EngineModel::Rep& x = M.reps[(size_t)r];
            while (!x.stopped && x.moves < (uint64_t)(e + 1) * E) M.step(x, (uint32_t)r);
            any_running |= !x.stopped;
        }
        epochs = e + 1;
        int sr = -1;
        for (int r = 0; r < R; ++r) {
            const EngineModel::Rep& x = M.reps[(size_t)r];
            if (x.solved && (sr < 0 || x.solve_move < M.reps[(size_t)sr].solve_move)) sr = r;
        }
        if (sr >= 0) {
            out.solved = 1;
            out.solve_replica = sr;
            out.solve_move = M.reps[(size_t)sr].solve_move;
            break;
        }
        if (M.tempering) {
            std::vector<int> owner((size_t)R);
            for (int r = 0; r < R; ++r) owner[(size_t)M.reps[(size_t)r].rung] = r;
            for (int q = 0; q + 1 < R; ++q) {
                const int wa = owner[(size_t)q], wb = owner[(size_t)q + 1];
                const double Ta = M.temps[(size_t)q], Tb = M.temps[(size_t)q + 1];
                const double Sa = M.reps[(size_t)wa].score, Sb = M.reps[(size_t)wb].score;
                const double arg = (1.0 / Ta - 1.0 / Tb) * (Sa - Sb);
                uint32_t w[4];
                const uint32_t ctr[4] = {(uint32_t)e, (uint32_t)((uint64_t)e >> 32), (uint32_t)q,
                                         0x40000000u};
                orc_philox4x32_10(ctr, M.key, w);
                if (arg >= 0.0 || du53(w[0], w[1]) < orc_xexp(arg)) {
                    M.reps[(size_t)wa].rung = q + 1;
                    M.reps[(size_t)wb].rung = q;
                    std::swap(owner[(size_t)q], owner[(size_t)q + 1]);
                    ++exchanges;
                }
            }
        }
        const int b = M.best_rep();
        if (M.reps[(size_t)b].best_score < M.gb_score) {
            const EngineModel::Rep& x = M.reps[(size_t)b];
            M.gb_score = x.best_score;
            M.gbP = x.bestP;
            M.gbQ = x.bestQ;
            M.gbXt = x.bestXt;
            M.gb_offnz = orc_offdiag_nonzero(M.gbP.data(), n);
        }
        if (!any_running) break;
    }
    const int b = M.best_rep();
    out.best_replica = out.solved && out.solve_replica >= 0 ? out.solve_replica : b;
    const EngineModel::Rep& win = M.reps[(size_t)out.best_replica];
    const bool cur = out.solved;
    out.best_score = cur ? win.score : win.best_score;
    out.epochs = epochs;
    out.exchanges = exchanges;
    for (int r = 0; r < R; ++r) {
        const EngineModel::Rep& x = M.reps[(size_t)r];
        out.total_moves += (int64_t)x.moves;
        out.accepts += x.accepts;
        out.refreshes += x.refreshes;
        out.checks += x.checks;
        out.recomputes += x.recomputes;
        if (P_out) std::copy(x.P.begin(), x.P.end(), P_out + (size_t)r * nn);
        if (Q_out) std::copy(x.Q.begin(), x.Q.end(), Q_out + (size_t)r * nn);
        if (Xt_out) std::copy(x.Xt.begin(), x.Xt.end(), Xt_out + (size_t)r * nn);
        if (dscal_out) {
            double* o = dscal_out + (size_t)r * 4;
            o[0] = x.score;
            o[1] = x.best_score;
            o[2] = x.local_best;
            o[3] = x.T;
        }
        if (scal_out) {
            int64_t* o = scal_out + (size_t)r * 8;
            o[0] = x.offnz;
            o[1] = (int64_t)x.moves;
            o[2] = x.best_move;
            o[3] = x.rung;
            o[4] = x.accepts;
            o[5] = x.refreshes;
            o[6] = x.checks;
            o[7] = x.msi;
        }
        if (log_out && log_count) {
            std::copy(x.log.begin(), x.log.end(), log_out + (size_t)r * log_cap);
            log_count[r] = (int64_t)x.log.size();
        }
    }
    if (best_P) {
        const auto& src = cur ? win.P : win.bestP;
        std::copy(src.begin(), src.end(), best_P);
    }
    if (best_Q) {
        const auto& src = cur ? win.Q : win.bestQ;
        std::copy(src.begin(), src.end(), best_Q);
    }
    if (best_Xt) {
        const auto& src = cur ? win.Xt : win.bestXt;
        std::copy(src.begin(), src.end(), best_Xt);
    }
    if (res) *res = out;
    return out.solved;
}

// ---------------------------------------------------------------------------
// Reference-draw mode of the dual model: one worker of xdual's annealer
// ---------------------------------------------------------------------------
// The same loop body (DualModelT::step_with) with the reference's arithmetic
// (RefArith), its draws (RefDraw, seeded with the worker seed) and its
// publish and maintenance rules, for worker t of a pool of `threads`: its
// diversity offsets, and either its fixed tempering rung (no exchange) or
// the cooling schedule. With one worker this is run_xdual's whole run
// (xdual_main.cpp:378-412 over run_annealer, congruence_anneal.hpp:422-444,
// tempering off below two workers), which is deterministic. The dual start is
// the reference's inverse of P, the weight lambda scaled by the ratio of the
// start's primal and dual off-diagonal norms (:389-392).
extern "C" int32_t orc_dual_ref_run(const int64_t* P_start, const int64_t* Xt_start, int32_t n,
                                    const int32_t* active, int32_t n_active,
                                    const orc_ref_params* rp, const orc_dual_params* dpp,
                                    double lambda, int64_t* P_out, double* Q_out,
                                    int64_t* Xt_out, int64_t* gb_P, double* gb_Q,
                                    int64_t* gb_Xt, orc_log_entry* log_out, int64_t* log_count,
                                    int64_t log_cap, orc_dual_ref_result* res) {
    DualModelT<RefArith> M;
    M.n = n;
    M.p = rp->base;
    M.dp = *dpp;
    M.log_cap = log_cap;
    if (active && n_active > 0) {
        M.act.assign(active, active + n_active);
    } else {
        M.act.resize((size_t)n);
        for (int k = 0; k < n; ++k) M.act[(size_t)k] = k;
    }
    M.N = (int)M.act.size();
    const int threads = std::max(1, rp->base.replicas);
    M.tempering = rp->base.tempering && threads >= 2;
    M.rung_T = rp->rung_T;
    const size_t nn = (size_t)n * n;
    const std::vector<int64_t> P0(P_start, P_start + nn), X0(Xt_start, Xt_start + nn);
    std::vector<double> Q0(nn);
    orc_dual_ref_result out{};
    if (n < 1 || !orc_invert(P0.data(), n, Q0.data())) {
        if (res) *res = out;
        return -1;
    }
    const long double sp = RefArith::frob2_p(P0.data(), n), sq = RefArith::frob2_q(Q0.data(), n);
    long double c = (long double)lambda;
    if (sp > 0.0L && sq > 0.0L && std::isfinite((double)sq)) c *= sp / sq;
    M.c = c;
    const long double score0 = RefArith::score(P0.data(), Q0.data(), n, c);
    // suggest_t_init (xdual_main.cpp:212-219)
    M.t_init = rp->base.t_init > 0.0
                   ? rp->base.t_init
                   : std::max(1.0, 2.0 * (double)(score0 / (long double)((int64_t)n * n)));
    M.reps.resize(1);
    auto& x = M.reps[0];
    x.P = P0;
    x.Q = Q0;
    x.Xt = X0;
    x.rm.resize((size_t)n);
    orc_compute_row_mass(x.P.data(), n, x.rm.data());
    x.score = x.local_best = x.best_score = score0;
    x.offnz = orc_offdiag_nonzero(P0.data(), n);
    x.T = M.t_init;
    x.cooling = rp->base.cooling;
    orc_fractions(&rp->base, rp->thread_id, &x.gf, &x.tf);
    M.gb_score = score0;
    M.gb_offnz = x.offnz;
    M.gbP = P0;
    M.gbQ = Q0;
    M.gbXt = X0;
    if (n >= 2 && M.N >= 2) {
        RefDraw D(rp->worker_seed, M.N);
        while (!x.stopped) {
            const uint64_t m = ++x.moves;
            M.step_with(x, D, m);
        }
    }
    out.solved = x.solved;
    out.moves = (int64_t)x.moves;
    out.accepts = x.accepts;
    out.refreshes = x.refreshes;
    out.checks = x.checks;
    out.recomputes = x.recomputes;
    out.score = (double)x.score;
    out.gb_score = (double)M.gb_score;
    out.c = (double)c;
    if (P_out) std::copy(x.P.begin(), x.P.end(), P_out);
    if (Q_out) std::copy(x.Q.begin(), x.Q.end(), Q_out);
    if (Xt_out) std::copy(x.Xt.begin(), x.Xt.end(), Xt_out);
    if (gb_P) std::copy(M.gbP.begin(), M.gbP.end(), gb_P);
    if (gb_Q) std::copy(M.gbQ.begin(), M.gbQ.end(), gb_Q);
    if (gb_Xt) std::copy(M.gbXt.begin(), M.gbXt.end(), gb_Xt);
    if (log_out && log_count) {
        std::copy(x.log.begin(), x.log.end(), log_out);
        *log_count = (int64_t)x.log.size();
    }
    if (res) *res = out;
    return out.solved;
}
