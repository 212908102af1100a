// oracle_dual.cpp -- CPU oracle for the xdual primal/dual objective.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h). Restates, in code of our own, the
// move primitives of xdual's DualObjective (src/xdual_main.cpp:67-364) in the
// engine's double specification, and the reference's double Gauss-Jordan
// inverse (invert_to, src/mat_io.hpp:297-337) exactly:
//   * every sum over the columns k of a row is the pairwise (binary tree) sum
//     of its terms in column order, zero-padded to 128 columns;
//   * a primal term nv^2 - old^2 is formed exactly in 128-bit integers and
//     rounded once to double; a dual term is nv*nv - old*old in double;
//   * products and sums are single IEEE operations (no fused multiply-add).
// The reference accumulates in long double; tests/test_oracle_dual.py pins
// these functions against its outputs to a relative tolerance and the inverse
// bit for bit.
#include <cmath>
#include <cstdint>
#include <vector>

#include "oracle.h"

namespace {

constexpr int kPad = 128;

// pairwise sum of x[0..n) zero-padded to 128 terms
double pairsum(const double* x, int n) {
    double buf[kPad];
    for (int k = 0; k < kPad; ++k) buf[k] = k < n ? x[k] : 0.0;
    for (int w = kPad / 2; w >= 1; w /= 2)
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
        tp[(size_t)k] = (double)(nv * nv - (__int128)old * old);
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

// Off-diagonal squared Frobenius norms: per row the pairwise sum of its
// off-diagonal squares (primal squares exact in 128 bits, rounded once), the
// rows summed in order.
double orc_frob2_off_p(const int64_t* P, int32_t n) {
    std::vector<double> t((size_t)n);
    double s = 0.0;
    for (int r = 0; r < n; ++r) {
        for (int k = 0; k < n; ++k) {
            const __int128 v = P[(size_t)r * n + k];
            t[(size_t)k] = k == r ? 0.0 : (double)(v * v);
        }
        s += pairsum(t.data(), n);
    }
    return s;
}

double orc_frob2_off_q(const double* Q, int32_t n) {
    std::vector<double> t((size_t)n);
    double s = 0.0;
    for (int r = 0; r < n; ++r) {
        for (int k = 0; k < n; ++k) {
            const double v = Q[(size_t)r * n + k];
            t[(size_t)k] = k == r ? 0.0 : v * v;
        }
        s += pairsum(t.data(), n);
    }
    return s;
}

// Sampled residual of the dual (DualObjective::periodic_maintenance,
// xdual_main.cpp:316-333): max over rows r of |(P Q)_rk - delta_rk|, each
// product row a pairwise sum over t of P_rt * Q_tk.
double orc_dual_residual(const int64_t* P, const double* Q, int32_t n, int32_t k) {
    std::vector<double> t((size_t)n);
    double worst = 0.0;
    for (int r = 0; r < n; ++r) {
        for (int u = 0; u < n; ++u) t[(size_t)u] = (double)P[(size_t)r * n + u] * Q[(size_t)u * n + k];
        double acc = pairsum(t.data(), n);
        if (r == k) acc -= 1.0;
        const double a = std::fabs(acc);
        if (a > worst) worst = a;
    }
    return worst;
}

}  // extern "C"
