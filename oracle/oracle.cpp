// oracle.cpp -- CPU oracle for the xtax L1 congruence annealer.
//
// TEST INFRASTRUCTURE ONLY: the parity tests, smoke() and bench.py's
// cpu_baseline leg use it as the checker. The product never links it.
//
// Every move primitive below restates the reference's arithmetic (file:line
// cited per function) in code of our own; the trajectory model restates the
// engine's per-move algorithm (the reference's anneal_worker loop with Philox
// draws, see the model section).
#include "oracle.h"
#include "ref_draw.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <random>
#include <utility>
#include <array>
#include <vector>

namespace {

constexpr int64_t LIM = ORC_MAGNITUDE_LIMIT;
constexpr int64_t CAP = ORC_SHEAR_CAP;

inline int64_t iabs(int64_t v) { return v < 0 ? -v : v; }
inline bool over(int64_t v) { return v > LIM || v < -LIM; }

}  // namespace

// ---------------------------------------------------------------------------
// Primitives
// ---------------------------------------------------------------------------

// Nearest integer to a/b, ties away from zero (mat_io.hpp:41-48).
int64_t orc_rounded_div(int64_t a, int64_t b) {
    int64_t q = a / b;
    const int64_t r = a - q * b;
    if (2 * iabs(r) >= iabs(b)) q += ((a < 0) == (b < 0)) ? 1 : -1;
    return q;
}

// 2*sum|A_ij| - sum|A_ii| (mat_io.hpp:229-235).
int64_t orc_score(const int64_t* A, int32_t n) {
    int64_t full = 0, diag = 0;
    for (int64_t t = 0; t < (int64_t)n * n; ++t) full += iabs(A[t]);
    for (int32_t i = 0; i < n; ++i) diag += iabs(A[(int64_t)i * n + i]);
    return 2 * full - diag;
}

// Nonzero pairs above the diagonal (mat_io.hpp:219-225).
int64_t orc_offdiag_nonzero(const int64_t* A, int32_t n) {
    int64_t c = 0;
    for (int32_t i = 0; i < n; ++i)
        for (int32_t j = i + 1; j < n; ++j) c += A[(int64_t)i * n + j] != 0;
    return c;
}

// Evaluate Add(pivot i, target j, s) without mutation (main.cpp:46-91).
// Bound check first on the transform row when requested, then on the new row j
// of A (k != j), then on the new diagonal; delta = 4*offdiag + diag change.
int32_t orc_try_add(const int64_t* A, const int64_t* Xt, int32_t n, int32_t i, int32_t j,
                    int64_t s, int64_t* new_a_row, int64_t* new_xt_row, int64_t* delta,
                    int64_t* d_nonzero) {
    if (new_xt_row) {
        const int64_t* xj = Xt + (int64_t)j * n;
        const int64_t* xi = Xt + (int64_t)i * n;
        for (int32_t r = 0; r < n; ++r) {
            const int64_t v = xj[r] + s * xi[r];
            if (over(v)) return 0;
            new_xt_row[r] = v;
        }
    }
    const int64_t* aj = A + (int64_t)j * n;
    const int64_t* ai = A + (int64_t)i * n;
    int64_t off = 0, dnz = 0;
    for (int32_t k = 0; k < n; ++k) {
        if (k == j) continue;
        const int64_t old = aj[k];
        const int64_t nv = old + s * ai[k];
        if (over(nv)) return 0;
        new_a_row[k] = nv;
        off += iabs(nv) - iabs(old);
        if (old == 0) {
            if (nv != 0) ++dnz;
        } else if (nv == 0) {
            --dnz;
        }
    }
    const int64_t a_ii = ai[i];
    const int64_t a_ji = aj[i];
    const int64_t old_jj = aj[j];
    const int64_t new_jj = old_jj + 2 * s * a_ji + s * s * a_ii;
    if (over(new_jj)) return 0;
    new_a_row[j] = new_jj;
    *delta = 4 * off + (iabs(new_jj) - iabs(old_jj));
    *d_nonzero = dnz;
    return 1;
}

// Proposed row j of Xt for an accepted Add (main.cpp:97-108).
int32_t orc_add_transform_row(const int64_t* Xt, int32_t n, int32_t i, int32_t j, int64_t s,
                              int64_t* out) {
    const int64_t* xj = Xt + (int64_t)j * n;
    const int64_t* xi = Xt + (int64_t)i * n;
    for (int32_t r = 0; r < n; ++r) {
        const int64_t v = xj[r] + s * xi[r];
        if (over(v)) return 0;
        out[r] = v;
    }
    return 1;
}

// Write row j and its mirror column into A, row j into Xt, and keep the
// per-row off-diagonal mass current (main.cpp:112-133).
void orc_commit_add(int64_t* A, int64_t* Xt, int32_t n, int32_t j, const int64_t* new_a_row,
                    const int64_t* new_xt_row, int64_t* row_mass) {
    int64_t* aj = A + (int64_t)j * n;
    int64_t mass = 0;
    for (int32_t k = 0; k < n; ++k) {
        if (k == j) continue;
        const int64_t old = aj[k];
        const int64_t nv = new_a_row[k];
        aj[k] = nv;
        A[(int64_t)k * n + j] = nv;
        const int64_t na = iabs(nv);
        row_mass[k] += na - iabs(old);
        mass += na;
    }
    aj[j] = new_a_row[j];
    row_mass[j] = mass;
    std::memcpy(Xt + (int64_t)j * n, new_xt_row, sizeof(int64_t) * (size_t)n);
}

// row_mass[r] = sum_{k != r} |A_rk| (main.cpp:136-144).
void orc_compute_row_mass(const int64_t* A, int32_t n, int64_t* row_mass) {
    for (int32_t r = 0; r < n; ++r) {
        int64_t m = 0;
        for (int32_t k = 0; k < n; ++k)
            if (k != r) m += iabs(A[(int64_t)r * n + k]);
        row_mass[r] = m;
    }
}

// Score change of Add(i,j,s); 2^62 when some entry would leave the bound
// (main.cpp:150-168).
int64_t orc_score_delta_l1(const int64_t* A, int32_t n, int32_t i, int32_t j, int64_t s) {
    const int64_t* aj = A + (int64_t)j * n;
    const int64_t* ai = A + (int64_t)i * n;
    int64_t off = 0;
    for (int32_t k = 0; k < n; ++k) {
        if (k == j) continue;
        const int64_t old = aj[k];
        const int64_t nv = old + s * ai[k];
        if (over(nv)) return (int64_t)1 << 62;
        off += iabs(nv) - iabs(old);
    }
    const int64_t a_ii = ai[i];
    const int64_t a_ji = aj[i];
    const int64_t old_jj = aj[j];
    const int64_t new_jj = old_jj + 2 * s * a_ji + s * s * a_ii;
    if (over(new_jj)) return (int64_t)1 << 62;
    return 4 * off + (iabs(new_jj) - iabs(old_jj));
}

// Exact best integer shear under L1 (main.cpp:181-242): the |a_ik|-weighted
// lower median of the breakpoints -a_jk/a_ik (as doubles, sorted), its floor
// and ceil, then -round(a_ji/a_ii) and its two neighbours; every candidate
// clamped to +/-64 and deduplicated in that order; the first candidate with the
// strictly smallest negative delta wins, else 0.
int64_t orc_best_shear_l1(const int64_t* A, int32_t n, int32_t i, int32_t j) {
    const int64_t* ai = A + (int64_t)i * n;
    const int64_t* aj = A + (int64_t)j * n;
    std::vector<std::pair<double, int64_t>> keys;
    keys.reserve((size_t)n);
    int64_t total = 0;
    for (int32_t k = 0; k < n; ++k) {
        if (k == j || ai[k] == 0) continue;
        keys.emplace_back(-(double)aj[k] / (double)ai[k], iabs(ai[k]));
        total += iabs(ai[k]);
    }
    double med = 0.0;
    if (!keys.empty()) {
        std::sort(keys.begin(), keys.end(),
                  [](const auto& x, const auto& y) { return x.first < y.first; });
        med = keys.back().first;
        int64_t acc = 0;
        for (const auto& kv : keys) {
            acc += kv.second;
            if (2 * acc >= total) {
                med = kv.first;
                break;
            }
        }
    }
    int64_t cand[8];
    int nc = 0;
    auto push = [&](double x) {
        int64_t v = (int64_t)std::llround(x);
        v = std::min<int64_t>(CAP, std::max<int64_t>(-CAP, v));
        for (int t = 0; t < nc; ++t)
            if (cand[t] == v) return;
        if (nc < 8) cand[nc++] = v;
    };
    push(std::floor(med));
    push(std::ceil(med));
    const int64_t a_ii = ai[i];
    if (a_ii != 0) {
        const int64_t q = -orc_rounded_div(aj[i], a_ii);
        push((double)q);
        push((double)(q + 1));
        push((double)(q - 1));
    }
    int64_t best_s = 0, best_d = 0;
    for (int t = 0; t < nc; ++t) {
        if (cand[t] == 0) continue;
        const int64_t d = orc_score_delta_l1(A, n, i, j, cand[t]);
        if (d < best_d) {
            best_d = d;
            best_s = cand[t];
        }
    }
    return best_s;
}

// X^T A X in int64 without overflow checks (mat_io.hpp:340-364).
void orc_congruence_of(const int64_t* A, const int64_t* X, int32_t n, int64_t* out) {
    std::vector<int64_t> AX((size_t)n * n, 0);
    for (int32_t i = 0; i < n; ++i)
        for (int32_t k = 0; k < n; ++k) {
            const int64_t a = A[(int64_t)i * n + k];
            if (!a) continue;
            for (int32_t j = 0; j < n; ++j) AX[(size_t)i * n + j] += a * X[(int64_t)k * n + j];
        }
    std::fill(out, out + (int64_t)n * n, 0);
    for (int32_t i = 0; i < n; ++i)
        for (int32_t k = 0; k < n; ++k) {
            const int64_t x = X[(int64_t)k * n + i];
            if (!x) continue;
            for (int32_t j = 0; j < n; ++j) out[(int64_t)i * n + j] += x * AX[(size_t)k * n + j];
        }
}

// A = L L^T for m basis rows in dimension d (mat_io.hpp:545-559).
void orc_gram_of(const int64_t* L, int32_t m, int32_t d, int64_t* out) {
    for (int32_t i = 0; i < m; ++i)
        for (int32_t j = i; j < m; ++j) {
            int64_t s = 0;
            for (int32_t k = 0; k < d; ++k) s += L[(int64_t)i * d + k] * L[(int64_t)j * d + k];
            out[(int64_t)i * m + j] = s;
            out[(int64_t)j * m + i] = s;
        }
}

// R = X^T L (mat_io.hpp:564-580).
void orc_xt_times(const int64_t* X, const int64_t* L, int32_t m, int32_t d, int64_t* out) {
    std::fill(out, out + (int64_t)m * d, 0);
    for (int32_t i = 0; i < m; ++i)
        for (int32_t p = 0; p < m; ++p) {
            const int64_t x = X[(int64_t)p * m + i];
            if (!x) continue;
            for (int32_t k = 0; k < d; ++k) out[(int64_t)i * d + k] += x * L[(int64_t)p * d + k];
        }
}

// ---------------------------------------------------------------------------
// Philox4x32-10 and the deterministic exp
// ---------------------------------------------------------------------------

void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int round = 0; round < 10; ++round) {
        if (round) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c0 = n0;
        c1 = (uint32_t)p1;
        c2 = n2;
        c3 = (uint32_t)p0;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

// exp(x) for x <= 0 as the engine specifies it: 0 below -40 (e^-40 < 2^-53,
// the resolution of the uniform it is compared with); otherwise k = rint(x/ln2),
// r = x - k ln2 in two fma steps, a degree-13 Taylor polynomial in Horner form
// (all fma), scaled by 2^k. Every step is an IEEE operation with one rounding,
// so CPU and GPU give the same bits.
double orc_xexp(double x) {
    if (!(x > -40.0)) return 0.0;
    if (x > 0.0) x = 0.0;
    const double inv_ln2 = 1.4426950408889634;
    const double ln2_hi = 0x1.62e42fefa39efp-1;
    const double ln2_lo = 0x1.abc9e3b39803fp-56;
    const double kd = std::nearbyint(x * inv_ln2);
    double r = std::fma(-kd, ln2_hi, x);
    r = std::fma(-kd, ln2_lo, r);
    // 1/k! for k = 13 .. 1
    static const double c[13] = {
        1.0 / 6227020800.0, 1.0 / 479001600.0, 1.0 / 39916800.0, 1.0 / 3628800.0,
        1.0 / 362880.0,     1.0 / 40320.0,     1.0 / 5040.0,     1.0 / 720.0,
        1.0 / 120.0,        1.0 / 24.0,        1.0 / 6.0,        1.0 / 2.0,
        1.0};
    double p = c[0];
    for (int t = 1; t < 13; ++t) p = std::fma(p, r, c[t]);
    p = std::fma(p, r, 1.0);
    return std::ldexp(p, (int)kd);
}

// ---------------------------------------------------------------------------
// Exact verification (multi-modular, CRT-sized)
// ---------------------------------------------------------------------------
namespace {

typedef unsigned __int128 u128;

uint64_t mulmod(uint64_t a, uint64_t b, uint64_t p) { return (uint64_t)((u128)a * b % p); }

uint64_t powmod(uint64_t a, uint64_t e, uint64_t p) {
    uint64_t r = 1 % p;
    a %= p;
    while (e) {
        if (e & 1) r = mulmod(r, a, p);
        a = mulmod(a, a, p);
        e >>= 1;
    }
    return r;
}

bool is_prime_u64(uint64_t n) {
    if (n < 2) return false;
    static const uint64_t small[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    for (uint64_t q : small) {
        if (n % q == 0) return n == q;
    }
    uint64_t d = n - 1;
    int s = 0;
    while (!(d & 1)) {
        d >>= 1;
        ++s;
    }
    for (uint64_t a : small) {
        uint64_t x = powmod(a, d, n);
        if (x == 1 || x == n - 1) continue;
        bool comp = true;
        for (int t = 1; t < s; ++t) {
            x = mulmod(x, x, n);
            if (x == n - 1) {
                comp = false;
                break;
            }
        }
        if (comp) return false;
    }
    return true;
}

const std::vector<uint64_t>& primes61() {
    static std::vector<uint64_t> ps = [] {
        std::vector<uint64_t> v;
        for (uint64_t c = ((uint64_t)1 << 61) - 1; v.size() < 256; c -= 2)
            if (is_prime_u64(c)) v.push_back(c);
        return v;
    }();
    return ps;
}

uint64_t red(int64_t v, uint64_t p) {
    const int64_t r = v % (int64_t)p;
    return (uint64_t)(r < 0 ? r + (int64_t)p : r);
}

// det(M) mod p by elimination over Z/p.
uint64_t det_mod(std::vector<uint64_t> M, int n, uint64_t p) {
    uint64_t det = 1;
    for (int c = 0; c < n; ++c) {
        int piv = -1;
        for (int r = c; r < n; ++r)
            if (M[(size_t)r * n + c]) {
                piv = r;
                break;
            }
        if (piv < 0) return 0;
        if (piv != c) {
            for (int k = 0; k < n; ++k) std::swap(M[(size_t)piv * n + k], M[(size_t)c * n + k]);
            det = (p - det) % p;
        }
        const uint64_t pv = M[(size_t)c * n + c];
        det = mulmod(det, pv, p);
        const uint64_t inv = powmod(pv, p - 2, p);
        for (int r = c + 1; r < n; ++r) {
            const uint64_t f = mulmod(M[(size_t)r * n + c], inv, p);
            if (!f) continue;
            for (int k = c; k < n; ++k) {
                const uint64_t sub = mulmod(f, M[(size_t)c * n + k], p);
                uint64_t v = M[(size_t)r * n + k];
                M[(size_t)r * n + k] = v >= sub ? v - sub : v + p - sub;
            }
        }
    }
    return det;
}

}  // namespace

int32_t orc_verify(const int64_t* A0, const int64_t* X, const int64_t* A, int32_t n,
                   orc_verify_report* rep) {
    orc_verify_report out{};
    const int64_t nn = (int64_t)n * n;
    // A diagonal?
    for (int32_t i = 0; i < n; ++i)
        for (int32_t j = 0; j < n; ++j)
            if (i != j && A[(int64_t)i * n + j] != 0) out.status |= 2;
    // Bound on |(X^T A0 X)_ij| + |A_ij|, in log2.
    long double mx = 1, ma0 = 1, ma = 1;
    for (int64_t t = 0; t < nn; ++t) {
        mx = std::max(mx, (long double)iabs(X[t]));
        ma0 = std::max(ma0, (long double)iabs(A0[t]));
        ma = std::max(ma, (long double)iabs(A[t]));
    }
    const long double lb =
        std::log2((long double)n * n * mx * mx * ma0 + ma) + 2.0L;  // Prod p > 2*bound
    // Hadamard bound on |det X|.
    long double hb = 0;
    for (int32_t c = 0; c < n; ++c) {
        long double s2 = 0;
        for (int32_t r = 0; r < n; ++r) {
            const long double v = (long double)X[(int64_t)r * n + c];
            s2 += v * v;
        }
        hb += 0.5L * std::log2(std::max(s2, (long double)1));
    }
    const auto& ps = primes61();
    const int need_cong = (int)std::ceil(lb / 60.0L);
    const int need_det = (int)std::ceil((hb + 2.0L) / 60.0L);
    const int need = std::max(need_cong, need_det);
    const int use = std::min<int>(need, (int)ps.size());
    int det_pos = 1, det_neg = 1;
    for (int t = 0; t < use; ++t) {
        const uint64_t p = ps[(size_t)t];
        std::vector<uint64_t> a0((size_t)nn), x((size_t)nn), ax((size_t)nn, 0);
        for (int64_t q = 0; q < nn; ++q) {
            a0[(size_t)q] = red(A0[q], p);
            x[(size_t)q] = red(X[q], p);
        }
        if (t < need_cong) {
            for (int32_t i = 0; i < n; ++i)
                for (int32_t k = 0; k < n; ++k) {
                    const uint64_t a = a0[(size_t)i * n + k];
                    if (!a) continue;
                    for (int32_t j = 0; j < n; ++j)
                        ax[(size_t)i * n + j] =
                            (ax[(size_t)i * n + j] + mulmod(a, x[(size_t)k * n + j], p)) % p;
                }
            for (int32_t i = 0; i < n && !(out.status & 1); ++i)
                for (int32_t j = 0; j < n; ++j) {
                    uint64_t s = 0;
                    for (int32_t k = 0; k < n; ++k)
                        s = (s + mulmod(x[(size_t)k * n + i], ax[(size_t)k * n + j], p)) % p;
                    if (s != red(A[(int64_t)i * n + j], p)) {
                        out.status |= 1;
                        break;
                    }
                }
        }
        if (t < need_det) {
            const uint64_t d = det_mod(x, n, p);
            if (d != 1) det_pos = 0;
            if (d != p - 1) det_neg = 0;
        }
    }
    out.primes_used = use;
    if (need > (int)ps.size()) out.status |= 8;
    if (!det_pos && !det_neg) out.status |= 4;
    out.det_sign = det_pos ? 1 : (det_neg ? -1 : 0);
    if (rep) *rep = out;
    return out.status;
}

// ---------------------------------------------------------------------------
// Move-log replay
// ---------------------------------------------------------------------------
int64_t orc_replay(int64_t* A, int64_t* Xt, int32_t n, const orc_log_entry* log, int64_t count) {
    std::vector<int64_t> row((size_t)n), xrow((size_t)n), mass((size_t)n);
    orc_compute_row_mass(A, n, mass.data());
    for (int64_t e = 0; e < count; ++e) {
        const int32_t i = log[e].i, j = log[e].j;
        const int64_t s = log[e].s;
        if (i < 0 || j < 0 || i >= n || j >= n || i == j) return e;
        int64_t d = 0, dnz = 0;
        if (!orc_try_add(A, Xt, n, i, j, s, row.data(), xrow.data(), &d, &dnz)) return e;
        orc_commit_add(A, Xt, n, j, row.data(), xrow.data(), mass.data());
    }
    return count;
}

// ---------------------------------------------------------------------------
// Trajectory model of the engine
// ---------------------------------------------------------------------------
//
// Per replica r and move m (1-based), the engine draws Philox blocks
//   W(c) = philox4x32_10(key = {seed_lo, seed_hi}, ctr = {q_lo, q_hi, r, 16 t + c})
// with q = (m - 1) / 4, t = (m - 1) mod 4, for slots c = 0 .. 15 (one block
// call covers the 64 slots of four consecutive moves); a tournament of more
// than 16 samples takes slot c >= 16 from counter word 64 (c / 16) + 16 t +
// c mod 16, above every word of the first 16 slots. The words:
//   W(c).x  row-tournament candidate c       W(c).y  pivot-tournament cand. c
//   W(0).z  targeted-pick coin               W(0).w  greedy coin
//   W(1).z  sign of the +/-1 shear (bit 0)   W(1).w  Metropolis uniform, high
//   W(2).z  Metropolis uniform, low          W(2).w  uniform pivot position
//   W(3).z  uniform target position          W(3).w  fallback pivot position
//   W(4).z  stall sweep coin                 W(4).w  stall sweep pivot
// A position in the active list is mulhi32(w, N). The uniform pair and the
// fallback pivot draw the second index as (p + 1 + mulhi32(w, N-1)) mod N,
// i.e. uniformly among the others, in one draw.
// The replica-exchange sweep after epoch e draws, for adjacent rungs (q, q+1),
//   philox(key, {e_lo, e_hi, q, 0x40000000}).{x, y} as a 53-bit uniform.
namespace {

struct ModelRep {
    std::vector<int64_t> A, Xt, rm, bestA, bestXt;
    int64_t score = 0, offnz = 0;
    uint64_t moves = 0;
    int64_t local_best = 0, msi = 0;
    double T = 0, cooling = 0;
    int32_t cool_counter = 0, win_p = 0, win_a = 0;
    int64_t best_score = 0, best_move = 0;
    int32_t rung = 0;
    bool stopped = false, solved = false;
    int64_t solve_move = 0, accepts = 0;
    double gf = 0, tf = 0;
    std::vector<orc_log_entry> log;
};

inline uint32_t mulhi32(uint32_t w, uint32_t n) { return (uint32_t)(((uint64_t)w * n) >> 32); }
inline double u01(uint32_t w) { return (double)w * 0x1p-32; }
inline double u53(uint32_t hi, uint32_t lo) {
    return (double)(((uint64_t)(hi >> 5) << 26) | (lo >> 6)) * 0x1p-53;
}

// Planted mistakes, for the tests that show the pins catch them (0 = none):
//   1 auto-accept only strictly downhill moves (d < 0 instead of d <= 0)
//   2 the hot-row tournament keeps the lighter row (< instead of >)
//   3 the worker-diversity offset with the wrong sign
//   4 the exchange argument with the wrong sign
//   5 the Metropolis exponent with the wrong sign (uphill moves as likely
//     as their inverse)
int g_mutation = 0;

// The engine's draws for one move: Philox words of its slots (above).
struct PhiloxDraw {
    std::vector<std::array<uint32_t, 4>> W = std::vector<std::array<uint32_t, 4>>(16);
    uint32_t N = 0;
    int row = 0, piv = 0;  // tournament candidates drawn so far
    bool coin(double f) const { return u01(W[0][2]) < f; }
    uint32_t index() { return mulhi32(W[row++][0], N); }
    uint32_t pivot_index() { return mulhi32(W[piv++][1], N); }
    uint32_t fallback(uint32_t tpos) const {
        return (tpos + 1 + mulhi32(W[3][3], N - 1)) % N;
    }
    void pair(uint32_t& ip, uint32_t& jp) const {
        ip = mulhi32(W[2][3], N);
        jp = (ip + 1 + mulhi32(W[3][2], N - 1)) % N;
    }
    bool greedy_coin(double f) const { return u01(W[0][3]) < f; }
    int64_t sign() const { return (W[1][2] & 1u) ? 1 : -1; }
    // exp argument: d times -1/T (the reciprocal formed once per
    // temperature, as the device keeps it)
    bool metropolis(int64_t d, double T) const {
        return u53(W[1][3], W[2][2]) < orc_xexp((double)d * (-1.0 / T));
    }
    bool sweep_coin(double f) const { return u01(W[4][2]) < f; }
    uint32_t sweep_index() const { return mulhi32(W[4][3], N); }
};

struct Model {
    int n = 0, N = 0;
    std::vector<int> act;
    orc_params p{};
    bool tempering = false;
    double t_init = 0;
    std::vector<double> temps;
    uint32_t key[2] = {0, 0};
    std::vector<ModelRep> reps;
    int64_t gb_score = 0;  // instance best at epoch start
    std::vector<int64_t> gbA, gbXt;
    int64_t log_cap = 0;
    // reference-draw mode: one worker, the reference's draws and publish rule
    bool ref = false;
    double rung_T = 0;  // tempering: the worker's (fixed) rung temperature

    // Slot c of move m: Philox block at counter (q, r, 64 (c / 16) + 16 t +
    // c mod 16) with q = (m - 1) / 4 and t = (m - 1) mod 4, so one block
    // call can serve four consecutive moves (16 slots each).
    void block(uint64_t m, uint32_t r, uint32_t c, uint32_t w[4]) const {
        const uint64_t q = (m - 1) >> 2;
        const uint32_t t = (uint32_t)((m - 1) & 3);
        const uint32_t ctr[4] = {(uint32_t)q, (uint32_t)(q >> 32), r,
                                 64 * (c >> 4) + 16 * t + (c & 15)};
        orc_philox4x32_10(ctr, key, w);
    }

    void log_commit(ModelRep& R, uint64_t m, int i, int j, int64_t s, uint16_t flags) {
        if ((int64_t)R.log.size() < log_cap)
            R.log.push_back(orc_log_entry{(int64_t)m, (int16_t)i, (int16_t)j, (int16_t)s, flags});
    }

    // Evaluate + commit helper (reference: attempt_add's evaluate/commit,
    // congruence_anneal.hpp:232-240, and L1Objective::commit, main.cpp:393-399).
    bool commit_move(ModelRep& R, int i, int j, int64_t s, int64_t d, int64_t dnz,
                     const std::vector<int64_t>& row, uint64_t m, uint16_t flags) {
        std::vector<int64_t> xrow((size_t)n);
        if (!orc_add_transform_row(R.Xt.data(), n, i, j, s, xrow.data())) return false;
        orc_commit_add(R.A.data(), R.Xt.data(), n, j, row.data(), xrow.data(), R.rm.data());
        R.score += d;
        R.offnz += dnz;
        ++R.accepts;
        log_commit(R, m, i, j, s, flags);
        return true;
    }

    // Greedy reduction sweep from a pivot (congruence_anneal.hpp:246-260).
    bool sweep(ModelRep& R, int pivot, uint64_t m) {
        const int64_t before = R.score;
        // The reference's sweep ends as soon as the run is done (:250): in the
        // reference-draw mode the move cap stands for that stop, raised before
        // the capped move's stall handling.
        if (ref && p.max_moves && R.moves >= p.max_moves) return false;
        std::vector<int64_t> row((size_t)n);
        for (int q = 0; q < N; ++q) {
            const int j = act[(size_t)q];
            if (j == pivot) continue;
            const int64_t s = orc_best_shear_l1(R.A.data(), n, pivot, j);
            if (s == 0) continue;
            int64_t d = 0, dnz = 0;
            if (!orc_try_add(R.A.data(), nullptr, n, pivot, j, s, row.data(), nullptr, &d, &dnz))
                continue;
            if ((double)d > 0.0) continue;
            commit_move(R, pivot, j, s, d, dnz, row, m, 1);
        }
        return R.score < before;
    }

    void track_best(ModelRep& R, uint64_t m) {
        if (ref) return;  // the reference keeps no per-worker best state
        if (R.score < R.best_score) {
            R.best_score = R.score;
            R.best_move = (int64_t)m;
            R.bestA = R.A;
            R.bestXt = R.Xt;
        }
    }

    // One iteration of the worker loop (congruence_anneal.hpp:276-413) with
    // the engine's draws: Philox slots of (replica, move).
    void step(ModelRep& R, uint32_t r) {
        const uint64_t m = ++R.moves;
        PhiloxDraw D;
        D.N = (uint32_t)N;
        const int G = std::max(p.target_samples, 5);
        D.W.resize((size_t)G);
        for (int g = 0; g < G; ++g) block(m, r, (uint32_t)g, D.W[(size_t)g].data());
        step_with(R, D, m);
    }

    // The loop body, shared by both draw sources. `ref` selects the
    // reference's publish and solve rule (a strict new global best on an
    // accepted downhill move, congruence_anneal.hpp:309-352) in place of the
    // engine's (any replica at a diagonal; per-replica bests).
    template <class D>
    void step_with(ModelRep& R, D& draw, uint64_t m) {
        if (tempering) R.T = ref ? rung_T : temps[(size_t)R.rung];

        // pair: hot-row and pivot tournaments, else uniform (:211-228, :283-288)
        int i, j;
        if (draw.coin(R.tf)) {
            uint32_t tpos = draw.index();
            for (int c = 1; c < p.target_samples; ++c) {
                const uint32_t q = draw.index();
                const int64_t wq = R.rm[(size_t)act[q]], wt = R.rm[(size_t)act[tpos]];
                if (g_mutation == 2 ? wq < wt : wq > wt) tpos = q;
            }
            const int target = act[tpos];
            int ibest = -1;
            int64_t vbest = -1;
            for (int c = 0; c < p.target_samples; ++c) {
                const int cand = act[draw.pivot_index()];
                if (cand == target) continue;
                const int64_t v = iabs(R.A[(size_t)target * n + cand]);
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
        // shear: the exact best with probability greedy, +-1 otherwise or
        // when the best is 0 (:290-296)
        int64_t s = 0;
        if (draw.greedy_coin(R.gf)) s = orc_best_shear_l1(R.A.data(), n, i, j);
        if (s == 0) s = draw.sign();

        // attempt_add (:232-240): evaluate, Metropolis, commit
        int res = -1;
        {
            std::vector<int64_t> row((size_t)n);
            int64_t d = 0, dnz = 0;
            if (orc_try_add(R.A.data(), nullptr, n, i, j, s, row.data(), nullptr, &d, &dnz)) {
                const bool ok = (g_mutation == 1 ? d < 0 : d <= 0) ||
                                draw.metropolis(g_mutation == 5 ? -d : d, R.T);
                if (ok && commit_move(R, i, j, s, d, dnz, row, m, 0)) res = d < 0 ? 1 : 0;
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
        if (ref) {
            if (res == 1 && R.score < gb_score) {
                gb_score = R.score;
                gbA = R.A;
                gbXt = R.Xt;
                if (R.offnz == 0) {
                    R.solved = true;
                    R.stopped = true;
                    R.solve_move = (int64_t)m;
                    return;
                }
            }
        } else {
            track_best(R, m);
            if (R.offnz == 0) {
                R.solved = true;
                R.stopped = true;
                R.solve_move = (int64_t)m;
                return;
            }
        }

        if (tempering) {
            // stall: the optional reduction sweep only (:363-371)
            if (R.msi > p.stuck_threshold) {
                if (p.sweep_fraction > 0.0 && draw.sweep_coin(p.sweep_fraction)) {
                    sweep(R, act[draw.sweep_index()], m);
                    track_best(R, m);
                }
                R.local_best = R.score;
                R.msi = 0;
            }
        } else {
            // stall: sweep, reheat, reseed when far behind (:372-394); else cool (:396-411)
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
                    R.A = gbA;
                    R.Xt = gbXt;
                    R.score = orc_score(R.A.data(), n);
                    R.offnz = orc_offdiag_nonzero(R.A.data(), n);
                    orc_compute_row_mass(R.A.data(), n, R.rm.data());
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
        if (!ref && R.offnz == 0) {
            R.solved = true;
            R.stopped = true;
            R.solve_move = (int64_t)m;
            return;
        }
        if (p.max_moves && R.moves >= p.max_moves) R.stopped = true;
    }

    // Instance best: argmin over replicas of (best_score, best_move, replica).
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

}  // namespace

int64_t orc_exchange_sweep(const double* temps, const double* score, int32_t* rung, int32_t R,
                           uint64_t seed, uint64_t epoch) {
    // congruence_anneal.hpp:525-550: owner of each rung, then adjacent rungs
    // bottom up; the configuration carried up a rung is compared again at the
    // next pair. The uniform of pair q after epoch e is Philox (key = seed,
    // counter = {e, q, 0x40000000}).{x, y} as 53 bits.
    std::vector<int> owner((size_t)R);
    for (int r = 0; r < R; ++r) owner[(size_t)rung[r]] = r;
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    int64_t acc = 0;
    for (int q = 0; q + 1 < R; ++q) {
        const int wa = owner[(size_t)q], wb = owner[(size_t)q + 1];
        double arg = (1.0 / temps[q] - 1.0 / temps[q + 1]) * (score[wa] - score[wb]);
        if (g_mutation == 4) arg = -arg;
        uint32_t w[4];
        const uint32_t ctr[4] = {(uint32_t)epoch, (uint32_t)(epoch >> 32), (uint32_t)q,
                                 0x40000000u};
        orc_philox4x32_10(ctr, key, w);
        if (arg >= 0.0 || u53(w[0], w[1]) < orc_xexp(arg)) {
            rung[wa] = q + 1;
            rung[wb] = q;
            std::swap(owner[(size_t)q], owner[(size_t)q + 1]);
            ++acc;
        }
    }
    return acc;
}

void orc_set_mutation(int32_t which) { g_mutation = which; }

int32_t orc_get_mutation(void) { return g_mutation; }

int32_t orc_metropolis(int64_t d, double T, uint32_t u_hi, uint32_t u_lo) {
    // the step's test with the engine's uniform (Model::step_with, PhiloxDraw)
    PhiloxDraw D;
    D.W[1][3] = u_hi;
    D.W[2][2] = u_lo;
    return ((g_mutation == 1 ? d < 0 : d <= 0) || D.metropolis(g_mutation == 5 ? -d : d, T))
               ? 1
               : 0;
}

void orc_ladder(double top, double t_min, int32_t R, double* temps) {
    // congruence_anneal.hpp:454-468
    const double bot = std::min(t_min, top);
    if (R == 1) {
        temps[0] = top;
        return;
    }
    const double ratio = std::pow(top / std::max(bot, 1e-12), 1.0 / (double)(R - 1));
    double t = bot;
    for (int k = 0; k < R; ++k) {
        temps[k] = t;
        t *= ratio;
    }
}

void orc_fractions(const orc_params* p, int32_t r, double* greedy, double* target) {
    // congruence_anneal.hpp:200-207, with the replica index in place of the
    // worker id.
    double g = p->greedy_fraction, t = p->target_fraction;
    if (p->worker_diversity > 0.0 && p->replicas > 1) {
        double off = p->worker_diversity * ((double)r / (double)(p->replicas - 1) - 0.5);
        if (g_mutation == 3) off = -off;
        g = std::clamp(g + off, 0.0, 1.0);
        t = std::clamp(t + off, 0.0, 1.0);
    }
    *greedy = g;
    *target = t;
}

double orc_suggest_t_init(const int64_t* A, int32_t n) {
    // main.cpp:383-387
    const double avg = (double)orc_score(A, n) / (double)((int64_t)n * n);
    return std::max(1.0, 2.0 * avg);
}

int32_t orc_model_run(const int64_t* A_start, const int64_t* Xt_start, int32_t n,
                      const int32_t* active, int32_t n_active, const orc_params* pp,
                      int32_t max_epochs, int64_t* A_out, int64_t* Xt_out, int64_t* best_A,
                      int64_t* best_Xt, int64_t* scal_out, orc_log_entry* log_out,
                      int64_t* log_count, int64_t log_cap, orc_model_result* res) {
    Model M;
    M.n = n;
    M.p = *pp;
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
    const int64_t nn = (int64_t)n * n;
    std::vector<int64_t> A0(A_start, A_start + nn), X0(Xt_start, Xt_start + nn);
    const int64_t score0 = orc_score(A0.data(), n);
    M.t_init = pp->t_init > 0.0 ? pp->t_init : orc_suggest_t_init(A0.data(), n);
    if (M.tempering) {
        M.temps.resize((size_t)R);
        orc_ladder(M.t_init, pp->t_min, R, M.temps.data());
    }
    M.reps.resize((size_t)R);
    for (int r = 0; r < R; ++r) {
        ModelRep& x = M.reps[(size_t)r];
        x.A = A0;
        x.Xt = X0;
        x.rm.resize((size_t)n);
        orc_compute_row_mass(x.A.data(), n, x.rm.data());
        x.score = score0;
        x.offnz = orc_offdiag_nonzero(A0.data(), n);
        x.local_best = score0;
        x.T = M.t_init;
        x.cooling = pp->cooling;
        x.best_score = score0;
        x.best_move = 0;
        x.bestA = A0;
        x.bestXt = X0;
        x.rung = r;
        orc_fractions(pp, r, &x.gf, &x.tf);
    }
    M.gb_score = score0;
    M.gbA = A0;
    M.gbXt = X0;

    orc_model_result out{};
    const bool trivial = n < 2 || M.N < 2;
    const bool start_diag = orc_offdiag_nonzero(A0.data(), n) == 0;
    int epochs = 0;
    int64_t exchanges = 0;
    if (start_diag) {
        out.solved = 1;
        out.solve_replica = 0;
        out.solve_move = 0;
    }
    const uint64_t E = (uint64_t)std::max(1, pp->exchange_interval);
    for (int e = 0; !trivial && !start_diag && e < max_epochs; ++e) {
        bool any_running = false;
        for (int r = 0; r < R; ++r) {
            ModelRep& x = M.reps[(size_t)r];
            while (!x.stopped && x.moves < (uint64_t)(e + 1) * E) M.step(x, (uint32_t)r);
            any_running |= !x.stopped;
        }
        epochs = e + 1;
        // Solved? earliest (move, replica) wins.
        int sr = -1;
        for (int r = 0; r < R; ++r) {
            const ModelRep& x = M.reps[(size_t)r];
            if (x.solved && (sr < 0 || x.solve_move < M.reps[(size_t)sr].solve_move)) sr = r;
        }
        if (sr >= 0) {
            out.solved = 1;
            out.solve_replica = sr;
            out.solve_move = M.reps[(size_t)sr].solve_move;
            break;
        }
        if (M.tempering) {
            std::vector<int32_t> rung((size_t)R);
            std::vector<double> score((size_t)R);
            for (int r = 0; r < R; ++r) {
                rung[(size_t)r] = M.reps[(size_t)r].rung;
                score[(size_t)r] = (double)M.reps[(size_t)r].score;
            }
            exchanges += orc_exchange_sweep(M.temps.data(), score.data(), rung.data(), R,
                                            pp->seed, (uint64_t)e);
            for (int r = 0; r < R; ++r) M.reps[(size_t)r].rung = rung[(size_t)r];
        }
        const int b = M.best_rep();
        if (M.reps[(size_t)b].best_score < M.gb_score) {
            M.gb_score = M.reps[(size_t)b].best_score;
            M.gbA = M.reps[(size_t)b].bestA;
            M.gbXt = M.reps[(size_t)b].bestXt;
        }
        if (!any_running) break;
    }
    const int b = M.best_rep();
    out.best_replica = out.solved ? out.solve_replica : b;
    out.best_score = out.solved ? M.reps[(size_t)out.solve_replica].score : M.reps[(size_t)b].best_score;
    out.epochs = epochs;
    out.exchanges = exchanges;
    for (int r = 0; r < R; ++r) {
        const ModelRep& x = M.reps[(size_t)r];
        out.total_moves += (int64_t)x.moves;
        out.accepts += x.accepts;
        if (A_out) std::copy(x.A.begin(), x.A.end(), A_out + (int64_t)r * nn);
        if (Xt_out) std::copy(x.Xt.begin(), x.Xt.end(), Xt_out + (int64_t)r * nn);
        if (best_A) std::copy(x.bestA.begin(), x.bestA.end(), best_A + (int64_t)r * nn);
        if (best_Xt) std::copy(x.bestXt.begin(), x.bestXt.end(), best_Xt + (int64_t)r * nn);
        if (scal_out) {
            int64_t* o = scal_out + (int64_t)r * 8;
            o[0] = x.score;
            o[1] = x.offnz;
            o[2] = (int64_t)x.moves;
            o[3] = x.best_score;
            o[4] = x.best_move;
            o[5] = x.rung;
            o[6] = x.accepts;
            o[7] = x.local_best;
        }
        if (log_out && log_count) {
            std::copy(x.log.begin(), x.log.end(), log_out + (int64_t)r * log_cap);
            log_count[r] = (int64_t)x.log.size();
        }
    }
    if (res) *res = out;
    return out.solved;
}

// ---------------------------------------------------------------------------
// Reference-draw mode of the model: one worker of the reference's annealer
// ---------------------------------------------------------------------------
// The same loop body as the engine's model (Model::step_with), driven by the
// reference's draws (RefDraw) and its publish rule, for one worker `t` of a
// pool of `threads`: the worker's seed, its diversity offsets, and either its
// fixed tempering rung (no exchange) or the cooling schedule. With threads = 1
// this is run_annealer's whole single-threaded run (tempering is off below two
// workers, congruence_anneal.hpp:427), which is deterministic.

uint32_t orc_ref_worker_seed(uint64_t seed, int32_t t) {
    // congruence_anneal.hpp:484-488 (seed 0 draws from random_device there)
    std::mt19937 seed_rng((uint32_t)seed);
    uint32_t w = 0;
    for (int k = 0; k <= t; ++k) w = seed_rng();
    return w ^ ((uint32_t)t * 0x9e3779b9u);
}

int32_t orc_ref_worker_run(const int64_t* A_start, const int64_t* Xt_start, int32_t n,
                           const int32_t* active, int32_t n_active, const orc_ref_params* rp,
                           int64_t* A_out, int64_t* Xt_out, int64_t* gb_A, int64_t* gb_Xt,
                           orc_log_entry* log_out, int64_t* log_count, int64_t log_cap,
                           orc_ref_result* res) {
    Model M;
    M.n = n;
    M.p = rp->base;
    M.ref = true;
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
    const int64_t nn = (int64_t)n * n;
    std::vector<int64_t> A0(A_start, A_start + nn), X0(Xt_start, Xt_start + nn);
    const int64_t score0 = orc_score(A0.data(), n);
    M.t_init = rp->base.t_init > 0.0 ? rp->base.t_init : orc_suggest_t_init(A0.data(), n);
    M.reps.resize(1);
    ModelRep& x = M.reps[0];
    x.A = A0;
    x.Xt = X0;
    x.rm.resize((size_t)n);
    orc_compute_row_mass(x.A.data(), n, x.rm.data());
    x.score = score0;
    x.offnz = orc_offdiag_nonzero(A0.data(), n);
    x.local_best = score0;
    x.T = M.t_init;
    x.cooling = rp->base.cooling;
    x.best_score = score0;
    orc_fractions(&rp->base, rp->thread_id, &x.gf, &x.tf);
    M.gb_score = score0;
    M.gbA = A0;
    M.gbXt = X0;
    orc_ref_result out{};
    if (n >= 2 && M.N >= 2) {  // run_annealer returns the start otherwise (:435, :444)
        RefDraw D(rp->worker_seed, M.N);
        while (!x.stopped) {
            const uint64_t m = ++x.moves;
            M.step_with(x, D, m);
        }
    }
    out.solved = x.solved;
    out.moves = (int64_t)x.moves;
    out.accepts = x.accepts;
    out.score = x.score;
    out.gb_score = M.gb_score;
    if (A_out) std::copy(x.A.begin(), x.A.end(), A_out);
    if (Xt_out) std::copy(x.Xt.begin(), x.Xt.end(), Xt_out);
    if (gb_A) std::copy(M.gbA.begin(), M.gbA.end(), gb_A);
    if (gb_Xt) std::copy(M.gbXt.begin(), M.gbXt.end(), gb_Xt);
    if (log_out && log_count) {
        std::copy(x.log.begin(), x.log.end(), log_out);
        *log_count = (int64_t)x.log.size();
    }
    if (res) *res = out;
    return out.solved;
}
