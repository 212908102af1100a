// verify.cpp -- exact host verification of an annealer result (xtx_verify).
//
// Checks, with no floating point and no overflow:
//   1. X^T A0 X == A. With m_x = max|X|, m_0 = max|A0| the entries of A0 X are
//      bounded by n m_0 m_x and those of X^T A0 X by n^2 m_0 m_x^2; when that
//      bound is below 2^126 the product is formed directly in __int128.
//      Otherwise it is checked modulo k distinct primes p < 2^62 with
//      prod p > 2 (n^2 m_0 m_x^2 + max|A|), which decides equality (CRT).
//   2. A is diagonal.
//   3. |det X| == 1: det X mod the same kind of primes, with prod p above twice
//      Hadamard's bound prod_j ||X e_j||, all residues equal to +1 (or all -1).
// The reference's own in-binary check is the L-mode Gram comparison
// (main.cpp:853-857); this is the strict form the engine reports.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <mutex>
#include <vector>

#include "xtx.h"

namespace {

typedef __int128 i128;
typedef unsigned __int128 u128;

uint64_t mul_mod(uint64_t a, uint64_t b, uint64_t m) { return (uint64_t)((u128)a * b % m); }

uint64_t pow_mod(uint64_t b, uint64_t e, uint64_t m) {
    uint64_t r = 1;
    b %= m;
    while (e) {
        if (e & 1) r = mul_mod(r, b, m);
        b = mul_mod(b, b, m);
        e >>= 1;
    }
    return r;
}

// Deterministic Miller-Rabin for 64-bit integers (first 12 prime bases).
bool prime64(uint64_t n) {
    static const uint64_t bases[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    if (n < 2) return false;
    for (uint64_t b : bases)
        if (n % b == 0) return n == b;
    uint64_t d = n - 1;
    unsigned s = 0;
    for (; (d & 1) == 0; d >>= 1) ++s;
    for (uint64_t b : bases) {
        uint64_t x = pow_mod(b, d, n);
        if (x == 1 || x == n - 1) continue;
        bool witness = true;
        for (unsigned t = 1; t < s && witness; ++t) {
            x = mul_mod(x, x, n);
            witness = x != n - 1;
        }
        if (witness) return false;
    }
    return true;
}

// Primes just below 2^62, largest first (generated once).
const std::vector<uint64_t>& prime_table() {
    static std::vector<uint64_t> table;
    static std::once_flag once;
    std::call_once(once, [] {
        uint64_t c = ((uint64_t)1 << 62) - 57;  // odd start below 2^62
        while (table.size() < 160) {
            if (prime64(c)) table.push_back(c);
            c -= 2;
        }
    });
    return table;
}

inline uint64_t residue(int64_t v, uint64_t p) {
    const int64_t r = v % (int64_t)p;
    return (uint64_t)(r < 0 ? r + (int64_t)p : r);
}

// det M mod p, Gaussian elimination over GF(p).
uint64_t det_mod_p(std::vector<uint64_t> M, int n, uint64_t p) {
    uint64_t det = 1;
    for (int col = 0; col < n; ++col) {
        int piv = col;
        while (piv < n && M[(size_t)piv * n + col] == 0) ++piv;
        if (piv == n) return 0;
        if (piv != col) {
            std::swap_ranges(M.begin() + (ptrdiff_t)piv * n, M.begin() + (ptrdiff_t)(piv + 1) * n,
                             M.begin() + (ptrdiff_t)col * n);
            det = det ? p - det : 0;
        }
        const uint64_t pv = M[(size_t)col * n + col];
        det = mul_mod(det, pv, p);
        const uint64_t inv = pow_mod(pv, p - 2, p);
        for (int row = col + 1; row < n; ++row) {
            const uint64_t f = mul_mod(M[(size_t)row * n + col], inv, p);
            if (f == 0) continue;
            for (int k = col; k < n; ++k) {
                const uint64_t t = mul_mod(f, M[(size_t)col * n + k], p);
                const uint64_t v = M[(size_t)row * n + k];
                M[(size_t)row * n + k] = v >= t ? v - t : v + (p - t);
            }
        }
    }
    return det;
}

}  // namespace

extern "C" int xtx_verify(const int64_t* A0, const int64_t* X, const int64_t* A, int32_t n,
                          xtx_verify_report* report) {
    if (!A0 || !X || !A || n < 1 || !report) return XTX_E_ARG;
    xtx_verify_report rep{};
    const size_t nn = (size_t)n * (size_t)n;
    for (int32_t r = 0; r < n; ++r)
        for (int32_t c = 0; c < n; ++c)
            if (r != c && A[(size_t)r * n + c] != 0) rep.status |= 2;

    u128 mx = 1, m0 = 1, ma = 0;
    for (size_t q = 0; q < nn; ++q) {
        mx = std::max<u128>(mx, (u128)(X[q] < 0 ? -(i128)X[q] : (i128)X[q]));
        m0 = std::max<u128>(m0, (u128)(A0[q] < 0 ? -(i128)A0[q] : (i128)A0[q]));
        ma = std::max<u128>(ma, (u128)(A[q] < 0 ? -(i128)A[q] : (i128)A[q]));
    }
    const long double log_bound = std::log2((long double)n) * 2 + std::log2((long double)m0) +
                                  2 * std::log2((long double)mx);
    if (log_bound < 125.0L) {
        rep.exact_int128 = 1;
        std::vector<i128> AX(nn, 0);
        for (int32_t r = 0; r < n; ++r)
            for (int32_t k = 0; k < n; ++k) {
                const i128 a = A0[(size_t)r * n + k];
                if (a == 0) continue;
                for (int32_t c = 0; c < n; ++c) AX[(size_t)r * n + c] += a * X[(size_t)k * n + c];
            }
        for (int32_t r = 0; r < n && !(rep.status & 1); ++r)
            for (int32_t c = 0; c < n; ++c) {
                i128 s = 0;
                for (int32_t k = 0; k < n; ++k) s += (i128)X[(size_t)k * n + r] * AX[(size_t)k * n + c];
                if (s != (i128)A[(size_t)r * n + c]) {
                    rep.status |= 1;
                    break;
                }
            }
    }
    // Hadamard bound of det X, log2.
    long double log_had = 0;
    for (int32_t c = 0; c < n; ++c) {
        long double s2 = 0;
        for (int32_t r = 0; r < n; ++r) {
            const long double v = (long double)X[(size_t)r * n + c];
            s2 += v * v;
        }
        if (s2 > 1) log_had += 0.5L * std::log2(s2);
    }
    const auto& primes = prime_table();
    const long double lb_cong =
        std::log2(std::pow(2.0L, log_bound) + (long double)ma + 1.0L) + 1.0L;
    const int k_cong = rep.exact_int128 ? 0 : (int)std::ceil(lb_cong / 61.0L);
    const int k_det = (int)std::ceil((log_had + 1.0L) / 61.0L) + 1;
    const int k = std::max(k_cong, k_det);
    if (k > (int)primes.size()) {
        rep.status |= 8;
        *report = rep;
        return XTX_OK;
    }
    bool all_pos = true, all_neg = true;
    std::vector<uint64_t> xm(nn), am(nn), axm(nn);
    for (int t = 0; t < k; ++t) {
        const uint64_t p = primes[(size_t)t];
        for (size_t q = 0; q < nn; ++q) xm[q] = residue(X[q], p);
        if (t < k_cong && !(rep.status & 1)) {
            for (size_t q = 0; q < nn; ++q) am[q] = residue(A0[q], p);
            std::fill(axm.begin(), axm.end(), 0);
            for (int32_t r = 0; r < n; ++r)
                for (int32_t kk = 0; kk < n; ++kk) {
                    const uint64_t a = am[(size_t)r * n + kk];
                    if (!a) continue;
                    for (int32_t c = 0; c < n; ++c) {
                        const uint64_t v = axm[(size_t)r * n + c] + mul_mod(a, xm[(size_t)kk * n + c], p);
                        axm[(size_t)r * n + c] = v >= p ? v - p : v;
                    }
                }
            for (int32_t r = 0; r < n && !(rep.status & 1); ++r)
                for (int32_t c = 0; c < n; ++c) {
                    uint64_t s = 0;
                    for (int32_t kk = 0; kk < n; ++kk) {
                        s += mul_mod(xm[(size_t)kk * n + r], axm[(size_t)kk * n + c], p);
                        if (s >= p) s -= p;
                    }
                    if (s != residue(A[(size_t)r * n + c], p)) {
                        rep.status |= 1;
                        break;
                    }
                }
        }
        if (t < k_det) {
            const uint64_t d = det_mod_p(xm, n, p);
            all_pos &= d == 1;
            all_neg &= d == p - 1;
        }
        rep.primes_used = t + 1;
    }
    if (!all_pos && !all_neg) rep.status |= 4;
    rep.det_sign = all_pos ? 1 : (all_neg ? -1 : 0);
    *report = rep;
    return XTX_OK;
}
