// publish.cpp -- host-side parts of xtax around the annealer: the
// publish-time symmetric reorderings (--rcm, --centroid), the unimodularity
// test that gates strict deflation, and the pivot clearing of strict
// deflation (--deflate). Plain C++ on the host; the annealer itself is the
// gfx950 engine.
#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <vector>

#include "engine_types.h"
#include "xtx.h"

namespace {

using Perm = std::vector<int>;

int64_t at(const int64_t* M, int n, int r, int c) { return M[(size_t)r * n + c]; }

// Reverse Cuthill-McKee order of the nonzero pattern of a symmetric matrix
// (main.cpp:249-286): i and j are adjacent when A_ij != 0 (i != j). Each
// breadth-first pass starts at the unvisited index of least degree (the first
// such index), visits each vertex's unvisited neighbours by increasing degree
// (std::sort, as the reference does), and the concatenated visiting order is
// reversed. perm[a] is the old index placed at a.
Perm rcm(const int64_t* A, int n) {
    std::vector<int> degree((size_t)n, 0);
    for (int r = 0; r < n; ++r)
        for (int c = 0; c < n; ++c) degree[(size_t)r] += (c != r && at(A, n, r, c) != 0);
    std::vector<char> seen((size_t)n, 0);
    Perm order;
    order.reserve((size_t)n);
    std::vector<int> next;
    while ((int)order.size() < n) {
        int start = -1;
        for (int v = 0; v < n; ++v)
            if (!seen[(size_t)v] && (start < 0 || degree[(size_t)v] < degree[(size_t)start]))
                start = v;
        seen[(size_t)start] = 1;
        size_t head = order.size();
        order.push_back(start);
        for (; head < order.size(); ++head) {
            const int v = order[head];
            next.clear();
            for (int c = 0; c < n; ++c)
                if (c != v && at(A, n, v, c) != 0 && !seen[(size_t)c]) next.push_back(c);
            std::sort(next.begin(), next.end(),
                      [&](int a, int b) { return degree[(size_t)a] < degree[(size_t)b]; });
            for (int c : next) {
                seen[(size_t)c] = 1;
                order.push_back(c);
            }
        }
    }
    std::reverse(order.begin(), order.end());
    return order;
}

// Iterated row centre of mass order (main.cpp:292-324): up to eight passes;
// each computes, for every position a of the current order, the
// |entry|-weighted mean column position of its row in that order (extended
// precision sums over the positions in order, the mean rounded to double; a
// zero row keeps key a), stably sorts the positions by key, and stops when
// nothing moved.
Perm centroid(const int64_t* A, int n) {
    Perm order((size_t)n);
    for (int a = 0; a < n; ++a) order[(size_t)a] = a;
    if (n < 2) return order;
    std::vector<double> key((size_t)n);
    std::vector<int> idx((size_t)n);
    Perm moved((size_t)n);
    for (int pass = 0; pass < 8; ++pass) {
        for (int a = 0; a < n; ++a) {
            const int r = order[(size_t)a];
            long double wsum = 0.0L, psum = 0.0L;
            for (int b = 0; b < n; ++b) {
                const long double w = (long double)std::llabs(at(A, n, r, order[(size_t)b]));
                psum += (long double)b * w;
                wsum += w;
            }
            key[(size_t)a] = wsum > 0.0L ? (double)(psum / wsum) : (double)a;
        }
        for (int a = 0; a < n; ++a) idx[(size_t)a] = a;
        std::stable_sort(idx.begin(), idx.end(),
                         [&](int p, int q) { return key[(size_t)p] < key[(size_t)q]; });
        bool any = false;
        for (int a = 0; a < n; ++a) {
            moved[(size_t)a] = order[(size_t)idx[(size_t)a]];
            any |= idx[(size_t)a] != a;
        }
        order.swap(moved);
        if (!any) break;
    }
    return order;
}

// A <- A[perm][perm], X <- X[:, perm] (new index a takes old index perm[a]).
void apply(int64_t* A, int64_t* X, int n, const Perm& perm) {
    std::vector<int64_t> T((size_t)n * n);
    for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b)
            T[(size_t)a * n + b] = at(A, n, perm[(size_t)a], perm[(size_t)b]);
    std::copy(T.begin(), T.end(), A);
    if (X) {
        for (int r = 0; r < n; ++r)
            for (int a = 0; a < n; ++a) T[(size_t)r * n + a] = at(X, n, r, perm[(size_t)a]);
        std::copy(T.begin(), T.end(), X);
    }
}

// det(A) mod p, p < 2^31 prime, by elimination over Z/p (mat_io.hpp:597-630).
int64_t det_mod(const int64_t* A, int n, int64_t p) {
    std::vector<int64_t> m((size_t)n * n);
    for (size_t k = 0; k < m.size(); ++k) m[k] = ((A[k] % p) + p) % p;
    auto inv = [p](int64_t v) {  // v^(p-2) mod p
        int64_t r = 1, b = v % p;
        for (int64_t e = p - 2; e > 0; e >>= 1) {
            if (e & 1) r = r * b % p;
            b = b * b % p;
        }
        return r;
    };
    int64_t det = 1 % p;
    for (int c = 0; c < n; ++c) {
        int pr = c;
        while (pr < n && m[(size_t)pr * n + c] == 0) ++pr;
        if (pr == n) return 0;
        if (pr != c) {
            for (int k = 0; k < n; ++k) std::swap(m[(size_t)pr * n + k], m[(size_t)c * n + k]);
            det = (p - det) % p;
        }
        const int64_t pv = m[(size_t)c * n + c];
        det = det * pv % p;
        const int64_t iv = inv(pv);
        for (int r = c + 1; r < n; ++r) {
            const int64_t f = m[(size_t)r * n + c] * iv % p;
            if (!f) continue;
            for (int k = c; k < n; ++k) {
                int64_t v = (m[(size_t)r * n + k] - f * m[(size_t)c * n + k] % p) % p;
                m[(size_t)r * n + k] = v < 0 ? v + p : v;
            }
        }
    }
    return det % p;
}

bool within(__int128 v) { return v <= (__int128)xtx::kLimit && v >= -(__int128)xtx::kLimit; }

}  // namespace

extern "C" {

int xtx_reorder(int64_t* A, int64_t* X, int32_t n, int32_t flags, int32_t* perm) {
    if (!A || n < 1 || n > 4096 || (flags & ~3)) return XTX_E_ARG;
    Perm total((size_t)n);
    for (int a = 0; a < n; ++a) total[(size_t)a] = a;
    auto compose = [&](const Perm& p) {
        Perm t((size_t)n);
        for (int a = 0; a < n; ++a) t[(size_t)a] = total[(size_t)p[(size_t)a]];
        total.swap(t);
    };
    if (flags & XTX_REORDER_RCM) {
        const Perm p = rcm(A, n);
        apply(A, X, n, p);
        compose(p);
    }
    if (flags & XTX_REORDER_CENTROID) {
        const Perm p = centroid(A, n);
        apply(A, X, n, p);
        compose(p);
    }
    if (perm)
        for (int a = 0; a < n; ++a) perm[a] = total[(size_t)a];
    return XTX_OK;
}

int xtx_is_unimodular(const int64_t* A, int32_t n, int32_t* result) {
    if (!A || n < 1 || !result) return XTX_E_ARG;
    const int64_t primes[2] = {1000000007LL, 998244353LL};
    bool plus = true, minus = true;
    for (int64_t p : primes) {
        const int64_t d = det_mod(A, n, p);
        plus &= d == 1;
        minus &= d == p - 1;
    }
    *result = (plus || minus) ? 1 : 0;
    return XTX_OK;
}

int xtx_deflate_pivot(int64_t* M, int64_t* Xt, int32_t n, int32_t i, const int32_t* active,
                      int32_t n_active, int32_t* cleared) {
    if (!M || !Xt || !active || !cleared || n < 1 || i < 0 || i >= n) return XTX_E_ARG;
    const int64_t a_ii = at(M, n, i, i);
    if (a_ii != 1 && a_ii != -1) return XTX_E_ARG;
    for (int q = 0; q < n_active; ++q)
        if (active[q] < 0 || active[q] >= n) return XTX_E_ARG;
    const std::vector<int64_t> M0(M, M + (size_t)n * n), X0(Xt, Xt + (size_t)n * n);
    std::vector<int64_t> row((size_t)n), xrow((size_t)n);
    *cleared = 1;
    for (int q = 0; q < n_active; ++q) {
        const int j = active[q];
        if (j == i) continue;
        const int64_t a_ji = at(M, n, j, i);
        if (a_ji == 0) continue;
        const __int128 s = a_ii == 1 ? -(__int128)a_ji : (__int128)a_ji;  // -a_ji / a_ii
        bool ok = true;
        for (int r = 0; r < n && ok; ++r) {
            const __int128 v = (__int128)at(Xt, n, j, r) + s * at(Xt, n, i, r);
            ok = within(v);
            xrow[(size_t)r] = (int64_t)v;
        }
        for (int k = 0; k < n && ok; ++k) {
            if (k == j) continue;
            const __int128 v = (__int128)at(M, n, j, k) + s * at(M, n, i, k);
            ok = within(v);
            row[(size_t)k] = (int64_t)v;
        }
        const __int128 jj = (__int128)at(M, n, j, j) + 2 * s * a_ji + s * s * a_ii;
        ok = ok && within(jj);
        if (!ok) {
            std::copy(M0.begin(), M0.end(), M);
            std::copy(X0.begin(), X0.end(), Xt);
            *cleared = 0;
            return XTX_OK;
        }
        row[(size_t)j] = (int64_t)jj;
        for (int k = 0; k < n; ++k) {
            M[(size_t)j * n + k] = row[(size_t)k];
            M[(size_t)k * n + j] = row[(size_t)k];
        }
        std::copy(xrow.begin(), xrow.end(), Xt + (size_t)j * n);
    }
    return XTX_OK;
}

}  // extern "C"
