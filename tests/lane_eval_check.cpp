// lane_eval_check.cpp -- the one-lane proposal arithmetic of
// xtax_amd/csrc/lane_eval.h (host build) against the oracle's restatement of
// the reference's best_shear_l1 and try_add (oracle/oracle.cpp, pinned by the
// golden vectors), on seeded random matrices with entries below 2^16 (the
// rows the lane arithmetic takes): dense and sparse rows, entries up to
// 2^16 - 1, small entries under a heavy diagonal or beside a small one, every
// order to 48 padded to its kernel order (columns past n are zero, as in the
// kernel's LDS) and at 64, every ordered pair (i, j) of each
// matrix, the best shear and the score change of it, of +-1, of the cap and of
// a random shear.
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

#include "../oracle/oracle.h"
#include "lane_eval.h"

using namespace xtx::lane;

namespace {
int failures = 0;
long checks = 0;

void expect(bool ok, const char* what, int n, int i, int j, int64_t s) {
    ++checks;
    if (!ok && failures < 20)
        std::fprintf(stderr, "lane_eval_check: %s (n=%d i=%d j=%d s=%lld)\n", what, n, i, j,
                     (long long)s);
    failures += !ok;
}

template <int NMAX>
void check_pair(const std::vector<int64_t>& A, int n, int i, int j, std::mt19937_64& rng) {
    const int64_t* ri = A.data() + (size_t)i * n;
    const int64_t* rj = A.data() + (size_t)j * n;
    Pair<NMAX> r;
    for (int k = 0; k < NMAX; ++k) {
        const int32_t ai = k < n ? (int32_t)ri[k] : 0, aj = k < n ? (int32_t)rj[k] : 0;
        to_pq(ai, aj, r.p[k], r.qb[k]);
    }
    to_pq((int32_t)ri[j], (int32_t)rj[j], r.pj, r.qbj);
    r.a_ii = (int32_t)ri[i];
    r.a_ji = (int32_t)rj[i];
    r.a_jj = (int32_t)rj[j];
    int32_t tot = 0, sumq = 0;
    totals(r, tot, sumq);
    int64_t want_tot = 0, want_sumq = 0;
    for (int k = 0; k < n; ++k)
        if (k != j) {
            want_tot += ri[k] < 0 ? -ri[k] : ri[k];
            want_sumq += rj[k] < 0 ? -rj[k] : rj[k];
        }
    expect(tot == want_tot && sumq == want_sumq, "row totals", n, i, j, 0);

    const int64_t want = orc_best_shear_l1(A.data(), n, i, j);
    int32_t off = 0;
    const int32_t got = best_shear(r, tot, sumq, &off);
    expect(got == want, "best shear", n, i, j, want);
    // score changes: the best shear (through its off sum), +-1, the cap, a random s
    std::uniform_int_distribution<int> rs(-kCap, kCap);
    const int32_t ss[6] = {(int32_t)want, 1, -1, kCap, -kCap, rs(rng)};
    for (int t = 0; t < 6; ++t) {
        const int32_t s = ss[t];
        if (s == 0) continue;
        std::vector<int64_t> nrow((size_t)n);
        int64_t d = 0, dnz = 0;
        const bool ok = orc_try_add(A.data(), nullptr, n, i, j, s, nrow.data(), nullptr, &d, &dnz);
        expect(ok, "small rows stay in bound", n, i, j, s);
        const int32_t o = t == 0 ? off : abs_sum(r, s) - sumq;
        expect(delta_of(r, s, o) == d, "score change", n, i, j, s);
    }
}

template <int NMAX>
void run_order(int n, std::mt19937_64& rng, int mats) {
    for (int m = 0; m < mats; ++m) {
        const int fam = mats > 6 ? 6 : m % 6;
        int64_t scale = 20;
        double density = 1.0;
        if (fam == 1) scale = 3000;
        if (fam == 2) scale = kSmallEntry - 1;
        if (fam == 3) scale = 1;
        if (fam == 4) density = 0.2;
        if (fam == 5) scale = 3;
        if (fam == 6) scale = 6;
        std::uniform_int_distribution<int64_t> d(-scale, scale);
        std::uniform_int_distribution<int64_t> dd(fam == 6 ? -3 : -200, fam == 6 ? 3 : 200);
        std::uniform_real_distribution<double> u(0.0, 1.0);
        std::vector<int64_t> A((size_t)n * n, 0);
        for (int i = 0; i < n; ++i)
            for (int j = i; j < n; ++j) {
                int64_t v = u(rng) < density ? d(rng) : 0;
                // small off-diagonal entries under a heavy diagonal (5), or
                // entries to 6 under a diagonal to 3 (6): integer medians far
                // from -round(a_ji / a_ii), where the floor of the median is a
                // candidate of its own and wins now and then
                if (fam >= 5 && i == j) v = dd(rng);
                A[(size_t)i * n + j] = A[(size_t)j * n + i] = v;
            }
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j)
                if (i != j) check_pair<NMAX>(A, n, i, j, rng);
    }
}
}  // namespace

int main() {
    std::mt19937_64 rng(2026);
    for (int n = 2; n <= 8; ++n) run_order<8>(n, rng, 6);
    for (int n = 9; n <= 16; ++n) run_order<16>(n, rng, 6);
    for (int n = 17; n <= 24; n += 2) run_order<24>(n, rng, 6);
    for (int n = 25; n <= 32; n += 3) run_order<32>(n, rng, 6);
    for (int n = 33; n <= 40; n += 3) run_order<40>(n, rng, 6);
    for (int n = 41; n <= 48; n += 3) run_order<48>(n, rng, 6);
    run_order<64>(64, rng, 6);
    for (int n = 3; n <= 8; ++n) run_order<8>(n, rng, 4000);
    std::printf("lane_eval_check: %ld checks, %d failures\n", checks, failures);
    return failures ? 1 : 0;
}
