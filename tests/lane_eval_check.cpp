// lane_eval_check.cpp -- the one-lane proposal arithmetic of
// xtax_amd/csrc/lane_eval.h (host build) against the oracle's restatement of
// the reference's best_shear_l1 and try_add (oracle/oracle.cpp, pinned by the
// golden vectors), on seeded random matrices: small entries (the int32 path),
// entries up to and beyond the 2^48 bound (the int64 path), sparse rows, every
// order to 64 and every ordered pair (i, j) of each matrix.
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
    const int64_t a_ii = ri[i], a_ji = rj[i], a_jj = rj[j];
    auto row = [&](int k, int64_t& ai, int64_t& aj) {
        ai = ri[k];
        aj = rj[k];
    };
    const int64_t want = orc_best_shear_l1(A.data(), n, i, j);
    int64_t boff = 0;
    const int64_t got_big = shear_big(row, n, j, a_ii, a_ji, a_jj, &boff);
    expect(got_big == want, "int64 best shear", n, i, j, want);

    bool small = true;
    for (int k = 0; k < n; ++k)
        small &= abs64(ri[k]) < kSmallEntry && abs64(rj[k]) < kSmallEntry;
    int32_t p[NMAX], q[NMAX];
    int32_t sumq = 0;
    for (int k = 0; k < NMAX; ++k) {
        p[k] = q[k] = 0;
        if (k < n && k != j && small) {
            p[k] = (int32_t)abs64(ri[k]);
            q[k] = (int32_t)(ri[k] >= 0 ? rj[k] : -rj[k]);
            sumq += abs32(q[k]);
        }
    }
    int32_t soff = 0;
    if (small) {
        const int64_t got_small = shear_small<NMAX>(p, q, sumq, (int32_t)a_ii, (int32_t)a_ji,
                                                    (int32_t)a_jj, &soff);
        expect(got_small == want, "int32 best shear", n, i, j, want);
    }
    // evaluations: the best shear (with its known off sum), +-1, the cap, a random s
    std::uniform_int_distribution<int> rs(-kCap, kCap);
    const int64_t ss[6] = {want, 1, -1, kCap, -kCap, (int64_t)rs(rng)};
    for (int t = 0; t < 6; ++t) {
        const int64_t s = ss[t];
        if (s == 0) continue;
        std::vector<int64_t> nrow((size_t)n);
        int64_t d = 0, dnz = 0;
        const bool ok = orc_try_add(A.data(), nullptr, n, i, j, s, nrow.data(), nullptr, &d, &dnz);
        int64_t bd = 0, bo = 0;
        int32_t bz = 0;
        const bool bok = eval_big(row, n, j, s, a_ii, a_ji, a_jj, &bd, &bz, &bo);
        expect(bok == ok, "int64 feasibility", n, i, j, s);
        if (ok) expect(bd == d && bz == dnz, "int64 delta / nonzero change", n, i, j, s);
        if (t == 0 && ok && want != 0) expect(bo == boff, "int64 off sum of the best", n, i, j, s);
        if (small) {
            expect(ok, "small rows stay in bound", n, i, j, s);
            int64_t sd = 0;
            int32_t sz = 0, so = 0;
            const bool know = t == 0 && want != 0;
            eval_small<NMAX>(p, q, sumq, (int32_t)s, (int32_t)a_ii, (int32_t)a_ji, (int32_t)a_jj,
                             know, soff, &sd, &sz, &so);
            expect(sd == d && sz == dnz, "int32 delta / nonzero change", n, i, j, s);
        }
    }
}

template <int NMAX>
void run_order(int n, std::mt19937_64& rng, int mats) {
    for (int m = 0; m < mats; ++m) {
        const int fam = m % 6;
        int64_t scale = 20;
        double density = 1.0;
        if (fam == 1) scale = 3000;
        if (fam == 2) scale = kSmallEntry - 1;
        if (fam == 3) scale = (int64_t)1 << 30;
        if (fam == 4) scale = kLim + 1;  // some entries past the bound
        if (fam == 5) density = 0.2;
        std::uniform_int_distribution<int64_t> d(-scale, scale);
        std::uniform_real_distribution<double> u(0.0, 1.0);
        std::vector<int64_t> A((size_t)n * n, 0);
        for (int i = 0; i < n; ++i)
            for (int j = i; j < n; ++j) {
                int64_t v = u(rng) < density ? d(rng) : 0;
                if (fam == 4 && u(rng) < 0.5) v = (v % 1000);
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
    for (int n = 2; n <= 16; ++n) run_order<16>(n, rng, 6);
    for (int n = 17; n <= 32; n += 3) run_order<32>(n, rng, 6);
    for (int n = 33; n <= 48; n += 5) run_order<48>(n, rng, 6);
    run_order<64>(64, rng, 6);
    std::printf("lane_eval_check: %ld checks, %d failures\n", checks, failures);
    return failures ? 1 : 0;
}
