// host_check.cpp -- the host-only code of the engine (xtax_amd/csrc/verify.cpp,
// xtax_amd/csrc/publish.cpp) and the CPU oracle (oracle/oracle.cpp,
// oracle/oracle_dual.cpp) exercised on seeded random inputs, for a build with
// AddressSanitizer and UndefinedBehaviorSanitizer (xtax_amd/Makefile
// `sanitize`, run by tests/test_sanitized_host.py). Every out-of-bounds access,
// overflow or use after free in those files aborts the run; the checks below
// also compare the product's host functions with the oracle's.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../oracle/oracle.h"
#include "xtx.h"

namespace {

int failures = 0;

void expect(bool ok, const char* what, int seed) {
    if (!ok) {
        std::fprintf(stderr, "host_check: %s (case %d)\n", what, seed);
        ++failures;
    }
}

// A random symmetric matrix of order n with entries in [-scale, scale].
std::vector<int64_t> symmetric(std::mt19937_64& rng, int n, int64_t scale) {
    std::uniform_int_distribution<int64_t> d(-scale, scale);
    std::vector<int64_t> A((size_t)n * n);
    for (int i = 0; i < n; ++i)
        for (int j = i; j < n; ++j) A[(size_t)i * n + j] = A[(size_t)j * n + i] = d(rng);
    return A;
}

// A unimodular transform: a product of random +-1 shears.
std::vector<int64_t> unimodular(std::mt19937_64& rng, int n, int shears) {
    std::vector<int64_t> X((size_t)n * n, 0);
    for (int i = 0; i < n; ++i) X[(size_t)i * n + i] = 1;
    std::uniform_int_distribution<int> pick(0, n - 1);
    for (int t = 0; t < shears && n > 1; ++t) {
        const int a = pick(rng);
        int b = pick(rng);
        if (b == a) b = (a + 1) % n;
        const int64_t s = (rng() & 1) ? 1 : -1;
        for (int r = 0; r < n; ++r) X[(size_t)r * n + b] += s * X[(size_t)r * n + a];
    }
    return X;
}

std::vector<int64_t> congruence(const std::vector<int64_t>& A, const std::vector<int64_t>& X, int n) {
    std::vector<int64_t> out((size_t)n * n);
    orc_congruence_of(A.data(), X.data(), n, out.data());
    return out;
}

}  // namespace

int main() {
    std::mt19937_64 rng(12345);
    for (int c = 0; c < 400; ++c) {
        const int n = 1 + (int)(rng() % 24);
        const int64_t scale = (c % 4 == 0) ? ((int64_t)1 << 40) : 30;
        std::vector<int64_t> A0 = symmetric(rng, n, scale);
        std::vector<int64_t> X = unimodular(rng, n, 2 * n);
        std::vector<int64_t> A = congruence(A0, X, n);

        // exact verification: the product's and the oracle's agree
        xtx_verify_report vr{};
        orc_verify_report orr{};
        xtx_verify(A0.data(), X.data(), A.data(), n, &vr);
        orc_verify(A0.data(), X.data(), A.data(), n, &orr);
        expect((vr.status & ~2) == 0 && (orr.status & ~2) == 0, "verify of a true congruence", c);
        expect((vr.status & 2) == (orr.status & 2), "diagonal bit", c);
        if (n > 1) {
            std::vector<int64_t> B = A;
            B[1] += 1;  // breaks X^T A0 X == A
            xtx_verify(A0.data(), X.data(), B.data(), n, &vr);
            expect((vr.status & 1) != 0, "verify rejects a wrong A", c);
        }

        // publish-time reordering keeps score and nonzero count
        for (int flags = 1; flags <= 3; ++flags) {
            std::vector<int64_t> R = A, RX = X;
            std::vector<int32_t> perm((size_t)n);
            expect(xtx_reorder(R.data(), RX.data(), n, flags, perm.data()) == XTX_OK, "reorder",
                   c);
            expect(orc_score(R.data(), n) == orc_score(A.data(), n), "reorder keeps the score", c);
            expect(orc_offdiag_nonzero(R.data(), n) == orc_offdiag_nonzero(A.data(), n),
                   "reorder keeps the nonzero count", c);
            std::vector<int64_t> RA = congruence(A0, RX, n);
            expect(RA == R, "reordered X still maps A0 to A", c);
        }

        // unimodularity and strict-deflation pivot clearing
        int32_t uni = -1;
        xtx_is_unimodular(A.data(), n, &uni);
        expect(uni == 0 || uni == 1, "is_unimodular result", c);
        std::vector<int64_t> M = A, Xt((size_t)n * n);
        for (int r = 0; r < n; ++r)
            for (int k = 0; k < n; ++k) Xt[(size_t)k * n + r] = X[(size_t)r * n + k];
        for (int i = 0; i < n; ++i) {
            const int64_t d = M[(size_t)i * n + i];
            if (d != 1 && d != -1) continue;
            std::vector<int32_t> act((size_t)n);
            for (int k = 0; k < n; ++k) act[(size_t)k] = k;
            int32_t cleared = -1;
            xtx_deflate_pivot(M.data(), Xt.data(), n, i, act.data(), n, &cleared);
            std::vector<int64_t> Xn((size_t)n * n);
            for (int r = 0; r < n; ++r)
                for (int k = 0; k < n; ++k) Xn[(size_t)r * n + k] = Xt[(size_t)k * n + r];
            expect(congruence(A0, Xn, n) == M, "deflation keeps X^T A0 X = M", c);
            break;
        }

        // oracle primitives on random moves
        if (n >= 2) {
            std::vector<int64_t> Xt0((size_t)n * n, 0), row((size_t)n), xrow((size_t)n),
                mass((size_t)n);
            for (int i = 0; i < n; ++i) Xt0[(size_t)i * n + i] = 1;
            orc_compute_row_mass(A.data(), n, mass.data());
            for (int t = 0; t < 20; ++t) {
                const int i = (int)(rng() % n);
                const int j = (i + 1 + (int)(rng() % (n - 1))) % n;
                const int64_t s = orc_best_shear_l1(A.data(), n, i, j);
                int64_t d = 0, dnz = 0;
                const int64_t sv = s ? s : 1;
                if (orc_try_add(A.data(), Xt0.data(), n, i, j, sv, row.data(), xrow.data(), &d,
                                &dnz)) {
                    const int64_t before = orc_score(A.data(), n);
                    orc_commit_add(A.data(), Xt0.data(), n, j, row.data(), xrow.data(),
                                   mass.data());
                    expect(orc_score(A.data(), n) == before + d, "delta of a committed move", c);
                }
            }
        }
    }

    // the trajectory model in both draw modes, small runs
    for (int c = 0; c < 12; ++c) {
        const int n = 3 + c;
        std::vector<int64_t> A0 = symmetric(rng, n, 20), I((size_t)n * n, 0);
        for (int i = 0; i < n; ++i) I[(size_t)i * n + i] = 1;
        orc_params p{};
        p.t_init = 30;
        p.t_min = 0.5;
        p.cooling = 0.999;
        p.reheat = 1.0;
        p.reseed_factor = 1.25;
        p.greedy_fraction = 0.5;
        p.target_fraction = 0.5;
        p.target_accept = 0.44;
        p.worker_diversity = 0.4;
        p.sweep_fraction = 0.3;
        p.moves_per_cool = 50;
        p.stuck_threshold = 100;
        p.target_samples = 1 + c % 16;
        p.tempering = c % 2;
        p.adaptive_cooling = c % 3 == 0;
        p.exchange_interval = 200;
        p.replicas = 4;
        p.seed = (uint64_t)c;
        p.max_moves = 1500;
        orc_model_result mr{};
        const int64_t cap = 4000;
        std::vector<orc_log_entry> logs((size_t)p.replicas * cap);
        std::vector<int64_t> counts((size_t)p.replicas);
        orc_model_run(A0.data(), I.data(), n, nullptr, 0, &p, 10, nullptr, nullptr, nullptr,
                      nullptr, nullptr, logs.data(), counts.data(), cap, &mr);
        orc_ref_params rp{};
        rp.base = p;
        rp.worker_seed = orc_ref_worker_seed(7 + c, c % 4);
        rp.thread_id = c % 4;
        rp.rung_T = 5.0;
        std::vector<int64_t> Aout((size_t)n * n), Xout((size_t)n * n), gA((size_t)n * n),
            gX((size_t)n * n);
        std::vector<orc_log_entry> rlog(cap);
        int64_t rc = 0;
        orc_ref_result rr{};
        orc_ref_worker_run(A0.data(), I.data(), n, nullptr, 0, &rp, Aout.data(), Xout.data(),
                           gA.data(), gX.data(), rlog.data(), &rc, cap, &rr);
        expect(rr.moves > 0, "reference-draw run made moves", c);
    }
    // the dual trajectory model: the engine's mode, and the reference's mode
    // (cooling, or a fixed tempering rung), small runs with re-inversions and
    // residual checks
    for (int c = 0; c < 8; ++c) {
        const int n = 3 + c;
        const size_t nn = (size_t)n * n;
        std::vector<int64_t> P0 = symmetric(rng, n, 9), I(nn, 0);
        for (int i = 0; i < n; ++i) {
            I[(size_t)i * n + i] = 1;
            P0[(size_t)i * n + i] += 50 * n;  // diagonally dominant: full rank
        }
        std::vector<double> Q0(nn);
        if (!orc_invert(P0.data(), n, Q0.data())) {
            expect(false, "invertible dual start", c);
            continue;
        }
        orc_params p{};
        p.t_init = 0;
        p.t_min = 0.5;
        p.cooling = 0.999;
        p.reheat = 1.0;
        p.reseed_factor = 1.25;
        p.greedy_fraction = 0.5;
        p.target_fraction = 0.5;
        p.target_accept = 0.44;
        p.worker_diversity = 0.4;
        p.sweep_fraction = 0.3;
        p.moves_per_cool = 50;
        p.stuck_threshold = 100;
        p.target_samples = 1 + c % 16;
        p.tempering = c % 2;
        p.adaptive_cooling = c % 3 == 0;
        p.exchange_interval = 200;
        p.replicas = 4;
        p.seed = (uint64_t)c;
        p.max_moves = 1500;
        orc_dual_params dp{};
        dp.c = 0.5;
        dp.dual_tol = 1e-12;
        dp.dual_refresh = 700;
        dp.dual_check = 300;
        const int64_t cap = 4000;
        std::vector<orc_log_entry> logs((size_t)p.replicas * cap);
        std::vector<int64_t> counts((size_t)p.replicas);
        orc_dual_result dr{};
        orc_dual_model_run(P0.data(), Q0.data(), I.data(), n, nullptr, 0, &p, &dp, 8, nullptr,
                           nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                           logs.data(), counts.data(), cap, &dr);
        expect(dr.total_moves > 0, "dual model made moves", c);
        orc_ref_params rp{};
        rp.base = p;
        rp.worker_seed = orc_ref_worker_seed(11 + c, c % 4);
        rp.thread_id = c % 4;
        rp.rung_T = 5.0;
        std::vector<int64_t> Pf(nn), Xf(nn), gP(nn), gX(nn);
        std::vector<double> Qf(nn), gQ(nn);
        std::vector<orc_log_entry> rlog(cap);
        int64_t rc = 0;
        orc_dual_ref_result rr{};
        const int32_t ok = orc_dual_ref_run(P0.data(), I.data(), n, nullptr, 0, &rp, &dp, 1.0,
                                            Pf.data(), Qf.data(), Xf.data(), gP.data(),
                                            gQ.data(), gX.data(), rlog.data(), &rc, cap, &rr);
        expect(ok >= 0 && rr.moves > 0, "dual reference-mode run made moves", c);
    }
    std::printf("host_check: %s\n", failures ? "FAILED" : "ok");
    return failures ? 1 : 0;
}
