// xtax_cli.cpp -- `xtax`, the command-line drop-in over the gfx950 engine.
//
// Same input modes, flags, defaults, output files, log lines and exit codes as
// the reference's xtax front end (src/main.cpp:557-874), with the annealer
// call (`canneal::run_annealer`, main.cpp:716-718 and :844-846) replaced by
// xtx_solve_plain. `-t/--threads` sets the replica count (one tempering rung
// per replica, as one worker per thread in the reference); the GPU default is
// 64. Flags that only tune CPU threads (--use-hyperthreads, --no-pin) are
// accepted and have no effect. --deflate / --deflate-blocks run the
// reference's deflation outer loop over the engine; --rcm / --centroid reorder
// the published best. GPU-only additions: --device, --max-moves (proposals
// per replica), --max-epochs (epochs of exchange-interval moves; a run with
// it and no time limit is deterministic), --lookahead (proposals of a replica
// evaluated at once; speed only).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <csignal>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <functional>
#include <iostream>
#include <numeric>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include <unistd.h>

#include "cli_common.h"
#include "xtx.h"

namespace {

using namespace cli;

int64_t score_of(const Mat& A) {
    int64_t full = 0, diag = 0;
    for (auto v : A.v) full += v < 0 ? -v : v;
    for (int i = 0; i < A.rows; ++i) diag += std::llabs(A.at(i, i));
    return 2 * full - diag;
}

int64_t offnz_of(const Mat& A) {
    int64_t c = 0;
    for (int i = 0; i < A.rows; ++i)
        for (int j = i + 1; j < A.rows; ++j) c += A.at(i, j) != 0;
    return c;
}

constexpr int64_t kDeflateOffdiagCap = 1 << 13;  // main.cpp:30

// The deflation outer loop (solve_with_deflation, main.cpp:460-555): anneal
// the active sub-problem for a slice (through the engine with the active
// list, from the current transform), then lock pivots and drop them from the
// active set -- strict: a +-1 pivot whose column is at most 2^13 in
// magnitude is cleared exactly (xtx_deflate_pivot); relaxed: a nonzero pivot
// whose column is already zero off the diagonal -- until the active set has
// fewer than two indices, the matrix is diagonal, time runs out or the user
// interrupts. Starts from the identity transform; writes the CSVs on each
// report. Returns false (message printed) on an engine error.
bool solve_with_deflation(xtx_ctx* ctx, const Mat& A, const xtx_params& params, bool relaxed,
                          double slice, bool verbose, Mat& M, Mat& Xt) {
    const int n = A.rows;
    M = A;
    Xt = square(n);
    for (int i = 0; i < n; ++i) Xt.at(i, i) = 1;
    std::vector<int32_t> active((size_t)n);
    std::iota(active.begin(), active.end(), 0);
    const auto t0 = std::chrono::steady_clock::now();
    auto elapsed = [&]() {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    };
    std::cout << "[deflate] start mode=" << (relaxed ? "blocks" : "strict")
              << " score=" << score_of(M) << " n=" << n << " slice=" << slice
              << "s threads=" << params.replicas << "\n";
    xtx_params sp = params;
    sp.on_best = verbose ? print_new_best_l1 : nullptr;  // slices are quiet (main.cpp:482)
    const double report_every = std::max(params.save_interval, slice);
    double last_report = -1e18;
    Mat X = square(n), bA = square(n), bX = square(n);
    while (active.size() >= 2) {
        if (g_interrupted) break;
        const double el = elapsed();
        if (params.max_seconds > 0.0) {
            if (el >= params.max_seconds) break;
            sp.max_seconds = std::min(slice, params.max_seconds - el);
        } else {
            sp.max_seconds = slice;
        }
        X = transposed(Xt);
        xtx_stats st{};
        if (xtx_solve_plain(ctx, A.v.data(), X.v.data(), n, &sp, active.data(),
                            (int32_t)active.size(), bA.v.data(), bX.v.data(), &st) != XTX_OK) {
            std::cerr << "Error: " << xtx_last_error(ctx) << "\n";
            return false;
        }
        if (verbose)
            std::cout << "[deflate] slice moves=" << st.moves << " best=" << st.best_score
                      << " seconds=" << st.seconds << "\n";
        M = bA;
        Xt = transposed(bX);
        int locked = 0;
        const std::vector<int32_t> snapshot = active;
        for (int i : snapshot) {
            const int64_t a_ii = M.at(i, i);
            if (relaxed) {
                if (a_ii == 0) continue;
                bool orthogonal = true;
                for (int j : active)
                    if (j != i && M.at(j, i) != 0) {
                        orthogonal = false;
                        break;
                    }
                if (!orthogonal) continue;
            } else {
                if (std::llabs(a_ii) != 1) continue;
                int64_t max_off = 0;
                for (int j : active)
                    if (j != i) max_off = std::max<int64_t>(max_off, std::llabs(M.at(j, i)));
                if (max_off > kDeflateOffdiagCap) continue;
                int32_t cleared = 0;
                if (xtx_deflate_pivot(M.v.data(), Xt.v.data(), n, i, active.data(),
                                      (int32_t)active.size(), &cleared) != XTX_OK ||
                    !cleared)
                    continue;
            }
            active.erase(std::find(active.begin(), active.end(), i));
            ++locked;
        }
        const bool diagonal = offnz_of(M) == 0;
        const double now = elapsed();
        if (locked > 0 || diagonal || now - last_report >= report_every) {
            std::cout << "[deflate] t=" << now << "s score=" << score_of(M)
                      << " active=" << active.size() << " offdiag=" << offnz_of(M);
            if (locked > 0) std::cout << " (locked " << locked << ")";
            std::cout << "\n";
            write_csv(transposed(Xt), "best_X.csv");
            write_csv(M, "best_A.csv");
            last_report = now;
        }
        if (diagonal) break;
    }
    std::cout << "[deflate] done active=" << active.size() << " score=" << score_of(M)
              << " seconds=" << elapsed() << "\n";
    return true;
}

}  // namespace

int main(int argc, char** argv) {
    std::signal(SIGINT, on_sigint);
    xtx_params p;
    xtx_params_default(&p);
    std::string a_csv, l_csv, x_csv, gram_rows, lattice_dims;
    uint64_t modulus = 0;
    int device = 0;
    bool rcm = false, centroid = false, deflate = false, deflate_blocks = false, verbose = false;
    double deflate_slice = 0.5;
    std::vector<Opt> opts = {
        {{"-A"}, false, set_string(a_csv)},
        {{"-L", "--lattice"}, false, set_string(l_csv)},
        {{"-X"}, false, set_string(x_csv)},
        {{"--modulus"}, false, set_modulus(modulus)},
        {{"--gram-rows"}, false, set_string(gram_rows)},
        {{"--lattice-dims"}, false, set_string(lattice_dims)},
        {{"--rcm"}, true, [&](const std::string&) { rcm = true; }},
        {{"--centroid"}, true, [&](const std::string&) { centroid = true; }},
        {{"--deflate"}, true, [&](const std::string&) { deflate = true; }},
        {{"--deflate-blocks"}, true, [&](const std::string&) { deflate_blocks = true; }},
        {{"--deflate-slice"}, false, set_double(deflate_slice)},
        {{"--verbose"}, true, [&](const std::string&) { verbose = true; }},
    };
    for (auto& o : engine_options(p, device)) opts.push_back(std::move(o));
    const int rc_args =
        parse_args(argc, argv, opts, "xtax (gfx950 engine): see the reference README for the options\n");
    if (rc_args >= 0) return rc_args;
    if (p.replicas < 1) p.replicas = 1;
    seed_from_device(p);

    const bool have_A = !a_csv.empty(), have_L = !l_csv.empty();
    if (have_A == have_L) {
        std::cerr << "Error: provide exactly one of -A or -L\n";
        return 1;
    }
    if (deflate && deflate_blocks) {
        std::cerr << "Error: use at most one of --deflate or --deflate-blocks\n";
        return 1;
    }
    if (!gram_rows.empty() && (deflate || deflate_blocks)) {
        std::cerr << "Error: --gram-rows cannot be combined with --deflate or --deflate-blocks\n";
        return 1;
    }
    if (!gram_rows.empty() && (rcm || centroid)) {
        std::cerr << "Error: --gram-rows cannot be combined with --rcm or --centroid (reordering "
                     "would remap the row indices)\n";
        return 1;
    }
    if (!lattice_dims.empty() && !have_L) {
        std::cerr << "Error: --lattice-dims requires -L (it selects lattice columns for the Gram "
                     "calculation)\n";
        return 1;
    }

    Mat A, X, Lat;
    std::vector<int> dims, rows;
    try {
        if (have_A) {
            A = read_csv(a_csv, true);
            if (!x_csv.empty()) {
                X = read_csv(x_csv, true);
                if (X.rows != A.rows) throw std::runtime_error("X dimensions do not match A");
            }
            if (modulus > 0) {
                reduce_mod(A.v, modulus);
                std::cout << "[modulus] " << modulus_note(modulus) << "\n";
            }
        } else {
            Lat = read_csv(l_csv, false);
            if (Lat.rows < 2) throw std::runtime_error("lattice must have at least 2 basis vectors");
            if (Lat.rows > Lat.cols)
                std::cerr << "Warning: lattice has more vectors (" << Lat.rows
                          << ") than dimensions (" << Lat.cols
                          << "); the basis is linearly dependent and the Gram matrix is singular\n";
            if (modulus > 0) {
                reduce_mod(Lat.v, modulus);
                std::cout << "[modulus] " << modulus_note(modulus) << "\n";
            }
            if (!lattice_dims.empty()) {
                dims = parse_index_spec(lattice_dims, Lat.cols, "--lattice-dims");
                if (dims.empty()) throw std::runtime_error("--lattice-dims: no dimensions given");
            }
            A = gram(Lat, dims);
            std::cout << "[lattice] m=" << Lat.rows << " d=" << Lat.cols
                      << " score=" << score_of(A) << "\n";
            if (!dims.empty()) {
                // the restricted metric's report (main.cpp:797-809)
                std::cout << "[lattice-dims] Gram built from " << dims.size() << " of "
                          << Lat.cols << " lattice dimensions\n";
                if ((int)dims.size() < Lat.rows)
                    std::cout << "[lattice-dims] note: fewer selected dimensions ("
                              << dims.size() << ") than rows (" << Lat.rows
                              << "), so the restricted Gram is singular\n";
                int zero_diag = 0;
                for (int i = 0; i < Lat.rows; ++i) zero_diag += A.at(i, i) == 0;
                if (zero_diag > 0)
                    std::cout << "[lattice-dims] warning: " << zero_diag
                              << " row(s) are zero on the selected dimensions; the restricted "
                                 "Gram is singular\n";
            }
        }
        if (!gram_rows.empty()) {
            rows = parse_index_spec(gram_rows, A.rows, "--gram-rows");
            if (rows.size() < 2)
                throw std::runtime_error(
                    "--gram-rows: need at least 2 rows (a move uses a pivot and a target)");
            std::cout << "[gram-rows] restricting moves to " << rows.size() << " of " << A.rows
                      << " rows\n";
        }
    } catch (const std::exception& e) {
        std::cerr << "Error: " << e.what() << "\n";
        return 1;
    }

    if (deflate) {
        std::cout << (have_A ? "[deflate] verifying the matrix is unimodular (det = +/-1)...\n"
                             : "[deflate] verifying the Gram matrix is unimodular (det = +/-1)...\n");
        int32_t uni = 0;
        if (xtx_is_unimodular(A.v.data(), A.rows, &uni) != XTX_OK || !uni) {
            if (have_A)
                std::cerr << "Error: --deflate requires a unimodular matrix (det = +/-1); this one "
                             "is not.\n";
            else
                std::cerr << "Error: --deflate requires a unimodular matrix (det = +/-1); this Gram "
                             "is not\n       (a lattice Gram is unimodular only when the lattice is "
                             "unimodular).\n";
            std::cerr << "       Use --deflate-blocks for a relaxed (orthogonal-summand) deflation,\n"
                      << "       or drop --deflate to run the plain annealer.\n";
            return 1;
        }
    }

    if (xtx_open(device, &g_ctx) != XTX_OK) {
        std::cerr << "Error: " << xtx_last_error(nullptr) << "\n";
        return 1;
    }
    const int n = A.rows;
    Mat bA = square(n), bX = square(n);
    p.verify = 1;
    int64_t final_score = 0;
    if (deflate || deflate_blocks) {
        Mat Xt;
        if (!solve_with_deflation(g_ctx, A, p, deflate_blocks, deflate_slice, verbose, bA, Xt)) {
            xtx_close(g_ctx);
            return 1;
        }
        bX = transposed(Xt);
        final_score = score_of(bA);
    } else {
        std::cout << "[t=0s] start score=" << score_of(A) << " n=" << n << " threads="
                  << p.replicas << ((p.tempering && p.replicas >= 2) ? " (parallel tempering)" : "")
                  << "\n"
                  << std::flush;  // the run has started, also when stdout is a pipe
        xtx_stats st{};
        // mid-run saves of the best (publish_files, main.cpp:413-416), reordered
        // like every published best under --rcm/--centroid
        const int pub_flags = (rcm ? XTX_REORDER_RCM : 0) | (centroid ? XTX_REORDER_CENTROID : 0);
        BestReports reports;
        reports.ctx = g_ctx;
        reports.save_interval = p.save_interval;
        reports.save = [&]() {
            Mat sA = square(n), sX = square(n);
            if (xtx_best_so_far(g_ctx, sA.v.data(), nullptr, sX.v.data()) != XTX_OK) return;
            if (pub_flags) xtx_reorder(sA.v.data(), sX.v.data(), n, pub_flags, nullptr);
            write_csv(sX, "best_X.csv");
            write_csv(sA, "best_A.csv");
        };
        p.on_best = print_new_best_l1;
        p.on_best_user = &reports;
        const int rc = xtx_solve_plain(
            g_ctx, A.v.data(), (have_A && !x_csv.empty()) ? X.v.data() : nullptr, n, &p,
            rows.empty() ? nullptr : rows.data(), (int32_t)rows.size(), bA.v.data(), bX.v.data(),
            &st);
        if (rc != XTX_OK) {
            std::cerr << "Error: " << xtx_last_error(g_ctx) << "\n";
            xtx_close(g_ctx);
            return 1;
        }
        if (st.solved)
            std::cout << "[t=" << st.seconds << "s] matrix is diagonal, stopping (thread "
                      << st.solve_replica << ")\n";
        std::cout << "done moves=" << st.moves << " seconds=" << st.seconds << " moves_per_sec="
                  << (st.seconds > 0 ? (double)st.moves / st.seconds : 0.0) << "\n";
        // --rcm / --centroid reorder every published best (main.cpp:407-411,
        // congruence_anneal.hpp:319); the start state is never published, so
        // a run that found no better state returns it unordered
        const int flags = (rcm ? XTX_REORDER_RCM : 0) | (centroid ? XTX_REORDER_CENTROID : 0);
        const int64_t start_score = st.best_score;
        if (flags && start_score < score_of(have_A && !x_csv.empty() ? congruence(A, X) : A))
            xtx_reorder(bA.v.data(), bX.v.data(), n, flags, nullptr);
        final_score = st.best_score;
    }
    std::cout << "Final best score: " << final_score << "\n";
    try {
        if (have_A) {
            if (n <= 20) {
                std::cout << "A:\n";
                print_mat(bA);
                std::cout << "----\nX:\n";
                print_mat(bX);
            }
            if (g_interrupted)
                std::cout << "[interrupted] writing best_X.csv and best_A.csv before exit\n";
            write_csv(bX, "best_X.csv");
            write_csv(bA, "best_A.csv");
        } else {
            // final basis X^T L and the Gram consistency check (main.cpp:850-857)
            const Mat Lf = xt_times(bX, Lat);
            const Mat g2 = gram(Lf, dims);
            if (g2.v != bA.v)
                std::cout << "[warn] Gram(X^T L) != annealed Gram (transform mismatch, possible "
                             "int64 overflow)\n";
            if (g_interrupted)
                std::cout << "[interrupted] writing final_L.csv, best_A.csv and best_X.csv before "
                             "exit\n";
            write_csv(Lf, "final_L.csv");
            write_csv(bA, "best_A.csv");
            write_csv(bX, "best_X.csv");
            if (Lat.rows <= 20 && Lat.cols <= 20) {
                std::cout << "Final L:\n";
                print_mat(Lf);
                std::cout << "----\nGram:\n";
                print_mat(bA);
            }
        }
    } catch (const std::exception& e) {
        std::cerr << "Error: " << e.what() << "\n";
        xtx_close(g_ctx);
        return 1;
    }
    xtx_close(g_ctx);
    return 0;
}
