// xdual_cli.cpp -- `xdual`, the command-line drop-in for the reference's
// simultaneous primal/dual annealer (src/xdual_main.cpp:426-652) over the
// gfx950 engine: the same input modes (-A with optional -X, -L), flags,
// defaults, output files (best_X.csv, best_P.csv, best_Q.csv; final_L.csv in L
// mode), log lines and exit codes, with run_xdual (:378-424) replaced by
// xtx_dual_solve. `-t/--threads` sets the replica count (GPU default 64);
// --use-hyperthreads and --no-pin are accepted and have no effect. The lines
// the reference prints before its run ([xdual] lambda= ...) come from the
// solve's statistics. GPU-only additions: --device, --max-moves, --max-epochs.
#include <cmath>
#include <cstdint>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include "cli_common.h"
#include "xtx.h"

using namespace cli;

namespace {

// write_matrixd_csv (mat_io.hpp:418-430): 17 significant digits, atomic replace
void write_csv_d(const std::vector<double>& Q, int n, const std::string& path) {
    const std::string tmp = path + ".tmp";
    {
        std::ofstream out(tmp);
        if (!out) throw std::runtime_error("Failed to open output file: " + tmp);
        out.precision(17);
        for (int r = 0; r < n; ++r) {
            for (int c = 0; c < n; ++c) {
                out << Q[(size_t)r * n + c];
                if (c + 1 < n) out << ',';
            }
            out << '\n';
        }
        out.flush();
        if (!out) throw std::runtime_error("Failed while writing: " + tmp);
    }
    std::error_code ec;
    std::filesystem::rename(tmp, path, ec);
    if (ec) throw std::runtime_error("Failed to replace " + path + ": " + ec.message());
}

void print_d(const std::vector<double>& Q, int n) {
    for (int r = 0; r < n; ++r) {
        for (int c = 0; c < n; ++c) {
            std::cout << Q[(size_t)r * n + c];
            if (c + 1 < n) std::cout << ',';
        }
        std::cout << '\n';
    }
}

int64_t offnz(const Mat& A) {
    int64_t c = 0;
    for (int i = 0; i < A.rows; ++i)
        for (int j = i + 1; j < A.cols; ++j) c += A.at(i, j) != 0;
    return c;
}

}  // namespace

int main(int argc, char** argv) {
    std::signal(SIGINT, on_sigint);
    xtx_params p;
    xtx_params_default(&p);
    xtx_dual_params dp;
    xtx_dual_params_default(&dp);
    std::string a_csv, l_csv, x_csv, gram_rows;
    uint64_t modulus = 0;
    int device = 0;
    std::vector<Opt> opts = {
        {{"-A"}, false, set_string(a_csv)},
        {{"-L", "--lattice"}, false, set_string(l_csv)},
        {{"-X"}, false, set_string(x_csv)},
        {{"--modulus"}, false, set_modulus(modulus)},
        {{"--gram-rows"}, false, set_string(gram_rows)},
        {{"--lambda"}, false, set_double(dp.lambda)},
        {{"--dual-refresh"}, false, set_int(dp.dual_refresh)},
        {{"--dual-check"}, false, set_int(dp.dual_check)},
        {{"--dual-tol"}, false, set_double(dp.dual_tol)},
        {{"--lambda-ramp"}, false, set_double(dp.lambda_ramp)},
    };
    for (auto& o : engine_options(p, device)) opts.push_back(std::move(o));
    const int rc_args = parse_args(
        argc, argv, opts, "xdual (gfx950 engine): see the reference README for the options\n");
    if (rc_args >= 0) return rc_args;
    if (p.replicas < 1) p.replicas = 1;
    seed_from_device(p);

    const bool have_A = !a_csv.empty(), have_L = !l_csv.empty();
    if (have_A == have_L) {
        std::cerr << "Error: provide exactly one of -A or -L\n";
        return 1;
    }
    Mat A, X, Lat;
    std::vector<int> rows;
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
            if (Lat.rows < 2) {
                std::cerr << "Error: lattice must have at least 2 basis vectors\n";
                return 1;
            }
            if (Lat.rows > Lat.cols)
                std::cerr << "Warning: lattice has more vectors (" << Lat.rows
                          << ") than dimensions (" << Lat.cols
                          << "); the basis is linearly dependent and the Gram matrix is singular\n";
            if (modulus > 0) {
                reduce_mod(Lat.v, modulus);
                std::cout << "[modulus] " << modulus_note(modulus) << "\n";
                for (int i = 0; i < Lat.rows; ++i) {
                    bool zero = true;
                    for (int k = 0; k < Lat.cols && zero; ++k) zero = Lat.at(i, k) == 0;
                    if (zero)
                        std::cerr << "Warning: row " << i << " became zero under the modulus; "
                                  << "the Gram matrix is singular and the dual inversion will "
                                     "fail\n";
                }
            }
            A = gram(Lat, {});
            std::cout << "[lattice] m=" << Lat.rows << " d=" << Lat.cols << "\n";
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

    if (xtx_open(device, &g_ctx) != XTX_OK) {
        std::cerr << "Error: " << xtx_last_error(nullptr) << "\n";
        return 1;
    }
    const int n = A.rows;
    Mat bP = square(n), bX = square(n);
    std::vector<double> bQ((size_t)n * n);
    p.verify = 1;
    // mid-run saves of the best (DualObjective::publish_files,
    // xdual_main.cpp:351-355)
    BestReports reports;
    reports.ctx = g_ctx;
    reports.save_interval = p.save_interval;
    reports.save = [&]() {
        Mat sP = square(n), sX = square(n);
        std::vector<double> sQ((size_t)n * n);
        if (xtx_best_so_far(g_ctx, sP.v.data(), sQ.data(), sX.v.data()) != XTX_OK) return;
        write_csv(sX, "best_X.csv");
        write_csv(sP, "best_P.csv");
        write_csv_d(sQ, n, "best_Q.csv");
    };
    p.on_best = print_new_best_dual;
    p.on_best_user = &reports;
    std::cout << "[xdual] inverting the Gram (double-precision dual P^-1)...\n";
    xtx_dual_stats st{};
    const int rc = xtx_dual_solve(g_ctx, A.v.data(), (have_A && !x_csv.empty()) ? X.v.data() : nullptr,
                                  n, &p, &dp, rows.empty() ? nullptr : rows.data(),
                                  (int32_t)rows.size(), bP.v.data(), bQ.data(), bX.v.data(), &st);
    if (rc != XTX_OK) {
        std::cerr << "Error: " << xtx_last_error(g_ctx) << "\n";
        xtx_close(g_ctx);
        return 1;
    }
    {
        const Mat Pstart = (have_A && !x_csv.empty()) ? congruence(A, X) : A;
        std::cout << "[xdual] lambda=" << dp.lambda << " c=" << st.c
                  << " score=" << (dp.lambda_ramp > 0.0 ? st.start_primal : st.start_score)
                  << " primal=" << offnz(Pstart) << " dual=" << std::sqrt(st.start_dual) << "\n";
    }
    if (st.solved)
        std::cout << "[t=" << st.seconds << "s] matrix is diagonal, stopping (thread "
                  << st.solve_replica << ")\n";
    std::cout << "done moves=" << st.moves << " seconds=" << st.seconds << " moves_per_sec="
              << (st.seconds > 0 ? (double)st.moves / st.seconds : 0.0)
              << " refreshes=" << st.refreshes << " residual=" << st.dual_residual << "\n";
    std::cout << "Final best score: " << st.best_score << "\n";
    try {
        if (have_A) {
            if (n <= 20) {
                std::cout << "P:\n";
                print_mat(bP);
                std::cout << "----\nQ:\n";
                print_d(bQ, n);
                std::cout << "----\nX:\n";
                print_mat(bX);
            }
            if (g_interrupted)
                std::cout << "[interrupted] writing best_X.csv, best_P.csv and best_Q.csv before "
                             "exit\n";
            write_csv(bX, "best_X.csv");
            write_csv(bP, "best_P.csv");
            write_csv_d(bQ, n, "best_Q.csv");
        } else {
            const Mat Lf = xt_times(bX, Lat);
            if (gram(Lf, {}).v != bP.v)
                std::cout << "[warn] Gram(X^T L) != annealed P (transform mismatch, possible "
                             "int64 overflow)\n";
            if (g_interrupted)
                std::cout << "[interrupted] writing final_L.csv, best_P.csv and best_X.csv before "
                             "exit\n";
            write_csv(Lf, "final_L.csv");
            write_csv(bP, "best_P.csv");
            write_csv_d(bQ, n, "best_Q.csv");
            write_csv(bX, "best_X.csv");
            if (Lat.rows <= 20 && Lat.cols <= 20) {
                std::cout << "Final L:\n";
                print_mat(Lf);
                std::cout << "----\nPrimal Gram P:\n";
                print_mat(bP);
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
