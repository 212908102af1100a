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
// it and no time limit is deterministic).
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

#include "xtx.h"

namespace {

std::atomic<bool> g_interrupted{false};
xtx_ctx* g_ctx = nullptr;

const char kInterruptMsg[] =
    "\n[interrupt received, finishing and writing the best result so far...]\n";

// First Ctrl-C: stop flag + one acknowledgement; later ones are swallowed
// (stop_signal.hpp:52-64).
void on_sigint(int) {
    if (!g_interrupted.exchange(true)) {
        if (g_ctx) xtx_request_stop(g_ctx);
        ssize_t r = write(STDERR_FILENO, kInterruptMsg, sizeof(kInterruptMsg) - 1);
        (void)r;
    }
}

struct Mat {
    int rows = 0, cols = 0;
    std::vector<int64_t> v;
    int64_t& at(int r, int c) { return v[(size_t)r * cols + c]; }
    int64_t at(int r, int c) const { return v[(size_t)r * cols + c]; }
};

// CSV reader with the reference's rules (mat_io.hpp:432-489): blank lines
// skipped, comma-separated std::stoll tokens, equal row lengths.
Mat read_csv(const std::string& path, bool square) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("Cannot open " + path);
    Mat M;
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty()) continue;
        std::stringstream ss(line);
        std::string tok;
        int cols = 0;
        while (std::getline(ss, tok, ',')) {
            M.v.push_back((int64_t)std::stoll(tok));
            ++cols;
        }
        if (M.rows == 0) M.cols = cols;
        if (cols != M.cols) throw std::runtime_error("Bad CSV row length in " + path);
        ++M.rows;
    }
    if (M.rows == 0)
        throw std::runtime_error(std::string(square ? "Empty matrix file: " : "Empty lattice file: ") +
                                 path);
    if (square && M.rows != M.cols) throw std::runtime_error("Matrix must be square in " + path);
    return M;
}

// Atomic replace: write file.tmp, then rename (mat_io.hpp:389-402).
void write_csv(const Mat& M, const std::string& path) {
    const std::string tmp = path + ".tmp";
    {
        std::ofstream out(tmp);
        if (!out) throw std::runtime_error("Failed to open output file: " + tmp);
        for (int r = 0; r < M.rows; ++r) {
            for (int c = 0; c < M.cols; ++c) {
                out << M.at(r, c);
                if (c + 1 < M.cols) out << ',';
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

void print_mat(const Mat& M) {
    for (int r = 0; r < M.rows; ++r) {
        for (int c = 0; c < M.cols; ++c) {
            std::cout << M.at(r, c);
            if (c + 1 < M.cols) std::cout << ',';
        }
        std::cout << '\n';
    }
}

// --modulus (mat_io.hpp:75-88): balanced residues in (-m/2, m/2]; m = 1
// saturates nonzero entries to 1.
void reduce_mod(std::vector<int64_t>& v, uint64_t m) {
    if (m == 1) {
        for (auto& x : v) x = x != 0;
        return;
    }
    const int64_t mm = (int64_t)m, half = mm / 2;
    for (auto& x : v) {
        int64_t r = x % mm;
        if (r < 0) r += mm;
        if (r > half) r -= mm;
        x = r;
    }
}

std::string modulus_note(uint64_t m) {
    if (m == 1) return "saturated entries (nonzero -> 1, 0 stays 0)";
    return "reduced entries mod " + std::to_string(m) + " (balanced residues in (-m/2, m/2])";
}

// Index lists "0,3..6,9" (mat_io.hpp:97-141): sorted, deduplicated.
std::vector<int> parse_index_spec(const std::string& spec, int n, const std::string& opt) {
    auto trim = [](std::string s) {
        const auto b = s.find_first_not_of(" \t");
        if (b == std::string::npos) return std::string();
        const auto e = s.find_last_not_of(" \t");
        return s.substr(b, e - b + 1);
    };
    auto index = [&](const std::string& tok) -> int {
        size_t pos = 0;
        long long v = -1;
        try {
            v = std::stoll(tok, &pos);
        } catch (const std::exception&) {
            pos = 0;
        }
        if (tok.empty() || pos != tok.size() || v < 0)
            throw std::runtime_error(opt + ": invalid index '" + tok + "'");
        if (v >= n)
            throw std::runtime_error(opt + ": index " + tok + " out of range (valid indices are 0.." +
                                     std::to_string(n - 1) + ")");
        return (int)v;
    };
    std::vector<int> out;
    std::stringstream ss(spec);
    std::string tok;
    while (std::getline(ss, tok, ',')) {
        tok = trim(tok);
        if (tok.empty()) continue;
        const size_t dots = tok.find("..");
        if (dots == std::string::npos) {
            out.push_back(index(tok));
        } else {
            const int lo = index(trim(tok.substr(0, dots)));
            const int hi = index(trim(tok.substr(dots + 2)));
            if (lo > hi) throw std::runtime_error(opt + ": reversed range '" + tok + "'");
            for (int r = lo; r <= hi; ++r) out.push_back(r);
        }
    }
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
    return out;
}

Mat gram(const Mat& L, const std::vector<int>& dims) {
    Mat G;
    G.rows = G.cols = L.rows;
    G.v.assign((size_t)L.rows * L.rows, 0);
    for (int i = 0; i < L.rows; ++i)
        for (int j = i; j < L.rows; ++j) {
            int64_t s = 0;
            if (dims.empty())
                for (int k = 0; k < L.cols; ++k) s += L.at(i, k) * L.at(j, k);
            else
                for (int k : dims) s += L.at(i, k) * L.at(j, k);
            G.at(i, j) = G.at(j, i) = s;
        }
    return G;
}

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

Mat transposed(const Mat& M) {
    Mat T;
    T.rows = M.cols;
    T.cols = M.rows;
    T.v.assign(M.v.size(), 0);
    for (int r = 0; r < M.rows; ++r)
        for (int c = 0; c < M.cols; ++c) T.at(c, r) = M.at(r, c);
    return T;
}

Mat square(int n) {
    Mat M;
    M.rows = M.cols = n;
    M.v.assign((size_t)n * n, 0);
    return M;
}

// X^T A X (entries within the 2^48 bound: the engine has checked the start
// state), accumulated in 128 bits
Mat congruence(const Mat& A, const Mat& X) {
    const int n = A.rows;
    std::vector<__int128> AX((size_t)n * n, 0);
    for (int r = 0; r < n; ++r)
        for (int k = 0; k < n; ++k) {
            const int64_t a = A.at(r, k);
            if (!a) continue;
            for (int c = 0; c < n; ++c) AX[(size_t)r * n + c] += (__int128)a * X.at(k, c);
        }
    Mat C = square(n);
    for (int r = 0; r < n; ++r)
        for (int c = 0; c < n; ++c) {
            __int128 v = 0;
            for (int k = 0; k < n; ++k) v += (__int128)X.at(k, r) * AX[(size_t)k * n + c];
            C.at(r, c) = (int64_t)v;
        }
    return C;
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

// ---- minimal flag parser for the reference's option surface ----
struct Opt {
    std::vector<std::string> names;
    bool is_flag;
    std::function<void(const std::string&)> set;
};

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
    auto D = [](double& d) { return [&d](const std::string& s) { d = std::stod(s); }; };
    auto I = [](int32_t& d) { return [&d](const std::string& s) { d = (int32_t)std::stol(s); }; };
    auto S = [](std::string& d) { return [&d](const std::string& s) { d = s; }; };
    std::vector<Opt> opts = {
        {{"-A"}, false, S(a_csv)},
        {{"-L", "--lattice"}, false, S(l_csv)},
        {{"-X"}, false, S(x_csv)},
        {{"--modulus"}, false, [&](const std::string& s) {
             const long long v = std::stoll(s);
             if (v < 1) throw std::runtime_error("--modulus: value must be >= 1");
             modulus = (uint64_t)v;
         }},
        {{"--gram-rows"}, false, S(gram_rows)},
        {{"--lattice-dims"}, false, S(lattice_dims)},
        {{"-t", "--threads", "--replicas"}, false, I(p.replicas)},
        {{"--use-hyperthreads"}, true, [](const std::string&) {}},
        {{"--no-pin"}, true, [](const std::string&) {}},
        {{"--seed"}, false, [&](const std::string& s) { p.seed = std::stoull(s); }},
        {{"--stuck-threshold"}, false, I(p.stuck_threshold)},
        {{"--t-init"}, false, D(p.t_init)},
        {{"--t-min"}, false, D(p.t_min)},
        {{"--cooling"}, false, D(p.cooling)},
        {{"--moves-per-cool"}, false, I(p.moves_per_cool)},
        {{"--reheat"}, false, D(p.reheat)},
        {{"--greedy-fraction"}, false, D(p.greedy_fraction)},
        {{"--target-fraction"}, false, D(p.target_fraction)},
        {{"--target-samples"}, false, I(p.target_samples)},
        {{"--reseed-factor"}, false, D(p.reseed_factor)},
        {{"--sweep-fraction"}, false, D(p.sweep_fraction)},
        {{"--tempering"}, true, [&](const std::string&) { p.tempering = 1; }},
        {{"--no-tempering"}, true, [&](const std::string&) { p.tempering = 0; }},
        {{"--exchange-interval"}, false, I(p.exchange_interval)},
        {{"--adaptive-cooling"}, true, [&](const std::string&) { p.adaptive_cooling = 1; }},
        {{"--worker-diversity"}, false, D(p.worker_diversity)},
        {{"--max-seconds"}, false, D(p.max_seconds)},
        {{"--save-interval"}, false, D(p.save_interval)},
        {{"--rcm"}, true, [&](const std::string&) { rcm = true; }},
        {{"--centroid"}, true, [&](const std::string&) { centroid = true; }},
        {{"--deflate"}, true, [&](const std::string&) { deflate = true; }},
        {{"--deflate-blocks"}, true, [&](const std::string&) { deflate_blocks = true; }},
        {{"--deflate-slice"}, false, D(deflate_slice)},
        {{"--verbose"}, true, [&](const std::string&) { verbose = true; }},
        {{"--device"}, false, I(device)},
        {{"--max-moves"}, false, [&](const std::string& s) { p.max_moves = std::stoull(s); }},
        {{"--max-epochs"}, false, I(p.max_epochs)},
    };
    try {
        for (int a = 1; a < argc; ++a) {
            std::string arg = argv[a], val;
            const size_t eq = arg.find('=');
            bool has_eq = false;
            if (arg.rfind("--", 0) == 0 && eq != std::string::npos) {
                val = arg.substr(eq + 1);
                arg = arg.substr(0, eq);
                has_eq = true;
            }
            if (arg == "-h" || arg == "--help") {
                std::cout << "xtax (gfx950 engine): see the reference README for the options\n";
                return 0;
            }
            bool found = false;
            for (const auto& o : opts) {
                if (std::find(o.names.begin(), o.names.end(), arg) == o.names.end()) continue;
                found = true;
                if (o.is_flag) {
                    o.set("");
                } else {
                    if (!has_eq) {
                        if (a + 1 >= argc) throw std::runtime_error(arg + " requires an argument");
                        val = argv[++a];
                    }
                    o.set(val);
                }
                break;
            }
            if (!found) throw std::runtime_error("The following argument was not expected: " + arg);
        }
    } catch (const std::exception& e) {
        std::cerr << e.what() << "\nRun with --help for more information.\n";
        return 109;  // CLI11's parse-error exit code
    }
    if (p.replicas < 1) p.replicas = 1;

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
                  << "\n";
        xtx_stats st{};
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
            Mat Lf;
            Lf.rows = Lat.rows;
            Lf.cols = Lat.cols;
            Lf.v.assign(Lat.v.size(), 0);
            for (int i = 0; i < Lat.rows; ++i)
                for (int q = 0; q < Lat.rows; ++q) {
                    const int64_t x = bX.at(q, i);
                    if (!x) continue;
                    for (int k = 0; k < Lat.cols; ++k) Lf.at(i, k) += x * Lat.at(q, k);
                }
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
