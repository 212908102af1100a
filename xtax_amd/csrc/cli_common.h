// cli_common.h -- pieces shared by the command-line drop-ins (xtax, xdual):
// Ctrl-C handling, the reference's CSV formats and index lists, lattice
// Grams, and a flag parser for its option surface. Plain C++20 on the host.
#pragma once
#include <algorithm>
#include <atomic>
#include <csignal>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <functional>
#include <random>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include <unistd.h>

#include "xtx.h"

namespace cli {

inline std::atomic<bool> g_interrupted{false};
inline xtx_ctx* g_ctx = nullptr;

inline const char kInterruptMsg[] =
    "\n[interrupt received, finishing and writing the best result so far...]\n";

// First Ctrl-C: stop flag + one acknowledgement; later ones are swallowed
// (stop_signal.hpp:52-64).
inline void on_sigint(int) {
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
inline Mat read_csv(const std::string& path, bool square) {
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
inline void write_csv(const Mat& M, const std::string& path) {
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

inline void print_mat(const Mat& M) {
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
inline void reduce_mod(std::vector<int64_t>& v, uint64_t m) {
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

inline std::string modulus_note(uint64_t m) {
    if (m == 1) return "saturated entries (nonzero -> 1, 0 stays 0)";
    return "reduced entries mod " + std::to_string(m) + " (balanced residues in (-m/2, m/2])";
}

// Index lists "0,3..6,9" (mat_io.hpp:97-141): sorted, deduplicated.
inline std::vector<int> parse_index_spec(const std::string& spec, int n, const std::string& opt) {
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

inline Mat gram(const Mat& L, const std::vector<int>& dims) {
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

inline Mat transposed(const Mat& M) {
    Mat T;
    T.rows = M.cols;
    T.cols = M.rows;
    T.v.assign(M.v.size(), 0);
    for (int r = 0; r < M.rows; ++r)
        for (int c = 0; c < M.cols; ++c) T.at(c, r) = M.at(r, c);
    return T;
}

inline Mat square(int n) {
    Mat M;
    M.rows = M.cols = n;
    M.v.assign((size_t)n * n, 0);
    return M;
}

// X^T A X (entries within the 2^48 bound: the engine has checked the start
// state), accumulated in 128 bits
inline Mat congruence(const Mat& A, const Mat& X) {
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

// ---- flag parser for the reference's option surface (CLI11 conventions:
// "--name value" or "--name=value", unknown flags are an error) ----
struct Opt {
    std::vector<std::string> names;
    bool is_flag;
    std::function<void(const std::string&)> set;
};

inline std::function<void(const std::string&)> set_double(double& d) {
    return [&d](const std::string& s) { d = std::stod(s); };
}
inline std::function<void(const std::string&)> set_int(int32_t& d) {
    return [&d](const std::string& s) { d = (int32_t)std::stol(s); };
}
inline std::function<void(const std::string&)> set_string(std::string& d) {
    return [&d](const std::string& s) { d = s; };
}

// Applies argv to opts. Returns -1 to go on, else the exit code (0 after
// --help, 109 -- CLI11's parse-error code -- after a bad argument).
inline int parse_args(int argc, char** argv, const std::vector<Opt>& opts, const char* help) {
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
                std::cout << help;
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
        return 109;
    }
    return -1;
}

// --seed 0 (the default) seeds from std::random_device, as the reference's
// run_annealer does (congruence_anneal.hpp:484-485); the engine's own key 0
// is an ordinary key.
inline void seed_from_device(xtx_params& p) {
    if (p.seed != 0) return;
    std::random_device rd;
    while (p.seed == 0) p.seed = ((uint64_t)rd() << 32) | (uint64_t)rd();
}

// The engine options both drop-ins share (congruence_anneal.hpp:122-153 via
// main.cpp:583-640 and xdual_main.cpp:465-500), plus the GPU-only --device,
// --max-moves, --max-epochs and --lookahead (xtx_params.lookahead: proposals
// of a replica evaluated at once; speed only).
inline std::vector<Opt> engine_options(xtx_params& p, int& device) {
    return {
        {{"-t", "--threads", "--replicas"}, false, set_int(p.replicas)},
        {{"--use-hyperthreads"}, true, [](const std::string&) {}},
        {{"--no-pin"}, true, [](const std::string&) {}},
        {{"--seed"}, false, [&p](const std::string& s) { p.seed = std::stoull(s); }},
        {{"--stuck-threshold"}, false, set_int(p.stuck_threshold)},
        {{"--t-init"}, false, set_double(p.t_init)},
        {{"--t-min"}, false, set_double(p.t_min)},
        {{"--cooling"}, false, set_double(p.cooling)},
        {{"--moves-per-cool"}, false, set_int(p.moves_per_cool)},
        {{"--reheat"}, false, set_double(p.reheat)},
        {{"--greedy-fraction"}, false, set_double(p.greedy_fraction)},
        {{"--target-fraction"}, false, set_double(p.target_fraction)},
        {{"--target-samples"}, false, set_int(p.target_samples)},
        {{"--reseed-factor"}, false, set_double(p.reseed_factor)},
        {{"--sweep-fraction"}, false, set_double(p.sweep_fraction)},
        {{"--tempering"}, true, [&p](const std::string&) { p.tempering = 1; }},
        {{"--no-tempering"}, true, [&p](const std::string&) { p.tempering = 0; }},
        {{"--exchange-interval"}, false, set_int(p.exchange_interval)},
        {{"--adaptive-cooling"}, true, [&p](const std::string&) { p.adaptive_cooling = 1; }},
        {{"--worker-diversity"}, false, set_double(p.worker_diversity)},
        {{"--max-seconds"}, false, set_double(p.max_seconds)},
        {{"--save-interval"}, false, set_double(p.save_interval)},
        {{"--device"}, false, [&device](const std::string& s) { device = (int)std::stol(s); }},
        {{"--max-moves"}, false, [&p](const std::string& s) { p.max_moves = std::stoull(s); }},
        {{"--max-epochs"}, false, set_int(p.max_epochs)},
        {{"--lookahead"}, false, set_int(p.lookahead)},
    };
}

// What a "new best" report does besides printing: the reference rewrites the
// best_*.csv files on a new best at most every save_interval seconds
// (congruence_anneal.hpp:327-340); `save` writes them from the state
// xtx_best_so_far returns.
struct BestReports {
    xtx_ctx* ctx = nullptr;
    double save_interval = 2.0;
    double last_save = -1e300;
    std::function<void()> save;
};

inline void maybe_save(BestReports* b, double seconds) {
    if (b && b->save && seconds - b->last_save >= b->save_interval) {
        b->save();
        b->last_save = seconds;
    }
}

// The reference's "new best" progress line (congruence_anneal.hpp:331-334)
// from the engine's on_best report; the replica stands in for the thread.
inline void print_new_best_l1(void* user, double seconds, int64_t score, double, int32_t replica) {
    std::cout << "[t=" << seconds << "s] new best score=" << score << " (thread " << replica
              << ")\n";
    maybe_save(static_cast<BestReports*>(user), seconds);
}
inline void print_new_best_dual(void* user, double seconds, int64_t, double score,
                                int32_t replica) {
    std::cout << "[t=" << seconds << "s] new best score=" << score << " (thread " << replica
              << ")\n";
    maybe_save(static_cast<BestReports*>(user), seconds);
}

// --modulus value (>= 1)
inline std::function<void(const std::string&)> set_modulus(uint64_t& m) {
    return [&m](const std::string& s) {
        const long long v = std::stoll(s);
        if (v < 1) throw std::runtime_error("--modulus: value must be >= 1");
        m = (uint64_t)v;
    };
}

// Final basis X^T L (main.cpp:850-851, xdual_main.cpp:627-628).
inline Mat xt_times(const Mat& X, const Mat& L) {
    Mat R;
    R.rows = L.rows;
    R.cols = L.cols;
    R.v.assign(L.v.size(), 0);
    for (int i = 0; i < L.rows; ++i)
        for (int q = 0; q < L.rows; ++q) {
            const int64_t x = X.at(q, i);
            if (!x) continue;
            for (int k = 0; k < L.cols; ++k) R.at(i, k) += x * L.at(q, k);
        }
    return R;
}

}  // namespace cli
