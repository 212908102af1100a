// dual.cpp -- host side of the primal/dual objective (xdual,
// src/xdual_main.cpp): xtx_dual_solve runs the engine's plain form with the
// k_anneal_dual kernels. Per epoch: k_dual_maint (re-inversion, residual
// check, rescoring), k_anneal_dual (the moves), k_epoch_end (solve
// retirement, exchange, instance best); the host polls the instance between
// epochs and, with a dual-weight ramp, sets the weight of the next epoch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <vector>

#include "engine_types.h"
#include "host_internal.h"
#include "xtx.h"

namespace xtx {
hipError_t launch_anneal_dual(hipStream_t st, int G, int C, int K, int units_per_block,
                              const RepRef* table, int count, Inst* insts, int nmax,
                              int rep_bytes, const volatile int* stop);
hipError_t launch_dual_maint(hipStream_t st, Inst* insts, int count, int max_r);
hipError_t launch_dual_probe(hipStream_t st, int G, int C, const int64_t* P, const double* Q,
                             int n, double cw, const int32_t* qs, int nq, int64_t* out_i,
                             double* out_d, int64_t* rows_p, double* rows_q, double* scores);
hipError_t launch_epoch_end(hipStream_t st, Inst* insts, int count, int max_r,
                            const int* list = nullptr);
hipError_t launch_replicate(hipStream_t st, int64_t* dst, const int64_t* src, int64_t nn, int R);
}  // namespace xtx

using namespace xtx;
using namespace xtx::host;

namespace {

// Lane group and LDS bytes of a replica: P and Q upper triangles of the
// padded order, 5 staged Philox blocks per lane group, the active list, and
// for a replica over several waves two buffers of K 64-byte proposals and a
// poll word. The group width is fixed by n (G = 16, 32, 64 lanes; two columns
// per lane above 64), which the fold-then-pair column sum needs to give the
// same bits for every n; lookahead K (the most up to `want` that the order
// has a kernel for: 16 or 4 with 16 lanes, 8 or 2 with 32, 4 with 64) adds
// lane groups per replica.
struct DualLayout {
    int G = 0, C = 0, K = 1, nmax = 0, rep_bytes = 0;
};

DualLayout dual_layout(int n, int want) {
    DualLayout L;
    L.nmax = order_of(n);
    L.G = L.nmax <= 16 ? 16 : (L.nmax <= 32 ? 32 : 64);
    L.C = L.nmax <= 64 ? 1 : 2;
    if (L.C == 1) {
        if (L.G == 16) L.K = want >= 16 ? 16 : (want >= 4 ? 4 : 1);
        if (L.G == 32) L.K = want >= 8 ? 8 : (want >= 2 ? 2 : 1);
        if (L.G == 64) L.K = want >= 4 ? 4 : 1;
    }
    const int64_t tri = (int64_t)L.nmax * (L.nmax + 1) / 2;
    int64_t bytes = tri * 16 + L.K * 5 * 16 + L.nmax;
    if (L.G * L.K > 64) bytes = (bytes + 15) / 16 * 16 + 2 * L.K * 64 + 16;
    L.rep_bytes = (int)((bytes + 15) / 16 * 16);
    return L;
}

// invert_to (mat_io.hpp:297-337): Gauss-Jordan with partial pivoting (first
// row of largest magnitude) on [P | I] in double, the right half symmetrized
// by averaging. False, and Q untouched, if P is singular.
bool invert(const int64_t* P, int n, double* Q) {
    const int w = 2 * n;
    std::vector<double> m((size_t)n * w, 0.0);
    auto at = [&](int r, int c) -> double& { return m[(size_t)r * w + c]; };
    for (int r = 0; r < n; ++r) {
        for (int c = 0; c < n; ++c) at(r, c) = (double)P[(size_t)r * n + c];
        at(r, n + r) = 1.0;
    }
    for (int col = 0; col < n; ++col) {
        int piv = col;
        double big = std::fabs(at(col, col));
        for (int r = col + 1; r < n; ++r) {
            const double v = std::fabs(at(r, col));
            if (v > big) {
                big = v;
                piv = r;
            }
        }
        if (big == 0.0) return false;
        if (piv != col)
            for (int c = 0; c < w; ++c) std::swap(at(piv, c), at(col, c));
        const double inv = 1.0 / at(col, col);
        for (int c = 0; c < w; ++c) at(col, c) *= inv;
        for (int r = 0; r < n; ++r) {
            if (r == col) continue;
            const double f = at(r, col);
            if (f == 0.0) continue;
            for (int c = 0; c < w; ++c) {
                const double prod = f * at(col, c);
                at(r, c) = at(r, c) - prod;
            }
        }
    }
    for (int r = 0; r < n; ++r)
        for (int c = 0; c < n; ++c) {
            const double x = at(r, n + c), y = at(c, n + r);
            Q[(size_t)r * n + c] = r == c ? x : 0.5 * (r < c ? x + y : y + x);
        }
    return true;
}

// The engine's column sum: column k + 64 folded onto column k, then the
// adjacent-pair binary tree over the 64 results.
double pairsum(const double* x, int n) {
    double b[64];
    for (int k = 0; k < 64; ++k) b[k] = (k < n ? x[k] : 0.0) + (k + 64 < n ? x[k + 64] : 0.0);
    for (int w = 32; w >= 1; w /= 2)
        for (int k = 0; k < w; ++k) b[k] = b[2 * k] + b[2 * k + 1];
    return b[0];
}

// |offdiag P|^2 + c |offdiag Q|^2 as the device computes it: per column the
// rows in order, then pairsum over the columns.
double dual_score(const int64_t* P, const double* Q, int n, double c) {
    std::vector<double> sp((size_t)n, 0.0), sq((size_t)n, 0.0);
    for (int k = 0; k < n; ++k)
        for (int r = 0; r < n; ++r) {
            if (r == k) continue;
            const double v = (double)P[(size_t)r * n + k];
            const double x = Q[(size_t)r * n + k];
            sp[(size_t)k] += v * v;
            sq[(size_t)k] += x * x;
        }
    const double fp = pairsum(sp.data(), n), fq = pairsum(sq.data(), n);
    const double cq = c * fq;
    return fp + cq;
}

// offdiag_frob2 in long double, row-major order (mat_io.hpp:240-249, :280-289),
// for the weight normalization
long double frob2_ld(const int64_t* P, const double* Q, int n) {
    long double s = 0.0L;
    for (int r = 0; r < n; ++r)
        for (int k = 0; k < n; ++k) {
            if (r == k) continue;
            const long double v = P ? (long double)P[(size_t)r * n + k] : (long double)Q[(size_t)r * n + k];
            s += v * v;
        }
    return s;
}

double residual(const int64_t* P, const double* Q, int n) {
    double worst = 0.0;
    for (int r = 0; r < n; ++r)
        for (int k = 0; k < n; ++k) {
            long double acc = 0.0L;
            for (int t = 0; t < n; ++t) acc += (long double)P[(size_t)r * n + t] * Q[(size_t)t * n + k];
            if (r == k) acc -= 1.0L;
            worst = std::max(worst, (double)std::fabs(acc));
        }
    return worst;
}

void check_dual_params(const xtx_dual_params& d) {
    auto bad = [](const char* m) { throw Err{XTX_E_ARG, m}; };
    if (!(d.lambda >= 0.0) || !std::isfinite(d.lambda)) bad("lambda must be finite and >= 0");
    if (d.dual_refresh < 0 || d.dual_check < 0) bad("dual_refresh and dual_check must be >= 0");
    if (!(d.dual_tol >= 0.0)) bad("dual_tol must be >= 0");
    if (!(d.lambda_ramp >= 0.0)) bad("lambda_ramp must be >= 0");
}

}  // namespace

extern "C" {

void xtx_dual_params_default(xtx_dual_params* d) {
    if (!d) return;
    std::memset(d, 0, sizeof(*d));
    d->lambda = 1.0;           // xdual_main.cpp:367-373
    d->dual_tol = 1e-6;
    d->lambda_ramp = 0.0;
    d->dual_refresh = 200000;
    d->dual_check = 25000;
}

int xtx_dual_solve(xtx_ctx* ctx, const int64_t* A0, const int64_t* X0, int32_t n,
                   const xtx_params* params, const xtx_dual_params* dparams,
                   const int32_t* active, int32_t n_active, int64_t* P_out, double* Q_out,
                   int64_t* X_out, xtx_dual_stats* stats) {
    if (!ctx || !params || !dparams) return XTX_E_ARG;
    const auto t0 = clk::now();
    try {
        HIPCHK(hipSetDevice(ctx->device));
        const xtx_params& p = *params;
        const xtx_dual_params& dp = *dparams;
        check_params(p);
        check_dual_params(dp);
        ctx->plain_log_valid = false;
        ctx->dual.valid = false;
        ctx->plain_rs.clear();
        HostInst& h = ctx->plain;
        h = HostInst{};
        if (n > kMaxGroupN) throw Err{XTX_E_ARG, "xdual: n exceeds 128 (an LDS-resident replica)"};
        if (p.target_samples > kMaxSamples) throw Err{XTX_E_ARG, "xdual: target_samples must be in [1, 16]"};
        prepare(h, A0, X0, n, p.replicas, p.seed, active, n_active, p);
        const size_t nn = (size_t)n * n;

        // the dual and its weight (run_xdual, xdual_main.cpp:381-405)
        std::vector<double> Q0(nn);
        if (!invert(h.Astart.data(), n, Q0.data()))
            throw Err{XTX_E_ARG, "could not invert the Gram (matrix singular / basis not full rank)"};
        const long double sp0 = frob2_ld(h.Astart.data(), nullptr, n);
        const long double sq0 = frob2_ld(nullptr, Q0.data(), n);
        long double cl = (long double)dp.lambda;
        if (sp0 > 0.0L && sq0 > 0.0L && std::isfinite((double)sq0)) cl *= sp0 / sq0;
        const double c_full = (double)cl;
        const bool ramp = dp.lambda_ramp > 0.0;
        double c_now = ramp ? 0.0 : c_full;
        const double score0 = dual_score(h.Astart.data(), Q0.data(), n, c_now);
        // suggest_t_init (xdual_main.cpp:212-219)
        const double avg = score0 / (double)((int64_t)n * n);
        h.t_init = p.t_init > 0.0 ? p.t_init : std::max(1.0, 2.0 * avg);
        h.P.t_init = h.t_init;
        if (h.P.tempering) h.temps = ladder(h.t_init, p.t_min, h.R);
        const DualLayout L = dual_layout(
            n, p.lookahead != 0 ? p.lookahead : kLookahead);
        if (L.rep_bytes > g_lds_block) throw Err{XTX_E_ARG, "n too large for an LDS-resident replica"};

        xtx_dual_stats st{};
        st.verify_status = -1;
        st.c = c_full;
        st.start_score = dual_score(h.Astart.data(), Q0.data(), n, c_full);
        st.start_primal = dual_score(h.Astart.data(), Q0.data(), n, 0.0);
        st.start_dual = (double)sq0;
        st.solve_replica = -1;
        auto finish_trivial = [&]() {
            if (P_out) std::memcpy(P_out, h.Astart.data(), nn * 8);
            if (Q_out) std::memcpy(Q_out, Q0.data(), nn * 8);
            if (X_out)
                for (int r = 0; r < n; ++r)
                    for (int c = 0; c < n; ++c) X_out[(size_t)r * n + c] = h.Xt0[(size_t)c * n + r];
            st.solved = h.offnz0 == 0;
            st.best_score = st.start_score;
            st.dual_residual = residual(h.Astart.data(), Q0.data(), n);
            st.stop_reason = 4;
            if (p.verify) {
                std::vector<int64_t> X(nn);
                for (int r = 0; r < n; ++r)
                    for (int c = 0; c < n; ++c) X[(size_t)r * n + c] = h.Xt0[(size_t)c * n + r];
                xtx_verify_report rep{};
                xtx_verify(h.A0.data(), X.data(), h.Astart.data(), n, &rep);
                st.verify_status = st.solved ? rep.status : (rep.status & ~2);
            }
            st.seconds = std::chrono::duration<double>(clk::now() - t0).count();
            if (stats) *stats = st;
        };
        if (n < 2 || h.N < 2 || h.offnz0 == 0) {
            finish_trivial();
            return XTX_OK;
        }

        // device buffers
        const int R = h.R;
        const size_t rb = align256((size_t)R * nn * 8), mb = align256(nn * 8);
        const size_t bytes = 6 * rb + 3 * mb + align256((size_t)R * nn * 16) +
                             align256((size_t)R * sizeof(RepState)) + align256((size_t)R * 8) +
                             align256((size_t)n) + align256((size_t)R * h.log_cap * sizeof(LogEnt));
        ctx->plain_arena.reserve(bytes);
        char* base = ctx->plain_arena.alloc(bytes);
        if (!base) throw Err{XTX_E_NOMEM, "arena"};
        char* cur = base;
        auto take = [&](size_t b) {
            char* q = cur;
            cur += align256(b);
            return q;
        };
        h.dmem = base;
        Inst& I = h.rec;
        I = Inst{};
        I.n = n;
        I.N = h.N;
        I.R = R;
        I.status = ST_RUNNING;
        I.key0 = (uint32_t)h.seed;
        I.key1 = (uint32_t)(h.seed >> 32);
        I.log_on = h.log_on;
        I.A = (int64_t*)take((size_t)R * nn * 8);
        I.Xt = (int64_t*)take((size_t)R * nn * 8);
        I.bA = (int64_t*)take((size_t)R * nn * 8);
        I.bXt = (int64_t*)take((size_t)R * nn * 8);
        I.Q = (double*)take((size_t)R * nn * 8);
        I.bQ = (double*)take((size_t)R * nn * 8);
        I.gA = (int64_t*)take(nn * 8);
        I.gXt = (int64_t*)take(nn * 8);
        I.gQ = (double*)take(nn * 8);
        I.work = (double*)take((size_t)R * nn * 16);
        I.rs = (RepState*)take((size_t)R * sizeof(RepState));
        I.temps = (double*)take((size_t)R * 8);
        I.act = (uint8_t*)take((size_t)n);
        I.log = h.log_on ? (LogEnt*)take((size_t)R * h.log_cap * sizeof(LogEnt)) : nullptr;
        I.log_cap = h.log_cap;
        I.P = h.P;
        I.solve_key = kNoKey;
        I.gb_offnz = h.offnz0;
        I.solve_rep = -1;
        I.solve_move = -1;
        I.dual = 1;
        I.dual_refresh = dp.dual_refresh;
        I.dual_check = dp.dual_check;
        I.dual_tol = dp.dual_tol;
        I.c = c_now;
        I.gb_dscore = score0;

        std::vector<RepState> rs((size_t)R, RepState{});
        for (int r = 0; r < R; ++r) {
            RepState& x = rs[(size_t)r];
            x.offnz = h.offnz0;
            x.T = h.t_init;
            x.cooling = h.P.cooling;
            x.rung = r;
            x.best_move = 0;
            x.flags = RF_CUR_BEST;
            x.dscore = x.dlocal_best = x.dbest_score = score0;
            x.dc = c_now;
            fractions(p, r, R, &x.gf, &x.tf);
        }
        std::vector<uint8_t> act(h.act.begin(), h.act.end());
        hipStream_t sm = ctx->stream;
        HIPCHK(hipMemcpyAsync(I.gA, h.Astart.data(), nn * 8, hipMemcpyHostToDevice, sm));
        HIPCHK(hipMemcpyAsync(I.gXt, h.Xt0.data(), nn * 8, hipMemcpyHostToDevice, sm));
        HIPCHK(hipMemcpyAsync(I.gQ, Q0.data(), nn * 8, hipMemcpyHostToDevice, sm));
        HIPCHK(launch_replicate(sm, I.A, I.gA, (int64_t)nn, R));
        HIPCHK(launch_replicate(sm, I.Xt, I.gXt, (int64_t)nn, R));
        HIPCHK(launch_replicate(sm, (int64_t*)I.Q, (const int64_t*)I.gQ, (int64_t)nn, R));
        HIPCHK(hipMemcpyAsync(I.rs, rs.data(), (size_t)R * sizeof(RepState),
                              hipMemcpyHostToDevice, sm));
        if (!h.temps.empty())
            HIPCHK(hipMemcpyAsync(I.temps, h.temps.data(), (size_t)R * 8, hipMemcpyHostToDevice, sm));
        HIPCHK(hipMemcpyAsync(I.act, act.data(), act.size(), hipMemcpyHostToDevice, sm));
        if (!ctx->d_plain) HIPCHK(hipMalloc(&ctx->d_plain, sizeof(Inst)));
        HIPCHK(hipMemcpyAsync(ctx->d_plain, &I, sizeof(Inst), hipMemcpyHostToDevice, sm));

        // replica table: whole waves of one instance
        std::vector<RepRef> refs;
        for (int r = 0; r < R; ++r) refs.push_back(RepRef{0, r});
        const int per_wave = std::max(1, 64 / (L.G * L.K));  // replicas per wave
        while (refs.size() % (size_t)per_wave) refs.push_back(RepRef{0, -1});
        RepRef* d_refs = nullptr;
        Inst* pin = nullptr;
        // released on every exit, including an error thrown below
        struct Release {
            RepRef*& refs;
            Inst*& pin;
            ~Release() {
                if (refs) (void)hipFree(refs);
                if (pin) (void)hipHostFree(pin);
            }
        } release{d_refs, pin};
        HIPCHK(hipMalloc(&d_refs, refs.size() * sizeof(RepRef)));
        HIPCHK(hipMemcpyAsync(d_refs, refs.data(), refs.size() * sizeof(RepRef),
                              hipMemcpyHostToDevice, sm));
        int waves = p.waves_per_block > 0 && L.G * L.K <= 64 ? p.waves_per_block : 1;
        while (waves > 1 && (int64_t)waves * per_wave * L.rep_bytes > g_lds_block) --waves;
        const int gpb = waves * per_wave;  // replicas per block
        HIPCHK(hipHostMalloc((void**)&pin, sizeof(Inst), hipHostMallocDefault));
        HIPCHK(hipStreamSynchronize(sm));

        // the solving replica's state if solved, else the instance best
        // (argmin of best score, best move, replica): P, Q and X; returns the
        // replica
        auto pick = [&](const Inst& J, std::vector<RepState>& rsv, std::vector<int64_t>& Pv,
                        std::vector<double>& Qv, std::vector<int64_t>& Xv) -> int {
            rsv.resize((size_t)R);
            HIPCHK(hipMemcpy(rsv.data(), J.rs, (size_t)R * sizeof(RepState),
                             hipMemcpyDeviceToHost));
            int w = 0;
            bool cur = true;
            if (J.status == ST_SOLVED) {
                w = J.solve_rep;
            } else {
                for (int r = 1; r < R; ++r) {
                    const RepState& x = rsv[(size_t)r];
                    const RepState& y = rsv[(size_t)w];
                    if (okey(x.dbest_score) < okey(y.dbest_score) ||
                        (okey(x.dbest_score) == okey(y.dbest_score) && x.best_move < y.best_move))
                        w = r;
                }
                const RepState& x = rsv[(size_t)w];
                if (!(x.flags & RF_CUR_BEST) && (x.flags & RF_BEST_STALE))
                    throw Err{XTX_E_STATE, "internal: best replica has a stale snapshot"};
                cur = (x.flags & RF_CUR_BEST) != 0;
            }
            std::vector<int64_t> Xt(nn);
            const size_t off = (size_t)w * nn;
            HIPCHK(hipMemcpy(Pv.data(), (cur ? J.A : J.bA) + off, nn * 8, hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(Xt.data(), (cur ? J.Xt : J.bXt) + off, nn * 8, hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(Qv.data(), (cur ? J.Q : J.bQ) + off, nn * 8, hipMemcpyDeviceToHost));
            for (int r = 0; r < n; ++r)
                for (int c = 0; c < n; ++c) Xv[(size_t)r * n + c] = Xt[(size_t)c * n + r];
            return w;
        };
        ctx->best_so_far = [&](int64_t* Po, double* Qo, int64_t* Xo) {
            std::vector<RepState> r2;
            std::vector<int64_t> P2(nn), X2(nn);
            std::vector<double> Q2(nn);
            pick(*pin, r2, P2, Q2, X2);
            if (Po) std::memcpy(Po, P2.data(), nn * 8);
            if (Qo) std::memcpy(Qo, Q2.data(), nn * 8);
            if (Xo) std::memcpy(Xo, X2.data(), nn * 8);
        };
        struct Clear {
            xtx_ctx* c;
            ~Clear() { c->best_so_far = nullptr; }
        } clear_snapshot{ctx};

        double kernel_ms = 0;
        int epochs = 0;
        int reason = 3;
        double reported = score0;  // the start state is never reported
        const auto ramp_t0 = clk::now();
        {
            Deadline dl(ctx->stop_host, p.max_seconds);
            for (;;) {
                if (ramp && c_now != c_full) {
                    // the weight of this epoch from the host clock; the
                    // maintenance kernel rescores every replica for it
                    const double el = std::chrono::duration<double>(clk::now() - ramp_t0).count();
                    const double frac = el / dp.lambda_ramp;
                    c_now = frac >= 1.0 ? c_full : (double)((long double)c_full * frac);
                    HIPCHK(hipMemcpyAsync((char*)ctx->d_plain + offsetof(Inst, c), &c_now, 8,
                                          hipMemcpyHostToDevice, sm));
                }
                HIPCHK(launch_dual_maint(sm, ctx->d_plain, 1, R));
                HIPCHK(hipEventRecord(ctx->ev0, sm));
                HIPCHK(launch_anneal_dual(sm, L.G, L.C, L.K, gpb, d_refs, (int)refs.size(),
                                          ctx->d_plain, L.nmax, L.rep_bytes, ctx->stop_dev));
                HIPCHK(hipEventRecord(ctx->ev1, sm));
                HIPCHK(launch_epoch_end(sm, ctx->d_plain, 1, R));
                HIPCHK(hipMemcpyAsync(pin, ctx->d_plain, sizeof(Inst), hipMemcpyDeviceToHost, sm));
                HIPCHK(hipStreamSynchronize(sm));
                float ms = 0;
                HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
                kernel_ms += ms;
                ++epochs;
                if (p.on_best && pin->status != ST_SOLVED && pin->gb_dscore < reported) {
                    reported = pin->gb_dscore;
                    p.on_best(p.on_best_user,
                              std::chrono::duration<double>(clk::now() - t0).count(), 0, reported,
                              pin->gb_rep);
                }
                if (pin->status == ST_SOLVED) {
                    reason = 0;
                    break;
                }
                if (pin->status == ST_EXHAUSTED) {
                    reason = 3;
                    break;
                }
                if (*(volatile int*)ctx->stop_host) {
                    reason = stop_reason_of(ctx->stop_host);
                    break;
                }
                if (p.max_epochs > 0 && epochs >= p.max_epochs) {
                    reason = 3;
                    break;
                }
            }
        }
        I = *pin;

        // outcome: the solving replica's state, else the instance best
        std::vector<int64_t> Pw(nn), X(nn);
        std::vector<double> Qw(nn);
        const int win = pick(I, rs, Pw, Qw, X);
        if (I.status == ST_SOLVED) {
            st.solved = 1;
            st.solve_replica = win;
            st.solve_move = I.solve_move;
        }
        if (P_out) std::memcpy(P_out, Pw.data(), nn * 8);
        if (Q_out) std::memcpy(Q_out, Qw.data(), nn * 8);
        if (X_out) std::memcpy(X_out, X.data(), nn * 8);

        for (const RepState& x : rs) {
            st.moves += (int64_t)x.moves;
            st.accepts += x.accepts;
            st.greedy_calls += x.greedy;
            st.infeasible += x.infeasible;
            st.sweeps += x.sweeps;
            st.refreshes += x.refreshes;
            st.checks += x.checks;
            st.recomputes += x.recomputes;
        }
        st.exchanges_tried = I.exch_tried;
        st.exchanges_accepted = I.exch_acc;
        st.best_replica = win;
        // re-based at the full weight, as run_xdual reports it (:415-421)
        st.best_score = dual_score(Pw.data(), Qw.data(), n, c_full);
        st.dual_residual = residual(Pw.data(), Qw.data(), n);
        st.epochs = epochs;
        st.stop_reason = reason;
        st.kernel_ms = kernel_ms;
        if (p.verify) {
            xtx_verify_report rep{};
            xtx_verify(h.A0.data(), X.data(), Pw.data(), n, &rep);
            st.verify_status = st.solved ? rep.status : (rep.status & ~2);
        }
        st.seconds = std::chrono::duration<double>(clk::now() - t0).count();
        if (stats) *stats = st;

        ctx->plain_rs = rs;
        ctx->dual.valid = true;
        ctx->dual.c = c_full;
        ctx->dual.Q0 = Q0;
        ctx->dual.rs = rs;
        if (h.log_on) {
            ctx->plain_log.resize((size_t)R * h.log_cap);
            HIPCHK(hipMemcpy(ctx->plain_log.data(), I.log, ctx->plain_log.size() * sizeof(LogEnt),
                             hipMemcpyDeviceToHost));
            ctx->plain_log_count.resize((size_t)R);
            for (int r = 0; r < R; ++r) ctx->plain_log_count[(size_t)r] = rs[(size_t)r].log_count;
            ctx->plain_log_valid = true;
        }
        if (p.verify && st.verify_status != 0) {
            ctx->err = "result failed exact verification";
            return XTX_E_VERIFY;
        }
        return XTX_OK;
    } catch (const Err& e) {
        return fail(ctx, e);
    }
}

int xtx_dual_last_replicas(xtx_ctx* ctx, int64_t* P, double* Q, int64_t* Xt, double* dscal,
                           int64_t* scal) {
    if (!ctx) return XTX_E_ARG;
    try {
        const HostInst& h = ctx->plain;
        if (!ctx->dual.valid || h.dmem == nullptr) throw Err{XTX_E_STATE, "no dual solve"};
        const size_t nn = (size_t)h.n * h.n;
        if (P) HIPCHK(hipMemcpy(P, h.rec.A, (size_t)h.R * nn * 8, hipMemcpyDeviceToHost));
        if (Q) HIPCHK(hipMemcpy(Q, h.rec.Q, (size_t)h.R * nn * 8, hipMemcpyDeviceToHost));
        if (Xt) HIPCHK(hipMemcpy(Xt, h.rec.Xt, (size_t)h.R * nn * 8, hipMemcpyDeviceToHost));
        for (int r = 0; r < h.R; ++r) {
            const RepState& s = ctx->dual.rs[(size_t)r];
            if (dscal) {
                double* o = dscal + (size_t)r * 4;
                o[0] = s.dscore;
                o[1] = s.dbest_score;
                o[2] = s.dlocal_best;
                o[3] = s.T;
            }
            if (scal) {
                int64_t* o = scal + (size_t)r * 8;
                o[0] = s.offnz;
                o[1] = (int64_t)s.moves;
                o[2] = s.best_move;
                o[3] = s.rung;
                o[4] = s.accepts;
                o[5] = s.refreshes;
                o[6] = s.checks;
                o[7] = s.msi;
            }
        }
        return XTX_OK;
    } catch (const Err& e) {
        return fail(ctx, e);
    }
}

int xtx_dual_probe(xtx_ctx* ctx, const int64_t* P, const double* Q, int32_t n, int32_t n_mats,
                   double c, const int32_t* queries, int32_t n_queries, int64_t* out_i,
                   double* out_d, int64_t* rows_p, double* rows_q, double* scores) {
    if (!ctx || !P || !Q || !queries || !out_i || n < 2 || n > kMaxGroupN || n_mats < 1 ||
        n_queries < 0)
        return XTX_E_ARG;
    try {
        for (int q = 0; q < n_queries; ++q) {
            const int32_t* x = queries + 4 * q;
            if (x[0] < 0 || x[0] >= n_mats || x[1] < 0 || x[1] >= n || x[2] < 0 || x[2] >= n ||
                x[1] == x[2] || x[3] < -kShearCap || x[3] > kShearCap)
                throw Err{XTX_E_ARG, "bad probe query"};
        }
        const DualLayout L = dual_layout(n, 1);
        if ((size_t)(64 / L.G) * (size_t)L.nmax * (L.nmax + 1) / 2 * 16 > (size_t)g_lds_block)
            throw Err{XTX_E_ARG, "n too large for the probe"};
        HIPCHK(hipSetDevice(ctx->device));
        const size_t nn = (size_t)n * n * n_mats;
        const size_t nq = (size_t)std::max(n_queries, 1);
        int64_t *dP = nullptr, *doi = nullptr, *drp = nullptr;
        double *dQ = nullptr, *dod = nullptr, *drq = nullptr, *dsc = nullptr;
        int32_t* dq = nullptr;
        HIPCHK(hipMalloc(&dP, nn * 8));
        HIPCHK(hipMalloc(&dQ, nn * 8));
        HIPCHK(hipMalloc(&dq, nq * 16));
        HIPCHK(hipMalloc(&doi, nq * 24));
        HIPCHK(hipMalloc(&dod, nq * 8));
        HIPCHK(hipMalloc(&drp, nq * n * 8));
        HIPCHK(hipMalloc(&drq, nq * n * 8));
        HIPCHK(hipMalloc(&dsc, nq * 8));
        HIPCHK(hipMemcpy(dP, P, nn * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dQ, Q, nn * 8, hipMemcpyHostToDevice));
        if (n_queries)
            HIPCHK(hipMemcpy(dq, queries, (size_t)n_queries * 16, hipMemcpyHostToDevice));
        HIPCHK(launch_dual_probe(ctx->stream, L.G, L.C, dP, dQ, n, c, dq, n_queries, doi, dod, drp,
                                 drq, dsc));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        const size_t q = (size_t)n_queries;
        if (q) {
            HIPCHK(hipMemcpy(out_i, doi, q * 24, hipMemcpyDeviceToHost));
            if (out_d) HIPCHK(hipMemcpy(out_d, dod, q * 8, hipMemcpyDeviceToHost));
            if (rows_p) HIPCHK(hipMemcpy(rows_p, drp, q * n * 8, hipMemcpyDeviceToHost));
            if (rows_q) HIPCHK(hipMemcpy(rows_q, drq, q * n * 8, hipMemcpyDeviceToHost));
            if (scores) HIPCHK(hipMemcpy(scores, dsc, q * 8, hipMemcpyDeviceToHost));
        }
        for (void* b : {(void*)dP, (void*)dQ, (void*)dq, (void*)doi, (void*)dod, (void*)drp,
                        (void*)drq, (void*)dsc})
            (void)hipFree(b);
        return XTX_OK;
    } catch (const Err& e) {
        return fail(ctx, e);
    }
}

}  // extern "C"
