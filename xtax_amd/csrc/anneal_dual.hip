// anneal_dual.hip -- gfx950 kernels of the engine on xdual's primal/dual
// objective (DualObjective, src/xdual_main.cpp:169-364): minimize
// |offdiag P|^2 + c |offdiag Q|^2 over integer congruences P = X^T P0 X, with
// the double dual Q = P^-1 carried along (Q transforms by the inverse
// transpose of each shear: target and pivot swapped, shear negated).
//
// Mapping as in anneal.hip: one replica per lane group of G lanes, lane g
// owning columns k = g + c G. The replica's P (int64) and Q (double) sit in LDS
// as upper triangles, X^T in HBM, P's row masses in registers. The engine's
// double specification makes every floating-point result independent of the
// launch shape: a column sum folds column k + 64 onto column k and then adds
// adjacent lanes as a binary tree (Grp::fsum); a primal square difference is
// (nv - old)(nv + old) with both factors exact; no fused multiply-add.
//
// Kernels:
//   k_anneal_dual<G,C,K> one epoch of many replicas (K > 1: lookahead, as in
//                       anneal.hip's k_anneal)
//   k_dual_maint        per replica, before an epoch: dual re-inversion
//                       (Gauss-Jordan, invert_to, src/mat_io.hpp:297-337) and
//                       residual check when their periods were crossed, and
//                       a score recomputation after a re-inversion or a
//                       change of the dual weight
//   k_dual_probe<G,C>   the move primitives on given matrices (parity tests)
#include <hip/hip_runtime.h>

#include "device_common.h"
#include "engine_types.h"

namespace xtx {

// ---------------------------------------------------------------------------
// replica state
// ---------------------------------------------------------------------------

template <int G, int C>
struct DualReplica {
    int64_t* sP;        // LDS: upper triangle of P, row r at D(r) (less r)
    double* sQ;         // LDS: upper triangle of Q, same offsets
    GLOBAL int64_t* gX; // HBM: X^T of this replica, n x n row-major
    int n, g;
    int Dk[C];          // D(k) of the lane's columns
    int64_t rm[C];      // row masses of P for the lane's rows

    __device__ __forceinline__ int D(int r) const { return (r * (2 * n - r - 1)) >> 1; }
    __device__ __forceinline__ int at(int a, int b) const { return a <= b ? D(a) + b : D(b) + a; }
    __device__ __forceinline__ void init_lanes() {
#pragma unroll
        for (int c = 0; c < C; ++c) Dk[c] = D(g + c * G);
    }
    // entry (r, k) for the lane's column k = g + c G, given D(r)
    __device__ __forceinline__ int at_lane(int r, int Dr, int c) const {
        const int k = g + c * G;
        return k >= r ? Dr + k : Dk[c] + r;
    }

    __device__ __forceinline__ void load_rows(int i, int j, int64_t (&pi)[C], int64_t (&pj)[C],
                                              double (&qi)[C], double (&qj)[C]) const {
        const int Di = D(i), Dj = D(j);
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const bool in = g + c * G < n;
            const int ai = at_lane(i, Di, c), aj = at_lane(j, Dj, c);
            pi[c] = in ? sP[ai] : 0;
            pj[c] = in ? sP[aj] : 0;
            qi[c] = in ? sQ[ai] : 0.0;
            qj[c] = in ? sQ[aj] : 0.0;
        }
    }
    __device__ __forceinline__ void prefetch_xt(int i, int j, int64_t (&xi)[C],
                                                int64_t (&xj)[C]) const {
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const int k = g + c * G;
            xi[c] = k < n ? gX[i * n + k] : 0;
            xj[c] = k < n ? gX[j * n + k] : 0;
        }
    }
    __device__ __forceinline__ void compute_row_mass() {
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const int k = g + c * G;
            int64_t m = 0;
            if (k < n)
                for (int t = 0; t < n; ++t)
                    if (t != k) m += iabs64(sP[at_lane(t, D(t), c)]);
            rm[c] = m;
        }
    }
    __device__ __forceinline__ int64_t mass_of(int row) const {
        int64_t v = Grp<G>::gather(rm[0], row & (G - 1));
        if (C > 1) {
            const int64_t v1 = Grp<G>::gather(rm[C > 1 ? 1 : 0], row & (G - 1));
            v = (row >= G) ? v1 : v;
        }
        return v;
    }
    // P and Q as full row-major n x n matrices in HBM
    __device__ __forceinline__ void store(GLOBAL int64_t* dP, GLOBAL double* dQ) const {
        for (int row = 0; row < n; ++row) {
            const int Dr = D(row);
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const int k = g + c * G;
                if (k < n) {
                    dP[row * n + k] = sP[at_lane(row, Dr, c)];
                    dQ[row * n + k] = sQ[at_lane(row, Dr, c)];
                }
            }
        }
    }
    __device__ __forceinline__ void load(const GLOBAL int64_t* P, const GLOBAL double* Q) {
        for (int row = 0; row < n; ++row) {
            const int Dr = D(row);
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const int k = g + c * G;
                if (k >= row && k < n) {
                    sP[Dr + k] = P[row * n + k];
                    sQ[Dr + k] = Q[row * n + k];
                }
            }
        }
    }
    __device__ __forceinline__ void copy_cols(GLOBAL int64_t* dst, const GLOBAL int64_t* src) const {
        for (int row = 0; row < n; ++row) {
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const int k = g + c * G;
                if (k < n) dst[row * n + k] = src[row * n + k];
            }
        }
    }
    // DualObjective::commit's transform row (xdual_main.cpp:236-247): false
    // when an entry would leave the bound (the move is then skipped)
    __device__ __forceinline__ bool xt_row(int64_t s, const int64_t (&xi)[C],
                                           const int64_t (&xj)[C], int64_t (&nx)[C]) const {
        bool bad = false;
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const int k = g + c * G;
            nx[c] = 0;
            if (k < n) {
                nx[c] = xj[c] + s * xi[c];
                bad |= over(nx[c]);
            }
        }
        return !Grp<G>::any(bad);
    }

    // The squared off-diagonal norms as the specification sums them: per
    // column the rows in order, then the columns folded and paired.
    __device__ __forceinline__ double score(double c_w) const {
        double cp[C], cq[C];
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const int k = g + c * G;
            double ap = 0.0, aq = 0.0;
            if (k < n)
                for (int r = 0; r < n; ++r) {
                    if (r == k) continue;
                    const int a = at_lane(r, D(r), c);
                    const double v = (double)sP[a];
                    const double w = sQ[a];
                    ap += v * v;
                    aq += w * w;
                }
            cp[c] = ap;
            cq[c] = aq;
        }
        double fp = cp[0], fq = cq[0];
        if constexpr (C == 2) {
            fp = cp[0] + cp[1];
            fq = cq[0] + cq[1];
        }
        fp = Grp<G>::fsum(fp);
        fq = Grp<G>::fsum(fq);
        const double cw = c_w * fq;
        return fp + cw;
    }
};

// Add(i, j, s) on P and Add(j, i, -s) on Q, evaluated (xdual_main.cpp:67-158,
// :225-234).
template <int C>
struct DualEval {
    int64_t np[C];  // proposed P[j][k] (k == j: new P_jj)
    double nq[C];   // proposed Q[i][k] (k == i: new Q_ii)
    double d;       // d_off_p + c d_off_q
    int32_t dnz;
    bool bad;
};

template <int G, int C>
__device__ __forceinline__ void dual_evaluate(const int64_t (&pi)[C], const int64_t (&pj)[C],
                                              const double (&qi)[C], const double (&qj)[C],
                                              int g, int n, int i, int j, int64_t s,
                                              int64_t p_ii, int64_t p_ji, int64_t p_jj,
                                              double q_ii, double q_ij, double q_jj,
                                              double c_w, DualEval<C>& e) {
    bool bad = false;
    int32_t dnz = 0;
    double tp[C], tq[C];
    const int64_t new_jj = p_jj + 2 * s * p_ji + s * s * p_ii;
    const double sf = (double)(-s);
    const double new_ii = (q_ii + (2.0 * sf) * q_ij) + (sf * sf) * q_jj;
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const int k = g + c * G;
        const bool in = k < n;
        const int64_t nv = pj[c] + s * pi[c];
        tp[c] = 0.0;
        e.np[c] = new_jj;
        if (in && k != j) {
            bad |= over(nv);
            tp[c] = (double)(nv - pj[c]) * (double)(nv + pj[c]);
            dnz += (pj[c] == 0 && nv != 0) ? 1 : ((pj[c] != 0 && nv == 0) ? -1 : 0);
            e.np[c] = nv;
        }
        const double prod = sf * qj[c];
        const double nq = qi[c] + prod;
        tq[c] = 0.0;
        e.nq[c] = new_ii;
        if (in && k != i) {
            const double a = nq * nq, b = qi[c] * qi[c];
            tq[c] = a - b;
            e.nq[c] = nq;
        }
    }
    double fp = tp[0], fq = tq[0];
    if constexpr (C == 2) {
        fp = tp[0] + tp[1];
        fq = tq[0] + tq[1];
    }
    const double d_p = 2.0 * Grp<G>::fsum(fp);
    const double d_q = 2.0 * Grp<G>::fsum(fq);
    const double cq = c_w * d_q;
    e.d = d_p + cq;
    e.dnz = Grp<G>::sum32(dnz);
    e.bad = Grp<G>::any(bad) || over(new_jj);
}

// DualObjective::best_shear (xdual_main.cpp:254-296) in the engine's double
// specification: the real minimizer of f(s) = 2 (quad s^2 + 2 lin s), its
// floor and floor + 1 clamped to +-64, the one with the lower f if below 0.
template <int G, int C>
__device__ __forceinline__ int64_t dual_best_shear(const int64_t (&pi)[C], const int64_t (&pj)[C],
                                                   const double (&qi)[C], const double (&qj)[C],
                                                   int g, int n, int i, int j, double c_w) {
    double ta[C], tb[C], tc[C], td[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const int k = g + c * G;
        const bool in = k < n;
        const double pik = (double)pi[c];
        const bool pk = in && k != j, qk = in && k != i;
        ta[c] = pk ? pik * pik : 0.0;
        tb[c] = pk ? (double)pj[c] * pik : 0.0;
        tc[c] = qk ? qj[c] * qj[c] : 0.0;
        td[c] = qk ? qi[c] * qj[c] : 0.0;
    }
    double fa = ta[0], fb = tb[0], fc = tc[0], fd = td[0];
    if constexpr (C == 2) {
        fa = ta[0] + ta[1];
        fb = tb[0] + tb[1];
        fc = tc[0] + tc[1];
        fd = td[0] + td[1];
    }
    const double a2 = Grp<G>::fsum(fa), b2 = Grp<G>::fsum(fb);
    const double c2 = Grp<G>::fsum(fc), d2 = Grp<G>::fsum(fd);
    const double cc2 = c_w * c2;
    const double quad = a2 + cc2;
    if (!(quad > 0.0)) return 0;
    const double cd2 = c_w * d2;
    const double lin = b2 - cd2;
    double s_star = -lin / quad;
    if (s_star != s_star) return 0;
    s_star = s_star > 4e18 ? 4e18 : (s_star < -4e18 ? -4e18 : s_star);
    const int64_t lo0 = (int64_t)__builtin_floor(s_star);
    const int64_t lo = clamp_cap(lo0), hi = clamp_cap(lo0 + 1);
    int64_t best_s = 0;
    double best_f = 0.0;
    const double w2 = 2.0 * lin;
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const int64_t s = t == 0 ? lo : hi;
        if (s == 0) continue;
        const double sd = (double)s;
        const double u = quad * sd;
        const double v = u * sd;
        const double x = w2 * sd;
        const double f = 2.0 * (v + x);
        if (f < best_f) {
            best_f = f;
            best_s = s;
        }
    }
    return best_s;
}

// ---------------------------------------------------------------------------
// one epoch
// ---------------------------------------------------------------------------

// One lookahead proposal as the other waves of its replica read it (LDS).
struct alignas(16) DualProposal {
    double d;
    int64_t s;
    double u;           // the Metropolis draw
    int32_t i, j;
    uint32_t w4z, w4w;  // the stall words
    uint32_t flags;     // 1: over the bound, 2: passes at the step's temperature, 4: greedy
    uint32_t pad[3];
};
static_assert(sizeof(DualProposal) == 64, "two proposals per 128 bytes");

// K > 1: lookahead exactly as in k_anneal (anneal.hip): the replica's K lane
// groups evaluate its next K proposals on its current state and resolve them
// in move order; K G = 64 in one wave (values through v_readlane, a committed
// move's rows through ds_bpermute), K G > 64 over a workgroup of K G / 64 waves
// (proposals through LDS, a committed move evaluated again by every group,
// workgroup barriers before every write of the shared state).
template <int G, int C, int K>
__global__ void __launch_bounds__(256, 2)
    k_anneal_dual(const RepRef* __restrict__ table, int count, Inst* __restrict__ insts, int nmax,
                  int rep_bytes, const volatile int* __restrict__ stop) {
    constexpr int W = K == 1 ? 1 : G * K / 64;  // waves per replica
    static_assert(K == 1 || (G * K) % 64 == 0, "lookahead groups fill whole waves");
    static_assert(W <= 4, "a workgroup of at most 256 lanes");
    extern __shared__ __align__(16) unsigned char smem[];
    using Gp = Grp<G>;
    constexpr int U = G * K;  // lanes per replica
    const int g = (int)threadIdx.x & (G - 1);
    const int kk = K == 1 ? 0 : (int)(threadIdx.x % (unsigned)U) / G;  // lookahead index
    const int slot = (int)threadIdx.x / U;
    const int gid = (int)blockIdx.x * ((int)blockDim.x / U) + slot;
    if (gid >= count) return;
    const int wave_gid0 =
        __builtin_amdgcn_readfirstlane((int)blockIdx.x * ((int)blockDim.x / U) +
                                       (int)(threadIdx.x >> 6) * (64 / U));
    Inst* I = insts + __builtin_amdgcn_readfirstlane(table[wave_gid0].inst);
    if (I->status != ST_RUNNING) return;
    const int r = U >= 64 ? __builtin_amdgcn_readfirstlane(table[gid].rep) : table[gid].rep;
    if (r < 0) return;
    GLOBAL RepState* S = (GLOBAL RepState*)I->rs + r;
    if (S->flags & RF_STOPPED) return;

    const size_t tri = (size_t)nmax * (nmax + 1) / 2;
    DualReplica<G, C> rp;
    rp.sP = (int64_t*)(smem + (size_t)slot * rep_bytes);
    rp.sQ = (double*)(rp.sP + tri);
    // then five staged Philox blocks per group, the active list, and for
    // W > 1 two buffers of K proposals (by step parity) and the poll word
    uint4* const words0 = (uint4*)(rp.sQ + tri);
    uint4* const swords = words0 + 5 * kk;
    uint8_t* sact = (uint8_t*)(words0 + 5 * K);
    DualProposal* const props =
        (DualProposal*)((((uintptr_t)(sact + nmax)) + 15) & ~(uintptr_t)15);
    volatile uint32_t* const spoll = (volatile uint32_t*)(props + 2 * K);
    rp.n = I->n;
    rp.g = g;
    rp.init_lanes();
    const int n = rp.n, N = I->N;
    const int64_t nn = (int64_t)n * n;
    GLOBAL int64_t* const gPcur = (GLOBAL int64_t*)I->A + (int64_t)r * nn;
    GLOBAL double* const gQcur = (GLOBAL double*)I->Q + (int64_t)r * nn;
    rp.gX = (GLOBAL int64_t*)I->Xt + (int64_t)r * nn;
    rp.load(gPcur, gQcur);
    const GLOBAL uint8_t* gact = (const GLOBAL uint8_t*)I->act;
    bool nonid = false;
    for (int q = g; q < N; q += G) {
        const uint8_t a = gact[q];
        sact[q] = a;
        nonid |= a != q;
    }
    const bool ident = N == n && __ballot(nonid) == 0ull;
    auto rowat = [&](uint32_t p) -> int { return ident ? (int)p : (int)sact[p]; };
    rp.compute_row_mass();

    const uint32_t key0 = I->key0, key1 = I->key1;
    const uint64_t epoch_end = (uint64_t)(I->epoch + 1u) * (uint64_t)I->P.epoch_moves;
    const uint64_t cap = I->P.max_moves;
    const uint64_t stop_at = (cap != 0 && cap < epoch_end) ? cap : epoch_end;
    const bool tempering = I->P.tempering != 0;
    const int TS = I->P.target_samples;
    const int32_t stuck = I->P.stuck_threshold;
    const int32_t per_cool = I->P.moves_per_cool;
    const bool logging = I->log != nullptr;
    const double cw = I->c;
    const double gb = I->gb_dscore;  // instance best at the epoch start

    uint64_t moves = (uint64_t)uni<U>((int64_t)S->moves);
    double score = S->dscore, local_best = S->dlocal_best, best_score = S->dbest_score;
    int64_t offnz = uni<U>(S->offnz);
    uint32_t msi = (uint32_t)uni<U>((int32_t)S->msi);
    double T = tempering ? ((const GLOBAL double*)I->temps)[S->rung] : S->T;
    double nbeta = -1.0 / T;
    double cooling = S->cooling;
    int32_t cool_counter = uni<U>(S->cool_counter), win_p = uni<U>(S->win_p);
    int32_t win_a = uni<U>(S->win_a);
    int64_t best_move = uni<U>(S->best_move);
    int32_t flags = uni<U>(S->flags);
    const Coin greedy_coin(S->gf), target_coin(S->tf);
    uint32_t accepts = 0, greedy_n = 0, infeasible = 0, sweeps = 0, logged = 0, recomputes = 0;

    auto leave_best = [&]() {
        if (flags & RF_CUR_BEST) {
            const Inst* Iq = opaque(I);
            if (best_score <= Iq->gb_dscore) {
                const int64_t off = (int64_t)r * n * n;
                if (kk == 0) {  // lookahead groups hold the same state
                    rp.store((GLOBAL int64_t*)Iq->bA + off, (GLOBAL double*)Iq->bQ + off);
                    rp.copy_cols((GLOBAL int64_t*)Iq->bXt + off, rp.gX);
                }
                flags &= ~RF_BEST_STALE;
            } else {
                flags |= RF_BEST_STALE;
            }
            flags &= ~RF_CUR_BEST;
        }
    };
    auto log_move = [&](uint64_t m, int i, int j, int64_t s, uint16_t f) {
        if (logging) {
            const Inst* Iq = opaque(I);
            const int64_t at = opaque_v(S)->log_count + logged;
            if (at < Iq->log_cap && g == 0 && kk == 0) {
                GLOBAL LogEnt* e = (GLOBAL LogEnt*)Iq->log + (int64_t)r * Iq->log_cap + at;
                e->move = (int64_t)m;
                e->i = (int16_t)i;
                e->j = (int16_t)j;
                e->s = (int16_t)s;
                e->flags = f;
            }
            ++logged;
        }
    };
    // P row/column j with its row masses, Q row/column i, X^T row j
    auto write = [&](int i, int j, const int64_t (&pj)[C], const DualEval<C>& e,
                     const int64_t (&nx)[C]) {
        int64_t dm[C];
        int64_t off = 0;
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const int k = g + c * G;
            dm[c] = (k < n && k != j) ? iabs64(e.np[c]) - iabs64(pj[c]) : 0;
            off += dm[c];
        }
        off = Gp::sum(off);
        const int Dj = rp.D(j), Di = rp.D(i);
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const int k = g + c * G;
            if (k < n) {
                rp.sP[rp.at_lane(j, Dj, c)] = e.np[c];
                rp.sQ[rp.at_lane(i, Di, c)] = e.nq[c];
                rp.rm[c] += (k != j) ? dm[c] : off;
                rp.gX[j * n + k] = nx[c];
            }
        }
    };
    auto sweep = [&](int pv, uint64_t m) -> bool {
        const double before = score;
        ++sweeps;
        for (int q = 0; q < N; ++q) {
            const int jj = sact[q];
            if (jj == pv) continue;
            int64_t pi[C], pj[C];
            double qi[C], qj[C];
            rp.load_rows(pv, jj, pi, pj, qi, qj);
            int64_t s = dual_best_shear<G, C>(pi, pj, qi, qj, g, n, pv, jj, cw);
            // lookahead groups all run the same sweep: their decisions are
            // wave-uniform
            if constexpr (K > 1) s = uni<64>(s);
            ++greedy_n;
            if (s == 0) continue;
            DualEval<C> e;
            dual_evaluate<G, C>(pi, pj, qi, qj, g, n, pv, jj, s, rp.sP[rp.at(pv, pv)],
                                rp.sP[rp.at(jj, pv)], rp.sP[rp.at(jj, jj)],
                                rp.sQ[rp.at(pv, pv)], rp.sQ[rp.at(pv, jj)],
                                rp.sQ[rp.at(jj, jj)], cw, e);
            if constexpr (K > 1) {
                e.bad = uni<64>((int32_t)e.bad) != 0;
                e.d = uni_f64(e.d);
                e.dnz = uni<64>(e.dnz);
            }
            if (e.bad || e.d > 0.0) continue;
            int64_t xi[C], xj[C], nx[C];
            rp.prefetch_xt(pv, jj, xi, xj);
            bool fits = rp.xt_row(s, xi, xj, nx);
            if constexpr (K > 1) fits = uni<64>((int32_t)fits) != 0;
            if (!fits) continue;
            leave_best();
            if constexpr (W > 1) __syncthreads();  // every wave has read what changes
            write(pv, jj, pj, e, nx);
            score += e.d;
            offnz += e.dnz;
            ++accepts;
            log_move(m, pv, jj, s, 1);
        }
        return score < before;
    };
    auto track_best = [&](uint64_t m) {
        if (score < best_score) {
            best_score = score;
            best_move = (int64_t)m;
            flags |= RF_CUR_BEST;
            flags &= ~RF_BEST_STALE;
        }
    };

    // ---- the lookahead path's steps (the one-group loop below spells the
    // same steps out inline) ----
    // pair (congruence_anneal.hpp:283-288, :211-228)
    auto choose_pair = [&](const uint4& w, int base, int cs, int sub, uint32_t w0z,
                           uint32_t w2w, uint32_t w3z, uint32_t w3w, int& i,
                           int& j) __attribute__((always_inline)) {
        if (target_coin(w0z)) {
            const bool in = cs >= 0 && cs < TS;
            const uint32_t p = in ? mulhi32(w.x, (uint32_t)N) : 0u;
            const int row = rowat(p);
            const int64_t wt = rp.mass_of(row);
            const uint64_t mx =
                Gp::slot_max(in ? (((uint64_t)wt << 5) | 16u | (uint32_t)(15 - cs)) : 0ull, sub);
            const uint32_t tpos = Gp::bcast(p, base + 15 - (int)(mx & 15));
            const int target = rowat(tpos);
            const uint32_t pp = in ? mulhi32(w.y, (uint32_t)N) : 0u;
            const int cand = rowat(pp);
            const bool valid = in && cand != target;
            const int64_t v = valid ? iabs64(rp.sP[rp.at(target, cand)]) : 0;
            const uint64_t vmax =
                Gp::slot_max(valid ? (((uint64_t)v << 5) | 16u | (uint32_t)(15 - cs)) : 0ull, sub);
            int ibest;
            if (vmax == 0) {
                ibest = rowat((tpos + 1u + mulhi32(w3w, (uint32_t)(N - 1))) % (uint32_t)N);
            } else {
                ibest = Gp::bcast(cand, base + 15 - (int)(vmax & 15));
            }
            i = ibest;
            j = target;
        } else {
            const uint32_t ip = mulhi32(w2w, (uint32_t)N);
            const uint32_t jp = (ip + 1u + mulhi32(w3z, (uint32_t)(N - 1))) % (uint32_t)N;
            i = rowat(ip);
            j = rowat(jp);
        }
    };
    // the commit of an accepted move (i, j, s, e) whose transform row fits:
    // the reference's recompute before a new global best decides a new best
    // on the recomputed score, the state it leaves flushed first if it is the
    // replica's best; `before_write` runs between the flush and the writes
    auto commit = [&](uint64_t m, int i, int j, int64_t s, const int64_t (&pj)[C],
                      const DualEval<C>& e, const int64_t (&nx)[C],
                      auto before_write) __attribute__((always_inline)) {
        const double ns = score + e.d;
        const bool rec = e.d < 0.0 && ns < best_score && ns < gb;
        bool newbest = ns < best_score;
        if (rec || !newbest) leave_best();
        before_write();
        write(i, j, pj, e, nx);
        score = ns;
        offnz += e.dnz;
        if (rec) {
            score = rp.score(cw);
            if constexpr (K > 1) score = uni_f64(score);
            ++recomputes;
            newbest = score < best_score;
        }
        if (newbest) {
            best_score = score;
            best_move = (int64_t)m;
            flags |= RF_CUR_BEST;
            flags &= ~RF_BEST_STALE;
        }
        ++accepts;
        log_move(m, i, j, s, 0);
    };
    // bookkeeping after move m and the stall handling that follows it
    // (congruence_anneal.hpp:363-412); true when the replica is solved,
    // `changed` set when a sweep commit or a reseed changed the state
    auto finish = [&](uint64_t m, int res, uint32_t w4z, uint32_t w4w,
                      bool& changed) __attribute__((always_inline)) -> bool {
        ++win_p;
        if (res >= 0) ++win_a;
        if (score < local_best) {
            local_best = score;
            msi = 0;
        } else {
            ++msi;
        }
        if (offnz == 0) {
            flags |= RF_SOLVED | RF_STOPPED;
            return true;
        }
        const uint32_t accepts0 = accepts;
        if (tempering) {
            if ((int64_t)msi > stuck) {
                const InstParams& Qp = opaque(I)->P;
                if (Qp.sweep_fraction > 0.0 && u01(w4z) < Qp.sweep_fraction) {
                    sweep(sact[mulhi32(w4w, (uint32_t)N)], m);
                    track_best(m);
                }
                local_best = score;
                msi = 0;
            }
        } else {
            if ((int64_t)msi > stuck) {
                const Inst* Iq = opaque(I);
                if (Iq->P.sweep_fraction > 0.0 && u01(w4z) < Iq->P.sweep_fraction) {
                    if (sweep(sact[mulhi32(w4w, (uint32_t)N)], m)) {
                        local_best = score;
                        msi = 0;
                    }
                    track_best(m);
                }
                const Inst* Iv = opaque(I);
                T = Iv->P.reheat * Iv->P.t_init;
                nbeta = -1.0 / T;
                if (score > gb * Iv->P.reseed_factor) {
                    leave_best();
                    log_move(m, -1, -1, 0, 2);
                    if constexpr (W > 1) __syncthreads();
                    rp.load((const GLOBAL int64_t*)Iv->gA, (const GLOBAL double*)Iv->gQ);
                    rp.copy_cols(rp.gX, (const GLOBAL int64_t*)Iv->gXt);
                    rp.compute_row_mass();
                    score = gb;
                    offnz = Iv->gb_offnz;
                    track_best(m);
                    changed = true;
                }
                local_best = score;
                msi = 0;
                cool_counter = 0;
            } else if (++cool_counter >= per_cool) {
                const InstParams& Qp = opaque(I)->P;
                cool_counter = 0;
                if (Qp.adaptive && win_p > 0) {
                    const double acc = (double)win_a / (double)win_p;
                    if (acc > Qp.target_accept)
                        cooling *= 0.999;
                    else
                        cooling /= 0.999;
                    cooling = cooling < 0.90 ? 0.90 : (cooling > 0.99999 ? 0.99999 : cooling);
                }
                win_p = 0;
                win_a = 0;
                T *= cooling;
                if (T < Qp.t_min) T = Qp.t_min;
                nbeta = -1.0 / T;
            }
        }
        changed |= accepts != accepts0;
        if (offnz == 0) {
            flags |= RF_SOLVED | RF_STOPPED;
            return true;
        }
        return false;
    };
    bool halt = false;
    constexpr int SB = G / 16;  // moves per Philox block of a group
    // this group's proposal of move mg: pair, rows, shear, evaluation, the
    // Metropolis draw and the stall words
    struct Prop {
        int i, j;
        int64_t s;
        int64_t pi[C], pj[C], xi[C], xj[C];
        DualEval<C> e;
        double u;
        bool gr;
        uint32_t w4z, w4w;
    };
    auto propose = [&](uint64_t mg, bool fetch_xt, Prop& P) __attribute__((always_inline)) {
        const uint32_t t4 = (uint32_t)((mg - 1) & 3u);
        const int sub = (int)(t4 & (uint32_t)(SB - 1));
        const uint64_t q = (mg - 1) >> 2;
        // the key as scalars (its load may have landed in vector registers)
        uint32_t k0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)key0);
        uint32_t k1 = (uint32_t)__builtin_amdgcn_readfirstlane((int)key1);
        opaque_u32(k0);
        opaque_u32(k1);
        const uint4 w = philox10((uint32_t)q, (uint32_t)(q >> 32), (uint32_t)r,
                                 16u * (t4 & ~(uint32_t)(SB - 1)) + (uint32_t)g, k0, k1);
        const int base = 16 * sub;
        const int cs = g - base;
        if (cs >= 0 && cs < 5) swords[cs] = w;
        __builtin_amdgcn_wave_barrier();
        const uint4 b0 = swords[0], b1 = swords[1], b2 = swords[2], b3 = swords[3];
        const uint4 b4 = swords[4];
        __builtin_amdgcn_wave_barrier();
        choose_pair(w, base, cs, sub, b0.z, b2.w, b3.z, b3.w, P.i, P.j);
        double qi[C], qj[C];
        rp.load_rows(P.i, P.j, P.pi, P.pj, qi, qj);
        if (fetch_xt) rp.prefetch_xt(P.i, P.j, P.xi, P.xj);
        const int64_t pm = (b1.z & 1u) ? 1 : -1;
        P.s = pm;
        P.gr = greedy_coin(b0.w);
        if (P.gr) {
            P.s = dual_best_shear<G, C>(P.pi, P.pj, qi, qj, g, n, P.i, P.j, cw);
            if (P.s == 0) P.s = pm;
        }
        dual_evaluate<G, C>(P.pi, P.pj, qi, qj, g, n, P.i, P.j, P.s, rp.sP[rp.at(P.i, P.i)],
                            rp.sP[rp.at(P.j, P.i)], rp.sP[rp.at(P.j, P.j)],
                            rp.sQ[rp.at(P.i, P.i)], rp.sQ[rp.at(P.i, P.j)],
                            rp.sQ[rp.at(P.j, P.j)], cw, P.e);
        P.u = u53(b1.w, b2.z);
        P.w4z = b4.z;
        P.w4w = b4.w;
    };

    if constexpr (K == 1) {
        uint4 w = make_uint4(0u, 0u, 0u, 0u);
        bool have_w = false;
        for (uint32_t it = 0; moves < stop_at; ++it) {
            if ((it & 255u) == 0u) {
                const unsigned long long key =
                    __hip_atomic_load(&I->solve_key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (key != kNoKey && moves >= (key >> kRepBits)) break;
                if (*stop) {
                    halt = true;
                    break;
                }
            }
            const uint64_t m = ++moves;
            // Philox slots of move m as in anneal.hip: counter (q, r, 16 t + c)
            constexpr int SUB = G / 16;
            const uint32_t t4 = (uint32_t)((m - 1) & 3u);
            const int sub = (int)(t4 & (uint32_t)(SUB - 1));
            if (sub == 0 || !have_w) {
                const uint64_t q = (m - 1) >> 2;
                uint32_t k0 = key0, k1 = key1;
                opaque_u32(k0);
                opaque_u32(k1);
                w = philox10((uint32_t)q, (uint32_t)(q >> 32), (uint32_t)r,
                             16u * (t4 & ~(uint32_t)(SUB - 1)) + (uint32_t)g, k0, k1);
                have_w = true;
            }
            const int base = 16 * sub;
            const int cs = g - base;
            uint32_t w0z, w0w, w1z, w2w, w3z;
            uint32_t l1w = 0, l2z = 0, l3w = 0, l4z = 0, l4w = 0;
            if constexpr (G == 64) {
                w0z = readlane32(w.z, base + 0);
                w0w = readlane32(w.w, base + 0);
                w1z = readlane32(w.z, base + 1);
                w2w = readlane32(w.w, base + 2);
                w3z = readlane32(w.z, base + 3);
            } else {
                if (cs >= 0 && cs < 5) swords[cs] = w;
                __builtin_amdgcn_wave_barrier();
                const uint4 b0 = swords[0], b1 = swords[1], b2 = swords[2], b3 = swords[3];
                const uint4 b4 = swords[4];
                __builtin_amdgcn_wave_barrier();
                w0z = b0.z;
                w0w = b0.w;
                w1z = b1.z;
                l1w = b1.w;
                l2z = b2.z;
                w2w = b2.w;
                w3z = b3.z;
                l3w = b3.w;
                l4z = b4.z;
                l4w = b4.w;
            }
            auto word = [&](int slot_, bool zw, uint32_t staged) -> uint32_t {
                if constexpr (G == 64) return readlane32(zw ? w.w : w.z, base + slot_);
                return staged;
            };
            auto w1w = [&]() { return word(1, true, l1w); };
            auto w2z = [&]() { return word(2, false, l2z); };
            auto w3w = [&]() { return word(3, true, l3w); };
            auto w4z = [&]() { return word(4, false, l4z); };
            auto w4w = [&]() { return word(4, true, l4w); };

            // ---- pair (congruence_anneal.hpp:283-288, :211-228) ----
            int i, j;
            if (target_coin(w0z)) {
                const bool in = cs >= 0 && cs < TS;
                const uint32_t p = in ? mulhi32(w.x, (uint32_t)N) : 0u;
                const int row = rowat(p);
                const int64_t wt = rp.mass_of(row);
                const uint64_t mx =
                    Gp::slot_max(in ? (((uint64_t)wt << 5) | 16u | (uint32_t)(15 - cs)) : 0ull, sub);
                const uint32_t tpos = Gp::bcast(p, base + 15 - (int)(mx & 15));
                const int target = rowat(tpos);
                const uint32_t pp = in ? mulhi32(w.y, (uint32_t)N) : 0u;
                const int cand = rowat(pp);
                const bool valid = in && cand != target;
                const int64_t v = valid ? iabs64(rp.sP[rp.at(target, cand)]) : 0;
                const uint64_t vmax =
                    Gp::slot_max(valid ? (((uint64_t)v << 5) | 16u | (uint32_t)(15 - cs)) : 0ull, sub);
                int ibest;
                if (vmax == 0) {
                    ibest = rowat((tpos + 1u + mulhi32(w3w(), (uint32_t)(N - 1))) % (uint32_t)N);
                } else {
                    ibest = Gp::bcast(cand, base + 15 - (int)(vmax & 15));
                }
                i = ibest;
                j = target;
            } else {
                const uint32_t ip = mulhi32(w2w, (uint32_t)N);
                const uint32_t jp = (ip + 1u + mulhi32(w3z, (uint32_t)(N - 1))) % (uint32_t)N;
                i = rowat(ip);
                j = rowat(jp);
            }

            int64_t pi[C], pj[C], xi[C], xj[C];
            double qi[C], qj[C];
            rp.load_rows(i, j, pi, pj, qi, qj);
            rp.prefetch_xt(i, j, xi, xj);

            // ---- shear (congruence_anneal.hpp:290-296) ----
            const int64_t pm = (w1z & 1u) ? 1 : -1;
            int64_t s = pm;
            if (greedy_coin(w0w)) {
                s = dual_best_shear<G, C>(pi, pj, qi, qj, g, n, i, j, cw);
                ++greedy_n;
                if (s == 0) s = pm;
            }

            // ---- attempt_add (congruence_anneal.hpp:232-240) ----
            DualEval<C> e;
            dual_evaluate<G, C>(pi, pj, qi, qj, g, n, i, j, s, rp.sP[rp.at(i, i)],
                                rp.sP[rp.at(j, i)], rp.sP[rp.at(j, j)], rp.sQ[rp.at(i, i)],
                                rp.sQ[rp.at(i, j)], rp.sQ[rp.at(j, j)], cw, e);
            int res = -1;
            if (e.bad) {
                ++infeasible;
            } else if (e.d <= 0.0 || u53(w1w(), w2z()) < xexp(e.d * nbeta)) {
                int64_t nx[C];
                if (rp.xt_row(s, xi, xj, nx)) {
                    // the reference recomputes the running score before it
                    // publishes a new global best (congruence_anneal.hpp:310-315);
                    // whether the move is a new best of the replica is then
                    // decided on the recomputed score, so the state it leaves is
                    // flushed first if it is the replica's best
                    const double ns = score + e.d;
                    const bool rec = e.d < 0.0 && ns < best_score && ns < gb;
                    bool newbest = ns < best_score;
                    if (rec || !newbest) leave_best();
                    write(i, j, pj, e, nx);
                    score = ns;
                    offnz += e.dnz;
                    if (rec) {
                        score = rp.score(cw);
                        ++recomputes;
                        newbest = score < best_score;
                    }
                    if (newbest) {
                        best_score = score;
                        best_move = (int64_t)m;
                        flags |= RF_CUR_BEST;
                        flags &= ~RF_BEST_STALE;
                    }
                    ++accepts;
                    log_move(m, i, j, s, 0);
                    res = e.d < 0.0 ? 1 : 0;
                }
            }
            ++win_p;
            if (res >= 0) ++win_a;
            if (score < local_best) {
                local_best = score;
                msi = 0;
            } else {
                ++msi;
            }
            if (offnz == 0) {
                flags |= RF_SOLVED | RF_STOPPED;
                break;
            }

            // ---- stall handling (congruence_anneal.hpp:363-412) ----
            if (tempering) {
                if ((int64_t)msi > stuck) {
                    const InstParams& Qp = opaque(I)->P;
                    if (Qp.sweep_fraction > 0.0 && u01(w4z()) < Qp.sweep_fraction) {
                        sweep(sact[mulhi32(w4w(), (uint32_t)N)], m);
                        track_best(m);
                    }
                    local_best = score;
                    msi = 0;
                }
            } else {
                if ((int64_t)msi > stuck) {
                    const Inst* Iq = opaque(I);
                    if (Iq->P.sweep_fraction > 0.0 && u01(w4z()) < Iq->P.sweep_fraction) {
                        if (sweep(sact[mulhi32(w4w(), (uint32_t)N)], m)) {
                            local_best = score;
                            msi = 0;
                        }
                        track_best(m);
                    }
                    const Inst* Iv = opaque(I);
                    T = Iv->P.reheat * Iv->P.t_init;
                    nbeta = -1.0 / T;
                    if (score > gb * Iv->P.reseed_factor) {
                        leave_best();
                        log_move(m, -1, -1, 0, 2);
                        if constexpr (W > 1) __syncthreads();
                        rp.load((const GLOBAL int64_t*)Iv->gA, (const GLOBAL double*)Iv->gQ);
                        rp.copy_cols(rp.gX, (const GLOBAL int64_t*)Iv->gXt);
                        rp.compute_row_mass();
                        score = gb;
                        offnz = Iv->gb_offnz;
                        track_best(m);
                    }
                    local_best = score;
                    msi = 0;
                    cool_counter = 0;
                } else if (++cool_counter >= per_cool) {
                    const InstParams& Qp = opaque(I)->P;
                    cool_counter = 0;
                    if (Qp.adaptive && win_p > 0) {
                        const double acc = (double)win_a / (double)win_p;
                        if (acc > Qp.target_accept)
                            cooling *= 0.999;
                        else
                            cooling /= 0.999;
                        cooling = cooling < 0.90 ? 0.90 : (cooling > 0.99999 ? 0.99999 : cooling);
                    }
                    win_p = 0;
                    win_a = 0;
                    T *= cooling;
                    if (T < Qp.t_min) T = Qp.t_min;
                    nbeta = -1.0 / T;
                }
            }
            if (offnz == 0) {
                flags |= RF_SOLVED | RF_STOPPED;
                break;
            }
        }
    } else if constexpr (W == 1) {
        // lookahead in one wave: group kk proposes move m0 + kk
        for (uint32_t it = 0; moves < stop_at; ++it) {
            if ((it & 255u) == 0u) {
                const unsigned long long key =
                    __hip_atomic_load(&I->solve_key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (key != kNoKey && moves >= (key >> kRepBits)) break;
                if (*stop) {
                    halt = true;
                    break;
                }
            }
            const uint64_t m0 = moves + 1;
            Prop P;
            propose(m0 + (uint64_t)kk, true, P);
            // Metropolis at the step's temperature; a move after a cooling
            // step inside this step is decided again at the new one
            const double nb0 = nbeta;
            const bool pass = !P.e.bad && (P.e.d <= 0.0 || P.u < xexp(P.e.d * nb0));
            const int kn = (stop_at - moves) < (uint64_t)K ? (int)(stop_at - moves) : K;
            const uint64_t live = kn == K ? ~0ull : ((1ull << (kn * G)) - 1ull);
            if ((__ballot(pass) & live) == 0ull && score >= local_best &&
                (int64_t)msi + kn <= (int64_t)stuck && (tempering || cool_counter + kn < per_cool)) {
                // every resolvable move is rejected and none reaches a stall or
                // a cooling step: the sequential bookkeeping, kn times
                const uint64_t lead = __ballot(g == 0) & live;
                greedy_n += (uint32_t)__popcll(__ballot(P.gr) & lead);
                infeasible += (uint32_t)__popcll(__ballot(P.e.bad) & lead);
                moves += (uint64_t)kn;
                win_p += kn;
                msi += (uint32_t)kn;
                if (!tempering) cool_counter += kn;
                continue;
            }
            int64_t nx[C];
            bool feas = false, have_nx = false, solved = false;
            for (int k = 0; k < kn; ++k) {
                const int L = k * G;  // first lane of group k
                const uint64_t m = m0 + (uint64_t)k;
                moves = m;
                if (readlane32((uint32_t)P.gr, L)) ++greedy_n;
                int res = -1;
                bool changed = false;
                if (readlane32((uint32_t)P.e.bad, L)) {
                    ++infeasible;
                } else {
                    const double d = readlane_f64(P.e.d, L);
                    bool acc = readlane32((uint32_t)pass, L) != 0;
                    if (nbeta != nb0) acc = d <= 0.0 || readlane_f64(P.u, L) < xexp(d * nbeta);
                    if (acc && !have_nx) {
                        feas = rp.xt_row(P.s, P.xi, P.xj, nx);  // every group's, once
                        have_nx = true;
                    }
                    if (acc && readlane32((uint32_t)feas, L)) {
                        // group k's move, written by every group
                        const int src = L + g;
                        int64_t cj[C], cx[C];
                        DualEval<C> ce;
#pragma unroll
                        for (int c = 0; c < C; ++c) {
                            cj[c] = bperm64(P.pj[c], src);
                            ce.np[c] = bperm64(P.e.np[c], src);
                            ce.nq[c] = __builtin_bit_cast(
                                double, bperm64(__builtin_bit_cast(int64_t, P.e.nq[c]), src));
                            cx[c] = bperm64(nx[c], src);
                        }
                        ce.d = d;
                        ce.dnz = (int32_t)readlane32((uint32_t)P.e.dnz, L);
                        ce.bad = false;
                        commit(m, (int)readlane32((uint32_t)P.i, L), (int)readlane32((uint32_t)P.j, L),
                               readlane64(P.s, L), cj, ce, cx, [] {});
                        res = d < 0.0 ? 1 : 0;
                        changed = true;
                    }
                }
                if (finish(m, res, readlane32(P.w4z, L), readlane32(P.w4w, L), changed)) {
                    solved = true;
                    break;
                }
                if (changed) break;
            }
            if (solved) break;
        }
    } else {
        // lookahead over W waves: group kk proposes move m0 + kk and publishes it
        const int ln = (int)(threadIdx.x & 63u);
        for (uint32_t it = 0; moves < stop_at; ++it) {
            const uint64_t m0 = moves + 1;
            Prop P;
            propose(m0 + (uint64_t)kk, false, P);
            const double nb0 = nbeta;
            const bool pass = !P.e.bad && (P.e.d <= 0.0 || P.u < xexp(P.e.d * nb0));
            DualProposal* const pb = props + (it & 1u) * K;
            if (g == 0) {
                DualProposal Q;
                Q.d = P.e.d;
                Q.s = P.s;
                Q.u = P.u;
                Q.i = P.i;
                Q.j = P.j;
                Q.w4z = P.w4z;
                Q.w4w = P.w4w;
                Q.flags = (P.e.bad ? 1u : 0u) | (pass ? 2u : 0u) | (P.gr ? 4u : 0u);
                pb[kk] = Q;
            }
            // the solve key and the stop flag, read once for every wave
            if ((it & 255u) == 0u && threadIdx.x == 0) {
                const unsigned long long key =
                    __hip_atomic_load(&I->solve_key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                *spoll = (key != kNoKey && moves >= (key >> kRepBits)) ? 1u : (*stop ? 2u : 0u);
            }
            __syncthreads();
            if ((it & 255u) == 0u) {
                const uint32_t pr = (uint32_t)uni<64>((int32_t)*spoll);
                if (pr == 2u) halt = true;
                if (pr != 0u) break;
            }
            const int kn = (stop_at - moves) < (uint64_t)K ? (int)(stop_at - moves) : K;
            const uint32_t fl = ln < kn ? pb[ln].flags : 0u;  // proposal ln's flags in lane ln
            if (__ballot((fl & 2u) != 0u) == 0ull && score >= local_best &&
                (int64_t)msi + kn <= (int64_t)stuck && (tempering || cool_counter + kn < per_cool)) {
                greedy_n += (uint32_t)__popcll(__ballot((fl & 4u) != 0u));
                infeasible += (uint32_t)__popcll(__ballot((fl & 1u) != 0u));
                moves += (uint64_t)kn;
                win_p += kn;
                msi += (uint32_t)kn;
                if (!tempering) cool_counter += kn;
                continue;
            }
            bool solved = false;
            for (int k = 0; k < kn; ++k) {
                const DualProposal* const Q = pb + k;
                const uint32_t f = (uint32_t)uni<64>((int32_t)Q->flags);
                const uint64_t m = m0 + (uint64_t)k;
                moves = m;
                if (f & 4u) ++greedy_n;
                int res = -1;
                bool changed = false;
                if (f & 1u) {
                    ++infeasible;
                } else {
                    const double d = uni_f64(Q->d);
                    bool acc = (f & 2u) != 0u;
                    if (nbeta != nb0) acc = d <= 0.0 || uni_f64(Q->u) < xexp(d * nbeta);
                    if (acc) {
                        // every group evaluates move k again from the shared
                        // state: its rows, its X^T columns
                        const int ik = uni<64>(Q->i), jk = uni<64>(Q->j);
                        const int64_t sk = uni<64>(Q->s);
                        int64_t ci[C], cj[C], xi[C], xj[C], nx[C];
                        double qi[C], qj[C];
                        rp.load_rows(ik, jk, ci, cj, qi, qj);
                        rp.prefetch_xt(ik, jk, xi, xj);
                        DualEval<C> ce;
                        dual_evaluate<G, C>(ci, cj, qi, qj, g, n, ik, jk, sk, rp.sP[rp.at(ik, ik)],
                                            rp.sP[rp.at(jk, ik)], rp.sP[rp.at(jk, jk)],
                                            rp.sQ[rp.at(ik, ik)], rp.sQ[rp.at(ik, jk)],
                                            rp.sQ[rp.at(jk, jk)], cw, ce);
                        ce.d = d;
                        ce.dnz = uni<64>(ce.dnz);
                        if (uni<64>((int32_t)rp.xt_row(sk, xi, xj, nx)) != 0) {
                            // every wave has read what changes before any writes
                            commit(m, ik, jk, sk, cj, ce, nx, [] { __syncthreads(); });
                            res = d < 0.0 ? 1 : 0;
                            changed = true;
                        }
                    }
                }
                if (finish(m, res, (uint32_t)uni<64>((int32_t)Q->w4z),
                           (uint32_t)uni<64>((int32_t)Q->w4w), changed)) {
                    solved = true;
                    break;
                }
                if (changed) break;
            }
            if (solved) break;
        }
    }
    if (!(flags & RF_STOPPED) && !halt && cap != 0 && moves >= cap) flags |= RF_STOPPED | RF_CAPPED;

    if (kk != 0) return;  // lookahead: group 0 writes the replica back
    rp.store(gPcur, gQcur);
    if (g == 0) {
        S->moves = moves;
        S->dscore = score;
        S->offnz = offnz;
        S->dlocal_best = local_best;
        S->msi = (int64_t)msi;
        S->T = T;
        S->cooling = cooling;
        S->cool_counter = cool_counter;
        S->win_p = win_p;
        S->win_a = win_a;
        S->dbest_score = best_score;
        S->best_move = best_move;
        S->flags = flags;
        S->accepts += accepts;
        S->greedy += greedy_n;
        S->infeasible += infeasible;
        S->sweeps += sweeps;
        S->recomputes += recomputes;
        S->log_count += logged;
        if (flags & RF_SOLVED) S->solve_move = (int64_t)moves;
    }
    if ((flags & RF_SOLVED) && g == 0)
        atomicMin(&I->solve_key, ((unsigned long long)moves << kRepBits) | (unsigned)r);
}

// ---------------------------------------------------------------------------
// maintenance before an epoch (one workgroup of 256 threads per replica)
// ---------------------------------------------------------------------------

// The fold-and-pair column sum over x[0..n) in LDS (thread 0's result).
__device__ double block_pairsum(double* buf, int n) {
    const int t = (int)threadIdx.x;
    if (t < 64) {
        const double a = t < n ? buf[t] : 0.0;
        const double b = t + 64 < n ? buf[t + 64] : 0.0;
        buf[128 + t] = a + b;
    }
    __syncthreads();
    double v = 0.0;
    if (t == 0) {
        double* b = buf + 128;
        for (int w = 32; w >= 1; w /= 2)
            for (int k = 0; k < w; ++k) b[k] = b[2 * k] + b[2 * k + 1];
        v = b[0];
    }
    __syncthreads();
    return v;
}

__global__ void __launch_bounds__(256) k_dual_maint(Inst* __restrict__ insts) {
    __shared__ double s_buf[256];
    __shared__ double s_f[kMaxGroupN];
    __shared__ double s_max[256];
    __shared__ int s_arg[256];
    Inst* I = insts + blockIdx.y;
    if (!I->dual || I->status != ST_RUNNING) return;
    const int r = (int)blockIdx.x;
    if (r >= I->R) return;
    RepState* S = I->rs + r;
    if (S->flags & RF_STOPPED) return;
    const int t = (int)threadIdx.x;
    const int n = I->n;
    const int64_t nn = (int64_t)n * n;
    const uint64_t m0 = S->m_maint, m1 = S->moves;
    const uint64_t per_r = (uint64_t)(I->dual_refresh > 0 ? I->dual_refresh : 0);
    const uint64_t per_c = (uint64_t)(I->dual_check > 0 ? I->dual_check : 0);
    const double cw = I->c;
    bool refresh = per_r && m1 / per_r > m0 / per_r;
    const bool check = !refresh && per_c && m1 / per_c > m0 / per_c;
    bool rescore = S->dc != cw;
    const int64_t* P = I->A + (int64_t)r * nn;
    double* Q = I->Q + (int64_t)r * nn;

    if (check) {
        // max over rows of |(P Q)_rk - delta_rk|, each row summed in order
        const int k = (int)(m1 / per_c * per_c % (uint64_t)n);
        double a = 0.0;
        if (t < n) {
            double acc = 0.0;
            for (int u = 0; u < n; ++u) {
                const double pr = (double)P[(int64_t)t * n + u] * Q[(int64_t)u * n + k];
                acc += pr;
            }
            if (t == k) acc -= 1.0;
            a = __builtin_fabs(acc);
            if (!(a > 0.0)) a = 0.0;  // NaN never raises the maximum
        }
        s_max[t] = a;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if (t < o) s_max[t] = s_max[t + o] > s_max[t] ? s_max[t + o] : s_max[t];
            __syncthreads();
        }
        refresh = s_max[0] > I->dual_tol;
        __syncthreads();
    }
    if (refresh) {
        // Gauss-Jordan on [P | I] (invert_to, mat_io.hpp:297-337); each
        // element sees the reference's operations in the same order
        double* A = I->work + (int64_t)r * nn * 2;
        const int w = 2 * n;
        for (int q = t; q < n * w; q += 256) {
            const int row = q / w, col = q % w;
            A[q] = col < n ? (double)P[(int64_t)row * n + col] : (col - n == row ? 1.0 : 0.0);
        }
        bool singular = false;
        __syncthreads();
        for (int col = 0; col < n; ++col) {
            // pivot: first row of maximal |A(r, col)|, r >= col (a NaN is
            // never preferred to the current row)
            double v = -1.0;
            int arg = n;
            if (t >= col && t < n) {
                v = __builtin_fabs(A[(int64_t)t * w + col]);
                if (v != v) v = t == col ? __builtin_inf() : -1.0;
                arg = t;
            }
            s_max[t] = v;
            s_arg[t] = arg;
            __syncthreads();
            for (int o = 128; o > 0; o >>= 1) {
                if (t < o) {
                    const double v2 = s_max[t + o];
                    const int a2 = s_arg[t + o];
                    if (v2 > s_max[t] || (v2 == s_max[t] && a2 < s_arg[t])) {
                        s_max[t] = v2;
                        s_arg[t] = a2;
                    }
                }
                __syncthreads();
            }
            const int piv = s_arg[0];
            const double best = __builtin_fabs(A[(int64_t)piv * w + col]);
            __syncthreads();
            if (best == 0.0) {  // singular: the dual stays as it is (every thread)
                singular = true;
                break;
            }
            if (piv != col)
                for (int q = t; q < w; q += 256) {
                    const double a = A[(int64_t)piv * w + q];
                    A[(int64_t)piv * w + q] = A[(int64_t)col * w + q];
                    A[(int64_t)col * w + q] = a;
                }
            __syncthreads();
            const double inv = 1.0 / A[(int64_t)col * w + col];
            __syncthreads();
            for (int q = t; q < w; q += 256) A[(int64_t)col * w + q] *= inv;
            __syncthreads();
            if (t < n) s_f[t] = A[(int64_t)t * w + col];
            __syncthreads();
            for (int q = t; q < n * w; q += 256) {
                const int row = q / w, c = q % w;
                const double f = s_f[row];
                if (row == col || f == 0.0) continue;
                const double prod = f * A[(int64_t)col * w + c];
                A[q] = A[q] - prod;
            }
            __syncthreads();
        }
        if (!singular) {
            // the current state may be the replica's unflushed best: the
            // best keeps the dual it was found with, so flush it first
            const int32_t fl = S->flags;
            if (fl & RF_CUR_BEST) {
                const bool keep = S->dbest_score <= I->gb_dscore;
                if (keep) {
                    const int64_t off = (int64_t)r * nn;
                    for (int q = t; q < (int)nn; q += 256) {
                        I->bA[off + q] = P[q];
                        I->bQ[off + q] = Q[q];
                        I->bXt[off + q] = I->Xt[off + q];
                    }
                }
                __syncthreads();
                if (t == 0)
                    S->flags = (fl & ~(RF_CUR_BEST | RF_BEST_STALE)) | (keep ? 0 : RF_BEST_STALE);
            }
            for (int q = t; q < (int)nn; q += 256) {
                const int a = q / n, b = q % n;
                const double x = A[(int64_t)a * w + n + b];
                const double y = A[(int64_t)b * w + n + a];
                Q[q] = a == b ? x : 0.5 * (a < b ? x + y : y + x);
            }
            if (t == 0) ++S->refreshes;
            rescore = true;
        }
        __syncthreads();
    }
    if (rescore) {
        // |offdiag P|^2 and |offdiag Q|^2: per column the rows in order, then
        // fold and pair the columns
        double sp = 0.0, sq = 0.0;
        if (t < n)
            for (int u = 0; u < n; ++u) {
                if (u == t) continue;
                const double v = (double)P[(int64_t)u * n + t];
                const double x = Q[(int64_t)u * n + t];
                sp += v * v;
                sq += x * x;
            }
        if (t < 128) s_buf[t] = sp;
        __syncthreads();
        const double fp = block_pairsum(s_buf, n);
        if (t < 128) s_buf[t] = sq;
        __syncthreads();
        const double fq = block_pairsum(s_buf, n);
        if (t == 0) {
            const double cq = cw * fq;
            S->dscore = fp + cq;
            S->dc = cw;
        }
    }
    if (t == 0) {
        S->m_maint = m1;
        if (check) ++S->checks;
    }
}

// ---------------------------------------------------------------------------
// probe (parity tests of the primitives)
// ---------------------------------------------------------------------------

// One query per group: q = {mat, i, j, s}; P/Q row-major matrices; out per
// query: {feasible, dnz, best_shear(i, j)} as int64 and {d} as double, then
// the proposed P row j and Q row i.
template <int G, int C>
__global__ void __launch_bounds__(64)
    k_dual_probe(const int64_t* __restrict__ Pm, const double* __restrict__ Qm, int n, double cw,
                 const int32_t* __restrict__ qs, int nq, int64_t* __restrict__ out_i,
                 double* __restrict__ out_d, int64_t* __restrict__ rows_p,
                 double* __restrict__ rows_q, double* __restrict__ scores) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int g = (int)threadIdx.x & (G - 1);
    const int slot = (int)threadIdx.x / G;
    const int q = (int)blockIdx.x * (64 / G) + slot;
    const int nmax = (n + 7) / 8 * 8;
    const size_t tri = (size_t)nmax * (nmax + 1) / 2;
    DualReplica<G, C> rp;
    rp.sP = (int64_t*)(smem + (size_t)slot * tri * 16);
    rp.sQ = (double*)(rp.sP + tri);
    rp.n = n;
    rp.g = g;
    rp.init_lanes();
    if (q >= nq) return;
    const int32_t* x = qs + 4 * q;
    const int mat = x[0], i = x[1], j = x[2];
    const int64_t s0 = x[3];
    const int64_t off = (int64_t)mat * n * n;
    rp.load((const GLOBAL int64_t*)(Pm + off), (const GLOBAL double*)(Qm + off));
    __builtin_amdgcn_wave_barrier();
    int64_t pi[C], pj[C];
    double qi[C], qj[C];
    rp.load_rows(i, j, pi, pj, qi, qj);
    const int64_t best = dual_best_shear<G, C>(pi, pj, qi, qj, g, n, i, j, cw);
    DualEval<C> e;
    dual_evaluate<G, C>(pi, pj, qi, qj, g, n, i, j, s0, rp.sP[rp.at(i, i)], rp.sP[rp.at(j, i)],
                        rp.sP[rp.at(j, j)], rp.sQ[rp.at(i, i)], rp.sQ[rp.at(i, j)],
                        rp.sQ[rp.at(j, j)], cw, e);
    const double sc = rp.score(cw);
    if (g == 0) {
        out_i[3 * q + 0] = e.bad ? 0 : 1;
        out_i[3 * q + 1] = e.dnz;
        out_i[3 * q + 2] = best;
        out_d[q] = e.d;
        scores[q] = sc;
    }
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const int k = g + c * G;
        if (k < n) {
            rows_p[(int64_t)q * n + k] = e.np[c];
            rows_q[(int64_t)q * n + k] = e.nq[c];
        }
    }
}

// ---------------------------------------------------------------------------
// host entry points
// ---------------------------------------------------------------------------

using DualKernel = void (*)(const RepRef*, int, Inst*, int, int, const volatile int*);
struct DualVariant {
    int G, C, K;
    DualKernel fn;
};

// The group width is fixed by the order (the column sums' bits depend on it);
// lookahead adds K groups per replica: K G = 64 in one wave, above in a
// workgroup of K G / 64 waves.
static const DualVariant kDualVariants[] = {
    {16, 1, 1, k_anneal_dual<16, 1, 1>},   {32, 1, 1, k_anneal_dual<32, 1, 1>},
    {64, 1, 1, k_anneal_dual<64, 1, 1>},   {64, 2, 1, k_anneal_dual<64, 2, 1>},
    {16, 1, 4, k_anneal_dual<16, 1, 4>},   {32, 1, 2, k_anneal_dual<32, 1, 2>},
    {16, 1, 16, k_anneal_dual<16, 1, 16>}, {32, 1, 8, k_anneal_dual<32, 1, 8>},
    {64, 1, 4, k_anneal_dual<64, 1, 4>},
};

bool dual_variant_exists(int G, int C, int K) {
    for (const auto& v : kDualVariants)
        if (v.G == G && v.C == C && v.K == K) return true;
    return false;
}

hipError_t init_dual_kernel_attrs(int max_lds_per_block) {
    for (const auto& v : kDualVariants) {
        const hipError_t e = hipFuncSetAttribute(
            (const void*)v.fn, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds_per_block);
        if (e != hipSuccess) return e;
    }
    const void* probes[] = {(const void*)&k_dual_probe<16, 1>, (const void*)&k_dual_probe<32, 1>,
                            (const void*)&k_dual_probe<64, 1>, (const void*)&k_dual_probe<64, 2>};
    for (const void* f : probes) {
        const hipError_t e =
            hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds_per_block);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// units_per_block: table entries (replicas, or padding) per block, G K lanes each
hipError_t launch_anneal_dual(hipStream_t st, int G, int C, int K, int units_per_block,
                              const RepRef* table, int count, Inst* insts, int nmax,
                              int rep_bytes, const volatile int* stop) {
    if (count <= 0) return hipSuccess;
    const DualVariant* v = nullptr;
    for (const auto& x : kDualVariants)
        if (x.G == G && x.C == C && x.K == K) v = &x;
    if (!v || units_per_block < 1 || units_per_block * G * K > 256) return hipErrorInvalidValue;
    const int threads = units_per_block * G * K;
    const int blocks = (count + units_per_block - 1) / units_per_block;
    const size_t lds = (size_t)units_per_block * rep_bytes;
    hipLaunchKernelGGL(v->fn, dim3(blocks), dim3(threads), lds, st, table, count, insts, nmax,
                       rep_bytes, stop);
    return hipGetLastError();
}

hipError_t launch_dual_maint(hipStream_t st, Inst* insts, int count, int max_r) {
    if (count <= 0 || max_r <= 0) return hipSuccess;
    k_dual_maint<<<dim3((unsigned)max_r, (unsigned)count), 256, 0, st>>>(insts);
    return hipGetLastError();
}

hipError_t launch_dual_probe(hipStream_t st, int G, int C, const int64_t* P, const double* Q,
                             int n, double cw, const int32_t* qs, int nq, int64_t* out_i,
                             double* out_d, int64_t* rows_p, double* rows_q, double* scores) {
    if (nq <= 0) return hipSuccess;
    const int per = 64 / G;
    const int blocks = (nq + per - 1) / per;
    const int nmax = (n + 7) / 8 * 8;
    const size_t lds = (size_t)per * nmax * (nmax + 1) / 2 * 16;
#define XTX_DUAL_PROBE(GG, CC)                                                                \
    k_dual_probe<GG, CC><<<blocks, 64, lds, st>>>(P, Q, n, cw, qs, nq, out_i, out_d, rows_p, \
                                                  rows_q, scores)
    if (G == 16 && C == 1)
        XTX_DUAL_PROBE(16, 1);
    else if (G == 32 && C == 1)
        XTX_DUAL_PROBE(32, 1);
    else if (G == 64 && C == 1)
        XTX_DUAL_PROBE(64, 1);
    else if (G == 64 && C == 2)
        XTX_DUAL_PROBE(64, 2);
    else
        return hipErrorInvalidValue;
#undef XTX_DUAL_PROBE
    return hipGetLastError();
}

}  // namespace xtx
