// lane_epoch.h -- the annealer with one proposal per lane (gfx950): one epoch
// of a wave's replicas, shared by the per-epoch launches (anneal_lane.hip) and
// the persistent batch kernel (anneal_persist.hip).
//
// A replica takes Q lanes (Q = 16: four replicas per wave). Each step, lane q
// of the replica makes and evaluates the replica's move m0 + q on its current
// state, serially over the columns -- the pair tournaments, the exact best
// shear and the score change (lane_eval.h), the Metropolis test -- with no
// reduction across lanes: at the batch's orders a proposal is a few thousand
// integer instructions of one lane instead of a chain of group reductions.
// The moves are then resolved in order exactly as the reference's worker loop
// decides them (congruence_anneal.hpp:276-413): rejected moves before the
// first accepted one (or before a stall or cooling step) are booked at once;
// the first state change (a commit, a sweep commit, a reseed) ends the step,
// and the proposals after it are made again on the new state. A cooling step
// inside a step re-decides the later proposals at the new temperature. Every
// resolved move is thus evaluated on the state, and decided at the
// temperature, of the sequential loop: the trajectory is the one-group
// kernel's and the model's, bit for bit.
//
// State: the replica's A stays exact (int64) in HBM, and its low 16 bits sit
// in LDS as the square NMAX x NMAX (row r at r NMAX, zero past n), with the
// row masses (int64), the number of entries of each row at or above 2^15
// (diagonal included) and the active list. A lane reads a row pair of small
// entries (below 2^15, so the 16-bit word is the entry) from LDS, eight
// columns per 16-byte read, with no address arithmetic per column; a proposal
// on a row with an entry of 2^15 or more (the batch's planted instances start
// below 2^14) is evaluated by the whole group from HBM (group_eval.h), and so
// is every pair of a stall sweep. 16-bit words halve the LDS of a replica
// against 32-bit ones, and LDS is what bounds the waves resident per compute
// unit. In HBM, entry (r, c)
// of A and of X^T is read and written by lane c % Q of the replica only (a
// commit's column-j writes go through that lane), so every lane reads back
// only global data it wrote itself.
#pragma once
#include <hip/hip_runtime.h>

#include "device_common.h"
#include "engine_types.h"
#include "group_eval.h"
#include "lane_eval.h"

namespace xtx {

// LDS byte offsets of one replica at kernel order NMAX (WIDE: no square).
template <int NMAX, bool WIDE = false>
struct LaneLds {
    static constexpr int kSquare = 0;                         // int16 [NMAX][NMAX]
    static constexpr int kMass = WIDE ? 0 : 2 * NMAX * NMAX;  // int64 [NMAX]
    static constexpr int kBig = kMass + 8 * NMAX;  // int32 [NMAX]
    static constexpr int kAct = kBig + 4 * NMAX;   // uint8 [NMAX]
    static constexpr int kBytes = (kAct + NMAX + 15) / 16 * 16;
};

// A proposal's outcome: the shear (the greedy best, or the given +-1 when that
// is 0 or not asked for), its score change and off-diagonal change; ok = false
// when it leaves the bound.
struct EvalOut {
    int32_t s;
    int64_t delta, off;
    bool ok;
};

template <int Q>
__device__ __forceinline__ uint32_t group_bits(bool pred, int gbase) {
    const uint64_t b = __ballot(pred);
    if constexpr (Q == 32)
        return (uint32_t)(b >> gbase);
    else
        return (uint32_t)(b >> gbase) & ((1u << Q) - 1u);
}
__device__ __forceinline__ int32_t from_lane32(int32_t v, int src) {
    return __builtin_amdgcn_ds_bpermute(src << 2, v);
}
__device__ __forceinline__ int64_t from_lane64(int64_t v, int src) {
    const int lo = __builtin_amdgcn_ds_bpermute(src << 2, (int)(uint32_t)(uint64_t)v);
    const int hi = __builtin_amdgcn_ds_bpermute(src << 2, (int)(uint32_t)((uint64_t)v >> 32));
    return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// The replica: its LDS and HBM state and the lane's place in its group.
// WIDE (orders above 128): no LDS square; every row counts as holding large
// entries, so every proposal is evaluated by the lane group from HBM.
template <int NMAX, int Q, bool WIDE = false>
struct LaneRep {
    static constexpr int CQ = (NMAX + Q - 1) / Q;  // columns per lane: qd + c Q
    SHARED int16_t* sS;
    SHARED int64_t* sM;
    SHARED int32_t* sBig;
    SHARED uint8_t* sAct;
    GLOBAL int64_t* gA;  // A, exact, row-major n x n
    int n, N, qd, gbase;
    bool ident;

    __device__ __forceinline__ int act(uint32_t p) const { return ident ? (int)p : (int)sAct[p]; }
    __device__ __forceinline__ int32_t s32(int a, int b) const { return sS[a * NMAX + b]; }
    // entries below this magnitude are exact in the LDS words
    static constexpr int64_t kSmall = (int64_t)1 << 15;
    __device__ __forceinline__ bool small(int r) const {
        if constexpr (WIDE) return false;
        return sBig[r] == 0;
    }

    // LDS from gA: the low words (zero past n), row masses and big counts;
    // lane qd reads columns qd + c Q. Group-uniform.
    __device__ __forceinline__ void load_lds() {
        for (int r = 0; r < NMAX; ++r) {
            int64_t mass = 0;
            int32_t big = 0;
#pragma unroll
            for (int c = 0; c < CQ; ++c) {
                const int k = qd + c * Q;
                const int64_t v = (r < n && k < n) ? gA[r * n + k] : 0;
                if (!WIDE && k < NMAX) sS[r * NMAX + k] = (int16_t)v;
                mass += k != r ? iabs64(v) : 0;
                big += iabs64(v) >= kSmall;
            }
            mass = Grp<Q>::sum(mass);
            big = Grp<Q>::sum32(big);
            if (qd == 0) {
                sM[r] = mass;
                sBig[r] = big;
            }
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }

    // rows i, j of small entries for one lane (lane_eval.h): eight columns
    // per 16-byte read, two per 32-bit word (low half first)
    __device__ __forceinline__ void pair_small(int i, int j, lane::Pair<NMAX>& pr) const {
        typedef int32_t v4i __attribute__((ext_vector_type(4)));
        const SHARED v4i* ri = (const SHARED v4i*)(sS + i * NMAX);
        const SHARED v4i* rj = (const SHARED v4i*)(sS + j * NMAX);
        auto two = [&](int k, int32_t wa, int32_t wb) __attribute__((always_inline)) {
            lane::to_pq((int32_t)(int16_t)wa, (int32_t)(int16_t)wb, pr.p[k], pr.qb[k]);
            lane::to_pq(wa >> 16, wb >> 16, pr.p[k + 1], pr.qb[k + 1]);
        };
#pragma unroll
        for (int v = 0; v < NMAX / 8; ++v) {
            const v4i a = ri[v], b = rj[v];
            two(8 * v, a.x, b.x);
            two(8 * v + 2, a.y, b.y);
            two(8 * v + 4, a.z, b.z);
            two(8 * v + 6, a.w, b.w);
        }
        pr.a_ii = s32(i, i);
        pr.a_ji = s32(j, i);
        pr.a_jj = s32(j, j);
        lane::to_pq(pr.a_ji, pr.a_jj, pr.pj, pr.qbj);
    }

    // entry (r, c) of gA, read by its lane and given to the whole group
    __device__ __forceinline__ int64_t entry64(int r, int c) const {
        const int64_t v = qd == (c & (Q - 1)) ? gA[r * n + c] : 0;
        return from_lane64(v, gbase + (c & (Q - 1)));
    }
};

// A pair evaluated by the lane group (lane qd: columns qd + c Q) from the
// replica's A in HBM, exact int64 and bound-checked (group_eval.h): proposals
// on rows with an entry of 2^16 or more, and the pairs of a stall sweep. Inline
// with no call inside: a call would make the kernel keep its live values in
// memory or in the registers the calling convention saves. Group-uniform.
template <int NMAX, int Q>
__device__ __forceinline__ EvalOut eval_group(const GLOBAL int64_t* gA, int n, int qd, int gbase,
                                           int i, int j, bool greedy, int32_t pm) {
    constexpr int CQ = LaneRep<NMAX, Q>::CQ;
    int64_t ai[CQ], aj[CQ];
#pragma unroll
    for (int c = 0; c < CQ; ++c) {
        const int k = qd + c * Q;
        ai[c] = k < n ? gA[i * n + k] : 0;
        aj[c] = k < n ? gA[j * n + k] : 0;
    }
    auto entry = [&](int r, int c) -> int64_t {
        const int64_t v = qd == (c & (Q - 1)) ? gA[r * n + c] : 0;
        return from_lane64(v, gbase + (c & (Q - 1)));
    };
    const int64_t a_ii = entry(i, i), a_ji = entry(j, i), a_jj = entry(j, j);
    int64_t s = 0, off = 0;
    if (greedy) s = best_shear_inl<Q, CQ>(ai, aj, qd, n, j, a_ii, a_ji, a_jj, &off);
    const bool know = s != 0;
    if (s == 0) s = pm;
    MoveEval<CQ> e;
    evaluate<Q, CQ>(ai, aj, qd, n, j, s, a_ii, a_ji, a_jj, e, know, off);
    EvalOut o;
    o.s = (int32_t)s;
    o.delta = e.delta;
    o.off = e.off;
    o.ok = !e.bad;
    return o;
}

// Waves per SIMD the kernel is compiled for: two, except orders 40 and 48 at
// 16 proposals per step, which would spill; at 32 proposals those orders hold
// one extra value in memory outside the step loop.
constexpr int lane_waves_per_simd(int nmax, int Q) { return nmax <= 32 || Q == 32 ? 2 : 1; }

// One epoch of a wave's replicas at kernel order NMAX, Q lanes (proposals)
// each: replica r (-1: none) of instance `inst`, the same instance for the
// whole wave, its LDS at `base`. Inline with no call: the kernels hold several
// orders.
template <int NMAX, int Q, bool WIDE>
__device__ __attribute__((always_inline)) inline void lane_epoch(
    Inst* __restrict__ insts, int inst, int r, SHARED unsigned char* base,
    const volatile int* __restrict__ stop) {
    static_assert(Q == 16 || Q == 32, "16 or 32 lanes per replica");
    static_assert(NMAX % 8 == 0, "orders are multiples of 8");
    static_assert(WIDE || NMAX <= 128, "orders above 128 run WIDE");
    using L = LaneLds<NMAX, WIDE>;
    constexpr int CQ = LaneRep<NMAX, Q, WIDE>::CQ;
    const int lane = (int)(threadIdx.x & 63u);
    const int qd = lane & (Q - 1);
    const int gbase = lane & ~(Q - 1);
    GLOBAL Inst* I = (GLOBAL Inst*)insts + inst;
    if (I->status != ST_RUNNING) return;
    if (r < 0) return;  // padding
    GLOBAL RepState* S = (GLOBAL RepState*)I->rs + r;
    if (S->flags & RF_STOPPED) return;

    LaneRep<NMAX, Q, WIDE> rp;
    rp.sS = (SHARED int16_t*)(base + L::kSquare);
    rp.sM = (SHARED int64_t*)(base + L::kMass);
    rp.sBig = (SHARED int32_t*)(base + L::kBig);
    rp.sAct = (SHARED uint8_t*)(base + L::kAct);
    rp.n = I->n;
    rp.N = I->N;
    rp.qd = qd;
    rp.gbase = gbase;
    const int n = rp.n, N = rp.N;
    const int64_t nn = (int64_t)n * n;
    rp.gA = (GLOBAL int64_t*)I->A + (int64_t)r * nn;
    GLOBAL int64_t* const gX = (GLOBAL int64_t*)I->Xt + (int64_t)r * nn;
    const GLOBAL uint8_t* gact = (const GLOBAL uint8_t*)I->act;
    bool nonid = false;
    for (int t = qd; t < N; t += Q) {
        const uint8_t a = gact[t];
        rp.sAct[t] = a;
        nonid |= a != t;
    }
    rp.ident = N == n && __ballot(nonid) == 0ull;  // the wave's instance: uniform
    rp.load_lds();

    const uint32_t k0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)I->key0);
    const uint32_t k1 = (uint32_t)__builtin_amdgcn_readfirstlane((int)I->key1);
    const uint64_t epoch_end = (uint64_t)(I->epoch + 1u) * (uint64_t)I->P.epoch_moves;
    const uint64_t cap = I->P.max_moves;
    const uint64_t stop_at = (cap != 0 && cap < epoch_end) ? cap : epoch_end;
    const bool tempering = I->P.tempering != 0;
    const int TS = I->P.target_samples;
    const int32_t stuck = I->P.stuck_threshold;
    const int32_t per_cool = I->P.moves_per_cool;
    const bool logging = I->log != nullptr;

    uint64_t moves = S->moves;
    int64_t score = S->score, offnz = S->offnz, local_best = S->local_best;
    uint32_t msi = (uint32_t)S->msi;
    double T = tempering ? ((const GLOBAL double*)I->temps)[S->rung] : S->T;
    double nbeta = -1.0 / T;
    double cooling = S->cooling;
    int32_t cool_counter = S->cool_counter, win_p = S->win_p, win_a = S->win_a;
    int64_t best_score = S->best_score, best_move = S->best_move;
    int32_t flags = S->flags;
    const Coin greedy_coin(S->gf), target_coin(S->tf);
    uint32_t accepts = 0, greedy_n = 0, infeasible = 0, sweeps = 0, logged = 0;

    auto leave_best = [&]() __attribute__((always_inline)) {
        if (flags & RF_CUR_BEST) {
            const GLOBAL Inst* Iq = I;
            if (best_score <= Iq->gb_score) {
                const int64_t off = (int64_t)r * nn;
                GLOBAL int64_t* bA = (GLOBAL int64_t*)Iq->bA + off;
                GLOBAL int64_t* bX = (GLOBAL int64_t*)Iq->bXt + off;
                for (int row = 0; row < n; ++row)
                    for (int k = qd; k < n; k += Q) {
                        bA[row * n + k] = rp.gA[row * n + k];
                        bX[row * n + k] = gX[row * n + k];
                    }
                flags &= ~RF_BEST_STALE;
            } else {
                flags |= RF_BEST_STALE;
            }
            flags &= ~RF_CUR_BEST;
        }
    };
    auto track_best = [&](uint64_t m) __attribute__((always_inline)) {
        if (score < best_score) {
            best_score = score;
            best_move = (int64_t)m;
            flags |= RF_CUR_BEST;
            flags &= ~RF_BEST_STALE;
        }
    };
    auto log_move = [&](uint64_t m, int i, int j, int64_t s, uint16_t f)
                        __attribute__((always_inline)) {
        if (logging) {
            const GLOBAL Inst* Iq = I;
            const int64_t at = opaque_v(S)->log_count + logged;
            if (at < Iq->log_cap && qd == 0) {
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
    // add_transform_row's bound (main.cpp:97-108) for Add(i, j, s) over the
    // group's columns; true when every new X^T entry is in bound
    auto xt_fits = [&](int i, int j, int64_t s) __attribute__((always_inline)) -> bool {
        bool bad = false;
        for (int k = qd; k < n; k += Q) bad |= over(gX[j * n + k] + s * gX[i * n + k]);
        return group_bits<Q>(bad, gbase) == 0u;
    };
    // commit_add (main.cpp:112-133) of Add(i, j, s), group-parallel: row and
    // column j of A (LDS low words, HBM), the row masses and big counts, row j
    // of X^T; `off` is the off-diagonal change (row j's new mass is its old one
    // plus it). Returns the change in off-diagonal nonzeros.
    auto commit = [&](int i, int j, int64_t s, int64_t off)
                      __attribute__((always_inline)) -> int32_t {
        const bool sm = rp.small(i) && rp.small(j);
        int64_t a_ii, a_ji, a_jj;
        if (sm) {
            a_ii = rp.s32(i, i);
            a_ji = rp.s32(j, i);
            a_jj = rp.s32(j, j);
        } else {
            a_ii = rp.entry64(i, i);
            a_ji = rp.entry64(j, i);
            a_jj = rp.entry64(j, j);
        }
        int64_t nv[CQ], dm[CQ];
        int32_t db[CQ];
        int32_t dnz = 0, dbj = 0;
#pragma unroll
        for (int c = 0; c < CQ; ++c) {
            const int k = qd + c * Q;
            nv[c] = 0;
            dm[c] = 0;
            db[c] = 0;
            if (k < n) {
                const int64_t aj = sm ? (int64_t)rp.s32(j, k) : rp.gA[j * n + k];
                const int64_t ai = sm ? (int64_t)rp.s32(i, k) : rp.gA[i * n + k];
                nv[c] = k == j ? a_jj + 2 * s * a_ji + s * s * a_ii : aj + s * ai;
                db[c] = (int32_t)(iabs64(nv[c]) >= LaneRep<NMAX, Q, WIDE>::kSmall) -
                        (int32_t)(iabs64(aj) >= LaneRep<NMAX, Q, WIDE>::kSmall);
                dbj += db[c];
                if (k != j) {
                    dnz += (int32_t)(nv[c] != 0) - (int32_t)(aj != 0);
                    dm[c] = iabs64(nv[c]) - iabs64(aj);
                }
            }
        }
        __builtin_amdgcn_wave_barrier();  // every lane has read what changes
#pragma unroll
        for (int c = 0; c < CQ; ++c) {
            const int k = qd + c * Q;
            if (k < n) {
                if constexpr (!WIDE) rp.sS[j * NMAX + k] = (int16_t)nv[c];
                rp.gA[j * n + k] = nv[c];
                gX[j * n + k] = gX[j * n + k] + s * gX[i * n + k];
                if (k != j) {
                    if constexpr (!WIDE) rp.sS[k * NMAX + j] = (int16_t)nv[c];
                    rp.sM[k] += dm[c];
                    rp.sBig[k] += db[c];
                }
            }
        }
        // column j of gA through its lane
        const int jl = j & (Q - 1);
#pragma unroll
        for (int c = 0; c < CQ; ++c)
            for (int qq = 0; qq < Q; ++qq) {
                const int k = qq + c * Q;
                if (k < n && k != j) {
                    const int64_t v = from_lane64(nv[c], gbase + qq);
                    if (qd == jl) rp.gA[k * n + j] = v;
                }
            }
        dbj = Grp<Q>::sum32(dbj);
        if (qd == jl) {
            rp.sM[j] += off;
            rp.sBig[j] += dbj;
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        return Grp<Q>::sum32(dnz);
    };
    // greedy reduction sweep from a pivot (congruence_anneal.hpp:246-260);
    // group-uniform
    auto sweep = [&](int pv, uint64_t m) __attribute__((always_inline)) -> bool {
        const int64_t before = score;
        ++sweeps;
        for (int t = 0; t < N; ++t) {
            const int jj = rp.act((uint32_t)t);
            if (jj == pv) continue;
            const EvalOut e = eval_group<NMAX, Q>(rp.gA, n, qd, gbase, pv, jj, true, 0);
            ++greedy_n;
            if (e.s == 0 || !e.ok || e.delta > 0) continue;
            if (!xt_fits(pv, jj, e.s)) continue;
            leave_best();
            offnz += commit(pv, jj, e.s, e.off);
            score += e.delta;
            ++accepts;
            log_move(m, pv, jj, e.s, 1);
        }
        return score < before;
    };
    // bookkeeping after move m and the stall handling that follows it
    // (congruence_anneal.hpp:298-412); true when the replica is solved
    auto finish = [&](uint64_t m, int res, uint32_t w4z, uint32_t w4w, bool& changed)
                      __attribute__((always_inline)) -> bool {
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
                const auto& P = I->P;
                if (P.sweep_fraction > 0.0 && u01(w4z) < P.sweep_fraction) {
                    sweep(rp.act(mulhi32(w4w, (uint32_t)N)), m);
                    track_best(m);
                }
                local_best = score;
                msi = 0;
            }
        } else {
            if ((int64_t)msi > stuck) {
                const GLOBAL Inst* Iq = I;
                if (Iq->P.sweep_fraction > 0.0 && u01(w4z) < Iq->P.sweep_fraction) {
                    if (sweep(rp.act(mulhi32(w4w, (uint32_t)N)), m)) {
                        local_best = score;
                        msi = 0;
                    }
                    track_best(m);
                }
                const GLOBAL Inst* Iv = I;
                T = Iv->P.reheat * Iv->P.t_init;
                nbeta = -1.0 / T;
                const int64_t gb_epoch = Iv->gb_score;
                if ((double)score > (double)gb_epoch * Iv->P.reseed_factor) {
                    leave_best();
                    log_move(m, -1, -1, 0, 2);
                    const GLOBAL int64_t* srcA = (const GLOBAL int64_t*)Iv->gA;
                    const GLOBAL int64_t* srcX = (const GLOBAL int64_t*)Iv->gXt;
                    for (int row = 0; row < n; ++row)
                        for (int k = qd; k < n; k += Q) {
                            rp.gA[row * n + k] = srcA[row * n + k];
                            gX[row * n + k] = srcX[row * n + k];
                        }
                    __builtin_amdgcn_wave_barrier();
                    rp.load_lds();
                    score = gb_epoch;
                    offnz = Iv->gb_offnz;
                    track_best(m);
                    changed = true;
                }
                local_best = score;
                msi = 0;
                cool_counter = 0;
            } else if (++cool_counter >= per_cool) {
                const auto& P = I->P;
                cool_counter = 0;
                if (P.adaptive && win_p > 0) {
                    const double acc = (double)win_a / (double)win_p;
                    if (acc > P.target_accept)
                        cooling *= 0.999;
                    else
                        cooling /= 0.999;
                    cooling = cooling < 0.90 ? 0.90 : (cooling > 0.99999 ? 0.99999 : cooling);
                }
                win_p = 0;
                win_a = 0;
                T *= cooling;
                if (T < P.t_min) T = P.t_min;
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

    bool done = moves >= stop_at;
    bool halt = false;
    for (uint32_t it = 0;; ++it) {
        if (__ballot(!done) == 0ull) break;
        if ((it & 255u) == 0u) {
            // the solve key and the stop flag (one instance per wave)
            const unsigned long long key =
                __hip_atomic_load(&I->solve_key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (key != kNoKey && moves >= (key >> kRepBits)) done = true;
            if (*stop) {
                halt = true;
                break;
            }
        }
        if (done) continue;

        // ---- make and evaluate move m0 + qd (this lane's proposal) ----
        const uint64_t m0 = moves + 1;
        const uint64_t mg = m0 + (uint64_t)qd;
        const uint64_t qq = (mg - 1) >> 2;
        const uint32_t c3 = 16u * (uint32_t)((mg - 1) & 3u);
        const int kn = (stop_at - moves) < (uint64_t)Q ? (int)(stop_at - moves) : Q;
        const uint32_t live = kn >= 32 ? ~0u : ((1u << kn) - 1u);
        // the key in scalar registers, fresh every step (not carried across it)
        uint32_t key0 = k0, key1 = k1;
        opaque_u32(key0);
        opaque_u32(key1);
        // Philox slots 0..4 (the coins, sign, Metropolis uniform, uniform pair,
        // fallback pivot, stall words) and the tournament draws of slots < TS:
        // the hot-row tournament as the draws come, the pivot tournament's
        // positions kept a byte each
        uint32_t wz[5], ww[5], pk[kMaxSamples / 4];
        uint32_t tpos = 0;
        int64_t tw = -1;
        // hot-row tournament: first maximum of the row masses
        auto hot = [&](int c, uint32_t wx) __attribute__((always_inline)) {
            const uint32_t pos = mulhi32(wx, (uint32_t)N);
            const int64_t w = rp.sM[rp.act(pos)];
            if (c == 0 || w > tw) {
                tw = w;
                tpos = pos;
            }
        };
        if constexpr (WIDE) {
            // any number of samples: slots 0..4 for their words, the rest for
            // the hot-row draws only (the group redraws the pivot ones)
#pragma unroll
            for (int c = 0; c < 5; ++c) {
                const uint4 W = philox10((uint32_t)qq, (uint32_t)(qq >> 32), (uint32_t)r,
                                         slot_word(c3, c), key0, key1);
                wz[c] = W.z;
                ww[c] = W.w;
                if (c < TS) hot(c, W.x);
            }
            for (int c = 5; c < TS; ++c) {
                const uint4 W = philox10((uint32_t)qq, (uint32_t)(qq >> 32), (uint32_t)r,
                                         slot_word(c3, c), key0, key1);
                hot(c, W.x);
            }
        } else {
#pragma unroll
            for (int c = 0; c < kMaxSamples; ++c) {
                if (c % 4 == 0) pk[c / 4] = 0u;
                if (c < 5 || c < TS) {
                    const uint4 W = philox10((uint32_t)qq, (uint32_t)(qq >> 32), (uint32_t)r,
                                             c3 + (uint32_t)c, key0, key1);
                    if (c < 5) {
                        wz[c] = W.z;
                        ww[c] = W.w;
                    }
                    if (c < TS) {
                        hot(c, W.x);
                        pk[c / 4] |= mulhi32(W.y, (uint32_t)N) << (8 * (c % 4));
                    }
                }
            }
        }
        const bool targeted = target_coin(wz[0]);
        int i = -1, j;
        bool defer = false;  // rows with an entry of 2^16 or more: the group evaluates
        if (targeted) {
            j = rp.act(tpos);
            if (rp.small(j)) {
                // pivot tournament: first maximum of |A[j][c]| over draws c != j
                int32_t vbest = -1;
#pragma unroll
                for (int c = 0; c < kMaxSamples; ++c) {
                    if (c < TS) {
                        const int cand = rp.act((pk[c / 4] >> (8 * (c % 4))) & 0xffu);
                        if (cand != j) {
                            const int32_t v = lane::abs32(rp.s32(j, cand));
                            if (v > vbest) {
                                vbest = v;
                                i = cand;
                            }
                        }
                    }
                }
                if (i < 0) i = rp.act((tpos + 1u + mulhi32(ww[3], (uint32_t)(N - 1))) % (uint32_t)N);
                defer = !rp.small(i);
            } else {
                defer = true;
            }
        } else {
            const uint32_t ip = mulhi32(ww[2], (uint32_t)N);
            const uint32_t jp = (ip + 1u + mulhi32(wz[3], (uint32_t)(N - 1))) % (uint32_t)N;
            i = rp.act(ip);
            j = rp.act(jp);
            defer = !rp.small(i) || !rp.small(j);
        }
        const bool gr = greedy_coin(ww[0]);
        const int32_t pm = (wz[1] & 1u) ? 1 : -1;
        EvalOut ev;
        ev.s = 0;
        ev.delta = 0;
        ev.off = 0;
        ev.ok = true;
        if constexpr (!WIDE) if (!defer) {
            lane::Pair<NMAX> pr;
            rp.pair_small(i, j, pr);
            int32_t tot, sumq, off = 0, s = 0;
            lane::totals(pr, tot, sumq);
            if (gr) s = lane::best_shear(pr, tot, sumq, &off);
            if (s == 0) {
                s = pm;
                off = lane::abs_sum(pr, s) - sumq;
            }
            ev.s = s;
            ev.off = off;
            ev.delta = lane::delta_of(pr, s, off);
        }
        // the deferred proposals of the group, one at a time, by the group
        uint32_t dmask = group_bits<Q>(defer, gbase) & live;
        while (dmask) {
            const int b = __builtin_ctz(dmask);
            dmask &= dmask - 1u;
            const int src = gbase + b;
            const int bj = from_lane32(j, src);
            int bi = from_lane32(i, src);
            if (bi < 0 && WIDE) {
                // the pivot tournament of move m0 + b, its draws remade by the
                // group Q at a time: lane qd takes draw c = base + qd, reads
                // |A[bj][cand]| from the lane owning column cand (row bj held
                // C columns per lane), and the group keeps the first maximum
                // through the key |a| 2^12 + 4095 - c (c < 4096)
                const uint64_t mgb = m0 + (uint64_t)b;
                const uint64_t qqb = (mgb - 1) >> 2;
                const uint32_t c3b = 16u * (uint32_t)((mgb - 1) & 3u);
                int64_t rowj[CQ];
#pragma unroll
                for (int e = 0; e < CQ; ++e) {
                    const int k = qd + e * Q;
                    rowj[e] = k < n ? rp.gA[bj * n + k] : 0;
                }
                int64_t bkey = -1;
                int bcand = -1;
                for (int base = 0; base < TS; base += Q) {
                    const int c = base + qd;
                    int cand = 0;
                    if (c < TS) {
                        uint32_t kk0 = k0, kk1 = k1;
                        opaque_u32(kk0);
                        opaque_u32(kk1);
                        const uint4 W = philox10((uint32_t)qqb, (uint32_t)(qqb >> 32), (uint32_t)r,
                                                 slot_word(c3b, c), kk0, kk1);
                        cand = rp.act(mulhi32(W.y, (uint32_t)N));
                    }
                    const int owner = gbase + (cand & (Q - 1));
                    int64_t val = 0;
#pragma unroll
                    for (int e = 0; e < CQ; ++e) {
                        const int64_t t = from_lane64(rowj[e], owner);
                        if (e == cand / Q) val = t;
                    }
                    const int64_t key =
                        (c < TS && cand != bj) ? (iabs64(val) << 12) | (int64_t)(4095 - c) : -1;
                    if (key > bkey) {
                        bkey = key;
                        bcand = cand;
                    }
                }
                const int64_t mx = Grp<Q>::max(bkey);
                if (mx >= 0) {
                    const uint32_t hit = group_bits<Q>(bkey == mx, gbase);
                    bi = from_lane32(bcand, gbase + __builtin_ctz(hit));
                }
            } else if (bi < 0) {
                // the pivot tournament on a row with large entries
                int64_t vbest = -1;
#pragma unroll
                for (int c = 0; c < kMaxSamples; ++c) {
                    if (c < TS) {
                        const uint32_t w = (uint32_t)from_lane32((int32_t)pk[c / 4], src);
                        const int cand = rp.act((w >> (8 * (c % 4))) & 0xffu);
                        if (cand != bj) {
                            const int64_t v = iabs64(rp.entry64(bj, cand));
                            if (v > vbest) {
                                vbest = v;
                                bi = cand;
                            }
                        }
                    }
                }
            }
            if (bi < 0) {
                // no draw other than the target: the fallback pivot
                const uint32_t btp = (uint32_t)from_lane32((int32_t)tpos, src);
                const uint32_t bw3 = (uint32_t)from_lane32((int32_t)ww[3], src);
                bi = rp.act((btp + 1u + mulhi32(bw3, (uint32_t)(N - 1))) % (uint32_t)N);
            }
            const bool bgr = from_lane32(gr ? 1 : 0, src) != 0;
            const int32_t bpm = from_lane32(pm, src);
            const EvalOut e = eval_group<NMAX, Q>(rp.gA, n, qd, gbase, bi, bj, bgr, bpm);
            if (qd == b) {
                i = bi;
                ev = e;
            }
        }
        const int32_t s = ev.s;
        const int64_t delta = ev.delta, off = ev.off;
        const bool bad = !ev.ok;
        const double u = u53(ww[1], wz[2]);
        bool pass = !bad && (delta <= 0 || u < xexp((double)delta * nbeta));

        // ---- resolve moves m0 .. m0 + kn - 1 in order (per replica) ----
        uint32_t pmask = group_bits<Q>(pass, gbase) & live;
        const uint32_t grm = group_bits<Q>(gr, gbase) & live;
        const uint32_t badm = group_bits<Q>(bad, gbase) & live;
        int pos = 0;
        while (pos < kn) {
            const uint32_t rest = pmask >> pos;
            const int kp = rest ? pos + __builtin_ctz(rest) : kn;
            int ke = kp;
            const int ks = pos + (int)((int64_t)stuck - (int64_t)msi);  // the stall's move
            if (ks < ke) ke = ks;
            if (!tempering) {
                const int kc = pos + per_cool - cool_counter - 1;  // the cooling step's move
                if (kc < ke) ke = kc;
            }
            if (ke > pos) {
                // moves pos .. ke-1: rejected, no stall, no cooling step
                const int cnt = ke - pos;
                const uint32_t rng = (cnt >= 32 ? ~0u : ((1u << cnt) - 1u)) << pos;
                greedy_n += (uint32_t)__builtin_popcount(grm & rng);
                infeasible += (uint32_t)__builtin_popcount(badm & rng);
                moves += (uint64_t)cnt;
                win_p += cnt;
                msi += (uint32_t)cnt;
                if (!tempering) cool_counter += cnt;
                pos = ke;
            }
            if (pos >= kn) break;
            const uint64_t m = m0 + (uint64_t)pos;
            moves = m;
            const int src = gbase + pos;
            greedy_n += (grm >> pos) & 1u;
            int res = -1;
            bool changed = false;
            if ((badm >> pos) & 1u) {
                ++infeasible;
            } else if (pos == kp) {
                const int ik = from_lane32(i, src), jk = from_lane32(j, src);
                const int32_t sk = from_lane32(s, src);
                const int64_t dk = from_lane64(delta, src), ok = from_lane64(off, src);
                if (xt_fits(ik, jk, sk)) {
                    const int64_t ns = score + dk;
                    if (ns < best_score) {
                        best_score = ns;
                        best_move = (int64_t)m;
                        flags |= RF_CUR_BEST;
                        flags &= ~RF_BEST_STALE;
                    } else {
                        leave_best();
                    }
                    offnz += commit(ik, jk, sk, ok);
                    score = ns;
                    ++accepts;
                    log_move(m, ik, jk, sk, 0);
                    res = dk < 0 ? 1 : 0;
                    changed = true;
                }
            }
            const double nb0 = nbeta;
            const uint32_t z4 = (uint32_t)from_lane32((int32_t)wz[4], src);
            const uint32_t w4 = (uint32_t)from_lane32((int32_t)ww[4], src);
            if (finish(m, res, z4, w4, changed)) {
                done = true;
                break;
            }
            if (changed) break;
            if (nbeta != nb0) {
                // a cooling step: the later proposals are decided again
                pass = !bad && (delta <= 0 || u < xexp((double)delta * nbeta));
                pmask = group_bits<Q>(pass, gbase) & live;
            }
            ++pos;
        }
        if (moves >= stop_at) done = true;
    }
    // the move cap ends the loop through stop_at
    if (!(flags & RF_STOPPED) && !halt && cap != 0 && moves >= cap) flags |= RF_STOPPED | RF_CAPPED;

    if (qd == 0) {
        S->moves = moves;
        S->score = score;
        S->offnz = offnz;
        S->local_best = local_best;
        S->msi = (int64_t)msi;
        S->T = T;
        S->cooling = cooling;
        S->cool_counter = cool_counter;
        S->win_p = win_p;
        S->win_a = win_a;
        S->best_score = best_score;
        S->best_move = best_move;
        S->flags = flags;
        S->accepts += accepts;
        S->greedy += greedy_n;
        S->infeasible += infeasible;
        S->sweeps += sweeps;
        S->log_count += logged;
        if (flags & RF_SOLVED) S->solve_move = (int64_t)moves;
        if (flags & RF_SOLVED)
            atomicMin((unsigned long long*)&I->solve_key,
                      ((unsigned long long)moves << kRepBits) | (unsigned)r);
    }
}

// One epoch of the block's replicas from a launch table: entry gid of the
// table is replica slot threadIdx.x / Q of block blockIdx.x, its LDS at
// slot * rep_bytes; every replica of a wave belongs to one instance (the host
// pads each instance's replica list to whole waves).
template <int NMAX, int Q, bool WIDE>
__device__ __attribute__((always_inline)) inline void lane_epoch_table(
    const RepRef* __restrict__ table, int count, Inst* __restrict__ insts, int rep_bytes,
    const volatile int* __restrict__ stop) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int slot = (int)threadIdx.x / Q;
    const int gid = (int)blockIdx.x * ((int)blockDim.x / Q) + slot;
    if (gid >= count) return;
    const int wave_gid0 = __builtin_amdgcn_readfirstlane((int)blockIdx.x * ((int)blockDim.x / Q) +
                                                         (int)(threadIdx.x >> 6) * (64 / Q));
    const int inst = __builtin_amdgcn_readfirstlane(table[wave_gid0].inst);
    lane_epoch<NMAX, Q, WIDE>(insts, inst, table[gid].rep,
                              (SHARED unsigned char*)smem + (size_t)slot * rep_bytes, stop);
}

}  // namespace xtx
