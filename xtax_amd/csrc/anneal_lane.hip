// anneal_lane.hip -- the annealer with one proposal per lane (gfx950).
//
// A replica takes Q lanes (Q = 16: four replicas per wave). Each step, lane q
// of the replica makes and evaluates the replica's move m0 + q on its current
// state, serially over the columns -- the pair tournaments, the exact best
// shear and the evaluation (lane_eval.h), the Metropolis test -- with no
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
// LDS per replica: A as its upper triangle laid out by the kernel's order NMAX
// (row r holds columns r .. NMAX-1, A[a][b] at T(min) + max with
// T(r) = r (2 NMAX - r - 1) / 2), the row masses (NMAX int64) and the active
// list (NMAX bytes). X^T stays in HBM, read and written on a commit only, in
// columns owned by fixed lanes (lane q of the replica owns columns q, q + Q,
// ...), so every lane reads back only global data it wrote itself.
#include <hip/hip_runtime.h>

#include "device_common.h"
#include "engine_types.h"
#include "lane_eval.h"

namespace xtx {

template <int NMAX>
__device__ __forceinline__ constexpr int ltri(int r) {
    return r * (2 * NMAX - r - 1) / 2;
}

// The replica's LDS: A (upper triangle at order NMAX), row masses, active list.
template <int NMAX, int Q>
struct LaneRep {
    int64_t* sA;
    int64_t* sM;
    uint8_t* sAct;
    int n, N, qd;
    bool ident;

    __device__ __forceinline__ int at(int a, int b) const {
        return a <= b ? ltri<NMAX>(a) + b : ltri<NMAX>(b) + a;
    }
    __device__ __forceinline__ int64_t A(int a, int b) const { return sA[at(a, b)]; }
    __device__ __forceinline__ int act(uint32_t p) const { return ident ? (int)p : (int)sAct[p]; }

    // rows i and j as the small path's p, q (lane_eval.h); false when some
    // entry of the two rows reaches 2^16 (p, q are then not used)
    __device__ __forceinline__ bool load_pq(int i, int j, int32_t (&p)[NMAX], int32_t (&q)[NMAX],
                                            int32_t& sumq) const {
        const int Ti = ltri<NMAX>(i), Tj = ltri<NMAX>(j);
        bool big = false;
        sumq = 0;
#pragma unroll
        for (int k = 0; k < NMAX; ++k) {
            // eight columns' loads in flight at a time: all 2 NMAX 64-bit loads
            // at once would hold 4 NMAX registers
            if (k % 8 == 0) XTX_PHASE();
            int64_t ai = 0, aj = 0;
            if (k < n) {
                ai = sA[k >= i ? Ti + k : ltri<NMAX>(k) + i];
                aj = sA[k >= j ? Tj + k : ltri<NMAX>(k) + j];
            }
            big |= ai >= lane::kSmallEntry || ai <= -lane::kSmallEntry ||
                   aj >= lane::kSmallEntry || aj <= -lane::kSmallEntry;
            const bool use = k < n && k != j;
            p[k] = use ? (int32_t)(ai < 0 ? -ai : ai) : 0;
            q[k] = use ? (int32_t)(ai >= 0 ? aj : -aj) : 0;
            sumq += q[k] < 0 ? -q[k] : q[k];
        }
        XTX_PHASE();
        return !big;
    }

    // group-parallel (lane qd: columns qd, qd + Q, ...) moves between HBM
    // (row-major n x n) and the LDS layout
    __device__ __forceinline__ void load_A(const GLOBAL int64_t* src) {
        for (int row = 0; row < n; ++row)
            for (int k = row + qd; k < n; k += Q) sA[ltri<NMAX>(row) + k] = src[row * n + k];
    }
    __device__ __forceinline__ void store_A(GLOBAL int64_t* dst) const {
        for (int row = 0; row < n; ++row)
            for (int k = qd; k < n; k += Q) dst[row * n + k] = A(row, k);
    }
    __device__ __forceinline__ void copy_cols(GLOBAL int64_t* dst, const GLOBAL int64_t* src) const {
        for (int row = 0; row < n; ++row)
            for (int k = qd; k < n; k += Q) dst[row * n + k] = src[row * n + k];
    }
    __device__ __forceinline__ void row_masses() {
        for (int k = qd; k < n; k += Q) {
            int64_t m = 0;
            for (int t = 0; t < n; ++t)
                if (t != k) m += lane::abs64(A(k, t));
            sM[k] = m;
        }
    }
};

// A proposal's evaluation: the shear (the greedy best, or the given +-1 when
// that is 0 or not asked for), its score change, nonzero change and
// off-diagonal change; ok = false when it leaves the bound.
struct EvalOut {
    int64_t s, delta, off;
    int32_t dnz;
    bool ok;
};

// The pair's evaluation by one lane, small or int64 path, called where it is
// rare (rows with an entry of 2^16 or more; the stall sweep), so its code and
// registers stay out of the move loop.
template <int NMAX, int Q>
__device__ __noinline__ EvalOut eval_pair(const int64_t* sA, int n, int i, int j, bool greedy,
                                          int64_t pm) {
    LaneRep<NMAX, Q> rp;
    rp.sA = (int64_t*)sA;
    rp.n = n;
    EvalOut o;
    int32_t p[NMAX], q[NMAX], sumq;
    const int64_t a_ii = rp.A(i, i), a_ji = rp.A(j, i), a_jj = rp.A(j, j);
    if (rp.load_pq(i, j, p, q, sumq)) {
        int32_t soff = 0, o32 = 0;
        o.s = 0;
        if (greedy)
            o.s = lane::shear_small<NMAX>(p, q, sumq, (int32_t)a_ii, (int32_t)a_ji, (int32_t)a_jj,
                                          &soff);
        const bool know = o.s != 0;
        if (o.s == 0) o.s = pm;
        lane::eval_small<NMAX>(p, q, sumq, (int32_t)o.s, (int32_t)a_ii, (int32_t)a_ji,
                               (int32_t)a_jj, know, soff, &o.delta, &o.dnz, &o32);
        o.off = o32;
        o.ok = true;
        return o;
    }
    auto row = [&](int k, int64_t& ai, int64_t& aj) {
        ai = rp.A(i, k);
        aj = rp.A(j, k);
    };
    o.s = 0;
    if (greedy) {
        int64_t boff = 0;
        o.s = lane::shear_big(row, n, j, a_ii, a_ji, a_jj, &boff);
    }
    if (o.s == 0) o.s = pm;
    o.ok = lane::eval_big(row, n, j, o.s, a_ii, a_ji, a_jj, &o.delta, &o.dnz, &o.off);
    return o;
}

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

// Waves per SIMD the kernel is compiled for. A lane's proposal holds both rows
// in registers (2 NMAX) besides the replica's state, so above order 16 one
// wave per SIMD keeps everything in registers (the LDS of four replicas per
// wave would allow two at order 32: 4.5 KB per replica, one at 48: 9.9 KB).
constexpr int lane_waves_per_simd(int nmax) { return nmax <= 16 ? 2 : 1; }

// One epoch of many replicas, Q lanes (proposals) each.
template <int NMAX, int Q>
__global__ void __launch_bounds__(256, lane_waves_per_simd(NMAX))
    k_anneal_lane(const RepRef* __restrict__ table, int count, Inst* __restrict__ insts,
                  int nmax, int rep_bytes, const volatile int* __restrict__ stop) {
    static_assert(Q == 8 || Q == 16 || Q == 32, "8, 16 or 32 lanes per replica");
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = (int)(threadIdx.x & 63u);
    const int qd = (int)(threadIdx.x % (unsigned)Q);
    const int gbase = lane & ~(Q - 1);
    const int slot = (int)threadIdx.x / Q;
    const int gid = (int)blockIdx.x * ((int)blockDim.x / Q) + slot;
    if (gid >= count || nmax != NMAX) return;
    // every replica of a wave belongs to one instance (the host pads each
    // instance's replica list to whole waves)
    const int wave_gid0 = __builtin_amdgcn_readfirstlane((int)blockIdx.x * ((int)blockDim.x / Q) +
                                                         (int)(threadIdx.x >> 6) * (64 / Q));
    Inst* I = insts + __builtin_amdgcn_readfirstlane(table[wave_gid0].inst);
    if (I->status != ST_RUNNING) return;
    const int r = table[gid].rep;
    if (r < 0) return;  // padding
    GLOBAL RepState* S = (GLOBAL RepState*)I->rs + r;
    if (S->flags & RF_STOPPED) return;

    constexpr int kAbytes = NMAX * (NMAX + 1) / 2 * 8;
    LaneRep<NMAX, Q> rp;
    rp.sA = (int64_t*)(smem + (size_t)slot * rep_bytes);
    rp.sM = (int64_t*)(smem + (size_t)slot * rep_bytes + kAbytes);
    rp.sAct = (uint8_t*)(rp.sM + NMAX);
    rp.n = I->n;
    rp.N = I->N;
    rp.qd = qd;
    const int n = rp.n, N = rp.N;
    const int64_t nn = (int64_t)n * n;
    GLOBAL int64_t* const gAcur = (GLOBAL int64_t*)I->A + (int64_t)r * nn;
    GLOBAL int64_t* const gX = (GLOBAL int64_t*)I->Xt + (int64_t)r * nn;
    rp.load_A(gAcur);
    const GLOBAL uint8_t* gact = (const GLOBAL uint8_t*)I->act;
    bool nonid = false;
    for (int t = qd; t < N; t += Q) {
        const uint8_t a = gact[t];
        rp.sAct[t] = a;
        nonid |= a != t;
    }
    rp.ident = N == n && __ballot(nonid) == 0ull;  // the wave's instance: uniform
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    rp.row_masses();
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");

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

    auto leave_best = [&]() {
        if (flags & RF_CUR_BEST) {
            const Inst* Iq = I;
            if (best_score <= Iq->gb_score) {
                const int64_t off = (int64_t)r * nn;
                rp.store_A((GLOBAL int64_t*)Iq->bA + off);
                rp.copy_cols((GLOBAL int64_t*)Iq->bXt + off, gX);
                flags &= ~RF_BEST_STALE;
            } else {
                flags |= RF_BEST_STALE;
            }
            flags &= ~RF_CUR_BEST;
        }
    };
    auto track_best = [&](uint64_t m) {
        if (score < best_score) {
            best_score = score;
            best_move = (int64_t)m;
            flags |= RF_CUR_BEST;
            flags &= ~RF_BEST_STALE;
        }
    };
    auto log_move = [&](uint64_t m, int i, int j, int64_t s, uint16_t f) {
        if (logging) {
            const Inst* Iq = I;
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
    auto xt_fits = [&](int i, int j, int64_t s) -> bool {
        bool bad = false;
        for (int k = qd; k < n; k += Q) bad |= lane::over(gX[j * n + k] + s * gX[i * n + k]);
        return group_bits<Q>(bad, gbase) == 0u;
    };
    // commit_add (main.cpp:112-133) of Add(i, j, s), group-parallel: row j of
    // A (one LDS entry per column), the row masses, row j of X^T; `off` is
    // the off-diagonal change (row j's new mass is its old one plus it)
    auto commit = [&](int i, int j, int64_t s, int64_t off) {
        const int64_t a_ii = rp.A(i, i), a_ji = rp.A(j, i), a_jj = rp.A(j, j);
        int64_t nv[(NMAX + Q - 1) / Q], dm[(NMAX + Q - 1) / Q];
#pragma unroll
        for (int c = 0; c < (NMAX + Q - 1) / Q; ++c) {
            const int k = qd + c * Q;
            nv[c] = 0;
            dm[c] = 0;
            if (k < n) {
                const int64_t aj = rp.A(j, k);
                nv[c] = k == j ? a_jj + 2 * s * a_ji + s * s * a_ii : aj + s * rp.A(i, k);
                dm[c] = k == j ? 0 : lane::abs64(nv[c]) - lane::abs64(aj);
            }
        }
        __builtin_amdgcn_wave_barrier();  // every lane has read what changes
#pragma unroll
        for (int c = 0; c < (NMAX + Q - 1) / Q; ++c) {
            const int k = qd + c * Q;
            if (k < n) {
                rp.sA[rp.at(j, k)] = nv[c];
                rp.sM[k] += k == j ? off : dm[c];
                gX[j * n + k] = gX[j * n + k] + s * gX[i * n + k];
            }
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    };
    // evaluation of a move by this lane alone: the small path inline, the
    // int64 path (some entry of the rows at 2^16 or more) through eval_pair
    auto evaluate = [&](int i, int j, bool greedy, int64_t pm) -> EvalOut {
        int32_t p[NMAX], q[NMAX], sumq;
        if (rp.load_pq(i, j, p, q, sumq)) {
            const int64_t a_ii = rp.A(i, i), a_ji = rp.A(j, i), a_jj = rp.A(j, j);
            EvalOut o;
            int32_t soff = 0, o32 = 0;
            o.s = 0;
            if (greedy)
                o.s = lane::shear_small<NMAX>(p, q, sumq, (int32_t)a_ii, (int32_t)a_ji,
                                              (int32_t)a_jj, &soff);
            const bool know = o.s != 0;
            if (o.s == 0) o.s = pm;
            lane::eval_small<NMAX>(p, q, sumq, (int32_t)o.s, (int32_t)a_ii, (int32_t)a_ji,
                                   (int32_t)a_jj, know, soff, &o.delta, &o.dnz, &o32);
            o.off = o32;
            o.ok = true;
            return o;
        }
        return eval_pair<NMAX, Q>(rp.sA, n, i, j, greedy, pm);
    };
    // greedy reduction sweep from a pivot (congruence_anneal.hpp:246-260);
    // every lane of the group computes the same values
    auto sweep = [&](int pv, uint64_t m) -> bool {
        const int64_t before = score;
        ++sweeps;
        for (int t = 0; t < N; ++t) {
            const int jj = rp.act((uint32_t)t);
            if (jj == pv) continue;
            const EvalOut e = eval_pair<NMAX, Q>(rp.sA, n, pv, jj, true, 0);
            ++greedy_n;
            if (e.s == 0 || !e.ok || e.delta > 0) continue;
            if (!xt_fits(pv, jj, e.s)) continue;
            leave_best();
            commit(pv, jj, e.s, e.off);
            score += e.delta;
            offnz += e.dnz;
            ++accepts;
            log_move(m, pv, jj, e.s, 1);
        }
        return score < before;
    };
    // bookkeeping after move m and the stall handling that follows it
    // (congruence_anneal.hpp:298-412); true when the replica is solved
    auto finish = [&](uint64_t m, int res, uint32_t w4z, uint32_t w4w, bool& changed) -> bool {
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
                const InstParams& P = I->P;
                if (P.sweep_fraction > 0.0 && u01(w4z) < P.sweep_fraction) {
                    sweep(rp.act(mulhi32(w4w, (uint32_t)N)), m);
                    track_best(m);
                }
                local_best = score;
                msi = 0;
            }
        } else {
            if ((int64_t)msi > stuck) {
                const Inst* Iq = I;
                if (Iq->P.sweep_fraction > 0.0 && u01(w4z) < Iq->P.sweep_fraction) {
                    if (sweep(rp.act(mulhi32(w4w, (uint32_t)N)), m)) {
                        local_best = score;
                        msi = 0;
                    }
                    track_best(m);
                }
                const Inst* Iv = I;
                T = Iv->P.reheat * Iv->P.t_init;
                nbeta = -1.0 / T;
                const int64_t gb_epoch = Iv->gb_score;
                if ((double)score > (double)gb_epoch * Iv->P.reseed_factor) {
                    leave_best();
                    log_move(m, -1, -1, 0, 2);
                    __builtin_amdgcn_wave_barrier();
                    rp.load_A((const GLOBAL int64_t*)Iv->gA);
                    rp.copy_cols(gX, (const GLOBAL int64_t*)Iv->gXt);
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                    rp.row_masses();
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                    score = gb_epoch;
                    offnz = Iv->gb_offnz;
                    track_best(m);
                    changed = true;
                }
                local_best = score;
                msi = 0;
                cool_counter = 0;
            } else if (++cool_counter >= per_cool) {
                const InstParams& P = I->P;
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
        // the key in scalar registers, fresh every step (not carried across it)
        uint32_t key0 = k0, key1 = k1;
        opaque_u32(key0);
        opaque_u32(key1);
        // Philox slots 0..4 (the coins, sign, Metropolis uniform, uniform pair,
        // fallback pivot, stall words) and the tournament draws of slots < TS
        uint32_t wz[5], ww[5], ys[kMaxSamples];
        uint32_t tpos = 0;
        int64_t tw = -1;
#pragma unroll 16
        for (int c = 0; c < kMaxSamples; ++c) {
            ys[c] = 0u;
            if (c < 5 || c < TS) {
                const uint4 W = philox10((uint32_t)qq, (uint32_t)(qq >> 32), (uint32_t)r,
                                         c3 + (uint32_t)c, key0, key1);
                if (c < 5) {
                    wz[c] = W.z;
                    ww[c] = W.w;
                }
                ys[c] = W.y;
                if (c < TS) {
                    // hot-row tournament: first maximum of the row masses
                    const uint32_t pos = mulhi32(W.x, (uint32_t)N);
                    const int64_t w = rp.sM[rp.act(pos)];
                    if (c == 0 || w > tw) {
                        tw = w;
                        tpos = pos;
                    }
                }
            }
        }
        int i, j;
        if (target_coin(wz[0])) {
            // pivot tournament: first maximum of |A[target][c]| over draws c != target
            const int target = rp.act(tpos);
            int ibest = -1;
            int64_t vbest = -1;
#pragma unroll 16
            for (int c = 0; c < kMaxSamples; ++c) {
                if (c < TS) {
                    const int cand = rp.act(mulhi32(ys[c], (uint32_t)N));
                    if (cand != target) {
                        const int64_t v = lane::abs64(rp.A(target, cand));
                        if (v > vbest) {
                            vbest = v;
                            ibest = cand;
                        }
                    }
                }
            }
            if (ibest < 0)
                ibest = rp.act((tpos + 1u + mulhi32(ww[3], (uint32_t)(N - 1))) % (uint32_t)N);
            i = ibest;
            j = target;
        } else {
            const uint32_t ip = mulhi32(ww[2], (uint32_t)N);
            const uint32_t jp = (ip + 1u + mulhi32(wz[3], (uint32_t)(N - 1))) % (uint32_t)N;
            i = rp.act(ip);
            j = rp.act(jp);
        }
        const bool gr = greedy_coin(ww[0]);
        const EvalOut ev = evaluate(i, j, gr, (wz[1] & 1u) ? 1 : -1);
        const int64_t s = ev.s, delta = ev.delta, off = ev.off;
        const int32_t dnz = ev.dnz;
        const bool bad = !ev.ok;
        const double u = u53(ww[1], wz[2]);
        bool pass = !bad && (delta <= 0 || u < xexp((double)delta * nbeta));

        // ---- resolve moves m0 .. m0 + kn - 1 in order (per replica) ----
        const int kn = (stop_at - moves) < (uint64_t)Q ? (int)(stop_at - moves) : Q;
        const uint32_t live = kn >= 32 ? ~0u : ((1u << kn) - 1u);
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
                const int64_t sk = from_lane64(s, src), dk = from_lane64(delta, src);
                const int64_t ok = from_lane64(off, src);
                const int32_t zk = from_lane32(dnz, src);
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
                    commit(ik, jk, sk, ok);
                    score = ns;
                    offnz += zk;
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

    __builtin_amdgcn_wave_barrier();
    rp.store_A(gAcur);
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
            atomicMin(&I->solve_key, ((unsigned long long)moves << kRepBits) | (unsigned)r);
    }
}

// ---------------------------------------------------------------------------
// host-side launch helpers
// ---------------------------------------------------------------------------

typedef void (*LaneKernel)(const RepRef*, int, Inst*, int, int, const volatile int*);

struct LaneVariant {
    int nmax, Q;
    LaneKernel fn;
};
static const LaneVariant kLaneVariants[] = {
    {8, 16, k_anneal_lane<8, 16>},   {16, 16, k_anneal_lane<16, 16>},
    {24, 16, k_anneal_lane<24, 16>}, {32, 16, k_anneal_lane<32, 16>},
    {40, 16, k_anneal_lane<40, 16>}, {48, 16, k_anneal_lane<48, 16>},
};

static const LaneVariant* find_lane_variant(int nmax, int Q) {
    for (const auto& v : kLaneVariants)
        if (v.nmax == nmax && v.Q == Q) return &v;
    return nullptr;
}

bool lane_variant_exists(int nmax, int Q) { return find_lane_variant(nmax, Q) != nullptr; }

// LDS of one replica of the lane kernel at order nmax
int lane_rep_bytes(int nmax) {
    const int b = nmax * (nmax + 1) / 2 * 8 + nmax * 8 + nmax;
    return (b + 15) / 16 * 16;
}

hipError_t init_lane_kernel_attrs(int max_lds_per_block) {
    for (const auto& v : kLaneVariants) {
        hipFuncAttributes fa;
        hipError_t e = hipFuncGetAttributes(&fa, (const void*)v.fn);
        if (e != hipSuccess) return e;
        e = hipFuncSetAttribute((const void*)v.fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                max_lds_per_block - (int)fa.sharedSizeBytes);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// units_per_block replicas (table entries, padding included) of Q lanes each
hipError_t launch_anneal_lane(hipStream_t st, int nmax, int Q, int units_per_block,
                              const RepRef* table, int count, Inst* insts, int rep_bytes,
                              const volatile int* stop) {
    if (count <= 0) return hipSuccess;
    const LaneVariant* v = find_lane_variant(nmax, Q);
    if (!v || units_per_block < 1 || units_per_block * Q > 256 || (units_per_block * Q) % 64 != 0 ||
        rep_bytes < lane_rep_bytes(nmax))
        return hipErrorInvalidValue;
    const int threads = units_per_block * Q;
    const int blocks = (count + units_per_block - 1) / units_per_block;
    const size_t lds = (size_t)units_per_block * rep_bytes;
    hipLaunchKernelGGL(v->fn, dim3(blocks), dim3(threads), lds, st, table, count, insts, nmax,
                       rep_bytes, stop);
    return hipGetLastError();
}

}  // namespace xtx
