// anneal.hip -- gfx950 kernels of the xtax L1 congruence annealer.
//
// Mapping: one replica per lane group of G lanes (G = 16, 32 or 64 by n), lane g
// owning columns k = g + c*G (c < C). A replica's A sits in LDS as its upper
// triangle (half the footprint of the square, so twice the resident replicas),
// its X^T in HBM (read on a proposal, written on acceptance), its row masses
// (main.cpp:136-144) in registers of the lanes owning those rows. Every sum over columns is a group reduction; every choice
// (pair, shear, accept) is made identically by all G lanes from broadcast
// values, so the group runs the reference's sequential worker loop
// (congruence_anneal.hpp:276-413) with O(n/G) work per lane per step.
// Arithmetic is exact int64 throughout; the only floating point is the
// Metropolis test and the ladder, specified bit-exactly (IEEE ops + fma).
//
// Kernels:
//   k_anneal<G,C,TRI> one epoch (exchange_interval moves) of many replicas
//   k_epoch_end     per instance: solve retirement, replica exchange sweep,
//                   instance-best bookkeeping
//   k_probe<G,C>    the move primitives on given matrices (parity tests)
//   k_probe_rng     Philox words / exp values (parity tests)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "device_common.h"
#include "engine_types.h"
#include "group_eval.h"

namespace xtx {

// ---- small-entry path ----------------------------------------------------
// When every entry of rows i and j is below 2^16 in magnitude (checked per
// move), each new off-diagonal entry is below 2^23, per-group sums below 2^30
// and no off-diagonal entry can reach the 2^48 bound: the same quantities are
// computed exactly in int32 with 32-bit reductions. Only the diagonal term
// (s^2 a_ii up to 2^28) stays int64. The results are identical to the int64
// path; the two differ only in instruction count.
constexpr int32_t kSmall = 1 << 16;

template <int G, int C>
__device__ __forceinline__ bool rows_small(const int64_t (&ai)[C], const int64_t (&aj)[C]) {
    bool big = false;
#pragma unroll
    for (int c = 0; c < C; ++c)
        big |= ai[c] >= kSmall || ai[c] <= -kSmall || aj[c] >= kSmall || aj[c] <= -kSmall;
    return !Grp<G>::any(big);
}

__device__ __forceinline__ int32_t iabs32(int32_t v) { return v < 0 ? -v : v; }

__device__ __forceinline__ int32_t rounded_div32(int32_t a, int32_t b) {
    int32_t q = a / b;
    const int32_t r = a - q * b;
    if (2 * iabs32(r) >= iabs32(b)) q += ((a < 0) == (b < 0)) ? 1 : -1;
    return q;
}

// With `know`, the off-diagonal sum of s is already known (from the best-shear
// candidates) and is not summed again.
template <int G, int C>
__device__ __forceinline__ void evaluate_small(const int32_t (&ai)[C], const int32_t (&aj)[C],
                                               int g, int n, int j, int64_t s, int64_t a_ii,
                                               int64_t a_ji, int64_t a_jj, MoveEval<C>& e,
                                               bool know, int64_t known_off) {
    int32_t off = 0, dnz = 0;
    const int32_t s32 = (int32_t)s;
    // |a_ii|, |a_ji|, |a_jj| < 2^16 here, so |new_jj| < 2^29: int32, in bound
    const int32_t jj32 = (int32_t)a_jj +
                         __mul24(s32, 2 * (int32_t)a_ji + __mul24(s32, (int32_t)a_ii));
    const int64_t new_jj = jj32;
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const int k = g + c * G;
        e.nv[c] = new_jj;
        if (k < n && k != j) {
            const int32_t old = aj[c];
            const int32_t nv = old + __mul24(s32, ai[c]);
            off += iabs32(nv) - iabs32(old);
            dnz += (old == 0) ? (nv != 0) : -(int32_t)(nv == 0);
            e.nv[c] = nv;
        }
    }
    off = know ? (int32_t)known_off : Grp<G>::sum32(off);
    dnz = Grp<G>::sum32(dnz);
    e.off = off;
    e.dnz = dnz;
    e.bad = false;
    e.delta = 4 * (int64_t)off + (int64_t)(iabs32(jj32) - iabs32((int32_t)a_jj));
}

// best_shear on the small path. Each lane forms ceil and floor of its
// breakpoint b_k = -a_jk/a_ik once from a correctly rounded double quotient
// (exact: b_k is at least 2^-16 from any integer it is not equal to, the
// quotient is within 2^-37 of b_k) and clamps them to [-65, 65], which leaves
// every weight sum W(b <= t), t in [-64, 64], unchanged. ceil(m) comes from
// the replica's LDS histogram (bins t + 64, zero on entry and on return):
// every breakpoint adds its weight to the bin of its ceiling, each lane scans
// its consecutive bins, and min{t : 2 W(ce <= t) >= W} - (-64) is the number
// of bins whose prefix sum stays below W/2 (129 when none reaches it: the
// sentinel t = 65). Weights are below 2^16 and W below 2^23: all uint32.
// *off_out receives the off-diagonal change of the returned shear.
template <int G, int C>
__device__ int64_t best_shear_small(const int32_t (&ai)[C], const int32_t (&aj)[C], int g, int n,
                                    int j, int64_t a_ii, int64_t a_ji, int64_t a_jj,
                                    uint32_t* hist, int32_t* off_out) {
    int32_t w[C], ce[C], fl[C];
    int32_t wl = 0;
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const int k = g + c * G;
        const bool bp = k < n && k != j && ai[c] != 0;
        w[c] = bp ? iabs32(ai[c]) : 0;
        wl += w[c];
        ce[c] = 0;
        fl[c] = 0;
        if (bp) {
            const double b = -(double)aj[c] / (double)ai[c];
            const int32_t up = (int32_t)__builtin_ceil(b), dn = (int32_t)__builtin_floor(b);
            ce[c] = up < -65 ? -65 : (up > 65 ? 65 : up);
            fl[c] = dn < -65 ? -65 : (dn > 65 ? 65 : dn);
        }
    }
    const int32_t tot = Grp<G>::sum32(wl);
    int cf = 0, cc = 0;
    if (tot != 0) {
        constexpr int B = med_per_lane(G);
#pragma unroll
        for (int c = 0; c < C; ++c)
            if (w[c] != 0 && ce[c] <= kShearCap)
                __hip_atomic_fetch_add(&hist[(ce[c] < -kShearCap ? -kShearCap : ce[c]) + kShearCap],
                                       (uint32_t)w[c], __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WAVEFRONT);
        __builtin_amdgcn_wave_barrier();
        uint32_t h[B];
        if (g * B < kMedBins) {
#pragma unroll
            for (int b = 0; b < B; ++b) h[b] = hist[g * B + b];
#pragma unroll
            for (int b = 0; b < B; ++b) hist[g * B + b] = 0u;
        } else {
#pragma unroll
            for (int b = 0; b < B; ++b) h[b] = 0u;
        }
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int b = 1; b < B; ++b) h[b] += h[b - 1];
        const uint32_t excl = Grp<G>::scan32(h[B - 1]) - h[B - 1];
        int32_t below = 0;
#pragma unroll
        for (int b = 0; b < B; ++b)
            below += (g * B + b < kMedBins && 2u * (excl + h[b]) < (uint32_t)tot) ? 1 : 0;
        const int lo = Grp<G>::sum32(below) - kShearCap;
        if (lo <= -kShearCap) {
            cf = cc = -kShearCap;
        } else if (lo > kShearCap) {
            cf = cc = kShearCap;
        } else {
            cc = lo;
            int32_t acc = 0;
#pragma unroll
            for (int c = 0; c < C; ++c) acc += fl[c] < lo ? w[c] : 0;
            acc = Grp<G>::sum32(acc);
            cf = (2 * acc >= tot) ? lo - 1 : lo;
        }
    }
    const bool has_q = a_ii != 0;
    const int64_t q = has_q ? -(int64_t)rounded_div32((int32_t)a_ji, (int32_t)a_ii) : 0;
    Cands cd;
    make_cands(cf, cc, has_q, q, cd);
    int64_t best_s = 0, best_d = 0;
    int32_t best_off = 0;
#pragma unroll
    for (int t = 0; t < 5; ++t) {
        const int64_t s = cd.v[t];
        if (cd.ok[t] && s != 0) {
            const int32_t s32 = (int32_t)s;
            int32_t part = 0;
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const int k = g + c * G;
                if (k < n && k != j) part += iabs32(aj[c] + __mul24(s32, ai[c])) - iabs32(aj[c]);
            }
            const int32_t off = Grp<G>::sum32(part);
            const int32_t jj32 = (int32_t)a_jj +
                                 __mul24(s32, 2 * (int32_t)a_ji + __mul24(s32, (int32_t)a_ii));
            const int64_t d = 4 * (int64_t)off + (int64_t)(iabs32(jj32) - iabs32((int32_t)a_jj));
            if (d < best_d) {
                best_d = d;
                best_s = s;
                best_off = off;
            }
        }
    }
    *off_out = best_off;
    return best_s;
}

// Both paths behind one call: the int32 path when rows i and j are small.
template <int G, int C>
// *off_out: the off-diagonal change sum_{k != j} |new| - |old| of the returned
// shear (when it is not 0), so its evaluation needs no second sum.
__device__ __forceinline__ int64_t best_shear_any(const int64_t (&ai)[C], const int64_t (&aj)[C],
                                                  bool small, int g, int n, int j, int64_t a_ii,
                                                  int64_t a_ji, int64_t a_jj,
                                                  uint32_t* hist, int64_t* off_out) {
    if (small) {
        int32_t si[C], sj[C];
#pragma unroll
        for (int c = 0; c < C; ++c) {
            si[c] = (int32_t)ai[c];
            sj[c] = (int32_t)aj[c];
        }
        int32_t off = 0;
        const int64_t s = best_shear_small<G, C>(si, sj, g, n, j, a_ii, a_ji, a_jj, hist, &off);
        *off_out = off;
        return s;
    }
    return best_shear<G, C>(ai, aj, g, n, j, a_ii, a_ji, a_jj, off_out);
}

template <int G, int C>
__device__ __forceinline__ void evaluate_any(const int64_t (&ai)[C], const int64_t (&aj)[C],
                                             bool small, int g, int n, int j, int64_t s,
                                             int64_t a_ii, int64_t a_ji, int64_t a_jj,
                                             MoveEval<C>& e, bool know = false,
                                             int64_t known_off = 0) {
    if (small) {
        int32_t si[C], sj[C];
#pragma unroll
        for (int c = 0; c < C; ++c) {
            si[c] = (int32_t)ai[c];
            sj[c] = (int32_t)aj[c];
        }
        evaluate_small<G, C>(si, sj, g, n, j, s, a_ii, a_ji, a_jj, e, know, known_off);
    } else {
        evaluate<G, C>(ai, aj, g, n, j, s, a_ii, a_ji, a_jj, e, know, known_off);
    }
}

// ---------------------------------------------------------------------------
// the annealer
// ---------------------------------------------------------------------------

// A replica's A in LDS, in one of two layouts (a per-bucket choice that does
// not change any result):
//  * TRI: the upper triangle, row r holding columns r .. n-1, so that A[a][b]
//    sits at D(min(a, b)) + max(a, b) -- half the footprint, twice the
//    resident replicas where LDS bounds occupancy;
//  * square: rows of odd stride P = nmax + 1 (column accesses hit distinct
//    banks), both mirror entries written on commit -- fewer index operations
//    where occupancy is bound by registers instead.
// XL: X^T in LDS for the epoch (n x n row-major after the replica's other LDS
// arrays; loaded at the epoch start, written back at its end), else in HBM.
template <int G, int C, bool TRI, bool XL = false>
struct Replica {
    int64_t* sT;
    GLOBAL int64_t* gX; // HBM: X^T of this replica, n x n row-major (touched on accept)
    SHARED int64_t* sX; // XL: the LDS copy of X^T
    const uint8_t* act;
    int n, N, g, P;
    int Dk[C];          // TRI: D(k) of the lane's columns
    int64_t rm[C];      // row masses of the lane's rows

    // start of row r, less r for TRI (r (2n - r - 1) is even)
    __device__ __forceinline__ int D(int r) const {
        if constexpr (TRI) return (r * (2 * n - r - 1)) >> 1;
        return r * P;
    }
    __device__ __forceinline__ int at(int a, int b) const {
        if constexpr (TRI) return a <= b ? D(a) + b : D(b) + a;
        return a * P + b;
    }
    __device__ __forceinline__ void init_lanes() {
#pragma unroll
        for (int c = 0; c < C; ++c) Dk[c] = TRI ? D(g + c * G) : 0;
    }
    // A[r][k] for the lane's column k = g + c G, given D(r)
    __device__ __forceinline__ int at_lane(int r, int Dr, int c) const {
        const int k = g + c * G;
        if constexpr (TRI) return k >= r ? Dr + k : Dk[c] + r;
        return Dr + k;
    }

    __device__ __forceinline__ void load_rows(int i, int j, int64_t (&ai)[C],
                                              int64_t (&aj)[C]) const {
        const int Di = D(i), Dj = D(j);
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const int k = g + c * G;
            ai[c] = k < n ? sT[at_lane(i, Di, c)] : 0;
            aj[c] = k < n ? sT[at_lane(j, Dj, c)] : 0;
        }
    }
    __device__ __forceinline__ int64_t entry(int a, int b) const { return sT[at(a, b)]; }

    __device__ __forceinline__ int64_t xget(int r, int k) const {
        if constexpr (XL) return sX[r * n + k];
        return gX[r * n + k];
    }
    __device__ __forceinline__ void xset(int r, int k, int64_t v) {
        if constexpr (XL)
            sX[r * n + k] = v;
        else
            gX[r * n + k] = v;
    }

    // Rows i and j of X^T, issued early so the HBM latency overlaps the
    // evaluation of the move (only an accepted move consumes them).
    __device__ __forceinline__ void prefetch_xt(int i, int j, int64_t (&xi)[C],
                                                int64_t (&xj)[C]) const {
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const int k = g + c * G;
            xi[c] = k < n ? xget(i, k) : 0;
            xj[c] = k < n ? xget(j, k) : 0;
        }
    }

    // X^T (column-owned) to an n x n matrix in HBM, and from one
    __device__ __forceinline__ void x_to(GLOBAL int64_t* dst) const {
        for (int row = 0; row < n; ++row)
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const int k = g + c * G;
                if (k < n) dst[row * n + k] = xget(row, k);
            }
    }
    __device__ __forceinline__ void x_from(const GLOBAL int64_t* src) {
        for (int row = 0; row < n; ++row)
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const int k = g + c * G;
                if (k < n) xset(row, k, src[row * n + k]);
            }
    }

    __device__ __forceinline__ void compute_row_mass() {
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const int k = g + c * G;
            int64_t m = 0;
            if (k < n)
                for (int t = 0; t < n; ++t)
                    if (t != k) m += iabs64(sT[at_lane(t, D(t), c)]);
            rm[c] = m;
        }
    }

    // row_mass[row] as seen by any lane of the group
    __device__ __forceinline__ int64_t mass_of(int row) const {
        int64_t v = Grp<G>::gather(rm[0], row & (G - 1));
        if (C > 1) {
            const int64_t v1 = Grp<G>::gather(rm[C > 1 ? 1 : 0], row & (G - 1));
            v = (row >= G) ? v1 : v;
        }
        return v;
    }

    __device__ __forceinline__ void store_A(GLOBAL int64_t* dst) const {
        for (int row = 0; row < n; ++row) {
            const int Dr = D(row);
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const int k = g + c * G;
                if (k < n) dst[row * n + k] = sT[at_lane(row, Dr, c)];
            }
        }
    }

    // a full row-major n x n matrix in HBM (its upper triangle for TRI)
    __device__ __forceinline__ void load_A(const GLOBAL int64_t* src) {
        for (int row = 0; row < n; ++row) {
            const int Dr = D(row);
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const int k = g + c * G;
                if ((!TRI || k >= row) && k < n) sT[Dr + k] = src[row * n + k];
            }
        }
    }

    // Transform-row bound check of an evaluated Add(i, j, s)
    // (add_transform_row, main.cpp:97-108): false leaves the state untouched.
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

    // commit_add (main.cpp:112-133): row j of A and its mirror column (one
    // entry per lane in TRI), row j of X^T, row masses (row j's new mass is
    // its old mass plus the off-diagonal change).
    __device__ __forceinline__ void write(int j, const int64_t (&aj)[C], const MoveEval<C>& e,
                                          const int64_t (&nx)[C]) {
        const int Dj = D(j);
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const int k = g + c * G;
            if (k < n) {
                sT[at_lane(j, Dj, c)] = e.nv[c];
                if (!TRI && k != j) sT[k * P + j] = e.nv[c];
                rm[c] += (k != j) ? iabs64(e.nv[c]) - iabs64(aj[c]) : e.off;
                xset(j, k, nx[c]);
            }
        }
    }
};

// One lookahead proposal as the other waves of its replica read it (LDS).
struct alignas(16) Proposal {
    int64_t delta;
    int64_t s;
    double u;         // the Metropolis draw
    int32_t i, j;
    uint32_t w4z, w4w;  // the stall words
    uint32_t flags;     // 1: over the bound, 2: passes at the step's temperature, 4: greedy
    uint32_t pad[3];
};
static_assert(sizeof(Proposal) == 64, "two proposals per 128 bytes");

// One epoch of many replicas.
//
// K == 1: one replica per lane group, 64/G replicas per wave; four waves per
// SIMD (at most 128 VGPRs).
//
// K > 1 (G K == 64, lookahead): one replica per wave. Its K groups evaluate the
// replica's next K proposals (moves m+1 .. m+K) on its current state at once,
// and the moves are then resolved in order: each is decided exactly as the
// one-group loop decides it, and the first that changes the state (a
// committed move, a sweep commit, a reseed) ends the step; the proposals after
// it are made again, on the new state, in the next step. A move that leaves
// the state unchanged leaves every later proposal valid, so each resolved move
// is evaluated on exactly the state, and decided with exactly the temperature,
// that the sequential loop would use: the trajectory is the same, K moves per
// step while proposals are rejected. Used where replicas are too few to fill
// the device (the annealer rejects most proposals: 0.2% to 5% are accepted at
// the workloads' settings).
//
// K > 64 / G: the replica's K groups span W = G K / 64 waves of one workgroup
// (one wave per SIMD of a CU). Each group publishes its proposal to LDS; after
// a workgroup barrier every wave reads all K and resolves them identically; a
// committed move is evaluated again by every group from the shared state and
// written by all of them (each its own columns and row masses), after a
// barrier that lets every wave finish reading the state it changes. Stall
// sweeps and reseeds write the same way. Every wave thus takes the same
// branches and meets the same barriers.
//
// XL: the replica's X^T sits in LDS for the epoch (the last nmax^2 int64 of
// its LDS share) instead of being read and written in HBM on every proposal
// and commit.
template <int G, int C, bool TRI, int K, bool XL = false>
__global__ void __launch_bounds__(G * K > 256 ? 512 : 256, K == 1 ? 4 : 2)
    k_anneal(const RepRef* __restrict__ table, int count, Inst* __restrict__ insts, int nmax,
             int rep_bytes, const volatile int* __restrict__ stop) {
    constexpr int W = K == 1 ? 1 : G * K / 64;  // waves per replica
    static_assert(K == 1 || (G * K) % 64 == 0, "lookahead groups fill whole waves");
    static_assert(W <= 8, "a workgroup of at most 512 lanes");
    extern __shared__ __align__(16) unsigned char smem[];
    using Gp = Grp<G>;
    constexpr int U = G * K;  // lanes per replica
    const int g = (int)threadIdx.x & (G - 1);
    const int kk = K == 1 ? 0 : (int)(threadIdx.x % (unsigned)U) / G;  // lookahead index
    const int slot = (int)threadIdx.x / U;
    const int gid = (int)blockIdx.x * ((int)blockDim.x / U) + slot;
    if (gid >= count) return;
    // Every group of a wave belongs to the same instance (the host pads each
    // instance's replica list to whole waves), so the instance record is read
    // through scalar registers.
    const int wave_gid0 =
        __builtin_amdgcn_readfirstlane((int)blockIdx.x * ((int)blockDim.x / U) +
                                       (int)(threadIdx.x >> 6) * (64 / U));
    Inst* I = insts + __builtin_amdgcn_readfirstlane(table[wave_gid0].inst);
    if (I->status != ST_RUNNING) return;
    const int r = U >= 64 ? __builtin_amdgcn_readfirstlane(table[gid].rep) : table[gid].rep;
    if (r < 0) return;  // padding group
    GLOBAL RepState* S = (GLOBAL RepState*)I->rs + r;
    if (S->flags & RF_STOPPED) return;

    Replica<G, C, TRI, XL> rp;
    rp.sT = (int64_t*)(smem + (size_t)slot * rep_bytes);
    rp.sX = (SHARED int64_t*)((SHARED unsigned char*)smem + (size_t)slot * rep_bytes + rep_bytes -
                              (XL ? (size_t)nmax * nmax * 8 : 0));
    rp.P = nmax + 1;
    // A: nmax (nmax + 1) / 2 or nmax (nmax + 1) int64, a multiple of 32 bytes
    // for nmax % 8 == 0; then five staged Philox blocks and a median
    // histogram per group, and the active list
    uint4* const words0 = (uint4*)(rp.sT + (TRI ? (size_t)nmax * (nmax + 1) / 2
                                                : (size_t)nmax * (nmax + 1)));
    uint4* const swords = words0 + 5 * kk;
    uint8_t* sact = (uint8_t*)(words0 + 5 * K);
    uint32_t* shist = (uint32_t*)(sact + nmax) + kk * med_slots(G);
    // W > 1: two buffers of K proposals (by step parity) and the poll word
    Proposal* const props =
        (Proposal*)((((uintptr_t)((uint32_t*)(sact + nmax) + K * med_slots(G))) + 15) & ~(uintptr_t)15);
    volatile uint32_t* const spoll = (volatile uint32_t*)(props + 2 * K);
    rp.act = sact;
    rp.n = I->n;
    rp.N = I->N;
    rp.g = g;
    rp.init_lanes();
    const int n = rp.n, N = rp.N;
    const int64_t nn = (int64_t)n * n;
    GLOBAL int64_t* const gAcur = (GLOBAL int64_t*)I->A + (int64_t)r * nn;
    rp.gX = (GLOBAL int64_t*)I->Xt + (int64_t)r * nn;
    rp.load_A(gAcur);
    if constexpr (XL) rp.x_from(rp.gX);
    const GLOBAL uint8_t* gact = (const GLOBAL uint8_t*)I->act;
    bool nonid = false;
    for (int q = g; q < N; q += G) {
        const uint8_t a = gact[q];
        sact[q] = a;
        nonid |= a != q;
    }
    // every group of the wave holds the same list: wave-uniform
    const bool ident = N == n && __ballot(nonid) == 0ull;
    // index at position p of the active list
    auto rowat = [&](uint32_t p) -> int { return ident ? (int)p : (int)sact[p]; };
    for (int q = g; q < med_slots(G); q += G) shist[q] = 0u;
    rp.compute_row_mass();

    // Values read on every move stay in registers (scalar ones for a
    // whole-wave group); what only the stall handling, cooling steps, best
    // flushes, reseeds and the move log need is read from the instance record
    // where it is used.
    const uint32_t key0 = I->key0, key1 = I->key1;
    const uint64_t epoch_end = (uint64_t)(I->epoch + 1u) * (uint64_t)I->P.epoch_moves;
    const uint64_t cap = I->P.max_moves;  // 0: none
    const uint64_t stop_at = (cap != 0 && cap < epoch_end) ? cap : epoch_end;
    const bool tempering = I->P.tempering != 0;
    const int TS = I->P.target_samples;
    const int32_t stuck = I->P.stuck_threshold;
    const int32_t per_cool = I->P.moves_per_cool;
    const bool logging = I->log != nullptr;

    uint64_t moves = (uint64_t)uni<U>((int64_t)S->moves);
    int64_t score = uni<U>(S->score), offnz = uni<U>(S->offnz);
    int64_t local_best = uni<U>(S->local_best);
    // moves since improvement: reset once above the threshold (an int32)
    uint32_t msi = (uint32_t)uni<U>((int32_t)S->msi);
    double T = tempering ? ((const GLOBAL double*)I->temps)[S->rung] : S->T;
    double nbeta = -1.0 / T;  // the Metropolis exponent is delta * (-1/T)
    double cooling = S->cooling;
    int32_t cool_counter = uni<U>(S->cool_counter), win_p = uni<U>(S->win_p);
    int32_t win_a = uni<U>(S->win_a);
    int64_t best_score = uni<U>(S->best_score), best_move = uni<U>(S->best_move);
    int32_t flags = uni<U>(S->flags);
    const Coin greedy_coin(S->gf), target_coin(S->tf);
    // statistics counted from the epoch start (an epoch stays far below 2^32
    // events), added to the replica record at its end
    uint32_t accepts = 0, greedy_n = 0, infeasible = 0, sweeps = 0, logged = 0;

    // Lazy best flush: the pre-change state goes to bA/bXt only when it is the
    // replica's best and could still be the instance's best.
    auto leave_best = [&]() {
        if (flags & RF_CUR_BEST) {
            const Inst* Iq = opaque(I);
            if (best_score <= Iq->gb_score) {
                const int64_t off = (int64_t)r * n * n;
                if (kk == 0) {  // lookahead groups hold the same state
                    rp.store_A((GLOBAL int64_t*)Iq->bA + off);
                    rp.x_to((GLOBAL int64_t*)Iq->bXt + off);
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
    // Greedy reduction sweep from a pivot (congruence_anneal.hpp:246-260).
    auto sweep = [&](int pv, uint64_t m) __attribute__((always_inline)) -> bool {
        const int64_t before = score;
        ++sweeps;
        for (int q = 0; q < N; ++q) {
            const int jj = sact[q];
            if (jj == pv) continue;
            int64_t ai[C], aj[C];
            rp.load_rows(pv, jj, ai, aj);
            const int64_t a_ii = rp.entry(pv, pv);
            const int64_t a_ji = rp.entry(jj, pv);
            const int64_t a_jj = rp.entry(jj, jj);
            const bool small = rows_small<G, C>(ai, aj);
            int64_t soff = 0;
            int64_t s =
                best_shear_any<G, C>(ai, aj, small, g, n, jj, a_ii, a_ji, a_jj, shist, &soff);
            // lookahead groups all run the same sweep: their decisions are
            // wave-uniform
            if constexpr (K > 1) s = uni<64>(s);
            ++greedy_n;
            if (s == 0) continue;
            MoveEval<C> e;
            evaluate_any<G, C>(ai, aj, small, g, n, jj, s, a_ii, a_ji, a_jj, e, true, soff);
            if constexpr (K > 1) {
                e.bad = uni<64>((int32_t)e.bad) != 0;
                e.delta = uni<64>(e.delta);
                e.dnz = uni<64>(e.dnz);
            }
            if (e.bad || e.delta > 0) continue;
            int64_t xi[C], xj[C], nx[C];
            rp.prefetch_xt(pv, jj, xi, xj);
            bool fits = rp.xt_row(s, xi, xj, nx);
            if constexpr (K > 1) fits = uni<64>((int32_t)fits) != 0;
            if (!fits) continue;
            leave_best();
            if constexpr (W > 1) __syncthreads();  // every wave has read what changes
            rp.write(jj, aj, e, nx);
            score += e.delta;
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

    // Steps of the lookahead loop; the one-group loop spells the same steps
    // out inline.
    // ---- pair (congruence_anneal.hpp:283-288, :211-228) ----
    // `w` is this lane's Philox block; the move's slot c sits in lane base + c
    // (cs = g - base), in the group's 16-lane row `sub`.
    auto choose_pair = [&](const uint4& w, int base, int cs, int sub, uint32_t w0z,
                           uint32_t w2w, uint32_t w3z, auto w3w, int& i,
                           int& j) __attribute__((always_inline)) {
        if (target_coin(w0z)) {
            // hot-row tournament: first maximum of row_mass over the draws,
            // as one max over keys (mass << 5 | 16 | 15 - draw index; 0 for
            // no draw) within the 16-lane row holding this move's slots
            const bool in = cs >= 0 && cs < TS;
            const uint32_t p = in ? mulhi32(w.x, (uint32_t)N) : 0u;
            const int row = rowat(p);
            const int64_t wt = rp.mass_of(row);
            const uint64_t mx =
                Gp::slot_max(in ? (((uint64_t)wt << 5) | 16u | (uint32_t)(15 - cs)) : 0ull, sub);
            const uint32_t tpos = Gp::bcast(p, base + 15 - (int)(mx & 15));
            const int target = rowat(tpos);
            // pivot tournament: first maximum of |A[target][c]| over draws c != target
            const uint32_t pp = in ? mulhi32(w.y, (uint32_t)N) : 0u;
            const int cand = rowat(pp);
            const bool valid = in && cand != target;
            const int64_t v = valid ? iabs64(rp.entry(target, cand)) : 0;
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
    };

    // Bookkeeping after move m (res: -1 rejected, 0 committed, 1 improving)
    // and the stall handling that follows it (congruence_anneal.hpp:363-412).
    // True when the replica is solved; `changed` is set when a sweep commit or
    // a reseed changed the state.
    auto finish = [&](uint64_t m, int res, auto w4z, auto w4w,
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
                const InstParams& Q = opaque(I)->P;
                if (Q.sweep_fraction > 0.0 && u01(w4z()) < Q.sweep_fraction) {
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
                const int64_t gb_epoch = Iv->gb_score;
                if ((double)score > (double)gb_epoch * Iv->P.reseed_factor) {
                    leave_best();
                    log_move(m, -1, -1, 0, 2);
                    if constexpr (W > 1) __syncthreads();
                    rp.load_A((const GLOBAL int64_t*)Iv->gA);
                    rp.x_from((const GLOBAL int64_t*)Iv->gXt);
                    rp.compute_row_mass();
                    score = gb_epoch;
                    offnz = Iv->gb_offnz;
                    track_best(m);
                    changed = true;
                }
                local_best = score;
                msi = 0;
                cool_counter = 0;
            } else if (++cool_counter >= per_cool) {
                const InstParams& Q = opaque(I)->P;
                cool_counter = 0;
                if (Q.adaptive && win_p > 0) {
                    const double acc = (double)win_a / (double)win_p;
                    if (acc > Q.target_accept)
                        cooling *= 0.999;
                    else
                        cooling /= 0.999;
                    cooling = cooling < 0.90 ? 0.90 : (cooling > 0.99999 ? 0.99999 : cooling);
                }
                win_p = 0;
                win_a = 0;
                T *= cooling;
                if (T < Q.t_min) T = Q.t_min;
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
    // the solve key and the stop flag, every 256 steps
    auto poll = [&](uint32_t it) -> bool {
        if ((it & 255u) == 0u) {
            const unsigned long long key =
                __hip_atomic_load(&I->solve_key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (key != kNoKey && moves >= (key >> kRepBits)) return true;
            if (*stop) {
                halt = true;
                return true;
            }
        }
        return false;
    };
    // Philox slots of move m: counter (q, r, 16 t + c), q = (m-1)/4,
    // t = (m-1) mod 4, c < 16. A group of G lanes computes G slots per call
    // (lane = slot - 16 (t - sub)), i.e. one call serves G/16 moves; the
    // move's slot c sits in lane base + c, base = 16 sub.
    constexpr int SB = G / 16;

    if constexpr (K == 1) {
        // (the steps inline: composed from the helpers above, this loop
        // needs more registers than four waves per SIMD allow)
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
            // Philox slots of move m: counter (q, r, 16 t + c), q = (m-1)/4,
            // t = (m-1) mod 4, c < 16. A group of G lanes computes G slots per
            // call (lane = slot - 16 (t - sub)), i.e. one call serves G/16 moves;
            // this move's slot c sits in lane base + c.
            constexpr int S = G / 16;
            const uint32_t t4 = (uint32_t)((m - 1) & 3u);
            const int sub = (int)(t4 & (uint32_t)(S - 1));
            if (sub == 0 || !have_w) {
                const uint64_t q = (m - 1) >> 2;
                uint32_t k0 = key0, k1 = key1;  // round keys made here, not held
                opaque_u32(k0);
                opaque_u32(k1);
                w = philox10((uint32_t)q, (uint32_t)(q >> 32), (uint32_t)r,
                             16u * (t4 & ~(uint32_t)(S - 1)) + (uint32_t)g, k0, k1);
                have_w = true;
            }
            const int base = 16 * sub;
            const int cs = g - base;  // this lane's slot for move m (valid in [0, 16))
            // words of slots 0..4 to every lane of the group: v_readlane for a
            // whole-wave group (the ones only some moves need are read where they
            // are used), an LDS bounce otherwise
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
            auto word = [&](int slot, bool zw, uint32_t staged) -> uint32_t {
                if constexpr (G == 64) return readlane32(zw ? w.w : w.z, base + slot);
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
                // hot-row tournament: first maximum of row_mass over the draws,
                // as one max over keys (mass << 5 | 16 | 15 - draw index; 0 for
                // no draw) within the 16-lane row holding this move's slots
                const bool in = cs >= 0 && cs < TS;
                const uint32_t p = in ? mulhi32(w.x, (uint32_t)N) : 0u;
                const int row = rowat(p);
                const int64_t wt = rp.mass_of(row);
                const uint64_t mx =
                    Gp::slot_max(in ? (((uint64_t)wt << 5) | 16u | (uint32_t)(15 - cs)) : 0ull, sub);
                const uint32_t tpos = Gp::bcast(p, base + 15 - (int)(mx & 15));
                const int target = rowat(tpos);
                // pivot tournament: first maximum of |A[target][c]| over draws c != target
                const uint32_t pp = in ? mulhi32(w.y, (uint32_t)N) : 0u;
                const int cand = rowat(pp);
                const bool valid = in && cand != target;
                const int64_t v = valid ? iabs64(rp.entry(target, cand)) : 0;
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

            int64_t ai[C], aj[C], xi[C], xj[C];
            rp.load_rows(i, j, ai, aj);
            rp.prefetch_xt(i, j, xi, xj);
            const int64_t a_ii = rp.entry(i, i);
            const int64_t a_ji = rp.entry(j, i);
            const int64_t a_jj = rp.entry(j, j);

            // ---- shear (congruence_anneal.hpp:290-296) ----
            const bool small = rows_small<G, C>(ai, aj);
            const int64_t pm = (w1z & 1u) ? 1 : -1;
            int64_t s = pm, soff = 0;
            bool know = false;
            if (greedy_coin(w0w)) {
                s = best_shear_any<G, C>(ai, aj, small, g, n, j, a_ii, a_ji, a_jj, shist, &soff);
                ++greedy_n;
                know = s != 0;
                if (s == 0) s = pm;
            }

            // ---- attempt_add (congruence_anneal.hpp:232-240) ----
            MoveEval<C> e;
            evaluate_any<G, C>(ai, aj, small, g, n, j, s, a_ii, a_ji, a_jj, e, know, soff);
            int res = -1;
            if (e.bad) {
                ++infeasible;
            } else if (e.delta <= 0 || u53(w1w(), w2z()) < xexp((double)e.delta * nbeta)) {
                int64_t nx[C];
                if (rp.xt_row(s, xi, xj, nx)) {
                    const int64_t ns = score + e.delta;
                    if (ns < best_score) {
                        best_score = ns;
                        best_move = (int64_t)m;
                        flags |= RF_CUR_BEST;
                        flags &= ~RF_BEST_STALE;
                    } else {
                        leave_best();
                    }
                    rp.write(j, aj, e, nx);
                    score = ns;
                    offnz += e.dnz;
                    ++accepts;
                    log_move(m, i, j, s, 0);
                    res = e.delta < 0 ? 1 : 0;
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
                    const InstParams& Q = opaque(I)->P;
                    if (Q.sweep_fraction > 0.0 && u01(w4z()) < Q.sweep_fraction) {
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
                    const int64_t gb_epoch = Iv->gb_score;
                    if ((double)score > (double)gb_epoch * Iv->P.reseed_factor) {
                        leave_best();
                        log_move(m, -1, -1, 0, 2);
                        rp.load_A((const GLOBAL int64_t*)Iv->gA);
                        rp.x_from((const GLOBAL int64_t*)Iv->gXt);
                        rp.compute_row_mass();
                        score = gb_epoch;
                        offnz = Iv->gb_offnz;
                        track_best(m);
                    }
                    local_best = score;
                    msi = 0;
                    cool_counter = 0;
                } else if (++cool_counter >= per_cool) {
                    const InstParams& Q = opaque(I)->P;
                    cool_counter = 0;
                    if (Q.adaptive && win_p > 0) {
                        const double acc = (double)win_a / (double)win_p;
                        if (acc > Q.target_accept)
                            cooling *= 0.999;
                        else
                            cooling /= 0.999;
                        cooling = cooling < 0.90 ? 0.90 : (cooling > 0.99999 ? 0.99999 : cooling);
                    }
                    win_p = 0;
                    win_a = 0;
                    T *= cooling;
                    if (T < Q.t_min) T = Q.t_min;
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
            if (poll(it)) break;
            const uint64_t m0 = moves + 1;
            const uint64_t mg = m0 + (uint64_t)kk;
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

            int i, j;
            choose_pair(w, base, cs, sub, b0.z, b2.w, b3.z, [&]() { return b3.w; }, i, j);
            int64_t ai[C], aj[C], xi[C], xj[C];
            rp.load_rows(i, j, ai, aj);
            rp.prefetch_xt(i, j, xi, xj);
            const int64_t a_ii = rp.entry(i, i);
            const int64_t a_ji = rp.entry(j, i);
            const int64_t a_jj = rp.entry(j, j);
            const bool small = rows_small<G, C>(ai, aj);
            const int64_t pm = (b1.z & 1u) ? 1 : -1;
            int64_t s = pm, soff = 0;
            bool know = false;
            const bool gr = greedy_coin(b0.w);
            if (gr) {
                s = best_shear_any<G, C>(ai, aj, small, g, n, j, a_ii, a_ji, a_jj, shist, &soff);
                know = s != 0;
                if (s == 0) s = pm;
            }
            MoveEval<C> e;
            evaluate_any<G, C>(ai, aj, small, g, n, j, s, a_ii, a_ji, a_jj, e, know, soff);
            // Metropolis at the step's temperature; a move after a cooling
            // step inside this step is decided again at the new one
            const double u = u53(b1.w, b2.z);
            const double nb0 = nbeta;
            const bool pass = !e.bad && (e.delta <= 0 || u < xexp((double)e.delta * nb0));

            // moves m0 .. m0 + kn - 1 are resolvable (the epoch ends at stop_at)
            const int kn = (stop_at - moves) < (uint64_t)K ? (int)(stop_at - moves) : K;
            const uint64_t live = kn == K ? ~0ull : ((1ull << (kn * G)) - 1ull);
            if ((__ballot(pass) & live) == 0ull && score >= local_best &&
                (int64_t)msi + kn <= (int64_t)stuck && (tempering || cool_counter + kn < per_cool)) {
                // every resolvable move is rejected and none of them reaches a
                // stall or a cooling step: the sequential bookkeeping, kn times
                const uint64_t lead = __ballot(g == 0) & live;
                greedy_n += (uint32_t)__popcll(__ballot(gr) & lead);
                infeasible += (uint32_t)__popcll(__ballot(e.bad) & lead);
                moves += (uint64_t)kn;
                win_p += kn;
                msi += (uint32_t)kn;
                if (!tempering) cool_counter += kn;
                continue;
            }

            int64_t nx[C];
            bool feas = false, have_nx = false;
            bool solved = false;
            for (int k = 0; k < kn; ++k) {
                const int L = k * G;  // first lane of group k
                const uint64_t m = m0 + (uint64_t)k;
                moves = m;
                if (readlane32((uint32_t)gr, L)) ++greedy_n;
                int res = -1;
                bool changed = false;
                if (readlane32((uint32_t)e.bad, L)) {
                    ++infeasible;
                } else {
                    const int64_t d = readlane64(e.delta, L);
                    bool acc = readlane32((uint32_t)pass, L) != 0;
                    if (nbeta != nb0) acc = d <= 0 || readlane_f64(u, L) < xexp((double)d * nbeta);
                    if (acc && !have_nx) {
                        feas = rp.xt_row(s, xi, xj, nx);  // every group's, once
                        have_nx = true;
                    }
                    if (acc && readlane32((uint32_t)feas, L)) {
                        const int64_t ns = score + d;
                        if (ns < best_score) {
                            best_score = ns;
                            best_move = (int64_t)m;
                            flags |= RF_CUR_BEST;
                            flags &= ~RF_BEST_STALE;
                        } else {
                            leave_best();
                        }
                        // group k's move, written by every group: each keeps
                        // its own row masses and X^T columns
                        const int src = L + g;
                        const int jk = (int)readlane32((uint32_t)j, L);
                        int64_t cj[C], cx[C];
                        MoveEval<C> ce;
#pragma unroll
                        for (int c = 0; c < C; ++c) {
                            cj[c] = bperm64(aj[c], src);
                            ce.nv[c] = bperm64(e.nv[c], src);
                            cx[c] = bperm64(nx[c], src);
                        }
                        ce.off = readlane64(e.off, L);
                        rp.write(jk, cj, ce, cx);
                        score = ns;
                        offnz += (int32_t)readlane32((uint32_t)e.dnz, L);
                        ++accepts;
                        log_move(m, (int)readlane32((uint32_t)i, L), jk, readlane64(s, L), 0);
                        res = d < 0 ? 1 : 0;
                        changed = true;
                    }
                }
                auto k4z = [&]() { return readlane32(b4.z, L); };
                auto k4w = [&]() { return readlane32(b4.w, L); };
                if (finish(m, res, k4z, k4w, changed)) {
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
            const uint64_t mg = m0 + (uint64_t)kk;
            const uint32_t t4 = (uint32_t)((mg - 1) & 3u);
            const int sub = (int)(t4 & (uint32_t)(SB - 1));
            const uint64_t q = (mg - 1) >> 2;
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

            int i, j;
            choose_pair(w, base, cs, sub, b0.z, b2.w, b3.z, [&]() { return b3.w; }, i, j);
            int64_t ai[C], aj[C];
            rp.load_rows(i, j, ai, aj);
            const int64_t a_ii = rp.entry(i, i);
            const int64_t a_ji = rp.entry(j, i);
            const int64_t a_jj = rp.entry(j, j);
            const bool small = rows_small<G, C>(ai, aj);
            const int64_t pm = (b1.z & 1u) ? 1 : -1;
            int64_t s = pm, soff = 0;
            bool know = false;
            const bool gr = greedy_coin(b0.w);
            if (gr) {
                s = best_shear_any<G, C>(ai, aj, small, g, n, j, a_ii, a_ji, a_jj, shist, &soff);
                know = s != 0;
                if (s == 0) s = pm;
            }
            MoveEval<C> e;
            evaluate_any<G, C>(ai, aj, small, g, n, j, s, a_ii, a_ji, a_jj, e, know, soff);
            const double u = u53(b1.w, b2.z);
            const double nb0 = nbeta;
            const bool pass = !e.bad && (e.delta <= 0 || u < xexp((double)e.delta * nb0));
            Proposal* const pb = props + (it & 1u) * K;
            if (g == 0) {
                Proposal P;
                P.delta = e.delta;
                P.s = s;
                P.u = u;
                P.i = i;
                P.j = j;
                P.w4z = b4.z;
                P.w4w = b4.w;
                P.flags = (e.bad ? 1u : 0u) | (pass ? 2u : 0u) | (gr ? 4u : 0u);
                pb[kk] = P;
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

            // moves m0 .. m0 + kn - 1 are resolvable (the epoch ends at stop_at)
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
                const Proposal* const P = pb + k;
                const uint32_t f = (uint32_t)uni<64>((int32_t)P->flags);
                const uint64_t m = m0 + (uint64_t)k;
                moves = m;
                if (f & 4u) ++greedy_n;
                int res = -1;
                bool changed = false;
                if (f & 1u) {
                    ++infeasible;
                } else {
                    const int64_t d = uni<64>(P->delta);
                    bool acc = (f & 2u) != 0u;
                    if (nbeta != nb0) acc = d <= 0 || uni_f64(P->u) < xexp((double)d * nbeta);
                    if (acc) {
                        // every group evaluates move k again from the shared
                        // state: its rows, its X^T columns
                        const int ik = uni<64>(P->i), jk = uni<64>(P->j);
                        const int64_t sk = uni<64>(P->s);
                        int64_t ci[C], cj[C], xi[C], xj[C], nx[C];
                        rp.load_rows(ik, jk, ci, cj);
                        rp.prefetch_xt(ik, jk, xi, xj);
                        MoveEval<C> ce;
                        evaluate<G, C>(ci, cj, g, n, jk, sk, rp.entry(ik, ik), rp.entry(jk, ik),
                                       rp.entry(jk, jk), ce);
                        const bool fits = uni<64>((int32_t)rp.xt_row(sk, xi, xj, nx)) != 0;
                        if (fits) {
                            const int64_t ns = score + d;
                            if (ns < best_score) {
                                best_score = ns;
                                best_move = (int64_t)m;
                                flags |= RF_CUR_BEST;
                                flags &= ~RF_BEST_STALE;
                            } else {
                                leave_best();
                            }
                            __syncthreads();  // every wave has read what changes
                            rp.write(jk, cj, ce, nx);
                            score = ns;
                            offnz += uni<64>(ce.dnz);
                            ++accepts;
                            log_move(m, ik, jk, sk, 0);
                            res = d < 0 ? 1 : 0;
                            changed = true;
                        }
                    }
                }
                auto k4z = [&]() { return (uint32_t)uni<64>((int32_t)P->w4z); };
                auto k4w = [&]() { return (uint32_t)uni<64>((int32_t)P->w4w); };
                if (finish(m, res, k4z, k4w, changed)) {
                    solved = true;
                    break;
                }
                if (changed) break;
            }
            if (solved) break;
        }
    }
    // the move cap ends the loop through stop_at
    if (!(flags & RF_STOPPED) && !halt && cap != 0 && moves >= cap) flags |= RF_STOPPED | RF_CAPPED;

    if (kk != 0) return;  // lookahead: group 0 writes the replica back
    rp.store_A(gAcur);
    if constexpr (XL) rp.x_to(rp.gX);
    if (g == 0) {
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
    }
    if ((flags & RF_SOLVED) && g == 0)
        atomicMin(&I->solve_key, ((unsigned long long)moves << kRepBits) | (unsigned)r);
}

// ---------------------------------------------------------------------------
// epoch end: retirement, replica exchange, instance best
// ---------------------------------------------------------------------------

// Replica exchange between adjacent rungs (congruence_anneal.hpp:520-551):
// owner[q] is the replica on rung q; rungs q and q+1 swap when
// arg = (1/T_q - 1/T_{q+1}) (S_a - S_b) >= 0 or u < exp(arg), sweeping q
// upward so a configuration can travel several rungs in one sweep.
constexpr int kSweepChunk = 512;

// The instance left ST_RUNNING: its record and state are final (the kernels
// of its epoch have ended, and nothing writes them again). Everything written
// so far is made visible to the host and its copy engines, then the host's
// flag is raised.
__device__ __forceinline__ void raise_done(Inst* I) {
    int32_t* f = I->done_flag;
    if (f) {
        __threadfence_system();
        __hip_atomic_store(f, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// One block per instance: instance list[b] (b when list is null).
__global__ void __launch_bounds__(256) k_epoch_end(Inst* __restrict__ insts,
                                                   const int* __restrict__ list) {
    extern __shared__ __align__(16) unsigned char smem[];
    Inst* I = insts + (list ? list[blockIdx.x] : (int)blockIdx.x);
    if (I->status != ST_RUNNING) return;
    const int t = (int)threadIdx.x;
    const int R = I->R;
    const int n = I->n;
    __shared__ int s_running;
    __shared__ int64_t s_bs[256];
    __shared__ int64_t s_bm[256];
    __shared__ int s_br[256];
    if (t == 0) s_running = 0;
    __syncthreads();
    const unsigned long long key = I->solve_key;
    if (key != kNoKey) {
        if (t == 0) {
            I->status = ST_SOLVED;
            I->solve_rep = (int32_t)(key & ((1ull << kRepBits) - 1ull));
            I->solve_move = (int64_t)(key >> kRepBits);
            raise_done(I);
        }
        return;
    }
    int running = 0;
    for (int r = t; r < R; r += blockDim.x) running |= !(I->rs[r].flags & RF_STOPPED);
    if (running) atomicOr(&s_running, 1);

    // instance best: argmin (best_score, best_move, replica); a primal/dual
    // instance orders its double scores through okey
    const bool dual = I->dual != 0;
    int64_t bs = INT64_MAX, bm = INT64_MAX;
    int br = -1;
    for (int r = t; r < R; r += blockDim.x) {
        const RepState& x = I->rs[r];
        const int64_t xs = dual ? okey(x.dbest_score) : x.best_score;
        if (xs < bs || (xs == bs && x.best_move < bm)) {
            bs = xs;
            bm = x.best_move;
            br = r;
        }
    }
    s_bs[t] = bs;
    s_bm[t] = bm;
    s_br[t] = br;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (t < o) {
            const int64_t b2 = s_bs[t + o], m2 = s_bm[t + o];
            const int r2 = s_br[t + o];
            const bool better = r2 >= 0 && (b2 < s_bs[t] || (b2 == s_bs[t] && m2 < s_bm[t]) ||
                                            (b2 == s_bs[t] && m2 == s_bm[t] && r2 < s_br[t]));
            if (better) {
                s_bs[t] = b2;
                s_bm[t] = m2;
                s_br[t] = r2;
            }
        }
        __syncthreads();
    }
    const int bestr = s_br[0];
    const bool improve = s_bs[0] < (dual ? okey(I->gb_dscore) : I->gb_score);
    if (improve && !I->P.tempering) {
        const RepState& x = I->rs[bestr];
        const bool cur = (x.flags & RF_CUR_BEST) != 0;
        const int64_t nn = (int64_t)n * n;
        const int64_t* srcA = (cur ? I->A : I->bA) + (int64_t)bestr * nn;
        const int64_t* srcX = (cur ? I->Xt : I->bXt) + (int64_t)bestr * nn;
        for (int64_t q = t; q < nn; q += blockDim.x) {
            I->gA[q] = srcA[q];
            I->gXt[q] = srcX[q];
        }
        if (dual) {
            const double* srcQ = (cur ? I->Q : I->bQ) + (int64_t)bestr * nn;
            for (int64_t q = t; q < nn; q += blockDim.x) I->gQ[q] = srcQ[q];
        }
    }
    __syncthreads();
    if (t == 0 && improve) {
        if (dual)
            I->gb_dscore = I->rs[bestr].dbest_score;
        else
            I->gb_score = s_bs[0];
        I->gb_rep = bestr;
        // offnz of the snapshot: count it from the copied matrix
    }
    if (improve && !I->P.tempering) {
        __syncthreads();
        int64_t c = 0;
        for (int64_t q = t; q < (int64_t)n * n; q += blockDim.x) {
            const int row = (int)(q / n), col = (int)(q % n);
            c += (col > row && I->gA[q] != 0);
        }
        s_bs[t] = c;
        __syncthreads();
        for (int o = blockDim.x / 2; o > 0; o >>= 1) {
            if (t < o) s_bs[t] += s_bs[t + o];
            __syncthreads();
        }
        if (t == 0) I->gb_offnz = s_bs[0];
    }

    // exchange sweep: the configuration carried upward meets, at each rung
    // q + 1, the replica that started the sweep there, so everything the
    // sequential decisions read except the carried score is fixed in advance;
    // it is staged into LDS in chunks of kSweepChunk rungs by all threads and
    // thread 0 runs the decisions from LDS
    if (I->P.tempering && R >= 2) {
        int* owner = (int*)smem;  // rung -> replica; rungs below the chunk: final
        double* s_sb = (double*)(smem + ((size_t)R * sizeof(int) + 7) / 8 * 8);
        double* s_db = s_sb + kSweepChunk;
        double* s_u = s_db + kSweepChunk;
        double* s_lu = s_u + kSweepChunk;
        __shared__ int s_carry;
        __shared__ double s_sa;
        __shared__ int64_t s_acc;
        auto score_of = [&](int x) {
            return dual ? I->rs[x].dscore : (double)I->rs[x].score;
        };
        for (int r = t; r < R; r += blockDim.x) owner[I->rs[r].rung] = r;
        __syncthreads();
        if (t == 0) {
            s_carry = owner[0];
            s_sa = score_of(owner[0]);
            s_acc = 0;
        }
        const uint32_t e = I->epoch;
        for (int q0 = 0; q0 < R - 1; q0 += kSweepChunk) {
            const int qn = R - 1 - q0 < kSweepChunk ? R - 1 - q0 : kSweepChunk;
            for (int k = t; k < qn; k += blockDim.x) {
                const int q = q0 + k;
                s_sb[k] = score_of(owner[q + 1]);
                const double Ta = I->temps[q], Tb = I->temps[q + 1];
                s_db[k] = 1.0 / Ta - 1.0 / Tb;
                const uint4 w = philox10(e, 0u, (uint32_t)q, 0x40000000u, I->key0, I->key1);
                s_u[k] = u53(w.x, w.y);
                s_lu[k] = log(s_u[k]);
            }
            __syncthreads();
            if (t == 0) {
                int wa = s_carry;
                double Sa = s_sa;
                int64_t acc = 0;
                // 8 rungs' inputs in registers at a time, so the LDS latency
                // is paid once per 8 decisions instead of on every one
                constexpr int kU = 8;
                for (int k0 = 0; k0 < qn; k0 += kU) {
                    int wbv[kU];
                    double sbv[kU], dbv[kU], luv[kU];
#pragma unroll
                    for (int j = 0; j < kU; ++j) {
                        const int k = k0 + j < qn ? k0 + j : qn - 1;
                        wbv[j] = owner[q0 + k + 1];
                        sbv[j] = s_sb[k];
                        dbv[j] = s_db[k];
                        luv[j] = s_lu[k];
                    }
#pragma unroll
                    for (int j = 0; j < kU; ++j) {
                        if (k0 + j >= qn) break;
                        const double arg = dbv[j] * (Sa - sbv[j]);
                        // u < xexp(arg) decided from log u where that is
                        // certain: xexp is within 4e-16 relative of exp above
                        // -40 (0 at or below), log within a few ulps, so a
                        // margin of 1e-6 leaves the exact test only for the
                        // rare close call
                        bool take;
                        if (arg >= 0.0)
                            take = true;
                        else if (!(arg > -40.0))
                            take = false;
                        else if (arg > luv[j] + 1e-6)
                            take = true;
                        else if (arg < luv[j] - 1e-6)
                            take = false;
                        else
                            take = s_u[k0 + j] < xexp(arg);
                        if (take) {
                            owner[q0 + k0 + j] = wbv[j];  // wb moves down, wa goes on up
                            ++acc;
                        } else {
                            owner[q0 + k0 + j] = wa;
                            wa = wbv[j];
                            Sa = sbv[j];
                        }
                    }
                }
                s_carry = wa;
                s_sa = Sa;
                s_acc += acc;
            }
            __syncthreads();
        }
        if (t == 0) {
            owner[R - 1] = s_carry;
            I->exch_tried += R - 1;
            I->exch_acc += s_acc;
        }
        __syncthreads();
        for (int q = t; q < R; q += blockDim.x) I->rs[owner[q]].rung = q;
    }
    __syncthreads();
    if (t == 0) {
        I->epoch = I->epoch + 1u;
        if (!s_running || (I->P.max_epochs > 0 && I->epoch >= (uint32_t)I->P.max_epochs)) {
            I->status = ST_EXHAUSTED;
            raise_done(I);
        }
    }
}

// ---------------------------------------------------------------------------
// probes (parity tests of the building blocks)
// ---------------------------------------------------------------------------

// One query per group: q = {mat, i, j, s}; out = {feasible, delta, dnz,
// best_shear(i, j), xt_feasible, delta of the best shear evaluated with the
// off-diagonal sum its search returned (0 when it is 0)}.
template <int G, int C>
__global__ void __launch_bounds__(64)
    k_probe(const int64_t* __restrict__ A, const int64_t* __restrict__ Xt, int n,
            const int32_t* __restrict__ queries, int nq, int64_t* __restrict__ out) {
    __shared__ uint32_t hist[(64 / G) * med_slots(G)];
    const int g = (int)threadIdx.x & (G - 1);
    const int q = (int)blockIdx.x * (64 / G) + (int)threadIdx.x / G;
    for (int t = (int)threadIdx.x; t < (64 / G) * med_slots(G); t += 64) hist[t] = 0u;
    __syncthreads();
    if (q >= nq) return;
    const int mat = queries[4 * q + 0], i = queries[4 * q + 1], j = queries[4 * q + 2];
    const int64_t s = queries[4 * q + 3];
    const int64_t* M = A + (int64_t)mat * n * n;
    const int64_t* X = Xt + (int64_t)mat * n * n;
    int64_t ai[C], aj[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const int k = g + c * G;
        ai[c] = k < n ? M[i * n + k] : 0;
        aj[c] = k < n ? M[j * n + k] : 0;
    }
    const int64_t a_ii = M[i * n + i], a_ji = M[j * n + i], a_jj = M[j * n + j];
    const bool small = rows_small<G, C>(ai, aj);
    MoveEval<C> e;
    evaluate_any<G, C>(ai, aj, small, g, n, j, s, a_ii, a_ji, a_jj, e);
    int64_t boff = 0;
    const int64_t b = best_shear_any<G, C>(ai, aj, small, g, n, j, a_ii, a_ji, a_jj,
                                           hist + ((int)threadIdx.x / G) * med_slots(G), &boff);
    MoveEval<C> eb;
    eb.delta = 0;
    if (b != 0) evaluate_any<G, C>(ai, aj, small, g, n, j, b, a_ii, a_ji, a_jj, eb, true, boff);
    bool bx = false;
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const int k = g + c * G;
        if (k < n) bx |= over(X[j * n + k] + s * X[i * n + k]);
    }
    bx = Grp<G>::any(bx);
    if (g == 0) {
        out[6 * q + 0] = !e.bad;
        out[6 * q + 1] = e.delta;
        out[6 * q + 2] = e.dnz;
        out[6 * q + 3] = b;
        out[6 * q + 4] = !bx;
        out[6 * q + 5] = b != 0 ? eb.delta : 0;
    }
}

__global__ void k_probe_rng(const uint32_t* __restrict__ ck, int count, uint32_t* __restrict__ out,
                            const double* __restrict__ x, int nx, double* __restrict__ ex) {
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (t < count) {
        const uint32_t* c = ck + 6 * t;
        const uint4 w = philox10(c[0], c[1], c[2], c[3], c[4], c[5]);
        out[4 * t + 0] = w.x;
        out[4 * t + 1] = w.y;
        out[4 * t + 2] = w.z;
        out[4 * t + 3] = w.w;
    }
    if (t < nx) ex[t] = xexp(x[t]);
}

// ---------------------------------------------------------------------------
// host-side launch helpers
// ---------------------------------------------------------------------------


typedef void (*AnnealKernel)(const RepRef*, int, Inst*, int, int, const volatile int*);

// Every (G, C, layout, lookahead) the launcher can run: G lanes per group, C
// columns per lane (G C >= n), triangular or square A, K groups per replica.
struct AnnealVariant {
    int G, C;
    bool tri;
    int K;
    AnnealKernel fn;
    bool xl = false;  // X^T in LDS
};
static const AnnealVariant kVariants[] = {
    {16, 1, true, 1, k_anneal<16, 1, true, 1>},   {16, 1, false, 1, k_anneal<16, 1, false, 1>},
    {32, 1, true, 1, k_anneal<32, 1, true, 1>},   {32, 1, false, 1, k_anneal<32, 1, false, 1>},
    {32, 2, true, 1, k_anneal<32, 2, true, 1>},   {32, 2, false, 1, k_anneal<32, 2, false, 1>},
    {64, 1, true, 1, k_anneal<64, 1, true, 1>},   {64, 1, false, 1, k_anneal<64, 1, false, 1>},
    {64, 2, true, 1, k_anneal<64, 2, true, 1>},   {64, 2, false, 1, k_anneal<64, 2, false, 1>},
    {16, 1, true, 4, k_anneal<16, 1, true, 4>},   {16, 2, true, 4, k_anneal<16, 2, true, 4>},
    {32, 1, true, 2, k_anneal<32, 1, true, 2>},   {32, 2, true, 2, k_anneal<32, 2, true, 2>},
    {16, 1, true, 16, k_anneal<16, 1, true, 16>}, {16, 2, true, 16, k_anneal<16, 2, true, 16>},
    {16, 1, false, 16, k_anneal<16, 1, false, 16>}, {16, 2, false, 16, k_anneal<16, 2, false, 16>},
    {32, 2, true, 8, k_anneal<32, 2, true, 8>},
    {16, 1, true, 32, k_anneal<16, 1, true, 32>}, {16, 2, true, 32, k_anneal<16, 2, true, 32>},
    {64, 2, true, 4, k_anneal<64, 2, true, 4>},
    {64, 2, true, 4, k_anneal<64, 2, true, 4, true>, true},
};

static const AnnealVariant* find_variant(int G, int C, bool tri, int K, bool xl) {
    for (const auto& v : kVariants)
        if (v.G == G && v.C == C && v.tri == tri && v.K == K && v.xl == xl) return &v;
    return nullptr;
}

bool anneal_variant_exists(int G, int C, bool tri, int K, bool xl) {
    return find_variant(G, C, tri, K, xl) != nullptr;
}

// Dynamic LDS limits are raised once, outside any stream capture, to the
// per-block maximum of the device minus each kernel's static LDS.
hipError_t init_kernel_attrs(int max_lds_per_block) {
    std::vector<const void*> fns;
    for (const auto& v : kVariants) fns.push_back((const void*)v.fn);
    fns.push_back((const void*)k_epoch_end);
    for (const void* f : fns) {
        hipFuncAttributes fa;
        hipError_t e = hipFuncGetAttributes(&fa, f);
        if (e != hipSuccess) return e;
        const int dyn = max_lds_per_block - (int)fa.sharedSizeBytes;
        e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, dyn);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// units_per_block: table entries (replicas, or padding) per block, G K lanes each
hipError_t launch_anneal(hipStream_t st, int G, int C, bool tri, int K, bool xl,
                         int units_per_block, const RepRef* table, int count, Inst* insts,
                         int nmax, int rep_bytes, const volatile int* stop) {
    if (count <= 0) return hipSuccess;
    const AnnealVariant* v = find_variant(G, C, tri, K, xl);
    if (!v || units_per_block < 1 || units_per_block * G * K > (G * K > 256 ? 512 : 256))
        return hipErrorInvalidValue;
    const int threads = units_per_block * G * K;
    const int blocks = (count + units_per_block - 1) / units_per_block;
    const size_t lds = (size_t)units_per_block * rep_bytes;
    hipLaunchKernelGGL(v->fn, dim3(blocks), dim3(threads), lds, st, table, count, insts, nmax,
                       rep_bytes, stop);
    return hipGetLastError();
}

// list (nullable): the count instances to close, as indices into insts
hipError_t launch_epoch_end(hipStream_t st, Inst* insts, int count, int max_r, const int* list) {
    if (count <= 0) return hipSuccess;
    const size_t lds = ((size_t)max_r * sizeof(int) + 7) / 8 * 8 + 4 * kSweepChunk * sizeof(double);
    hipLaunchKernelGGL(k_epoch_end, dim3(count), dim3(256), lds, st, insts, list);
    return hipGetLastError();
}

// Outcomes of finished instances gathered into one staging buffer (one block
// per instance), so a host copies many instances' results at once. Block q
// writes at offs[q]: int32 rep, int32 status (0, or 1 when the best replica's
// snapshot is stale), then the R replica records, then the result replica's
// A and X^T: the solving replica when the instance is solved, else the
// instance best, argmin (best_score, best_move, replica), from its current
// state or its best snapshot.
__global__ void __launch_bounds__(256) k_gather_outcomes(const Inst* __restrict__ insts,
                                                          const int* __restrict__ slots,
                                                          const int64_t* __restrict__ offs,
                                                          unsigned char* __restrict__ out) {
    const Inst* I = insts + slots[blockIdx.x];
    unsigned char* dst = out + offs[blockIdx.x];
    const int R = I->R;
    const int64_t nn = (int64_t)I->n * I->n;
    __shared__ int s_rep;
    __shared__ int64_t s_bs[256], s_bm[256];
    __shared__ int s_br[256];
    const int t = (int)threadIdx.x;
    int64_t bs = INT64_MAX, bm = INT64_MAX;
    int br = -1;
    for (int r = t; r < R; r += blockDim.x) {
        const RepState& x = I->rs[r];
        if (x.best_score < bs || (x.best_score == bs && x.best_move < bm)) {
            bs = x.best_score;
            bm = x.best_move;
            br = r;
        }
    }
    s_bs[t] = bs;
    s_bm[t] = bm;
    s_br[t] = br;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            const int u = t + w;
            const bool take = s_br[u] >= 0 &&
                              (s_br[t] < 0 || s_bs[u] < s_bs[t] ||
                               (s_bs[u] == s_bs[t] &&
                                (s_bm[u] < s_bm[t] || (s_bm[u] == s_bm[t] && s_br[u] < s_br[t]))));
            if (take) {
                s_bs[t] = s_bs[u];
                s_bm[t] = s_bm[u];
                s_br[t] = s_br[u];
            }
        }
        __syncthreads();
    }
    if (t == 0) {
        int rep = s_br[0], err = 0;
        if (I->status == ST_SOLVED) {
            rep = I->solve_rep;
        } else {
            const int32_t f = I->rs[rep].flags;
            err = (!(f & RF_CUR_BEST) && (f & RF_BEST_STALE)) ? 1 : 0;
        }
        s_rep = rep;
        ((int32_t*)dst)[0] = rep;
        ((int32_t*)dst)[1] = err;
    }
    __syncthreads();
    const int rep = s_rep;
    // the replica records, as 8-byte words
    const int64_t* rsw = (const int64_t*)I->rs;
    int64_t* d = (int64_t*)(dst + 16);
    const int64_t nrs = (int64_t)R * (int64_t)(sizeof(RepState) / 8);
    for (int64_t q = t; q < nrs; q += blockDim.x) d[q] = rsw[q];
    d += nrs;
    const bool cur = I->status == ST_SOLVED || (I->rs[rep].flags & RF_CUR_BEST) != 0;
    const int64_t* sa = (cur ? I->A : I->bA) + (int64_t)rep * nn;
    const int64_t* sx = (cur ? I->Xt : I->bXt) + (int64_t)rep * nn;
    for (int64_t q = t; q < nn; q += blockDim.x) {
        d[q] = sa[q];
        d[nn + q] = sx[q];
    }
}

hipError_t launch_gather_outcomes(hipStream_t st, const Inst* insts, const int* slots,
                                  const int64_t* offs, int count, unsigned char* out) {
    if (count <= 0) return hipSuccess;
    static_assert(sizeof(RepState) % 8 == 0, "replica records copy as 8-byte words");
    hipLaunchKernelGGL(k_gather_outcomes, dim3(count), dim3(256), 0, st, insts, slots, offs, out);
    return hipGetLastError();
}

__global__ void k_replicate(int64_t* __restrict__ dst, const int64_t* __restrict__ src, int64_t nn,
                            int R) {
    const int64_t total = nn * R;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total;
         q += (int64_t)gridDim.x * blockDim.x)
        dst[q] = src[q % nn];
}

hipError_t launch_replicate(hipStream_t st, int64_t* dst, const int64_t* src, int64_t nn, int R) {
    const int64_t total = nn * R;
    if (total <= 0) return hipSuccess;
    const int blocks = (int)std::min<int64_t>((total + 255) / 256, 4096);
    hipLaunchKernelGGL(k_replicate, dim3(blocks), dim3(256), 0, st, dst, src, nn, R);
    return hipGetLastError();
}

hipError_t launch_probe(hipStream_t st, int G, int C, const int64_t* A, const int64_t* Xt, int n,
                        const int32_t* q, int nq, int64_t* out) {
    const int per = 64 / G;
    const dim3 grid((nq + per - 1) / per), block(64);
    if (G == 16 && C == 1)
        hipLaunchKernelGGL((k_probe<16, 1>), grid, block, 0, st, A, Xt, n, q, nq, out);
    else if (G == 32 && C == 1)
        hipLaunchKernelGGL((k_probe<32, 1>), grid, block, 0, st, A, Xt, n, q, nq, out);
    else if (G == 64 && C == 1)
        hipLaunchKernelGGL((k_probe<64, 1>), grid, block, 0, st, A, Xt, n, q, nq, out);
    else if (G == 64 && C == 2)
        hipLaunchKernelGGL((k_probe<64, 2>), grid, block, 0, st, A, Xt, n, q, nq, out);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_probe_rng(hipStream_t st, const uint32_t* ck, int count, uint32_t* out,
                            const double* x, int nx, double* ex) {
    const int tot = count > nx ? count : nx;
    if (tot <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_probe_rng, dim3((tot + 255) / 256), dim3(256), 0, st, ck, count, out, x,
                       nx, ex);
    return hipGetLastError();
}

}  // namespace xtx
