// anneal_persist.hip -- the batch in one persistent launch (gfx950).
//
// Every wave of the grid is a worker. Work is a queue of units (one wave's
// replicas of one instance for one epoch, PUnit): a wave takes the next slot,
// waits until the slot holds a unit, runs that unit's epoch with the lane
// kernel's code (lane_epoch.h) and counts it off its instance. The wave that
// finishes an instance's last unit of the epoch runs the instance's epoch end
// itself (solve retirement, instance best, replica exchange: what k_epoch_end
// does with a block, here with one wave) and, while the instance runs on,
// appends its next epoch's units to the queue. Instances thus advance epoch by
// epoch independently, with no device-wide join and no launch between epochs:
// a wave that finishes a unit takes whatever unit is ready, of its home order
// class first (orders 8-16, 24-32 or 40-48): the two compute units that share
// an instruction cache keep to one class's code while it has work. Measured,
// waves that took units of every order as they came ran the mixed batch 10%
// slower than batches of one order class.
//
// One queue per XCD. The device is eight dies, each with its own L2, and a
// hand-off between dies would have to write back and invalidate a die's whole
// L2 (what an agent-scope fence does here): measured, that made the launch
// 25x slower than the per-epoch launches, every wave's spill reloads and
// row reads missing the L2 it shares. A wave reads its die's id from the
// hardware (XCC_ID) and serves only that die's queue; the host places every
// instance on one die, so a unit's successor reads what it left through the
// same L2: the writer waits until its stores are complete (s_waitcnt) and the
// reader invalidates its compute unit's L1 (buffer_inv sc1, which leaves the
// L2 lines of device memory in place) -- no L2 write-back.
//
// The trajectory of every replica is the per-epoch launches' (the same
// lane_epoch on the same state; the epoch end decides the same exchanges from
// the same scores and Philox words).
//
// Termination: a wave takes a unit only from below a queue's tail (a slot an
// appending wave has reserved and fills next), so it never waits on work that
// may not come; with every queue of its XCD empty it leaves once the XCD's
// `live` count (instances that can still append units) is 0, or when the stop
// word is set (it then sets every `live` far below 0, so every other wave
// leaves too). Every unit in a queue belongs to an instance still counted
// live, so the waits end whatever the set of resident waves.
#include <hip/hip_runtime.h>

#include "lane_epoch.h"

namespace xtx {

namespace {

__device__ __forceinline__ double readlane_f64(double v, int k) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, k);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), k);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// What a wave leaves for a wave of another compute unit of the same XCD: its
// stores complete in the shared L2. Then the reader: its L1 drops every line,
// so it reads what the L2 holds.
__device__ __forceinline__ void xcd_release() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_waitcnt(0);
}
__device__ __forceinline__ void xcd_acquire() {
    // agent scope: the L1 (and L2 lines of MTYPE NC memory; device memory is
    // RW, whose L2 lines stay); the work-group scope form is a no-op outside
    // threadgroup-split mode
    asm volatile("buffer_inv sc1" ::: "memory");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// The die (XCD) the wave runs on.
__device__ __forceinline__ int xcc_id() {
    return __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20);  // hwreg(HW_REG_XCC_ID, 0, 4)
}
// The compute-unit pair the wave runs on, within its XCD: HW_ID bits 15:9
// (shader engine, shader array, CU id without its lowest bit; the two compute
// units of a pair share an instruction cache).
__device__ __forceinline__ int cu_pair_key() {
    return __builtin_amdgcn_s_getreg((6 << 11) | (9 << 6) | 4);  // hwreg(HW_REG_HW_ID, 9, 7)
}

// The host's done flag for an instance that left ST_RUNNING: every lane's
// writes are made visible to the host first.
__device__ __forceinline__ void raise_done_wave(GLOBAL Inst* I, int lane) {
    int32_t* f = I->done_flag;
    __threadfence_system();
    if (f && lane == 0) __hip_atomic_store(f, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// The outcome of an instance that left ST_RUNNING, written into host-mapped
// memory by the wave that ended it (the host then retires it with no copy and
// no kernel of its own, while the launch runs): {rep, err, status, 0,
// solve_move, 0, 0}, the R replica records, then the reported replica's A and
// X^T -- the solving replica, else the instance best (k_gather_outcomes'
// choice and layout after its 16-byte head). err: the instance best's state
// was not kept (cannot happen).
__device__ void write_outcome(GLOBAL Inst* I, unsigned char* out, int rep, int32_t status,
                              int64_t solve_move, int lane) {
    const int R = I->R;
    const int64_t nn = (int64_t)I->n * I->n;
    const GLOBAL RepState* rs = (const GLOBAL RepState*)I->rs;
    const int32_t f = rs[rep].flags;
    const bool cur = status == ST_SOLVED || (f & RF_CUR_BEST) != 0;
    const int err = (status != ST_SOLVED && !(f & RF_CUR_BEST) && (f & RF_BEST_STALE)) ? 1 : 0;
    int64_t* d = (int64_t*)out;
    if (lane == 0) {
        ((int32_t*)out)[0] = rep;
        ((int32_t*)out)[1] = err;
        ((int32_t*)out)[2] = status;
        ((int32_t*)out)[3] = 0;
        d[2] = solve_move;
        d[3] = 0;
    }
    d += 4;
    const GLOBAL int64_t* rsw = (const GLOBAL int64_t*)rs;
    const int64_t nrs = (int64_t)R * (int64_t)(sizeof(RepState) / 8);
    for (int64_t q = lane; q < nrs; q += 64) d[q] = rsw[q];
    d += nrs;
    const GLOBAL int64_t* sa = (const GLOBAL int64_t*)(cur ? I->A : I->bA) + (int64_t)rep * nn;
    const GLOBAL int64_t* sx = (const GLOBAL int64_t*)(cur ? I->Xt : I->bXt) + (int64_t)rep * nn;
    for (int64_t q = lane; q < nn; q += 64) {
        d[q] = sa[q];
        d[nn + q] = sx[q];
    }
}

// The epoch end of one instance (k_epoch_end's semantics, anneal.hip) run by
// one wave: solve retirement, the instance best (argmin of (best score, best
// move, replica)), the cooling-mode snapshot, the exchange sweep between
// adjacent rungs (congruence_anneal.hpp:520-551: rungs q and q+1 swap when
// arg = (1/T_q - 1/T_{q+1}) (S_a - S_b) >= 0 or u < exp(arg), sweeping q
// upward), the epoch count and exhaustion. `owner`: R ints of LDS. An
// instance that ends has its outcome written to `out` (host-mapped; null:
// none) before its flag is raised. Returns true when the instance still runs.
// Wave-uniform; L1 objective, R <= 256.
__device__ bool epoch_end_wave(GLOBAL Inst* I, SHARED int* owner, unsigned char* out) {
    const int lane = (int)(threadIdx.x & 63u);
    const int R = I->R;
    const int n = I->n;
    const unsigned long long key = I->solve_key;
    if (key != kNoKey) {
        const int rep = (int)(key & ((1ull << kRepBits) - 1ull));
        const int64_t sm = (int64_t)(key >> kRepBits);
        if (lane == 0) {
            I->status = ST_SOLVED;
            I->solve_rep = (int32_t)rep;
            I->solve_move = sm;
        }
        if (out) write_outcome(I, out, rep, ST_SOLVED, sm, lane);
        raise_done_wave(I, lane);
        return false;
    }
    GLOBAL RepState* rs = (GLOBAL RepState*)I->rs;
    bool run_l = false;
    int64_t bs = INT64_MAX, bm = INT64_MAX;
    int br = -1;
    for (int r = lane; r < R; r += 64) {
        const int32_t fl = rs[r].flags;
        run_l |= !(fl & RF_STOPPED);
        const int64_t xs = rs[r].best_score, xm = rs[r].best_move;
        if (xs < bs || (xs == bs && xm < bm)) {
            bs = xs;
            bm = xm;
            br = r;
        }
    }
    const bool running = __ballot(run_l) != 0ull;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const int64_t b2 = __shfl_xor(bs, o), m2 = __shfl_xor(bm, o);
        const int r2 = __shfl_xor(br, o);
        const bool better = r2 >= 0 && (br < 0 || b2 < bs || (b2 == bs && m2 < bm) ||
                                        (b2 == bs && m2 == bm && r2 < br));
        if (better) {
            bs = b2;
            bm = m2;
            br = r2;
        }
    }
    bs = (int64_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(uint64_t)bs) |
         ((int64_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)bs >> 32)) << 32);
    br = __builtin_amdgcn_readfirstlane(br);
    const bool tempering = I->P.tempering != 0;
    const bool improve = bs < I->gb_score;
    if (improve && !tempering) {
        // cooling mode: the instance best becomes the reseed snapshot
        const bool cur = (rs[br].flags & RF_CUR_BEST) != 0;
        const int64_t nn = (int64_t)n * n;
        const GLOBAL int64_t* srcA = (const GLOBAL int64_t*)(cur ? I->A : I->bA) + (int64_t)br * nn;
        const GLOBAL int64_t* srcX = (const GLOBAL int64_t*)(cur ? I->Xt : I->bXt) + (int64_t)br * nn;
        GLOBAL int64_t* gA = (GLOBAL int64_t*)I->gA;
        GLOBAL int64_t* gX = (GLOBAL int64_t*)I->gXt;
        int64_t c = 0;
        for (int64_t q = lane; q < nn; q += 64) {
            const int64_t a = srcA[q];
            gA[q] = a;
            gX[q] = srcX[q];
            c += ((int)(q % n) > (int)(q / n) && a != 0);
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
        if (lane == 0) I->gb_offnz = c;
    }
    if (improve && lane == 0) {
        I->gb_score = bs;
        I->gb_rep = br;
    }

    if (tempering && R >= 2) {
        for (int r = lane; r < R; r += 64) owner[rs[r].rung] = r;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const GLOBAL double* temps = (const GLOBAL double*)I->temps;
        const uint32_t e = I->epoch;
        const uint32_t k0 = I->key0, k1 = I->key1;
        int carry = owner[0];
        double Sa = (double)rs[carry].score;
        int64_t acc = 0;
        // lane k stages rung q0 + k (everything the sequential decisions read
        // except the carried score is fixed in advance: the configuration
        // carried upward meets, at rung q + 1, the replica that started the
        // sweep there); the decisions then run in step on every lane
        for (int q0 = 0; q0 < R - 1; q0 += 64) {
            const int qn = R - 1 - q0 < 64 ? R - 1 - q0 : 64;
            int wb = 0;
            double sb = 0.0, db = 0.0, u = 0.0, lu = 0.0;
            if (lane < qn) {
                const int q = q0 + lane;
                wb = owner[q + 1];
                sb = (double)rs[wb].score;
                const double Ta = temps[q], Tb = temps[q + 1];
                db = 1.0 / Ta - 1.0 / Tb;
                const uint4 w = philox10(e, 0u, (uint32_t)q, 0x40000000u, k0, k1);
                u = u53(w.x, w.y);
                lu = log(u);
            }
            int newo = 0;
            for (int k = 0; k < qn; ++k) {
                const int wbk = __builtin_amdgcn_readlane(wb, k);
                const double sbk = readlane_f64(sb, k), dbk = readlane_f64(db, k);
                const double luk = readlane_f64(lu, k), uk = readlane_f64(u, k);
                const double arg = dbk * (Sa - sbk);
                // u < xexp(arg) decided from log u where that is certain:
                // xexp is within 4e-16 relative of exp above -40 (0 at or
                // below), log within a few ulps, so a margin of 1e-6 leaves the
                // exact test only for the rare close call
                bool take;
                if (arg >= 0.0)
                    take = true;
                else if (!(arg > -40.0))
                    take = false;
                else if (arg > luk + 1e-6)
                    take = true;
                else if (arg < luk - 1e-6)
                    take = false;
                else
                    take = uk < xexp(arg);
                int o;
                if (take) {
                    o = wbk;  // wb moves down, the carried one goes on up
                    ++acc;
                } else {
                    o = carry;
                    carry = wbk;
                    Sa = sbk;
                }
                if (lane == k) newo = o;
            }
            // rungs q0 .. q0 + qn - 1 are final; the next chunk reads owner
            // above q0 + qn only
            __builtin_amdgcn_wave_barrier();
            if (lane < qn) owner[q0 + lane] = newo;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        if (lane == 0) {
            owner[R - 1] = carry;
            I->exch_tried += R - 1;
            I->exch_acc += acc;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int q = lane; q < R; q += 64) rs[owner[q]].rung = q;
    }
    const uint32_t ne = I->epoch + 1u;
    const bool exhausted = !running || (I->P.max_epochs > 0 && ne >= (uint32_t)I->P.max_epochs);
    if (lane == 0) {
        I->epoch = ne;
        if (exhausted) I->status = ST_EXHAUSTED;
    }
    if (exhausted) {
        if (out) write_outcome(I, out, br, ST_EXHAUSTED, -1, lane);
        raise_done_wave(I, lane);
    }
    return !exhausted;
}

// One unit at order NMAX (Q proposals per replica): replicas r0 .. r0 + 64/Q - 1
// of instance `inst` (those at R or above: none), LDS side by side from wl.
// A call of its own per order (once per unit, outside every loop): each order's
// code gets its own register allocation, as in the per-epoch kernels. Inlined
// into one function, the six orders' code spilled in the step loops of orders
// 8 to 32 (none there in the per-epoch kernels) and ran 25-40% slower; the
// call's price is the save and restore of the registers the convention keeps.
template <int NMAX, int Q>
__device__ __attribute__((noinline)) void run_unit(Inst* insts, int inst, int r0, int R,
                                                               SHARED unsigned char* wl,
                                                               const volatile int* stop) {
    const int lane = (int)(threadIdx.x & 63u);
    const int slot = lane / Q;
    const int r = r0 + slot < R ? r0 + slot : -1;
    lane_epoch<NMAX, Q, false>(insts, inst, r, wl + slot * LaneLds<NMAX>::kBytes, stop);
}

}  // namespace

// A queue slot: the epoch the unit runs (relative to the launch's first, at
// most kEpochNone - 2) above kUnitBits bits of unit id.
constexpr int kUnitBits = 24;
constexpr int kEpochNone = 127;

// `live` after a stop: below any count the epoch ends can reach
constexpr int32_t kPersistAbort = -(1 << 30);

// LDS of one wave: the largest of (64 / Q) replicas of any order the kernel runs
constexpr int kPersistWaveLds =
    (2 * LaneLds<48>::kBytes > 4 * LaneLds<32>::kBytes ? 2 * LaneLds<48>::kBytes
                                                       : 4 * LaneLds<32>::kBytes);
static_assert(kPersistWaveLds >= 256 * 4, "the epoch end's owner array fits");

__global__ void __launch_bounds__(256, 2)
    k_lane_persist(PersistCtl c, Inst* __restrict__ insts, const volatile int* __restrict__ stop) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = (int)(threadIdx.x & 63u);
    SHARED unsigned char* wl =
        (SHARED unsigned char*)smem + (size_t)(threadIdx.x >> 6) * kPersistWaveLds;
    const int x = c.xq[xcc_id() & (kMaxXcd - 1)];
    if (x < 0) return;
    const int home = c.home[x * kCuKeys + cu_pair_key()];
    for (;;) {
        // take the ready unit of the lowest epoch (the instance furthest
        // behind): from the home class while its lowest ready epoch is at
        // most one later than the lowest of all, else from the class and
        // epoch slot that hold the lowest (FIFO within a slot)
        int qsel = -1;
        uint32_t slot = 0;
        for (uint32_t spin = 0;; ++spin) {
            int emin = kEpochNone, ehome = kEpochNone, qmin = -1, qhome = -1;
            uint32_t hmin = 0, hhome = 0;
            for (int j = 0; j < kClasses; ++j)
                for (int b = 0; b < kEpochSlots; ++b) {
                    const int qi = (x * kClasses + j) * kEpochSlots + b;
                    const uint32_t h = (uint32_t)__builtin_amdgcn_readfirstlane((int)__hip_atomic_load(
                        c.head + qi * kLineWords, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
                    const uint32_t t = (uint32_t)__builtin_amdgcn_readfirstlane((int)__hip_atomic_load(
                        c.tail + qi * kLineWords, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
                    if (h >= t) continue;
                    // a slot below the tail not yet written counts as the
                    // newest epoch
                    const int w = __builtin_amdgcn_readfirstlane(__hip_atomic_load(
                        c.q + c.qbase[qi] + h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
                    const int e = w >= 0 ? (w >> kUnitBits) : kEpochNone - 1;
                    if (e < emin) {
                        emin = e;
                        qmin = qi;
                        hmin = h;
                    }
                    if (j == home && e < ehome) {
                        ehome = e;
                        qhome = qi;
                        hhome = h;
                    }
                }
            if (qmin >= 0) {
                if (qhome >= 0 && ehome <= emin + 1) {
                    qmin = qhome;
                    hmin = hhome;
                }
                uint32_t seen = hmin;
                bool won = false;
                if (lane == 0)
                    won = __hip_atomic_compare_exchange_strong(c.head + qmin * kLineWords, &seen,
                                                               hmin + 1u, __ATOMIC_RELAXED,
                                                               __ATOMIC_RELAXED,
                                                               __HIP_MEMORY_SCOPE_AGENT);
                if (__builtin_amdgcn_readfirstlane(won ? 1 : 0) != 0) {
                    qsel = qmin;
                    slot = hmin;
                    break;
                }
                continue;  // another wave took it: look again
            }
        const int live = __builtin_amdgcn_readfirstlane(
                __hip_atomic_load(c.live + x * kLineWords, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
            if (live <= 0) return;
            // the stop word is host memory: a waiting wave reads it every
            // 64th time round (about 50 us), as a working wave does every
            // 256 steps; a read per turn from a thousand waiting waves fills
            // the host link and stalls every wave's stop poll behind it
            if ((spin & 63u) == 63u && *stop) {
                // a stop ends the launch for every wave, also when the stop
                // word is cleared before they see it
                if (lane < c.nq)
                    __hip_atomic_store(c.live + lane * kLineWords, kPersistAbort, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                return;
            }
            __builtin_amdgcn_s_sleep(16);
        }
        // the slot lies below the tail: its writer reserved it and stores the
        // unit next
        const int32_t* qs = c.q + c.qbase[qsel] + slot;
        int v;
        for (;;) {
            v = __builtin_amdgcn_readfirstlane(
                __hip_atomic_load(qs, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
            if (v >= 0) break;
            __builtin_amdgcn_s_sleep(1);
        }
        v &= (1 << kUnitBits) - 1;
        xcd_acquire();
        const PUnit u = c.units[v];
        const int inst = __builtin_amdgcn_readfirstlane(u.inst);
        const int r0 = __builtin_amdgcn_readfirstlane((int)u.r0);
        const int nmax = __builtin_amdgcn_readfirstlane((int)u.nmax);
        GLOBAL Inst* I = (GLOBAL Inst*)insts + inst;
        const int R = I->R;
        switch (nmax) {
            case 8: run_unit<8, 16>(insts, inst, r0, R, wl, stop); break;
            case 16: run_unit<16, 16>(insts, inst, r0, R, wl, stop); break;
            case 24: run_unit<24, 16>(insts, inst, r0, R, wl, stop); break;
            case 32: run_unit<32, 16>(insts, inst, r0, R, wl, stop); break;
            case 40: run_unit<40, 32>(insts, inst, r0, R, wl, stop); break;
            default: run_unit<48, 32>(insts, inst, r0, R, wl, stop); break;
        }
        // the unit's writes, then its count
        xcd_release();
        int old = 0;
        if (lane == 0)
            old = __hip_atomic_fetch_add(c.pend + inst * kLineWords, -1, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
        old = __builtin_amdgcn_readfirstlane(old);
        if (old != 1) continue;
        // the instance's last unit of the epoch: its epoch end
        xcd_acquire();
        const uint32_t e0 = I->epoch;
        unsigned char* out = c.out ? c.out + c.out_off[inst] : nullptr;
        const bool more = epoch_end_wave(I, (SHARED int*)wl, out) && e0 + 1u < c.epoch_limit;
        if (more) {
            const int nu = c.unum[inst], ub = c.ubase[inst];
            // the unit id with the epoch it runs (relative to the launch's
            // first), by which waves take the lowest epoch first
            const int rel = (int)(e0 + 1u - c.epoch_base) < kEpochNone - 2
                                ? (int)(e0 + 1u - c.epoch_base) : kEpochNone - 2;
            const int qi = (x * kClasses + order_class(nmax)) * kEpochSlots + (rel & (kEpochSlots - 1));
            if (lane == 0)
                __hip_atomic_store(c.pend + inst * kLineWords, nu, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            xcd_release();
            uint32_t k = 0;
            if (lane == 0)
                k = __hip_atomic_fetch_add(c.tail + qi * kLineWords, (uint32_t)nu, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
            k = (uint32_t)__builtin_amdgcn_readfirstlane((int)k);
            int32_t* q = c.q + c.qbase[qi];
            for (int i = lane; i < nu; i += 64) {
                if (k + (uint32_t)i < c.qcap[qi])
                    __hip_atomic_store(q + k + i, rel << kUnitBits | (ub + i), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                else
                    __hip_atomic_store(c.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        } else if (lane == 0) {
            __hip_atomic_fetch_add(c.live + x * kLineWords, -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// The XCD and CU-pair key of each of `blocks` one-wave workgroups (XCD << 8 |
// key), for the host's placement.
__global__ void __launch_bounds__(64) k_xcc_probe(int32_t* __restrict__ out) {
    if (threadIdx.x == 0) out[blockIdx.x] = xcc_id() << 8 | cu_pair_key();
}

// ---------------------------------------------------------------------------
// host-side launch helpers
// ---------------------------------------------------------------------------

int lane_persist_wave_lds() { return kPersistWaveLds; }

hipError_t init_persist_kernel_attrs(int max_lds_per_block) {
    hipFuncAttributes fa;
    hipError_t e = hipFuncGetAttributes(&fa, (const void*)k_lane_persist);
    if (e != hipSuccess) return e;
    return hipFuncSetAttribute((const void*)k_lane_persist,
                               hipFuncAttributeMaxDynamicSharedMemorySize,
                               max_lds_per_block - (int)fa.sharedSizeBytes);
}

// Workgroups of four waves the device holds at once (per compute unit).
int lane_persist_blocks_per_cu() {
    int b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_lane_persist, 256,
                                                     (size_t)4 * kPersistWaveLds) != hipSuccess)
        return 0;
    return b;
}

hipError_t launch_xcc_probe(hipStream_t st, int32_t* out, int blocks) {
    hipLaunchKernelGGL(k_xcc_probe, dim3(blocks), dim3(64), 0, st, out);
    return hipGetLastError();
}

hipError_t launch_lane_persist(hipStream_t st, const PersistCtl& c, Inst* insts, int blocks,
                               const volatile int* stop) {
    if (blocks < 1 || !c.q || !c.units || !c.pend || !c.head || !c.tail || !c.live || !c.err ||
        c.nq < 1 || c.nq > kMaxXcd)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_lane_persist, dim3(blocks), dim3(256), (size_t)4 * kPersistWaveLds, st, c,
                       insts, stop);
    return hipGetLastError();
}

}  // namespace xtx
