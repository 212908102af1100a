// device_common.h -- gfx950 building blocks shared by the annealer kernels:
// Philox4x32-10, the deterministic exp, scalar-register helpers, integer
// coins, and reductions/broadcasts inside a lane group (DPP within 16-lane
// rows, v_readlane across rows).
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "engine_types.h"

namespace xtx {

// HBM pointers with their address space spelled out: loads and stores through
// them become global_* instructions (vector-memory counter only) instead of
// flat_* ones, which would also hold up every LDS wait.
#define GLOBAL __attribute__((address_space(1)))
// LDS pointers the same way: ds_* instructions even where the pointer passes
// through a struct or a call (a generic pointer there becomes flat_*, which
// waits on both counters and takes the vector-memory path)
#define SHARED __attribute__((address_space(3)))

// ---------------------------------------------------------------------------
// building blocks
// ---------------------------------------------------------------------------

__device__ __forceinline__ int64_t iabs64(int64_t v) { return v < 0 ? -v : v; }
__device__ __forceinline__ bool over(int64_t v) { return v > kLimit || v < -kLimit; }

// Philox4x32-10 (Salmon et al., SC'11).
__device__ __forceinline__ uint4 philox10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                          uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        if (round) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        // each 64-bit product is one v_mad_u64_u32 (a lo/hi multiply pair
        // would be two quarter-rate instructions)
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c0 = n0;
        c1 = (uint32_t)p1;
        c2 = n2;
        c3 = (uint32_t)p0;
    }
    return make_uint4(c0, c1, c2, c3);
}

__device__ __forceinline__ uint32_t mulhi32(uint32_t w, uint32_t n) { return __umulhi(w, n); }
// Last counter word of draw slot c of a move whose word for slot 0 is c3 =
// 16 ((m - 1) mod 4): 16 slots per move in the words below 64, and slots past
// 15 (tournaments of more than 16 samples) at 64 (c / 16) + c3 + c mod 16.
__device__ __forceinline__ uint32_t slot_word(uint32_t c3, int c) {
    return 64u * ((uint32_t)c >> 4) + c3 + ((uint32_t)c & 15u);
}
__device__ __forceinline__ double u01(uint32_t w) { return (double)w * 0x1p-32; }
__device__ __forceinline__ double u53(uint32_t hi, uint32_t lo) {
    return (double)(((uint64_t)(hi >> 5) << 26) | (lo >> 6)) * 0x1p-53;
}

// exp(x), x <= 0, with the same IEEE operation sequence on every device:
// 0 below -40; k = rint(x/ln2); r = x - k ln2 (two fma); degree-13 Taylor in
// Horner form (fma); times 2^k.
// A double constant materialized where it is used (two scalar moves) rather
// than hoisted into registers held across the caller's loop.
__device__ __forceinline__ double kconst(double c) {
    asm volatile("" : "+s"(c));
    return c;
}

__device__ __forceinline__ double xexp(double x) {
    if (!(x > -40.0)) return 0.0;
    if (x > 0.0) x = 0.0;
    const double kd = __builtin_rint(x * kconst(1.4426950408889634));
    double r = __builtin_fma(-kd, kconst(0x1.62e42fefa39efp-1), x);
    r = __builtin_fma(-kd, kconst(0x1.abc9e3b39803fp-56), r);
    double p = kconst(1.0 / 6227020800.0);
    p = __builtin_fma(p, r, kconst(1.0 / 479001600.0));
    p = __builtin_fma(p, r, kconst(1.0 / 39916800.0));
    p = __builtin_fma(p, r, kconst(1.0 / 3628800.0));
    p = __builtin_fma(p, r, kconst(1.0 / 362880.0));
    p = __builtin_fma(p, r, kconst(1.0 / 40320.0));
    p = __builtin_fma(p, r, kconst(1.0 / 5040.0));
    p = __builtin_fma(p, r, kconst(1.0 / 720.0));
    p = __builtin_fma(p, r, kconst(1.0 / 120.0));
    p = __builtin_fma(p, r, kconst(1.0 / 24.0));
    p = __builtin_fma(p, r, kconst(1.0 / 6.0));
    p = __builtin_fma(p, r, 1.0 / 2.0);
    p = __builtin_fma(p, r, 1.0);
    p = __builtin_fma(p, r, 1.0);
    return __builtin_ldexp(p, (int)kd);
}

// Nearest integer a/b, ties away from zero (mat_io.hpp:41-48).
__device__ __forceinline__ int64_t rounded_div(int64_t a, int64_t b) {
    int64_t q = a / b;
    const int64_t r = a - q * b;
    if (2 * iabs64(r) >= iabs64(b)) q += ((a < 0) == (b < 0)) ? 1 : -1;
    return q;
}

__device__ __forceinline__ int64_t clamp_cap(int64_t v) {
    return v > kShearCap ? kShearCap : (v < -kShearCap ? -kShearCap : v);
}

// The same pointer, opaque to the optimizer: loads through it stay in the
// branch that makes them instead of being hoisted out of the move loop, so
// rarely used values do not hold registers across it.
template <class T>
__device__ __forceinline__ T* opaque(T* p) {
    asm volatile("" : "+s"(p));
    return p;
}
__device__ __forceinline__ void opaque_u32(uint32_t& v) { asm volatile("" : "+s"(v)); }
// the same for a pointer that may differ between the groups of a wave
template <class T>
__device__ __forceinline__ T* opaque_v(T* p) {
    asm volatile("" : "+v"(p));
    return p;
}

// A wave-uniform value in scalar registers (groups of a whole wave or more).
template <int G>
__device__ __forceinline__ int32_t uni(int32_t v) {
    if constexpr (G >= 64) return __builtin_amdgcn_readfirstlane(v);
    return v;
}
template <int G>
__device__ __forceinline__ int64_t uni(int64_t v) {
    if constexpr (G >= 64) {
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(uint64_t)v);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)v >> 32));
        return (int64_t)(((uint64_t)hi << 32) | lo);
    }
    return v;
}
// a double every lane of the wave holds alike, in scalar registers
__device__ __forceinline__ double uni_f64(double v) {
    return __builtin_bit_cast(double, uni<64>(__builtin_bit_cast(int64_t, v)));
}

// Coin u01(w) < f as an integer test: w 2^-32 and f 2^32 are exact doubles,
// so for integer w, u01(w) < f <=> w < ceil(f 2^32); `all` covers f >= 1.
struct Coin {
    uint32_t thr;
    bool all;
    __device__ __forceinline__ explicit Coin(double f) {
        const double x = __builtin_ceil(f * 4294967296.0);
        all = !(x < 4294967296.0);
        thr = x > 0.0 && !all ? (uint32_t)x : 0u;
    }
    __device__ __forceinline__ bool operator()(uint32_t w) const { return all || w < thr; }
};

// ---- cross-lane building blocks (DPP within 16-lane rows, v_readlane across
// rows; ds_bpermute only where each lane reads a different, runtime lane) ----
constexpr int kDppXor1 = 0xB1;        // quad_perm [1,0,3,2]
constexpr int kDppXor2 = 0x4E;        // quad_perm [2,3,0,1]
constexpr int kDppHalfMirror = 0x141; // row_half_mirror
constexpr int kDppMirror = 0x140;     // row_mirror
constexpr int kDppShr1 = 0x111;       // row_shr:1 .. row_shr:8
constexpr int kDppShr2 = 0x112;
constexpr int kDppShr4 = 0x114;
constexpr int kDppShr8 = 0x118;
constexpr int kDppBcast15 = 0x142;    // row_bcast:15
constexpr int kDppBcast31 = 0x143;    // row_bcast:31

template <int CTRL>
__device__ __forceinline__ uint32_t dpp32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, false);
}
// shifted-in lanes (no source within the row) read 0
template <int CTRL>
__device__ __forceinline__ uint32_t dpp32z(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, true);
}
__device__ __forceinline__ uint32_t umax(uint32_t a, uint32_t b) { return a > b ? a : b; }
template <int CTRL>
__device__ __forceinline__ int64_t dpp64(int64_t v) {
    const uint32_t lo = dpp32<CTRL>((uint32_t)(uint64_t)v);
    const uint32_t hi = dpp32<CTRL>((uint32_t)((uint64_t)v >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ uint32_t readlane32(uint32_t v, int lane) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, lane);
}
__device__ __forceinline__ int64_t readlane64(int64_t v, int lane) {
    const uint32_t lo = readlane32((uint32_t)(uint64_t)v, lane);
    const uint32_t hi = readlane32((uint32_t)((uint64_t)v >> 32), lane);
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

// int64 value of lane `src` (runtime, per lane) of the wave, for every lane
__device__ __forceinline__ int64_t bperm64(int64_t v, int src) {
    const int lo = __builtin_amdgcn_ds_bpermute(src << 2, (int)(uint32_t)(uint64_t)v);
    const int hi = __builtin_amdgcn_ds_bpermute(src << 2, (int)(uint32_t)((uint64_t)v >> 32));
    return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}
__device__ __forceinline__ double readlane_f64(double v, int lane) {
    return __builtin_bit_cast(double, readlane64(__builtin_bit_cast(int64_t, v), lane));
}

struct OpAdd {
    template <class T>
    __device__ __forceinline__ T operator()(T a, T b) const { return a + b; }
};
struct OpMax {
    template <class T>
    __device__ __forceinline__ T operator()(T a, T b) const { return a > b ? a : b; }
};

// Reductions and broadcasts inside a lane group of G lanes (G = 16, 32, 64;
// a group is 1, 2 or 4 whole 16-lane rows of the wave).
template <int G>
struct Grp {
    static __device__ __forceinline__ int lane() { return (int)(threadIdx.x & 63u); }

    template <class Op>
    static __device__ __forceinline__ int64_t all64(int64_t v, Op op) {
        v = op(v, dpp64<kDppXor1>(v));
        v = op(v, dpp64<kDppXor2>(v));
        v = op(v, dpp64<kDppHalfMirror>(v));
        v = op(v, dpp64<kDppMirror>(v));
        if constexpr (G == 16) {
            return v;
        } else if constexpr (G == 32) {
            // the group's two rows trade places: even row's value op odd row's
            const uint32_t lo = (uint32_t)(uint64_t)v, hi = (uint32_t)((uint64_t)v >> 32);
            const auto pl = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
            const auto ph = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
            const int64_t ev = (int64_t)(((uint64_t)ph[0] << 32) | pl[0]);
            const int64_t od = (int64_t)(((uint64_t)ph[1] << 32) | pl[1]);
            return op(ev, od);
        } else {
            const int64_t r0 = readlane64(v, 0), r1 = readlane64(v, 16);
            const int64_t r2 = readlane64(v, 32), r3 = readlane64(v, 48);
            return op(op(r0, r1), op(r2, r3));
        }
    }
    template <class Op>
    static __device__ __forceinline__ int32_t all32(int32_t v, Op op) {
        v = op(v, (int32_t)dpp32<kDppXor1>((uint32_t)v));
        v = op(v, (int32_t)dpp32<kDppXor2>((uint32_t)v));
        v = op(v, (int32_t)dpp32<kDppHalfMirror>((uint32_t)v));
        v = op(v, (int32_t)dpp32<kDppMirror>((uint32_t)v));
        if constexpr (G == 16) {
            return v;
        } else if constexpr (G == 32) {
            const auto pr = __builtin_amdgcn_permlane16_swap((uint32_t)v, (uint32_t)v, false, false);
            return op((int32_t)pr[0], (int32_t)pr[1]);
        } else {
            const int32_t r0 = (int32_t)readlane32((uint32_t)v, 0);
            const int32_t r1 = (int32_t)readlane32((uint32_t)v, 16);
            const int32_t r2 = (int32_t)readlane32((uint32_t)v, 32);
            const int32_t r3 = (int32_t)readlane32((uint32_t)v, 48);
            return op(op(r0, r1), op(r2, r3));
        }
    }
    // Sum of a double over the group as the binary tree of adjacent lanes
    // (lanes 2l and 2l+1, then adjacent pairs of those, ...): the butterfly
    // steps pair exactly those partial sums, and IEEE addition commutes, so
    // every lane of the group ends with the same bits.
    static __device__ __forceinline__ double fsum(double v) {
        auto dppf = [](double x, auto ctrl) {
            return __builtin_bit_cast(double, dpp64<decltype(ctrl)::value>(
                                                  __builtin_bit_cast(int64_t, x)));
        };
        v = v + dppf(v, std::integral_constant<int, kDppXor1>());
        v = v + dppf(v, std::integral_constant<int, kDppXor2>());
        v = v + dppf(v, std::integral_constant<int, kDppHalfMirror>());
        v = v + dppf(v, std::integral_constant<int, kDppMirror>());
        if constexpr (G == 16) {
            return v;
        } else if constexpr (G == 32) {
            // even row + odd row of the group, in every lane
            const uint64_t b = __builtin_bit_cast(uint64_t, v);
            const uint32_t lo = (uint32_t)b, hi = (uint32_t)(b >> 32);
            const auto pl = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
            const auto ph = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
            const double ev = __builtin_bit_cast(double, ((uint64_t)ph[0] << 32) | pl[0]);
            const double od = __builtin_bit_cast(double, ((uint64_t)ph[1] << 32) | pl[1]);
            return ev + od;
        } else {
            const int64_t b = __builtin_bit_cast(int64_t, v);
            const double r0 = __builtin_bit_cast(double, readlane64(b, 0));
            const double r1 = __builtin_bit_cast(double, readlane64(b, 16));
            const double r2 = __builtin_bit_cast(double, readlane64(b, 32));
            const double r3 = __builtin_bit_cast(double, readlane64(b, 48));
            return (r0 + r1) + (r2 + r3);
        }
    }
    static __device__ __forceinline__ int64_t sum(int64_t v) { return all64(v, OpAdd()); }
    static __device__ __forceinline__ int32_t sum32(int32_t v) { return all32(v, OpAdd()); }
    static __device__ __forceinline__ int64_t max(int64_t v) { return all64(v, OpMax()); }
    static __device__ __forceinline__ bool any(bool p) {
        const unsigned long long b = __ballot(p);
        if constexpr (G == 64) {
            return b != 0ull;
        } else {
            const int base = lane() & ~(G - 1);
            return ((b >> base) & ((1ull << G) - 1ull)) != 0ull;
        }
    }
    // value of lane c (compile-time) of the caller's group
    template <int CC>
    static __device__ __forceinline__ uint32_t bcast_c(uint32_t v) {
        if constexpr (G == 64) {
            return readlane32(v, CC);
        } else if constexpr (G == 32) {
            const uint32_t a = readlane32(v, CC), b = readlane32(v, 32 + CC);
            return (lane() & 32) ? b : a;
        } else {
            const uint32_t a = readlane32(v, CC), b = readlane32(v, 16 + CC);
            const uint32_t c = readlane32(v, 32 + CC), d = readlane32(v, 48 + CC);
            const int row = lane() >> 4;
            return row == 0 ? a : (row == 1 ? b : (row == 2 ? c : d));
        }
    }
    // value of lane src (runtime, uniform in the group) of the caller's group
    template <class T>
    static __device__ __forceinline__ T bcast(T v, int src) {
        if constexpr (G == 64 && sizeof(T) == 4)
            return (T)readlane32((uint32_t)v, src);
        else
            return __shfl(v, src, G);
    }
    // value of a lane that differs per lane (always a permute)
    template <class T>
    static __device__ __forceinline__ T gather(T v, int src) {
        return __shfl(v, src, G);
    }

    // unsigned maximum over each 16-lane row, in every lane of the row
    static __device__ __forceinline__ uint32_t rowmax32(uint32_t v) {
        v = umax(v, dpp32<kDppXor1>(v));
        v = umax(v, dpp32<kDppXor2>(v));
        v = umax(v, dpp32<kDppHalfMirror>(v));
        v = umax(v, dpp32<kDppMirror>(v));
        return v;
    }
    // a row-wide value of the group's row `sub` (group lanes 16 sub .. 16 sub
    // + 15) in every lane of the group; for G == 64 `sub` is wave-uniform
    static __device__ __forceinline__ uint32_t from_row(uint32_t v, int sub) {
        if constexpr (G == 16) {
            return v;
        } else if constexpr (G == 64) {
            return readlane32(v, 16 * sub);
        } else {
            // odd rows of the first operand trade places with even rows of
            // the second: [0] holds the pair's even row, [1] its odd row
            const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
            return sub ? r[1] : r[0];
        }
    }
    // unsigned 64-bit maximum over the group's row `sub`, in every lane of
    // the group (high words first, then low words among the lanes that hold
    // the maximal high word)
    static __device__ __forceinline__ uint64_t slot_max(uint64_t key, int sub) {
        const uint32_t hi = (uint32_t)(key >> 32), lo = (uint32_t)key;
        const uint32_t mh = rowmax32(hi);
        const uint32_t ml = rowmax32(hi == mh ? lo : 0u);
        return ((uint64_t)from_row(mh, sub) << 32) | from_row(ml, sub);
    }
    // inclusive prefix sum over the group's lanes in lane order
    static __device__ __forceinline__ uint32_t scan32(uint32_t v) {
        v += dpp32z<kDppShr1>(v);
        v += dpp32z<kDppShr2>(v);
        v += dpp32z<kDppShr4>(v);
        v += dpp32z<kDppShr8>(v);
        if constexpr (G >= 32)  // last lane of rows 0 and 2 into rows 1 and 3
            v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kDppBcast15, 0xA, 0xF, false);
        if constexpr (G == 64)  // last lane of row 1 into rows 2 and 3
            v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kDppBcast31, 0xC, 0xF, false);
        return v;
    }
};

}  // namespace xtx
