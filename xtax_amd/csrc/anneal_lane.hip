// anneal_lane.hip -- the lane kernel's per-epoch launches (one proposal per
// lane; lane_epoch.h) and their host-side launch helpers.
#include <hip/hip_runtime.h>

#include "lane_epoch.h"

namespace xtx {

// One epoch of the replicas of two orders in one launch: table entries below
// `split` (a whole number of blocks) are of order NHI, the rest of order NLO.
// Every order of an epoch then runs concurrently on the device's few hardware
// queues, the short lower order filling the compute units the higher one
// leaves idle. `nmax`: the launch's largest order (NHI, or NLO when no entry
// is of order NHI).
//
// WIDE (NLO, NHI = 64, 128 or 192, 256: column bounds of 2, 4, 6 and 8 columns
// per lane; orders above 128, and every order when a call's tournaments take
// more than 16 samples): A lives in HBM only and every proposal is evaluated
// by the replica's lane group, one after another; the tournaments take any
// number of samples.
template <int NLO, int NHI, int Q, bool WIDE = false>
__global__ void __launch_bounds__(256, lane_waves_per_simd(NHI, Q))
    k_anneal_lane(const RepRef* __restrict__ table, int count, int split,
                  Inst* __restrict__ insts, int nmax, int rep_bytes,
                  const volatile int* __restrict__ stop) {
    if (nmax != NHI && nmax != NLO) return;
    if ((int)blockIdx.x * ((int)blockDim.x / Q) < split)
        lane_epoch_table<NHI, Q, WIDE>(table, count, insts, rep_bytes, stop);
    else
        lane_epoch_table<NLO, Q, WIDE>(table, count, insts, rep_bytes, stop);
}

// ---------------------------------------------------------------------------
// host-side launch helpers
// ---------------------------------------------------------------------------

typedef void (*LaneKernel)(const RepRef*, int, int, Inst*, int, int, const volatile int*);

struct LaneVariant {
    int lo, hi, Q;
    LaneKernel fn;
};
static const LaneVariant kLaneVariants[] = {
    {8, 16, 16, k_anneal_lane<8, 16, 16>},  {24, 32, 16, k_anneal_lane<24, 32, 16>},
    {40, 48, 16, k_anneal_lane<40, 48, 16>}, {8, 16, 32, k_anneal_lane<8, 16, 32>},
    {24, 32, 32, k_anneal_lane<24, 32, 32>}, {40, 48, 32, k_anneal_lane<40, 48, 32>},
    {64, 128, 32, k_anneal_lane<64, 128, 32, true>}, {192, 256, 32, k_anneal_lane<192, 256, 32, true>},
};

static const LaneVariant* find_lane_variant(int nmax, int Q) {
    for (const auto& v : kLaneVariants)
        if ((v.lo == nmax || v.hi == nmax) && v.Q == Q) return &v;
    return nullptr;
}

bool lane_variant_exists(int nmax, int Q) { return find_lane_variant(nmax, Q) != nullptr; }

// The higher order of the launch pair order nmax belongs to (0: none)
int lane_pair_hi(int nmax) {
    for (const auto& v : kLaneVariants)
        if (v.lo == nmax || v.hi == nmax) return v.hi;
    return 0;
}

// LDS of one replica of the lane kernel at order nmax (LaneLds)
int lane_rep_bytes(int nmax) {
    // the wide mode (every order above 48 this kernel runs) has no square
    const int b = (nmax > 48 ? 0 : 2 * nmax * nmax) + 8 * nmax + 4 * nmax + nmax;
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

// units_per_block replicas (table entries, padding included) of Q lanes each;
// nmax: the largest order in the table; the first `split` entries (a multiple
// of units_per_block) are of the pair's higher order, the rest of its lower
// one.
hipError_t launch_anneal_lane(hipStream_t st, int nmax, int Q, int units_per_block,
                              const RepRef* table, int count, int split, Inst* insts,
                              int rep_bytes, const volatile int* stop) {
    if (count <= 0) return hipSuccess;
    const LaneVariant* v = find_lane_variant(nmax, Q);
    if (!v || units_per_block < 1 || units_per_block * Q > 256 || (units_per_block * Q) % 64 != 0 ||
        rep_bytes < lane_rep_bytes(nmax))
        return hipErrorInvalidValue;
    const int threads = units_per_block * Q;
    const int blocks = (count + units_per_block - 1) / units_per_block;
    if (split < 0 || split > blocks * units_per_block || split % units_per_block != 0 ||
        (split > 0 && nmax != v->hi))
        return hipErrorInvalidValue;
    const size_t lds = (size_t)units_per_block * rep_bytes;
    hipLaunchKernelGGL(v->fn, dim3(blocks), dim3(threads), lds, st, table, count, split, insts,
                       nmax, rep_bytes, stop);
    return hipGetLastError();
}

}  // namespace xtx
