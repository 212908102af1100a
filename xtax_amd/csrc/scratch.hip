// scratch.hip -- the device scratch of the no-alloc plain solve
// (xtx_solve_plain_noalloc).
//
// Callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc) get a plain solve whose
// device state -- the instance record, the replica table and every replica's
// A, X^T, best snapshots, scalars, ladder, active list and move log -- is
// carved out of this one fixed array. Each GPU the library runs on holds its
// own copy, at its own address, until the process ends.
//
// In host code `scratch` names a host-side stand-in, not device memory: its
// device address is resolved once per device with hipGetSymbolAddress (from
// xtx_open, with that device current) and only that address is used.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <mutex>

namespace xtx {

// 32 MiB: e.g. order 24 with 256 replicas needs 4.8 MB, order 48 with 256
// replicas 19 MB (four R n^2 int64 arrays plus the small per-replica records).
constexpr size_t kScratchBytes = (size_t)32 << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
std::mutex g_scratch_mtx;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// Resolves this device's address of `scratch` (the caller has made `device`
// current). Idempotent per device.
hipError_t scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mtx);
    if (g_scratch_ptr[device]) return hipSuccess;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!ptr) return hipErrorInvalidSymbol;
    g_scratch_ptr[device] = ptr;
    return hipSuccess;
}

// The device address of `scratch` on `device`, or null before scratch_init.
void* scratch_address(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mtx);
    return g_scratch_ptr[device];
}

size_t scratch_bytes() { return kScratchBytes; }

}  // namespace xtx
