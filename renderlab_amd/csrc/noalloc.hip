// rq_trace without allocation, for callers that cannot call hipMalloc (code
// under HIP graph capture, tight serving loops, harnesses that forbid it).
//
// rq_trace keeps two pieces of device memory per HIP stream besides its inputs
// and outputs, allocated on first use: the persistent kernel's ray cursor and
// the per-lane stack spill area of its grid. This entry point takes both from
// one file-scope device array instead:
//
//   scratch = CURSOR_BYTES (the ray cursor: 32 partition counters 256 bytes
//             apart, cleared by every launch with a fill kernel)
//             + NOALLOC_LANES lanes x TRACE_SPILL_WORDS words,
//             = 8 KiB + 131,072 x 896 B = 112 MiB + 8 KiB,
//
// which every GPU the library runs this entry point on holds until the process
// ends. The grid is capped to NOALLOC_LANES lanes (512 workgroups of 256, two
// per CU on a 256-CU part); the ray count is not limited, since the persistent
// lanes loop until the rays run out. Every call on one GPU shares that GPU's
// array: calls that overlap on one GPU (two streams, two host threads) race on
// the cursor and the spill area and must be ordered by the caller.
#include <hip/hip_runtime.h>

#include <mutex>

#include "internal.h"

namespace {

constexpr size_t NOALLOC_LANES = 131072;
constexpr size_t SCRATCH_BYTES = rq::CURSOR_BYTES + NOALLOC_LANES * rq::TRACE_SPILL_WORDS * sizeof(uint32_t);

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

// `scratch` named in host code is a host-side stand-in, not a device address:
// each device's copy is looked up once with that device current and kept here.
constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
void* g_ptr[MAX_DEVICES] = {};

rq_status device_scratch(void** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= MAX_DEVICES) {
        rq::set_error(std::string("rq_trace_noalloc: hipGetDevice: ") + hipGetErrorString(e));
        return RQ_ERR_DEVICE;
    }
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[dev]) {
        void* ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            rq::set_error(std::string("rq_trace_noalloc: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return RQ_ERR_DEVICE;
        }
        g_ptr[dev] = ptr;
    }
    *out = g_ptr[dev];
    return RQ_OK;
}

}  // namespace

extern "C" size_t rq_trace_noalloc_scratch_bytes(void) { return SCRATCH_BYTES; }

extern "C" rq_status rq_trace_noalloc(rq_scene* scene, rq_layout layout, const rq_ray* rays, uint32_t count,
                                      uint32_t mode, rq_hit* hits, uint32_t* vis_bits, void* stream) {
    void* base = nullptr;
    const rq_status st = device_scratch(&base);
    if (st != RQ_OK) return st;
    rq::FixedScratch f;
    f.counter = reinterpret_cast<uint32_t*>(base);
    f.spill = reinterpret_cast<uint32_t*>(reinterpret_cast<unsigned char*>(base) + rq::CURSOR_BYTES);
    f.lanes = NOALLOC_LANES;
    return rq::launch_trace(scene, layout, rays, count, mode, hits, vis_bits, stream, nullptr, nullptr, &f);
}
