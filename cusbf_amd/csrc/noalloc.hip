// noalloc.hip -- the batched query over a fixed, library-owned scratch array, for
// callers that cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc). No reference counterpart.
//
// cusbf_contains_batched_async keeps its per-call scratch -- the work-queue and
// status words and the record-start bitmap (one bit per symbol) -- in buffers the
// filter allocates and grows. This entry point uses `scratch`, a file-scope
// device array of kScratchBytes instead, and allocates nothing. Each GPU holds
// its own copy of the array (until the process ends) at its own address, which
// hipGetSymbolAddress returns with that device current; host code only ever uses
// that address, never the array's name (a host-side stand-in, not a device
// address). Every call on one GPU shares that GPU's array: calls that overlap
// on one GPU race. The prepass clears the status words and the bitmap on the
// stream before the kernels use them.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <mutex>

#include "../../include/cusbf_hip.h"
#include "filter_internal.hpp"
#include "launch.hpp"

// 32 MiB: 256 bytes of status words, then the bitmap of up to
// CUSBF_NOALLOC_MAX_SYMBOLS symbols.
constexpr uint64_t kScratchBytes = 32ull << 20;
static __device__ __align__(256) uint64_t scratch[kScratchBytes / 8];

namespace {

namespace ci = cusbf_internal;

constexpr uint64_t kStateBytes = 256;
static_assert(sbf::start_words(CUSBF_NOALLOC_MAX_SYMBOLS) * 8 + kStateBytes <= kScratchBytes,
              "the bitmap of the largest batch must fit the scratch");

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
uint64_t* g_scratch[kMaxDevices] = {};

// The scratch address on `device`, which must be current. The first call on a
// device looks it up; later calls reuse it.
int scratch_on(int device, uint64_t** out) {
    if (device < 0 || device >= kMaxDevices) {
        return ci::set_error(CUSBF_ERR_INVALID_ARGUMENT, "device index out of range for the no-allocation scratch");
    }
    std::lock_guard<std::mutex> lock(g_mutex);
    if (g_scratch[device] == nullptr) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return ci::set_hip_error(e, __FILE__, __LINE__);
        g_scratch[device] = static_cast<uint64_t*>(p);
    }
    *out = g_scratch[device];
    return CUSBF_OK;
}

}  // namespace

namespace cusbf_internal {

// Called by cusbf_create with the filter's device current: the lookup happens
// at set-up, once per device.
int noalloc_setup(int device) {
    uint64_t* p = nullptr;
    return scratch_on(device, &p);
}

}  // namespace cusbf_internal

extern "C" int cusbf_contains_batched_noalloc_async(const cusbf_filter* f, const char* d_symbols,
                                                    uint64_t num_symbols, const uint64_t* d_offsets,
                                                    uint64_t num_records, uint8_t* d_output,
                                                    uint64_t d_output_len, void* stream) {
    if (!f) return ci::set_error(CUSBF_ERR_INVALID_ARGUMENT, "null filter");
    if (num_symbols > CUSBF_NOALLOC_MAX_SYMBOLS) {
        char buf[160];
        std::snprintf(buf, sizeof(buf),
                      "batched sequence exceeds the no-allocation scratch (at most %llu symbols)",
                      static_cast<unsigned long long>(CUSBF_NOALLOC_MAX_SYMBOLS));
        return ci::set_error(CUSBF_ERR_INVALID_ARGUMENT, buf);
    }
    int current = 0;
    const hipError_t e = hipGetDevice(&current);
    if (e != hipSuccess) return ci::set_hip_error(e, __FILE__, __LINE__);
    if (current != ci::device(f)) {
        return ci::set_error(CUSBF_ERR_INVALID_ARGUMENT, "the filter's device is not the current device");
    }
    uint64_t* base = nullptr;
    const int rc = scratch_on(current, &base);
    if (rc != CUSBF_OK) return rc;
    uint32_t* state = reinterpret_cast<uint32_t*>(base);
    uint64_t* starts = base + kStateBytes / 8;
    return ci::contains_batched(f, d_symbols, num_symbols, d_offsets, num_records, d_output,
                                d_output_len, state, starts, static_cast<hipStream_t>(stream));
}
