// baseline.hip -- the plain per-record launch loop the batched form is compared
// against (bench-support library libcusbf_synth.so; not part of the filter's
// boundary).
//
// One insert_sequence / contains_sequence call per record (reference:
// include/cusbf/filter.cuh:179-190, 464-474), issued from native code so the
// loop itself adds no interpreter overhead. The entry points are passed in as
// function pointers (the C ABI of libcusbf_hip.so), so this library does not
// link against the filter library. Query answers land where the batched form
// puts them: record r's windows at d_out[offsets[r] ..].
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

using InsertFn = int (*)(void*, const char*, uint64_t, void*, uint64_t*);
using ContainsFn = int (*)(const void*, const char*, uint64_t, uint8_t*, uint64_t, void*);

}  // namespace

extern "C" {

// op 0: insert each record; op 1: query each record into d_out. Returns the
// first non-zero status of the called entry point (0 when every call succeeded).
int cusbf_bench_each_record(void* fn, int op, void* filter, const char* d_symbols, uint32_t width,
                            const uint64_t* h_offsets, uint64_t num_records, uint32_t k,
                            uint8_t* d_out, void* stream) {
    for (uint64_t r = 0; r < num_records; ++r) {
        const uint64_t a = h_offsets[r];
        const uint64_t b = h_offsets[r + 1];
        const uint64_t len = b - a;
        if (len < k) continue;
        int rc;
        if (op == 0) {
            rc = reinterpret_cast<InsertFn>(fn)(filter, d_symbols + a * width, len * width, stream,
                                                nullptr);
        } else {
            rc = reinterpret_cast<ContainsFn>(fn)(filter, d_symbols + a * width, len * width,
                                                  d_out + a, len - k + 1, stream);
        }
        if (rc != 0) return rc;
    }
    return 0;
}

}  // extern "C"
