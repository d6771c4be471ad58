// launch.hpp -- host-side launchers of the Super Bloom kernels (internal to libcusbf_hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "sbf_common.hpp"

namespace sbf {

// Input of one launch: either a byte sequence (`seq`, symbol_width bytes per
// symbol) or a dense packed sequence (`words`, LSB-first, 64/symbol_bits
// symbols per word).
//
// Batched form: `offsets` (num_records + 1 non-decreasing symbol offsets,
// offsets[0] = 0, offsets[num_records] = num_symbols) cut the byte sequence
// into records concatenated without separators; a window that crosses a
// record start is invalid. The prepass turns the offsets into `starts`, a
// bitmap with bit i set when symbol i starts a record, and checks the offsets
// into `state`.
struct SequenceInput {
    const char* seq = nullptr;
    const uint64_t* words = nullptr;
    uint64_t num_words = 0;
    uint64_t num_symbols = 0;
    bool dense = false;
    const uint64_t* offsets = nullptr;
    uint64_t num_records = 0;
    const uint64_t* starts = nullptr;  // record-start bitmap, start_words(num_symbols) words
    uint32_t* state = nullptr;         // [0] work-queue counter, [1] malformed-offsets flag
};

struct FilterView {
    uint64_t* words = nullptr;  // num_shards * 4
    uint64_t num_shards = 0;
};

// Launch shape of the persistent batched kernels.
struct BatchLaunch {
    uint32_t blocks_per_cu = 4;
    uint32_t compute_units = 256;
    uint32_t claim = 1;  // tiles per work-queue claim; 0 = waves stride over the tiles
};

using InsertFn = hipError_t (*)(const SequenceInput&, FilterView, hipStream_t);
using ContainsFn = hipError_t (*)(const SequenceInput&, FilterView, uint8_t*, hipStream_t);
using BatchedInsertFn = hipError_t (*)(const SequenceInput&, FilterView, const BatchLaunch&,
                                       hipStream_t);
using BatchedContainsFn = hipError_t (*)(const SequenceInput&, FilterView, const BatchLaunch&,
                                         uint8_t*, hipStream_t);

// Lane-strip kernels for one compile-time (k, s, m, h, alphabet).
struct StripKernelSet {
    uint32_t k, s, m, h, alphabet;
    bool dense;             // plain kernels also read dense packed input
    uint32_t wave_windows;  // windows per wave tile
    InsertFn insert;
    ContainsFn contains;
    BatchedInsertFn batched_insert[2];      // [0]: built for 3 waves per SIMD, [1]: for 4 (or null)
    BatchedContainsFn batched_contains[2];
};

// Returns the compiled lane-strip kernel set or nullptr.
const StripKernelSet* find_strip_kernels(uint32_t k, uint32_t s, uint32_t m, uint32_t h,
                                         uint32_t alphabet);

// Generic kernels: any alphabet / configuration, one window per thread; in the
// batched form they test the record-start bitmap (or, without it, search the
// offsets) for each window.
hipError_t generic_insert(const KernelConfig& c, const SequenceInput& in, FilterView f,
                          hipStream_t stream);
hipError_t generic_contains(const KernelConfig& c, const SequenceInput& in, FilterView f,
                            uint8_t* out, hipStream_t stream);
// d_counts[0] += valid windows, d_counts[1] += shard runs (caller zeroes).
hipError_t generic_count(const KernelConfig& c, const SequenceInput& in, uint64_t num_shards,
                         uint64_t* d_counts, hipStream_t stream);
hipError_t pack_dense(const KernelConfig& c, const char* seq, uint64_t num_symbols,
                      uint64_t* words, hipStream_t stream);

// Words of the record-start bitmap for num_symbols symbols (one spare word, so
// a 16-bit lane chunk just past the end still reads inside the bitmap).
__host__ __device__ inline uint64_t start_words(uint64_t num_symbols) { return num_symbols / 64 + 2; }

// Batched form set-up, in stream order before the batched kernels (fill
// kernels only, so the sequence can be captured in a graph): zero the state
// and the bitmap, then set one bit per record start and check the offsets.
hipError_t batch_prepass(const SequenceInput& in, uint64_t* starts, uint32_t* state,
                         hipStream_t stream);

// Server report counts: request r owns windows [req[2r], req[2r] + req[2r+1])
// of a batch (sorted, disjoint); counts[2r] += valid windows, counts[2r+1] +=
// hits equal to 1 (caller zeroes counts).
hipError_t request_counts(const KernelConfig& c, const char* seq, uint64_t num_symbols,
                          const uint8_t* hits, const uint64_t* req, uint64_t n_req,
                          uint64_t* counts, hipStream_t stream);

// Reductions.
hipError_t popcount_words(const uint64_t* words, uint64_t n, uint64_t* d_out, hipStream_t stream);
hipError_t count_ones(const uint8_t* hits, uint64_t n, uint64_t* d_out, hipStream_t stream);
hipError_t count_ones_per_record(const uint8_t* hits, const uint64_t* offsets,
                                 const uint64_t* counts, uint64_t num_records, uint64_t* d_out,
                                 hipStream_t stream);
hipError_t fill_u64(uint64_t* dst, uint64_t value, uint64_t n, hipStream_t stream);
// *d_out = sum over i of mix(words[i] ^ i * golden) mod 2^64 (filter image checksum).
hipError_t words_checksum(const uint64_t* words, uint64_t n, uint64_t* d_out, hipStream_t stream);

}  // namespace sbf
