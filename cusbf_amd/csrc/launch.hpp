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
    // 1 while every word is known to be zero: set by a clear, reset to 0 by
    // thread 0 of block 0 of every insert kernel at its start (null: no flag).
    uint64_t* zero = nullptr;
};

// Every insert kernel calls this first: the words may change from here on.
__device__ __forceinline__ void mark_words_written(const FilterView& f) {
    if (f.zero != nullptr && blockIdx.x == 0 && threadIdx.x == 0) *f.zero = 0;
}

// Launch shape of the persistent batched kernels.
struct BatchLaunch {
    uint32_t blocks_per_cu = 4;
    uint32_t compute_units = 256;
    uint32_t claim = 1;  // tiles per work-queue claim; 0 = waves stride over the tiles
};

// Sorted insert: where the lane-strip insert kernels write their runs instead
// of applying them with atomics. Wave w of the grid (block * 4 + wave) owns
// region w, slots [w * cap, (w + 1) * cap): slot i holds the pair
// (shard << 32) | i and the run's OR mask (4 words at masks + 4 i); the wave
// writes its record count to counts[w]. Runs past the region's capacity use
// atomics, and a wave whose region is full sets *overflow to 1. The launch
// uses at most regions / 4 blocks.
struct RecordSink {
    uint64_t* pairs = nullptr;
    uint64_t* masks = nullptr;
    uint32_t* counts = nullptr;
    uint32_t* overflow = nullptr;
    uint32_t cap = 0;
    uint32_t regions = 0;
};

// Every launcher takes the filter's configuration first: the lane strips
// compiled for one (k, s, m, h) ignore it, those with runtime spans read k, s
// and m from it.
using InsertFn = hipError_t (*)(const KernelConfig&, const SequenceInput&, FilterView,
                                const RecordSink*, hipStream_t);
using ContainsFn = hipError_t (*)(const KernelConfig&, const SequenceInput&, FilterView, uint8_t*,
                                  hipStream_t);
using BatchedInsertFn = hipError_t (*)(const KernelConfig&, const SequenceInput&, FilterView,
                                       const BatchLaunch&, const RecordSink*, hipStream_t);
using BatchedContainsFn = hipError_t (*)(const KernelConfig&, const SequenceInput&, FilterView,
                                         const BatchLaunch&, uint8_t*, hipStream_t);

// Lane-strip kernels for one compile-time (k, s, m, h, alphabet), or (runtime
// spans) for every (k, s, m) of one alphabet and h within the kernels' limits.
struct StripKernelSet {
    uint32_t k, s, m, h, alphabet;  // k = s = m = 0: runtime spans
    bool dense;             // plain kernels also read dense packed input
    uint32_t wave_windows;  // windows per wave tile
    InsertFn insert;
    ContainsFn contains;
    BatchedInsertFn batched_insert[2];      // [0]: built for 3 waves per SIMD, [1]: for 4 (or null)
    BatchedContainsFn batched_contains[2];
};

// Returns the lane-strip kernel set compiled for this configuration, else the
// runtime-span set that covers it, else nullptr (the generic kernels).
const StripKernelSet* find_strip_kernels(uint32_t k, uint32_t s, uint32_t m, uint32_t h,
                                         uint32_t alphabet);
// The runtime-span set for (alphabet, h) when (k, s, m) is within its limits.
const StripKernelSet* find_wide_kernels(uint32_t k, uint32_t s, uint32_t m, uint32_t h,
                                        uint32_t alphabet);

// Blocks of the plain lane-strip kernels (a dispatch stays below 2^32
// work-items) and of the persistent batched kernels.
constexpr uint64_t kStripMaxBlocks = 1ull << 16;
constexpr uint32_t kStripWavesPerBlock = 4;
inline uint32_t batched_grid(const BatchLaunch& b, uint32_t num_tiles) {
    uint64_t blocks = static_cast<uint64_t>(b.blocks_per_cu) * b.compute_units;
    const uint64_t need = (static_cast<uint64_t>(num_tiles) + kStripWavesPerBlock - 1) / kStripWavesPerBlock;
    if (blocks > need) blocks = need;
    if (blocks > kStripMaxBlocks) blocks = kStripMaxBlocks;
    if (blocks == 0) blocks = 1;
    return static_cast<uint32_t>(blocks);
}

// Sorted insert (sorted_insert.hip): one device allocation holding the record
// buffer of `regions` x `cap` slots (pairs twice, for the radix passes, and
// masks), the region counts, the radix histograms and scan partials, and the
// slice bounds. The partition groups the pairs by shard >> slice_bits, in
// ceil((log2(num_shards) - slice_bits) / 8) passes (at least one).
struct SortedLayout {
    uint32_t regions = 0;
    uint32_t cap = 0;
    uint64_t slots = 0;
    uint32_t slice_bits = 0;        // shards per slice = 2^slice_bits (32 B each in LDS)
    uint32_t slice_index_bits = 0;  // log2(num_shards) - slice_bits
    uint32_t num_slices = 0;
    uint32_t passes = 0;
    uint64_t hist_len = 0;
    uint32_t partials_len = 0;
    uint64_t off_pairs[2] = {0, 0};
    uint64_t off_masks = 0, off_counts = 0, off_hist = 0, off_partials = 0, off_totals = 0,
             off_bounds = 0, off_groups = 0;
    uint64_t bytes = 0;
};

namespace sorted {
// regions * cap must stay below 2^31; slice_bits in [1, 12].
SortedLayout sorted_layout(uint32_t regions, uint32_t cap, uint64_t num_shards, uint32_t slice_bits);
RecordSink record_sink(const SortedLayout& l, void* base);
// Allows the apply kernel its LDS slice; call before launching (outside any capture).
hipError_t prepare();
// First step of a sorted insert, in stream order before the insert kernels:
// zeroes the region counts and the overflow flag, and moves f.zero into the
// workspace (f.zero becomes 0), where the apply reads it.
hipError_t sorted_begin(const SortedLayout& l, void* base, FilterView f, hipStream_t stream);
// Partition the records the insert kernels wrote and OR them into the filter,
// in stream order after those kernels.
hipError_t sorted_apply(const SortedLayout& l, void* base, FilterView f, hipStream_t stream);
}  // namespace sorted

// Generic kernels: any alphabet / configuration, one window per thread; in the
// batched form they test the record-start bitmap (or, without it, search the
// offsets) for each window. With a record sink the insert writes its runs there
// (regions per wave, as for the lane-strip kernels; a wave takes 64 windows of a
// block tile at a time) for a sorted insert.
constexpr uint32_t kGenericWaveWindows = 64;
// Whether the generic kernels use hashed tiles for this configuration (the
// sorted insert's build always uses direct tiles).
bool generic_hashed_tiles(const KernelConfig& c);
hipError_t generic_insert(const KernelConfig& c, const SequenceInput& in, FilterView f,
                          const RecordSink* sink, hipStream_t stream);
hipError_t generic_contains(const KernelConfig& c, const SequenceInput& in, FilterView f,
                            uint8_t* out, hipStream_t stream);
// d_counts[0] += valid windows, d_counts[1] += shard runs (caller zeroes).
hipError_t generic_count(const KernelConfig& c, const SequenceInput& in, uint64_t num_shards,
                         uint64_t* d_counts, hipStream_t stream);
hipError_t pack_dense(const KernelConfig& c, const char* seq, uint64_t num_symbols,
                      uint64_t* words, hipStream_t stream);

// Words of the record-start bitmap for num_symbols symbols (one spare word, so
// a 16-bit lane chunk just past the end still reads inside the bitmap).
__host__ __device__ constexpr uint64_t start_words(uint64_t num_symbols) { return num_symbols / 64 + 2; }

// Batched form set-up, in stream order before the batched kernels (fill
// kernels only, so the sequence can be captured in a graph): zero the state
// and the bitmap, then set one bit per record start and check the offsets.
hipError_t batch_prepass(const SequenceInput& in, uint64_t* starts, uint32_t* state,
                         hipStream_t stream);

// Records back to back (num_records + 1 symbol offsets) -> the same records joined
// by one separator symbol each (num_symbols + num_records - 1 symbols).
hipError_t join_records(const char* in, uint64_t num_symbols, const uint64_t* offsets,
                        uint64_t num_records, uint32_t symbol_width, uint8_t separator, char* out,
                        hipStream_t stream);

// Server report counts: request r owns windows [req[2r], req[2r] + req[2r+1])
// of a batch (sorted, disjoint); counts[2r] += valid windows, counts[2r+1] +=
// hits equal to 1 (caller zeroes counts).
hipError_t request_counts(const KernelConfig& c, const char* seq, uint64_t num_symbols,
                          const uint8_t* hits, const uint64_t* req, uint64_t n_req,
                          uint64_t* counts, hipStream_t stream);

// Hit bytes (0 or 1) packed into bits, LSB first: bits[i / 64] bit i % 64 = hits[i];
// (n + 63) / 64 words.
hipError_t pack_hits(const uint8_t* hits, uint64_t n, uint64_t* bits, hipStream_t stream);

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
