// launch.hpp -- host-side launchers of the Super Bloom kernels (internal to libcusbf_hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "sbf_common.hpp"

namespace sbf {

// Input of one launch: either a byte sequence (`seq`, symbol_width bytes per
// symbol) or a dense packed sequence (`words`, LSB-first, 64/symbol_bits
// symbols per word).
struct SequenceInput {
    const char* seq = nullptr;
    const uint64_t* words = nullptr;
    uint64_t num_words = 0;
    uint64_t num_symbols = 0;
    bool dense = false;
};

struct FilterView {
    uint64_t* words = nullptr;  // num_shards * 4
    uint64_t num_shards = 0;
};

// Fast DNA kernels for one compile-time (k, s, m, h).
struct DnaKernelSet {
    uint32_t k, s, m, h;
    hipError_t (*insert)(const SequenceInput&, FilterView, hipStream_t);
    hipError_t (*contains)(const SequenceInput&, FilterView, uint8_t*, hipStream_t);
};

// Returns the compiled fast kernel set for (k, s, m, h) or nullptr.
const DnaKernelSet* find_dna_kernels(uint32_t k, uint32_t s, uint32_t m, uint32_t h);

// Generic kernels: any alphabet / configuration, one window per thread.
hipError_t generic_insert(const KernelConfig& c, const SequenceInput& in, FilterView f,
                          hipStream_t stream);
hipError_t generic_contains(const KernelConfig& c, const SequenceInput& in, FilterView f,
                            uint8_t* out, hipStream_t stream);
// d_counts[0] += valid windows, d_counts[1] += shard runs (caller zeroes).
hipError_t generic_count(const KernelConfig& c, const SequenceInput& in, uint64_t num_shards,
                         uint64_t* d_counts, hipStream_t stream);
hipError_t pack_dense(const KernelConfig& c, const char* seq, uint64_t num_symbols,
                      uint64_t* words, hipStream_t stream);

// Reductions.
hipError_t popcount_words(const uint64_t* words, uint64_t n, uint64_t* d_out, hipStream_t stream);
hipError_t count_ones(const uint8_t* hits, uint64_t n, uint64_t* d_out, hipStream_t stream);
hipError_t count_ones_per_record(const uint8_t* hits, const uint64_t* offsets,
                                 const uint64_t* counts, uint64_t num_records, uint64_t* d_out,
                                 hipStream_t stream);
hipError_t fill_u64(uint64_t* dst, uint64_t value, uint64_t n, hipStream_t stream);

}  // namespace sbf
