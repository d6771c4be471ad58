// filter_internal.hpp -- pieces of the filter object shared by the C ABI sources
// (capi.hip) and the request server (serve.hip). Internal to libcusbf_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/cusbf_hip.h"

namespace cusbf_internal {

// Record a failure as the calling thread's last error; returns `status`.
int set_error(int status, const char* message);
int set_hip_error(hipError_t err, const char* file, int line);

int device(const cusbf_filter* f);
uint32_t symbol_width(const cusbf_filter* f);
uint8_t separator(const cusbf_filter* f);
uint64_t window_count(const cusbf_filter* f, uint64_t bytes);
// normalized_record_batch.hpp:87-110 for one record.
uint64_t valid_kmers(const cusbf_filter* f, const char* seq, uint64_t bytes);
int validate_host_ranges(const cusbf_filter* f, uint64_t bytes, const cusbf_record_range* ranges,
                         uint64_t n);
// Query kernel over a device byte sequence (fast or generic family).
int launch_contains_bytes(const cusbf_filter* f, const char* d_seq, uint64_t bytes, uint8_t* d_out,
                          hipStream_t stream);

// Report counts of a batch's requests on the device (zeroes d_counts first):
// request r owns windows [d_req[2r], d_req[2r] + d_req[2r+1]); d_counts[2r] =
// valid windows, d_counts[2r+1] = positive windows.
int launch_request_counts(const cusbf_filter* f, const char* d_seq, uint64_t bytes,
                          const uint8_t* d_hits, const uint64_t* d_req, uint64_t n_req,
                          uint64_t* d_counts, hipStream_t stream);

// cusbf_contains_batched_async over the given scratch: d_state (16 words) and
// d_starts (sbf::start_words(num_symbols) words); both null = the filter's own
// scratch, allocated or grown here.
int contains_batched(const cusbf_filter* f, const char* d_symbols, uint64_t num_symbols,
                     const uint64_t* d_offsets, uint64_t num_records, uint8_t* d_output,
                     uint64_t d_output_len, uint32_t* d_state, uint64_t* d_starts,
                     hipStream_t stream);

// Looks up the no-allocation scratch of `device` (current) once; cusbf_create calls it.
int noalloc_setup(int device);

uint32_t kmer_length(const cusbf_filter* f);
const cusbf_config& config(const cusbf_filter* f);
// FASTX chunk capacity override in bytes (0 = from fill_fraction).
uint64_t fastx_chunk_bytes(const cusbf_filter* f);
// Error::fastx_parse (error.hpp:123-129): "source:line:column: detail", location kept.
int set_fastx_error(const char* source, uint32_t line, uint32_t column, const char* detail);

}  // namespace cusbf_internal
