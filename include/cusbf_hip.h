/*
 * cusbf_hip.h -- C ABI of the MI355X-native Super Bloom filter (libcusbf_hip.so).
 *
 * Plain pointers and sizes only. Device pointers are HIP device allocations on
 * the filter's device; streams are hipStream_t passed as void* (NULL = the
 * default stream). Every entry point returns a cusbf_status; on failure the
 * calling thread's last error (category + message) is readable through
 * cusbf_last_error().
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the reference checkout; all in include/cusbf/).
 */
#ifndef CUSBF_HIP_H
#define CUSBF_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CUSBF_ABI_VERSION 1

/* Status codes. A failing status minus one is the reference's ErrorCategory
 * index (error.hpp:67-73, pinned by :139-143): cuda=0, io=1, fastx_parse=2,
 * invalid_argument=3, resource=4. */
typedef enum cusbf_status {
    CUSBF_OK = 0,
    CUSBF_ERR_DEVICE = 1,           /* ErrorCategory::cuda (a HIP runtime error here) */
    CUSBF_ERR_IO = 2,               /* ErrorCategory::io */
    CUSBF_ERR_FASTX_PARSE = 3,      /* ErrorCategory::fastx_parse */
    CUSBF_ERR_INVALID_ARGUMENT = 4, /* ErrorCategory::invalid_argument */
    CUSBF_ERR_RESOURCE = 5          /* ErrorCategory::resource */
} cusbf_status;

/* Alphabets: Alphabet.cuh:113-142 (DNA), :150-178 (DnaTriplet), :189-218
 * (Protein); CUSTOM = any width-1 alphabet given as a byte->symbol table
 * (the Alphabet concept of Alphabet.cuh:94-107 evaluated for all 256 bytes). */
typedef enum cusbf_alphabet {
    CUSBF_ALPHABET_DNA = 0,
    CUSBF_ALPHABET_DNA_TRIPLET = 1,
    CUSBF_ALPHABET_PROTEIN = 2,
    CUSBF_ALPHABET_CUSTOM = 3
} cusbf_alphabet;

/* Runtime form of Config<K,S,M,HashCount,CudaBlockSize,Alphabet> (config.cuh:28-118).
 * symbol_count/invalid_symbol/separator/lut are read only for CUSTOM. */
typedef struct cusbf_config {
    uint32_t k;
    uint32_t s;
    uint32_t m;
    uint32_t hash_count;
    uint32_t alphabet; /* cusbf_alphabet */
    uint32_t symbol_count;
    uint8_t invalid_symbol;
    uint8_t separator;
    uint8_t reserved[2];
    uint8_t lut[256];
} cusbf_config;

/* RecordRange (Fastx.hpp:24-29). */
typedef struct cusbf_record_range {
    uint64_t sequence_offset;
    uint64_t sequence_bytes;
} cusbf_record_range;

/* FastxInsertReport (Fastx.hpp:150-158). */
typedef struct cusbf_insert_report {
    uint64_t records_indexed;
    uint64_t indexed_bases;
    uint64_t inserted_kmers;
} cusbf_insert_report;

/* FastxQueryReport (Fastx.hpp:161-170). */
typedef struct cusbf_query_report {
    uint64_t records_queried;
    uint64_t queried_bases;
    uint64_t queried_kmers;
    uint64_t positive_kmers;
} cusbf_query_report;

typedef struct cusbf_filter cusbf_filter;

/* --- errors --------------------------------------------------------------- */

/* Error::category()/message() of the last failure on this thread
 * (error.hpp:76-87). Returns the status of that failure (CUSBF_OK if none);
 * copies the NUL-terminated message into buf when buf_len > 0. */
int cusbf_last_error(char* buf, size_t buf_len);
int cusbf_abi_version(void);

/* --- configuration helpers (config.cuh:28-118) ------------------------------ */

/* Fills *cfg for a built-in alphabet; validates the static_asserts of
 * config.cuh:74-117 (returns CUSBF_ERR_INVALID_ARGUMENT when violated). */
int cusbf_config_init(cusbf_config* cfg, uint32_t k, uint32_t s, uint32_t m,
                      uint32_t hash_count, uint32_t alphabet);
/* Derived constants: symbolWidth (:49), symbolBits (:51). */
int cusbf_config_symbol_width(const cusbf_config* cfg, uint32_t* out);
int cusbf_config_symbol_bits(const cusbf_config* cfg, uint32_t* out);

/* --- lifetime (filter.cuh:108-130) ----------------------------------------- */

/* filter<Config>::filter(requestedFilterBits) (filter.cuh:108-120): shards =
 * bit_ceil(max(1, ceil(bits/256))), zeroed. Uses the current HIP device. */
int cusbf_create(const cusbf_config* cfg, uint64_t requested_filter_bits, cusbf_filter** out);
/* ~filter() (filter.cuh:130). NULL is a no-op. */
void cusbf_destroy(cusbf_filter* f);

/* clear(stream) (filter.cuh:793-803): zero all shards, synchronise. The words
 * are cleared by a fill kernel (no memset node); on a stream that is capturing
 * a graph the clear is recorded without the synchronize. */
int cusbf_clear(cusbf_filter* f, void* stream);
/* load_factor() (filter.cuh:810-823). */
int cusbf_load_factor(const cusbf_filter* f, float* out);
/* filter_bits() / num_shards() (filter.cuh:826-833). */
uint64_t cusbf_filter_bits(const cusbf_filter* f);
uint64_t cusbf_num_shards(const cusbf_filter* f);
/* ref().blocks() (filter.cuh:137-142, filter_ref.cuh:56-64): device address of
 * shard 0; shard i occupies words [4i, 4i+4). The caller may write the words
 * through it, so from the first call on the filter no longer assumes its words
 * are zero after a clear (a sorted insert then always reads the slices it
 * updates); the first call waits for the null stream. */
int cusbf_words_device_ptr(const cusbf_filter* f, uint64_t** out);
/* Parity dump of all num_shards*4 filter words into host memory (synchronous). */
int cusbf_copy_words_to_host(const cusbf_filter* f, uint64_t* host_words, uint64_t num_words);

/* --- byte-sequence path ---------------------------------------------------- */

/* insert_sequence_async(device_span<const char>, stream) (filter.cuh:179-190):
 * returns all k-mer windows (invalid ones are skipped by the kernel). */
int cusbf_insert_sequence_async(cusbf_filter* f, const char* d_sequence, uint64_t bytes,
                                void* stream, uint64_t* inserted);
/* insert_sequence(std::string_view, stream) (filter.cuh:154-167): host bytes,
 * staged through library-owned device memory, synchronises. */
int cusbf_insert_sequence(cusbf_filter* f, const char* h_sequence, uint64_t bytes, void* stream,
                          uint64_t* inserted);
/* contains_sequence_async(device_span<const char>, device_span<uint8_t>, stream)
 * (filter.cuh:464-474): one byte per window, 1 = present, 0 = absent/invalid.
 * Like the reference it does not check d_output_len against the window count;
 * the caller provides at least that many bytes. */
int cusbf_contains_sequence_async(const cusbf_filter* f, const char* d_sequence, uint64_t bytes,
                                  uint8_t* d_output, uint64_t d_output_len, void* stream);
/* contains_sequence(std::string_view, device_span<uint8_t>, stream) (filter.cuh:486-501):
 * host bytes staged to the device, answers written to device memory, synchronises. */
int cusbf_contains_sequence_host_to_device(const cusbf_filter* f, const char* h_sequence,
                                           uint64_t bytes, uint8_t* d_output,
                                           uint64_t d_output_len, void* stream);
/* contains_sequence(std::string_view, std::vector out) (filter.cuh:514-533):
 * host bytes in, host hits out (h_output must hold the window count), synchronises. */
int cusbf_contains_sequence(const cusbf_filter* f, const char* h_sequence, uint64_t bytes,
                            uint8_t* h_output, uint64_t h_output_len, void* stream);

/* --- dense packed path (dense_packed.hpp, filter.cuh:207-312) --------------- */

/* dense_packed_word_count / dense_packed_kmer_count (dense_packed.cuh:19-28). */
uint64_t cusbf_dense_packed_word_count(const cusbf_config* cfg, uint64_t num_symbols);
uint64_t cusbf_dense_packed_kmer_count(const cusbf_config* cfg, uint64_t num_symbols);
/* pack_dense_sequence_device (dense_packed.hpp:57-73). */
int cusbf_pack_dense_sequence_async(const cusbf_config* cfg, const char* d_sequence,
                                    uint64_t num_symbols, uint64_t* d_words, uint64_t num_words,
                                    void* stream);
/* insert_dense_packed_async (filter.cuh:207-219). */
int cusbf_insert_dense_packed_async(cusbf_filter* f, const uint64_t* d_words, uint64_t num_words,
                                    uint64_t num_symbols, void* stream, uint64_t* inserted);
/* contains_dense_packed_async (filter.cuh:248-259). */
int cusbf_contains_dense_packed_async(const cusbf_filter* f, const uint64_t* d_words,
                                      uint64_t num_words, uint64_t num_symbols, uint8_t* d_output,
                                      uint64_t d_output_len, void* stream);

/* insert_dense_packed(std::span<const uint64_t>, num_symbols, stream) (filter.cuh:226-240). */
int cusbf_insert_dense_packed(cusbf_filter* f, const uint64_t* h_words, uint64_t num_words,
                              uint64_t num_symbols, void* stream, uint64_t* inserted);
/* contains_dense_packed(std::span, num_symbols, device_span<uint8_t>) (filter.cuh:267-282). */
int cusbf_contains_dense_packed_host_to_device(const cusbf_filter* f, const uint64_t* h_words,
                                               uint64_t num_words, uint64_t num_symbols,
                                               uint8_t* d_output, uint64_t d_output_len,
                                               void* stream);
/* contains_dense_packed(std::span, num_symbols) -> std::vector (filter.cuh:289-312):
 * h_output must hold the window count. */
int cusbf_contains_dense_packed(const cusbf_filter* f, const uint64_t* h_words, uint64_t num_words,
                                uint64_t num_symbols, uint8_t* h_output, uint64_t h_output_len,
                                void* stream);

/* --- device record batches (filter.cuh:385-400, 603-618) ------------------- */

/* insert_record_batch_async: separators already embedded; ranges validated
 * (filter.cuh:848-876); returns the sum of per-record window counts. */
int cusbf_insert_record_batch_async(cusbf_filter* f, const char* d_sequence, uint64_t bytes,
                                    const cusbf_record_range* ranges, uint64_t num_ranges,
                                    void* stream, uint64_t* inserted);
/* contains_record_batch_async: output must hold every window of d_sequence. */
int cusbf_contains_record_batch_async(const cusbf_filter* f, const char* d_sequence,
                                      uint64_t bytes, const cusbf_record_range* ranges,
                                      uint64_t num_ranges, uint8_t* d_output,
                                      uint64_t d_output_len, void* stream);

/* --- host record batches with separator injection --------------------------- */

/* insert_record_batch(RecordBatchView) (filter.cuh:338-364): normalises
 * (normalized_record_batch.hpp:288-304), stages, inserts, synchronises. */
int cusbf_insert_record_batch(cusbf_filter* f, const char* h_sequence, uint64_t bytes,
                              const cusbf_record_range* ranges, uint64_t num_ranges,
                              void* stream, cusbf_insert_report* report);
/* query_record_batch(RecordBatchView) aggregate form (filter.cuh:549-552, 951-990). */
int cusbf_query_record_batch(const cusbf_filter* f, const char* h_sequence, uint64_t bytes,
                             const cusbf_record_range* ranges, uint64_t num_ranges, void* stream,
                             cusbf_query_report* report);

/* query_record_batch(RecordBatchView, Consumer) (filter.cuh:566-582, 1007-1131):
 * like cusbf_query_record_batch, plus one byte per window of the normalised
 * batch in h_hits (record r's hits start at its normalised offset,
 * query_layout.cuh:41-72); synchronises. */
int cusbf_query_record_batch_hits(const cusbf_filter* f, const char* h_sequence, uint64_t bytes,
                                  const cusbf_record_range* ranges, uint64_t num_ranges,
                                  void* stream, uint8_t* h_hits, uint64_t hits_len,
                                  cusbf_query_report* report);

/* --- counting and diagnostics ---------------------------------------------- */

/* thrust::count(hits == 1) (count_positive_kmers.cuh:47-62) as a HIP reduction;
 * d_count receives one uint64 (device memory). */
int cusbf_count_positive_async(const uint8_t* d_hits, uint64_t num_hits, uint64_t* d_count,
                               void* stream);
/* Per-record positive counts (count_positive_kmers.cuh:21-45, 69-90):
 * d_offsets/d_counts are per-record hit offsets and hit counts. */
int cusbf_count_positive_per_record_async(const uint8_t* d_hits, const uint64_t* d_offsets,
                                          const uint64_t* d_counts, uint64_t num_records,
                                          uint64_t* d_positive, void* stream);
/* No reference counterpart: counts valid windows and shard runs (maximal runs
 * of consecutive valid windows sharing a shard) of a device sequence, for the
 * byte model of the roofline. d_counts receives two uint64 (device memory). */
int cusbf_count_windows_runs_async(const cusbf_filter* f, const char* d_sequence, uint64_t bytes,
                                   uint64_t* d_counts, void* stream);

/* Name of the kernel family the filter dispatches to: "lane_strip" (compiled for this
 * configuration), "lane_strip_wide" (runtime spans) or "generic". */
const char* cusbf_kernel_family(const cusbf_filter* f);

/* --- insert path (no reference counterpart) ---------------------------------- */
/*
 * The lane-strip inserts apply a shard run either with one 64-bit device-scope
 * atomic OR per non-zero shard word (ATOMIC: the reference's method,
 * filter_ref.cuh:94-108; on MI355X each atomic is a memory-side
 * read-modify-write), or SORTED: the insert kernel writes each run as a
 * record (shard, 32-byte OR mask), radix passes group the records by filter
 * slice (2^slice_bits shards), and one block per touched slice ORs its records
 * into the slice in LDS and writes it back. Both give identical filter words.
 * SORTED needs a workspace of about (48 bytes x reserved records) that the
 * filter keeps (grow-only; freed by cusbf_destroy); runs beyond a wave's
 * reservation, and beyond the 256 LDS run slots of a wave tile, still use
 * atomics. AUTO (the default) takes SORTED when a cost model says it is faster
 * and the workspace fits in half the free device memory, else ATOMIC. Inserts
 * recorded into a stream capture always use ATOMIC, and so do batch plans on
 * the generic kernels; plain and batched inserts on the generic kernels (any
 * alphabet or configuration without lane-strip kernels) take SORTED as one
 * sorted insert (no window ranges) when forced, and under AUTO when the cost
 * model prefers it
 * (it weighs the sorted build's direct tiles against the atomic insert's hashed
 * tiles where the generic kernels hash whole tiles).
 * Sorted inserts on one filter share its workspace: one on another stream
 * waits for the previous one (stream-ordered, no host wait).
 */
#define CUSBF_INSERT_AUTO 0u
#define CUSBF_INSERT_ATOMIC 1u
#define CUSBF_INSERT_SORTED 2u
/* slice_bits: 1..12 (0 = 12); reserve_per_tile: records reserved per wave tile,
 * 1..1024 (0 = 2.5 / (k - m + 2) of the tile's windows, at most the tile's windows). */
int cusbf_set_insert_path(cusbf_filter* f, uint32_t path, uint32_t slice_bits,
                          uint32_t reserve_per_tile);
/* Current settings and the path the last insert call took (ATOMIC or SORTED). */
int cusbf_get_insert_path(const cusbf_filter* f, uint32_t* path, uint32_t* slice_bits,
                          uint32_t* reserve_per_tile, uint32_t* last_path);
/* Workspace limit of the sorted insert (no reference counterpart), in bytes;
 * 0 (the default) allows half of the free device memory. A plain insert (byte
 * or dense packed input) whose records would not fit the limit, or would need
 * 2^31 record slots or more, runs as several sorted inserts over consecutive
 * window ranges of whole wave tiles, one after another in stream order; the
 * words are the same as from one insert (OR is idempotent). Each range after
 * the first loads and stores the slices it touches, which the cost model of
 * AUTO counts against the atomic path. A limit below the workspace the filter
 * holds frees it (after the last sorted insert has finished). Batched inserts
 * run as one sorted insert within the limit, else on the atomic path. */
int cusbf_set_sorted_workspace_limit(cusbf_filter* f, uint64_t bytes);
/* The limit (0: default), the bytes the workspace holds now, and the number of
 * sorted inserts the last plain insert ran as (0 when it took the atomic path). */
int cusbf_get_sorted_workspace(const cusbf_filter* f, uint64_t* limit, uint64_t* allocated,
                               uint32_t* last_chunks);

/* --- batched record form (no single-call counterpart in the reference) -------- */
/*
 * Many records of uneven length in ONE launch, without the separator
 * normalisation of the record-batch path: d_symbols holds the records
 * concatenated back to back (num_symbols symbols, symbol_width bytes each) and
 * d_offsets (device memory, num_records + 1 entries) the symbol offset of each
 * record start, non-decreasing from 0 to num_symbols. Every window that lies
 * inside one record is inserted / queried exactly as insert_sequence /
 * contains_sequence would on that record alone (filter.cuh:179-190, 464-474);
 * a window that crosses a record start is skipped on insert and answered 0 on
 * query, the answer the reference gives for a window across a separator
 * (normalized_record_batch.hpp:208-232). Query output: one byte per window
 * start of the concatenation (num_symbols - k + 1 bytes); record r's answers
 * are d_output[d_offsets[r] .. d_offsets[r+1] - k].
 *
 * Launch sequence (stream order): fill kernels that zero the work-queue state
 * and a record-start bitmap (num_symbols / 8 bytes of library scratch; fill
 * kernels rather than memsets, so the sequence can be captured in a graph), a
 * prepass that sets one bit per record start and checks the offsets, and one
 * persistent kernel whose waves take tiles from the work queue (or stride over
 * them). Malformed offsets make the kernel do nothing; cusbf_batched_status /
 * cusbf_batch_plan_status report it.
 * Launch shape per record-length bucket comes from the filter's tuning table.
 * Calls on one filter share its batch scratch: issue them on one stream.
 */
#define CUSBF_BATCH_BUCKETS 4 /* mean record length < 256, < 4096, < 65536, >= 65536 symbols */
/* Table entries: CUSBF_BATCH_INSERT is the shape of a batched insert on the
 * atomic path, CUSBF_BATCH_INSERT_SORTED of one on the sorted path (see
 * cusbf_set_insert_path); plans take INSERT or CONTAINS as their op. */
typedef enum cusbf_batch_op {
    CUSBF_BATCH_INSERT = 0,
    CUSBF_BATCH_CONTAINS = 1,
    CUSBF_BATCH_INSERT_SORTED = 2
} cusbf_batch_op;

typedef struct cusbf_batch_tuning {
    uint32_t waves_per_simd; /* register budget of the kernel variant: 3 or 4 (4: DNA only) */
    uint32_t blocks_per_cu;  /* persistent blocks (4 waves each) per compute unit, 1..32 */
    uint32_t claim;          /* tiles per work-queue claim, 1..64; 0 = static striding */
    uint32_t reserved;
} cusbf_batch_tuning;

/* Batched kernels the filter's configuration has: bit 0 = lane-strip kernels
 * built for 3 waves per SIMD, bit 1 = for 4; 0 = the generic kernels (the
 * tuning table is then not used). */
uint32_t cusbf_batch_variants(const cusbf_filter* f);
/* Bucket of a batch: by mean record length num_symbols / num_records. */
uint32_t cusbf_batch_bucket(uint64_t num_symbols, uint64_t num_records);
int cusbf_batch_tuning_set(cusbf_filter* f, uint32_t bucket, uint32_t op,
                           const cusbf_batch_tuning* tuning);
int cusbf_batch_tuning_get(const cusbf_filter* f, uint32_t bucket, uint32_t op,
                           cusbf_batch_tuning* out);
int cusbf_insert_batched_async(cusbf_filter* f, const char* d_symbols, uint64_t num_symbols,
                               const uint64_t* d_offsets, uint64_t num_records, void* stream);
int cusbf_contains_batched_async(const cusbf_filter* f, const char* d_symbols,
                                 uint64_t num_symbols, const uint64_t* d_offsets,
                                 uint64_t num_records, uint8_t* d_output, uint64_t d_output_len,
                                 void* stream);
/* Synchronises the stream; CUSBF_ERR_INVALID_ARGUMENT if the last batched call's
 * offsets were malformed. */
int cusbf_batched_status(const cusbf_filter* f, void* stream);

/* No counterpart in the reference, whose device record batches take the records
 * joined already (insert_record_batch_async / contains_record_batch_async,
 * filter.cuh:385-400, 603-618): the records of a batched call (back to back,
 * num_records + 1 non-decreasing symbol offsets from 0 to num_symbols) joined by
 * one separator symbol of the filter's alphabet each, on the device -- the layout
 * the record-batch calls read (normalized_record_batch.hpp:208-232). d_out_len
 * must hold (num_symbols + num_records - 1) symbols. Malformed offsets give an
 * unspecified layout (no access outside d_symbols and d_out). */
int cusbf_join_records_async(const cusbf_filter* f, const char* d_symbols, uint64_t num_symbols,
                             const uint64_t* d_offsets, uint64_t num_records, char* d_out,
                             uint64_t d_out_len, void* stream);

/* No counterpart in the reference. For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): cusbf_contains_batched_async with the same arguments and answers,
 * whose scratch (status words and the record-start bitmap) lives in a fixed 32 MiB
 * device array of the library instead of buffers the filter allocates. It calls no
 * allocating routine. Batches of more than CUSBF_NOALLOC_MAX_SYMBOLS symbols fail
 * with CUSBF_ERR_INVALID_ARGUMENT. Every call on one GPU shares that GPU's array:
 * calls that overlap on one GPU (other streams, other threads, other filters) race
 * and must be ordered by the caller. The filter's device must be current.
 * cusbf_batched_status reports this call's malformed offsets until the next
 * batched call. */
#define CUSBF_NOALLOC_MAX_SYMBOLS ((1ull << 28) - 4096ull)
int cusbf_contains_batched_noalloc_async(const cusbf_filter* f, const char* d_symbols,
                                         uint64_t num_symbols, const uint64_t* d_offsets,
                                         uint64_t num_records, uint8_t* d_output,
                                         uint64_t d_output_len, void* stream);

/* A batched insert or query captured once into a hipGraph (reset, prepass,
 * persistent kernel) over fixed buffers and replayed with one launch. The plan
 * owns its scratch; `tuning` NULL = the filter's table entry for the batch's
 * bucket at creation. */
typedef struct cusbf_batch_plan cusbf_batch_plan;
int cusbf_batch_plan_create(cusbf_filter* f, uint32_t op, const char* d_symbols,
                            uint64_t num_symbols, const uint64_t* d_offsets, uint64_t num_records,
                            uint8_t* d_output, uint64_t d_output_len,
                            const cusbf_batch_tuning* tuning, cusbf_batch_plan** out);
/* Replays the plan on `stream`. An insert plan runs after the filter's previous
 * insert on any stream and before its next one, like a direct insert (skipped
 * while `stream` is capturing). */
int cusbf_batch_plan_launch(cusbf_batch_plan* p, void* stream);
/* Synchronises the stream; CUSBF_ERR_INVALID_ARGUMENT if the offsets were malformed. */
int cusbf_batch_plan_status(cusbf_batch_plan* p, void* stream);
void cusbf_batch_plan_destroy(cusbf_batch_plan* p);
/* No reference counterpart: valid windows and shard runs of a batched input
 * (windows across a record start are not valid), for the byte model.
 * d_counts receives two uint64 (device memory). */
int cusbf_count_batched_windows_runs_async(const cusbf_filter* f, const char* d_symbols,
                                           uint64_t num_symbols, const uint64_t* d_offsets,
                                           uint64_t num_records, uint64_t* d_counts,
                                           void* stream);

/* --- FASTA / FASTQ ingest (filter.cuh:402-470, 620-800; Fastx.hpp) ------------ */
/*
 * Parses FASTA or FASTQ text (the first header decides; CRLF line ends and
 * wrapped sequence / quality lines accepted; Fastx.hpp:240-460) and inserts or
 * queries every record, chunk by chunk: a chunk is flushed before it would pass
 * fill_fraction of free device memory (minus 64 MiB) spread over its buffers,
 * or the filter's chunk override (0 = one record per chunk); a record larger
 * than that is processed in a chunk of its own. Parse errors are CUSBF_ERR_FASTX_PARSE with the message
 * "source:line:column: detail" (cusbf_last_error_location gives the parts);
 * chunks before the error have been applied. Reports follow FastxInsertReport /
 * FastxQueryReport: records, sequence bytes, valid windows (no invalid symbol,
 * not across records), positive windows. Windows never span two records. The
 * sources: a host buffer ("<stream>" when source_name is NULL), a read callback,
 * or a file path (gzip when it starts with the gzip magic bytes, else mmap'd).
 */
typedef struct cusbf_fastx_record_view {
    uint64_t record_index;  /* in the input */
    const char* header;     /* without the leading '>' or '@' */
    uint64_t header_len;
    const char* sequence;
    uint64_t sequence_len;
    uint64_t queried_bases;
    uint64_t queried_kmers;  /* valid windows */
    uint64_t positive_kmers;
    const uint8_t* hits;     /* one byte per window (1 present), valid during the call */
    uint64_t hits_len;
} cusbf_fastx_record_view;

/* FastxRecordConsumer (Fastx.hpp:230-238): called once per record, in input
 * order; returning non-zero stops the call (it then fails with CUSBF_ERR_IO,
 * "FASTX record consumer stopped the query"; the filter stays usable). */
typedef int (*cusbf_fastx_consumer)(const cusbf_fastx_record_view* record, void* user);
/* Reads up to `capacity` bytes into dst: the count, 0 at the end, -1 on error. */
typedef int64_t (*cusbf_read_fn)(void* user, char* dst, uint64_t capacity);

/* insert_fastx(std::istream&) (filter.cuh:416-422) over a host buffer / a reader. */
int cusbf_insert_fastx_buffer(cusbf_filter* f, const char* text, uint64_t bytes,
                              const char* source_name, double fill_fraction, void* stream,
                              cusbf_insert_report* report);
int cusbf_insert_fastx_reader(cusbf_filter* f, cusbf_read_fn read, void* user,
                              const char* source_name, double fill_fraction, void* stream,
                              cusbf_insert_report* report);
/* insert_fastx_file (filter.cuh:437-453). */
int cusbf_insert_fastx_file(cusbf_filter* f, const char* path, double fill_fraction, void* stream,
                            cusbf_insert_report* report);
/* query_fastx / query_fastx_records / query_fastx_detailed (filter.cuh:632-760):
 * consumer may be NULL (aggregate report only). */
int cusbf_query_fastx_buffer(const cusbf_filter* f, const char* text, uint64_t bytes,
                             const char* source_name, double fill_fraction, void* stream,
                             cusbf_fastx_consumer consumer, void* user, cusbf_query_report* report);
int cusbf_query_fastx_reader(const cusbf_filter* f, cusbf_read_fn read, void* user_read,
                             const char* source_name, double fill_fraction, void* stream,
                             cusbf_fastx_consumer consumer, void* user, cusbf_query_report* report);
/* query_fastx_file / _records / _detailed (filter.cuh:652-800). */
int cusbf_query_fastx_file(const cusbf_filter* f, const char* path, double fill_fraction,
                           void* stream, cusbf_fastx_consumer consumer, void* user,
                           cusbf_query_report* report);
/* FastxReader alone (Fastx.hpp:302-460): records to the consumer (counts 0, no
 * hits); no device work. */
int cusbf_parse_fastx_buffer(const char* text, uint64_t bytes, const char* source_name,
                             cusbf_fastx_consumer consumer, void* user, uint64_t* records);
/* The same records grouped as the FASTX pipeline groups a FASTA text held in
 * memory (no reference counterpart): chunks of about chunk_bytes of text (0 =
 * 64 MiB) and at most max_records records (0 = no limit), a record larger than
 * a chunk taken alone; FASTA text is split by host threads, FASTQ text parsed as
 * by cusbf_parse_fastx_buffer. No device work. */
int cusbf_parse_fastx_buffer_chunked(const char* text, uint64_t bytes, const char* source_name,
                                     uint64_t chunk_bytes, uint64_t max_records,
                                     cusbf_fastx_consumer consumer, void* user, uint64_t* records);
/* No reference counterpart: fixed FASTX chunk capacity in bytes (0 = from fill_fraction). */
int cusbf_fastx_set_chunk_bytes(cusbf_filter* f, uint64_t bytes);
/* SourceLocation of the last error on this thread when it is a FASTX parse error
 * (error.hpp:20-41, 123-129): file, 1-based line and column (empty / 0 otherwise).
 * Returns the status of that error. */
int cusbf_last_error_location(char* file, size_t file_len, uint32_t* line, uint32_t* column);

/* --- filter images (no reference counterpart: filters live in device memory only) */
/*
 * A filter image is a 336-byte header (magic "CUSBFIMG", version 1, the
 * cusbf_config, shard count, filter bits, word count, checksum of the words)
 * followed by the num_shards * 4 words. cusbf_load recreates the filter on the
 * current device and checks the checksum on the device (CUSBF_ERR_IO "filter
 * image checksum mismatch: <path>"); a file that is not an image is
 * CUSBF_ERR_INVALID_ARGUMENT.
 */
int cusbf_save(const cusbf_filter* f, const char* path);
/* Header only: the image's configuration and filter bits. */
int cusbf_image_config(const char* path, cusbf_config* config, uint64_t* filter_bits);
int cusbf_load(const char* path, cusbf_filter** out);

/* --- request server (no reference counterpart) ------------------------------- */
/*
 * Serving form of the query path: the filter stays resident; a scheduler
 * thread admits queued requests into batch slots as slots free up
 * (continuous batching), each slot owning one HIP stream plus device and
 * pinned host staging carved from `staging_bytes`; batches in different slots
 * overlap on the device (copies and kernels). Every request is answered
 * exactly as cusbf_query_record_batch would answer it alone: one byte per
 * window of its normalised record batch (records joined by the alphabet
 * separator, filter.cuh:930-990 layout). Not thread-safe for one ticket; any
 * thread may submit or wait.
 */
typedef struct cusbf_server cusbf_server;

typedef struct cusbf_server_options {
    uint64_t staging_bytes; /* device (and pinned host) staging budget, split over the slots */
    uint32_t slots;         /* batches in flight */
    uint32_t max_wait_us;   /* extra coalescing wait for a batch that is not full (0 = none) */
} cusbf_server_options;

typedef struct cusbf_server_stats {
    uint64_t requests;
    uint64_t batches;
    uint64_t max_requests_per_batch;
    uint64_t bytes;
    uint64_t windows;
    /* Where a batch's time goes, summed over batches: the scheduler's normalisation
     * into pinned staging and its enqueue of copies and kernels (host), the batch's
     * copies and kernels on its stream (device, from events), and the completion
     * thread's copies of answers into the callers' buffers (host). */
    double fill_seconds;
    double launch_seconds;
    double device_seconds;
    double copy_seconds;
} cusbf_server_stats;

int cusbf_server_create(const cusbf_filter* f, const cusbf_server_options* options,
                        cusbf_server** out);
/* Finishes every queued request, then frees the server. */
void cusbf_server_destroy(cusbf_server* s);
/* Queues one request: a host record batch (ranges into h_sequence). h_hits receives
 * one byte per window of the normalised batch (hits_len must hold them). The
 * caller keeps h_sequence and h_hits alive until cusbf_server_wait returns.
 * A request larger than one slot's staging is rejected (CUSBF_ERR_RESOURCE). */
int cusbf_server_submit(cusbf_server* s, const char* h_sequence, uint64_t bytes,
                        const cusbf_record_range* ranges, uint64_t num_ranges, uint8_t* h_hits,
                        uint64_t hits_len, uint64_t* ticket);
/* Window count of the normalised batch a request will produce. */
uint64_t cusbf_server_request_windows(const cusbf_server* s, const cusbf_record_range* ranges,
                                      uint64_t num_ranges);
/* Blocks until the request is answered; fills its report and its latency
 * (submit to results in h_hits), then forgets the ticket. */
int cusbf_server_wait(cusbf_server* s, uint64_t ticket, cusbf_query_report* report,
                      double* latency_seconds);
int cusbf_server_stats_get(const cusbf_server* s, cusbf_server_stats* out);

#ifdef __cplusplus
}
#endif
#endif /* CUSBF_HIP_H */
