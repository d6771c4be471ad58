// capi.hip -- extern "C" boundary of libcusbf_hip.so (declared in include/cusbf_hip.h).
//
// Mirrors the hot-path methods of the reference's cusbf::filter<Config>
// (include/cusbf/filter.cuh:68-1848): same counting conventions, same argument
// checks and error strings, Result<T> categories mapped onto status codes.
// Kernel dispatch: configurations with a compiled lane-strip kernel set use
// the wave64 lane-strip kernels (kernels_strip.hip), everything else the
// generic kernels (kernels_generic.hip). There is no CPU fallback.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/cusbf_hip.h"
#include "filter_internal.hpp"
#include "host_pool.hpp"
#include "launch.hpp"
#include "sbf_common.hpp"

namespace {

thread_local int g_last_status = CUSBF_OK;
thread_local std::string g_last_message;
thread_local std::string g_last_file;  // FASTX parse errors: source, 1-based line and column
thread_local uint32_t g_last_line = 0;
thread_local uint32_t g_last_column = 0;

int fail(int status, std::string message) {
    g_last_status = status;
    g_last_message = std::move(message);
    g_last_file.clear();
    g_last_line = g_last_column = 0;
    return status;
}

// Error::cuda (error.hpp:109-117): "<site>: <runtime error string>".
int fail_hip(hipError_t err, const char* file, int line) {
    char buf[512];
    std::snprintf(buf, sizeof(buf), "%s:%d: %s", file, line, hipGetErrorString(err));
    return fail(CUSBF_ERR_DEVICE, buf);
}

#define CUSBF_HIP_TRY(expr)                                               \
    do {                                                                  \
        const hipError_t _e = (expr);                                     \
        if (_e != hipSuccess) return fail_hip(_e, __FILE__, __LINE__);    \
    } while (0)

uint32_t bit_width_u32(uint32_t x) {
    uint32_t w = 0;
    while (x) {
        ++w;
        x >>= 1;
    }
    return w;
}

struct Derived {
    uint32_t symbol_width;
    uint32_t symbol_count;
    uint32_t symbol_bits;
    uint8_t invalid_symbol;
    uint8_t separator;
};

// Alphabet.cuh:113-218 constants; config.cuh:51 symbolBits.
int derive(const cusbf_config* cfg, Derived* d) {
    switch (cfg->alphabet) {
        case CUSBF_ALPHABET_DNA:
            *d = Derived{1, 4, 0, 0xFF, 'N'};
            break;
        case CUSBF_ALPHABET_DNA_TRIPLET:
            *d = Derived{3, 64, 0, 0xFF, 'N'};
            break;
        case CUSBF_ALPHABET_PROTEIN:
            *d = Derived{1, 26, 0, 0xFF, '*'};
            break;
        case CUSBF_ALPHABET_CUSTOM:
            if (cfg->symbol_count == 0 || cfg->symbol_count > 255) {
                return fail(CUSBF_ERR_INVALID_ARGUMENT, "alphabet symbolCount must be in [1, 255]");
            }
            *d = Derived{1, cfg->symbol_count, 0, cfg->invalid_symbol, cfg->separator};
            break;
        default:
            return fail(CUSBF_ERR_INVALID_ARGUMENT, "unknown alphabet");
    }
    d->symbol_bits = bit_width_u32(d->symbol_count - 1);
    return CUSBF_OK;
}

// config.cuh:74-117, as runtime checks with the static_assert messages.
int validate(const cusbf_config* cfg, const Derived& d) {
    if (cfg->k == 0) return fail(CUSBF_ERR_INVALID_ARGUMENT, "k must be positive");
    if (cfg->m == 0 || cfg->m > cfg->k) return fail(CUSBF_ERR_INVALID_ARGUMENT, "m must satisfy 0 < m <= k");
    if (cfg->s == 0 || cfg->s > cfg->k) return fail(CUSBF_ERR_INVALID_ARGUMENT, "s must satisfy 0 < s <= k");
    if (d.symbol_bits == 0) return fail(CUSBF_ERR_INVALID_ARGUMENT, "alphabet needs at least two symbols");
    if (cfg->k * d.symbol_bits > 64) return fail(CUSBF_ERR_INVALID_ARGUMENT, "k-mer must fit in one packed uint64_t");
    if (cfg->hash_count == 0) return fail(CUSBF_ERR_INVALID_ARGUMENT, "At least one Bloom hash is required");
    if (cfg->hash_count > 16) return fail(CUSBF_ERR_INVALID_ARGUMENT, "This implementation provides 16 multiplicative salts");
    if (cfg->hash_count < 4) return fail(CUSBF_ERR_INVALID_ARGUMENT, "Sectorized layout requires hashCount >= blockWordCount");
    if (cfg->hash_count % 4 != 0) return fail(CUSBF_ERR_INVALID_ARGUMENT, "Hash count must distribute evenly across shard words");
    return CUSBF_OK;
}

hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace

// Pinned ring of the host-memory forms: buffers of kRingBytes; copies below
// kRingDirect bytes go straight from the caller's memory.
constexpr uint32_t kRingSlots = 3;
constexpr uint64_t kRingBytes = 32ull << 20;
constexpr uint64_t kRingDirect = 4ull << 20;

struct cusbf_filter {
    cusbf_config cfg{};
    Derived d{};
    sbf::KernelConfig kc{};
    const sbf::StripKernelSet* strip = nullptr;
    uint64_t num_shards = 0;
    uint64_t filter_bits = 0;
    uint64_t* d_words = nullptr;
    int device = 0;
    // grow-only staging (fastx_pipeline_state.cuh:145-201)
    char* d_stage = nullptr;
    uint64_t stage_cap = 0;
    uint8_t* d_result = nullptr;
    uint64_t result_cap = 0;
    uint64_t* d_scalar = nullptr;  // reductions
    uint64_t* d_rec = nullptr;     // host record batches: per-record requests and counts
    uint64_t rec_cap = 0;          // words
    // batched record form: tuning table, work-queue state, record-start bitmap
    uint32_t compute_units = 256;
    cusbf_batch_tuning tuning[CUSBF_BATCH_BUCKETS][3] = {};
    uint32_t* d_batch_state = nullptr;
    uint64_t* d_starts = nullptr;
    uint64_t starts_cap = 0;  // words
    bool batch_check_pending = false;
    const uint32_t* batch_check_state = nullptr;  // state words of the last batched call
    uint64_t fastx_chunk_bytes = 0;  // FASTX chunk capacity override (0 = from fill_fraction)
    // insert path (atomic or sorted) and the sorted insert's workspace
    uint32_t insert_path = CUSBF_INSERT_AUTO;
    uint32_t slice_bits = 12;
    uint32_t reserve_per_tile = 0;  // 0 = from the minimizer span
    uint32_t last_path = CUSBF_INSERT_ATOMIC;
    void* d_sorted = nullptr;
    uint64_t sorted_bytes = 0;
    uint64_t sorted_limit = 0;  // workspace limit in bytes (0: half of the free device memory)
    uint32_t last_chunks = 0;   // sorted inserts the last plain insert took (0: atomic)
    hipEvent_t sorted_done = nullptr;  // recorded after the last sorted insert's kernels
    // Device flag, 1 while every word is known to be zero (a clear sets it,
    // every insert kernel resets it): a sorted insert then need not read the
    // words. Once the words pointer has been handed out the library cannot see
    // every write, so clears stop setting it.
    uint64_t* d_zero = nullptr;
    mutable bool words_shared = false;
    hipEvent_t insert_done = nullptr;  // recorded after the last insert (outside captures)
    // Host forms: host threads copy the caller's bytes into one pinned buffer
    // while the DMA engine reads (or fills) another.
    char* h_ring[kRingSlots] = {};
    hipEvent_t ring_done[kRingSlots] = {};  // recorded after the last copy that used the buffer
    bool ring_pending[kRingSlots] = {};
    uint32_t ring_next = 0;
    bool insert_recorded = false;
    bool sorted_used = false;
};

struct cusbf_batch_plan {
    cusbf_filter* f = nullptr;
    uint32_t op = 0;
    bool empty = true;
    uint32_t* d_state = nullptr;
    uint64_t* d_starts = nullptr;
    void* d_sorted = nullptr;  // sorted insert workspace of an insert plan (or null)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

uint64_t symbols_of(const cusbf_filter* f, uint64_t bytes) { return bytes / f->d.symbol_width; }

uint64_t windows_of(const cusbf_filter* f, uint64_t bytes) {
    const uint64_t sym = symbols_of(f, bytes);
    return sym < f->cfg.k ? 0 : sym - f->cfg.k + 1;
}

int ensure_stage(cusbf_filter* f, uint64_t bytes) {
    if (bytes <= f->stage_cap) return CUSBF_OK;
    if (f->d_stage) CUSBF_HIP_TRY(hipFree(f->d_stage));
    f->d_stage = nullptr;
    f->stage_cap = 0;
    const hipError_t e = hipMalloc(&f->d_stage, bytes);
    if (e != hipSuccess) return fail_hip(e, __FILE__, __LINE__);
    f->stage_cap = bytes;
    return CUSBF_OK;
}

int ensure_result(cusbf_filter* f, uint64_t bytes) {
    if (bytes <= f->result_cap) return CUSBF_OK;
    if (f->d_result) CUSBF_HIP_TRY(hipFree(f->d_result));
    f->d_result = nullptr;
    f->result_cap = 0;
    const hipError_t e = hipMalloc(&f->d_result, bytes);
    if (e != hipSuccess) return fail_hip(e, __FILE__, __LINE__);
    f->result_cap = bytes;
    return CUSBF_OK;
}

// Host copy of `bytes` bytes, in parts of 4 MiB run on the host pool.
void parallel_copy(void* dst, const void* src, uint64_t bytes) {
    constexpr uint64_t kPart = 4ull << 20;
    const uint64_t parts = (bytes + kPart - 1) / kPart;
    if (parts <= 1) {
        std::memcpy(dst, src, bytes);
        return;
    }
    cusbf_internal::host_pool().run(parts, [&](size_t i) {
        const uint64_t at = i * kPart;
        std::memcpy(static_cast<char*>(dst) + at, static_cast<const char*>(src) + at,
                    std::min(kPart, bytes - at));
    });
}

// The next ring buffer, once the copy that last used it has finished.
int ring_slot(cusbf_filter* f, uint32_t* slot) {
    const uint32_t s = f->ring_next;
    if (f->h_ring[s] == nullptr) {
        if (f->ring_done[s] == nullptr) {
            CUSBF_HIP_TRY(hipEventCreateWithFlags(&f->ring_done[s], hipEventDisableTiming));
        }
        CUSBF_HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&f->h_ring[s]), kRingBytes, hipHostMallocDefault));
    } else if (f->ring_pending[s]) {
        CUSBF_HIP_TRY(hipEventSynchronize(f->ring_done[s]));
    }
    f->ring_pending[s] = false;
    f->ring_next = (s + 1) % kRingSlots;
    *slot = s;
    return CUSBF_OK;
}

// Host bytes into device memory, in stream order (does not wait): host
// threads fill one pinned buffer while the DMA engine reads the previous one.
int copy_in(cusbf_filter* f, void* d_dst, const void* h_src, uint64_t bytes, hipStream_t stream) {
    if (bytes < kRingDirect) {
        CUSBF_HIP_TRY(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, stream));
        return CUSBF_OK;
    }
    for (uint64_t at = 0; at < bytes; at += kRingBytes) {
        const uint64_t n = std::min(kRingBytes, bytes - at);
        uint32_t s = 0;
        const int rc = ring_slot(f, &s);
        if (rc != CUSBF_OK) return rc;
        parallel_copy(f->h_ring[s], static_cast<const char*>(h_src) + at, n);
        CUSBF_HIP_TRY(hipMemcpyAsync(static_cast<char*>(d_dst) + at, f->h_ring[s], n,
                                     hipMemcpyHostToDevice, stream));
        CUSBF_HIP_TRY(hipEventRecord(f->ring_done[s], stream));
        f->ring_pending[s] = true;
    }
    return CUSBF_OK;
}

// Device bytes (ready in stream order) into host memory, and waits: the DMA
// engine fills pinned buffers ahead while host threads empty the oldest one.
int copy_out(cusbf_filter* f, void* h_dst, const void* d_src, uint64_t bytes, hipStream_t stream) {
    if (bytes < kRingDirect) {
        CUSBF_HIP_TRY(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, stream));
        CUSBF_HIP_TRY(hipStreamSynchronize(stream));
        return CUSBF_OK;
    }
    const uint64_t chunks = (bytes + kRingBytes - 1) / kRingBytes;
    uint32_t slot_of[kRingSlots] = {};
    auto size_of = [&](uint64_t c) { return std::min(kRingBytes, bytes - c * kRingBytes); };
    auto issue = [&](uint64_t c) -> int {
        uint32_t s = 0;
        const int rc = ring_slot(f, &s);
        if (rc != CUSBF_OK) return rc;
        CUSBF_HIP_TRY(hipMemcpyAsync(f->h_ring[s], static_cast<const char*>(d_src) + c * kRingBytes,
                                     size_of(c), hipMemcpyDeviceToHost, stream));
        CUSBF_HIP_TRY(hipEventRecord(f->ring_done[s], stream));
        f->ring_pending[s] = true;
        slot_of[c % kRingSlots] = s;
        return CUSBF_OK;
    };
    for (uint64_t c = 0; c < std::min<uint64_t>(kRingSlots, chunks); ++c) {
        const int rc = issue(c);
        if (rc != CUSBF_OK) return rc;
    }
    for (uint64_t c = 0; c < chunks; ++c) {
        const uint32_t s = slot_of[c % kRingSlots];
        CUSBF_HIP_TRY(hipEventSynchronize(f->ring_done[s]));
        f->ring_pending[s] = false;
        parallel_copy(static_cast<char*>(h_dst) + c * kRingBytes, f->h_ring[s], size_of(c));
        if (c + kRingSlots < chunks) {  // the buffer just emptied is the next in the ring
            const int rc = issue(c + kRingSlots);
            if (rc != CUSBF_OK) return rc;
        }
    }
    return CUSBF_OK;
}

void release_ring(cusbf_filter* f) {
    for (uint32_t s = 0; s < kRingSlots; ++s) {
        if (f->ring_pending[s]) (void)hipEventSynchronize(f->ring_done[s]);
        if (f->h_ring[s]) (void)hipHostFree(f->h_ring[s]);
        if (f->ring_done[s]) (void)hipEventDestroy(f->ring_done[s]);
        f->h_ring[s] = nullptr;
        f->ring_done[s] = nullptr;
        f->ring_pending[s] = false;
    }
}

// fastx_pipeline_state.cuh:230-234 (release_all).
void release_staging(cusbf_filter* f) {
    if (f->d_stage) (void)hipFree(f->d_stage);
    if (f->d_result) (void)hipFree(f->d_result);
    f->d_stage = nullptr;
    f->d_result = nullptr;
    f->stage_cap = f->result_cap = 0;
}

sbf::FilterView view(const cusbf_filter* f) { return sbf::FilterView{f->d_words, f->num_shards, f->d_zero}; }

// Lane-strip kernels read byte input, and dense packed input for DNA.
bool strip_reads(const cusbf_filter* f, const sbf::SequenceInput& in) {
    return f->strip != nullptr && (!in.dense || f->strip->dense);
}

// --- sorted insert -------------------------------------------------------------

// Records per wave tile the sorted insert reserves: 2.5 / (w + 1) of the
// tile's windows for a minimizer span of w = k - m + 1 m-mers (a random
// sequence has 2 / (w + 1) shard runs per window), at most the tile's windows
// (a tile has at most one run per window).
uint32_t reserve_per_tile(const cusbf_filter* f) {
    if (f->reserve_per_tile) return f->reserve_per_tile;
    const double w = static_cast<double>(f->cfg.k - f->cfg.m + 1);
    const double r = std::ceil(f->strip->wave_windows * 2.5 / (w + 1.0));
    return static_cast<uint32_t>(std::min<double>(f->strip->wave_windows, std::max(1.0, r)));
}

// Cost model of AUTO, from MI355X measurements.
//
// Hot shards: a window's minimizer is the least hash among its w = k - m + 1
// m-mers, so the runs fall mostly on the ~sigma^m / w m-mers whose hashes are
// smallest (sigma = symbol count), at most every shard. Atomics on one shard
// serialise at ~1.2e8 per hot shard per second (fit to DNA (3, 2, 2, 4),
// (5, 4, 3, 4) and (31, 28, m, h) for m = 4..7 at 1 G windows).
double hot_shards(const cusbf_filter* f) {
    const double w = static_cast<double>(f->cfg.k - f->cfg.m + 1);
    const double mmers = std::pow(static_cast<double>(f->d.symbol_count), f->cfg.m);
    return std::max(1.0, std::min(static_cast<double>(f->num_shards), mmers / w));
}

// Shard runs of a random sequence: 2 / (w + 1) per window.
double expected_runs(const cusbf_filter* f, uint64_t windows) {
    return static_cast<double>(windows) * 2.0 / static_cast<double>(f->cfg.k - f->cfg.m + 2);
}

// Seconds the hot shards add to an atomic insert's four atomic ORs per run.
double contention_seconds(const cusbf_filter* f, double runs) {
    return runs * 4.0 / (1.2e8 * hot_shards(f));
}

// Atomic insert: four memory-side atomic ORs per run at ~25.6 G/s, plus the
// hot shards' serialisation.
double atomic_seconds(const cusbf_filter* f, double runs) {
    return runs * 4.0 / 25.6e9 + contention_seconds(f, runs);
}

// Sorted insert: ~45 ps of record traffic per run and, in each of its
// `chunks` sorted inserts, a load and store of every slice it touches, plus
// about a dozen launches.
double sorted_seconds(double runs, const sbf::SortedLayout& l, uint64_t chunks) {
    const double slices = static_cast<double>(l.num_slices);
    const double n = static_cast<double>(chunks);
    const double touched = slices * (1.0 - std::exp(-runs / n / slices));
    const double slice_bytes = 32.0 * static_cast<double>(1u << l.slice_bits);
    return runs * 45e-12 + n * (touched * slice_bytes * 2.0 / 4.0e12 + 80e-6);
}

bool sorted_pays(const cusbf_filter* f, uint64_t windows, const sbf::SortedLayout& l, uint64_t chunks) {
    const double runs = expected_runs(f, windows);
    return sorted_seconds(runs, l, chunks) < 0.8 * atomic_seconds(f, runs);
}

// Generic kernels. Where the atomic insert uses direct tiles both paths do the
// same per-window work and sorted_pays decides. Where it uses hashed tiles
// (~25 ps per window, atomics included), the sorted insert's direct tiles cost
// ~0.8 ps per m-mer and ~0.3 ps per s-mer hash of a window, so the sorted path
// pays only where hot shards slow the atomics by more than that.
bool generic_sorted_pays(const cusbf_filter* f, uint64_t windows, const sbf::SortedLayout& l) {
    if (!sbf::generic_hashed_tiles(f->kc)) return sorted_pays(f, windows, l, 1);
    const double n = static_cast<double>(windows);
    const double runs = expected_runs(f, windows);
    const double direct = 0.8e-12 * static_cast<double>(f->cfg.k - f->cfg.m + 1) +
                          0.3e-12 * static_cast<double>(f->cfg.k - f->cfg.s + 1) * f->cfg.hash_count;
    const double atomic_s = n * 25e-12 + contention_seconds(f, runs);
    const double sorted_s = n * direct + sorted_seconds(runs, l, 1);
    return sorted_s < 0.8 * atomic_s;
}

bool capturing(hipStream_t stream) {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(stream, &st) == hipSuccess && st != hipStreamCaptureStatusNone;
}

// Whether a sorted workspace of `bytes` is allowed: within the filter's limit,
// or, with no limit set, already held or within half of the free device memory.
bool workspace_fits(const cusbf_filter* f, uint64_t bytes) {
    if (f->sorted_limit) return bytes <= f->sorted_limit;
    if (bytes <= f->sorted_bytes) return true;
    size_t free_b = 0, total_b = 0;
    return hipMemGetInfo(&free_b, &total_b) == hipSuccess &&
           bytes <= (static_cast<uint64_t>(free_b) + f->sorted_bytes) / 2;
}

// Layout of a lane-strip insert over `tiles` wave tiles run by `blocks` blocks
// (4 regions each; `slack` scales the tiles a region reserves for). *ok is
// false when the records would need 2^31 slots or more.
sbf::SortedLayout region_layout(const cusbf_filter* f, uint64_t tiles, uint32_t blocks, double slack,
                                bool* ok) {
    const uint64_t regions = 4ull * blocks;
    const uint64_t per_region = static_cast<uint64_t>(
        std::ceil(static_cast<double>((tiles + regions - 1) / regions) * slack));
    const uint64_t cap = std::max<uint64_t>(1, per_region * reserve_per_tile(f));
    *ok = regions * cap < (1ull << 31);
    if (!*ok) return sbf::SortedLayout{};
    return sbf::sorted::sorted_layout(static_cast<uint32_t>(regions), static_cast<uint32_t>(cap),
                                      f->num_shards, f->slice_bits);
}

// Decides the path of a lane-strip insert over `tiles` wave tiles run by
// `blocks` blocks, as one sorted insert; *layout is filled when the sorted
// path is taken.
int choose_sorted(const cusbf_filter* f, uint64_t windows, uint64_t tiles, uint32_t blocks,
                  double slack, hipStream_t stream, sbf::SortedLayout* layout, bool* use) {
    *use = false;
    if (f->insert_path == CUSBF_INSERT_ATOMIC || windows == 0 || capturing(stream)) return CUSBF_OK;
    const bool forced = f->insert_path == CUSBF_INSERT_SORTED;
    bool ok = false;
    const sbf::SortedLayout l = region_layout(f, tiles, blocks, slack, &ok);
    if (!ok) {
        if (forced) return fail(CUSBF_ERR_RESOURCE, "insert is too large for one sorted insert");
        return CUSBF_OK;
    }
    if (!forced && !sorted_pays(f, windows, l, 1)) return CUSBF_OK;
    if (!workspace_fits(f, l.bytes)) {
        if (forced) return fail(CUSBF_ERR_RESOURCE, "sorted insert workspace does not fit in device memory");
        return CUSBF_OK;
    }
    *layout = l;
    *use = true;
    return CUSBF_OK;
}

// Blocks of a sorted plain insert: 16 per compute unit at most, so that every
// region reserves for a few hundred tiles; a smaller insert takes no more blocks
// than give each wave 16 tiles (four per compute unit at least), so that its
// regions still hold a whole partition wave's items.
uint32_t plain_sorted_blocks(const cusbf_filter* f, uint64_t tiles) {
    uint64_t blocks = (tiles + sbf::kStripWavesPerBlock - 1) / sbf::kStripWavesPerBlock;
    blocks = std::min<uint64_t>(blocks, 16ull * f->compute_units);
    blocks = std::min<uint64_t>(blocks, std::max<uint64_t>(4ull * f->compute_units, (tiles + 63) / 64));
    blocks = std::min<uint64_t>(blocks, sbf::kStripMaxBlocks);
    return static_cast<uint32_t>(std::max<uint64_t>(blocks, 1));
}

sbf::SortedLayout plain_layout(const cusbf_filter* f, uint64_t windows, bool* ok) {
    const uint64_t tiles = (windows + f->strip->wave_windows - 1) / f->strip->wave_windows;
    return region_layout(f, tiles, plain_sorted_blocks(f, tiles), 1.0, ok);
}

// A plain insert on the sorted path, as `chunks` sorted inserts over
// consecutive ranges of `chunk_windows` windows (the last may be shorter), so
// that one range's records fit the workspace limit and 2^31 slots.
struct SortedChunks {
    bool use = false;
    uint64_t chunk_windows = 0;
    uint64_t chunks = 0;
    uint64_t bytes = 0;  // workspace of the first (largest) range
};

// Range starts are multiples of whole wave tiles and of the symbols of a dense
// packed word, so a range of dense input starts on a word.
uint64_t chunk_align(const cusbf_filter* f) {
    const uint64_t a = f->strip->wave_windows;
    const uint64_t b = f->d.symbol_bits ? 64u / f->d.symbol_bits : 1u;
    return a / std::gcd(a, b) * b;
}

int plan_plain_sorted(const cusbf_filter* f, uint64_t windows, hipStream_t stream, SortedChunks* c) {
    *c = SortedChunks{};
    if (f->insert_path == CUSBF_INSERT_ATOMIC || windows == 0 || capturing(stream)) return CUSBF_OK;
    const bool forced = f->insert_path == CUSBF_INSERT_SORTED;
    const uint64_t align = chunk_align(f);
    size_t free_b = 0, total_b = 0;
    const double limit = f->sorted_limit ? static_cast<double>(f->sorted_limit)
                         : hipMemGetInfo(&free_b, &total_b) == hipSuccess
                             ? (static_cast<double>(free_b) + static_cast<double>(f->sorted_bytes)) / 2.0
                             : 0.0;
    uint64_t chunks = 1;
    for (;;) {
        uint64_t per = (windows + chunks - 1) / chunks;
        if (chunks > 1) per = (per + align - 1) / align * align;
        bool ok = false;
        const sbf::SortedLayout l = plain_layout(f, per, &ok);
        if (ok && workspace_fits(f, l.bytes)) {
            const uint64_t n = (windows + per - 1) / per;
            if (!forced && !sorted_pays(f, windows, l, n)) return CUSBF_OK;
            c->use = true;
            c->chunk_windows = per;
            c->chunks = n;
            c->bytes = l.bytes;
            return CUSBF_OK;
        }
        if (per <= align) break;
        // next try: scaled by how far this one is over its bounds
        double over = 2.0;
        if (ok && limit > 0.0) over = std::max(1.0, static_cast<double>(l.bytes) / limit);
        if (!ok) {
            bool fine = false;
            const sbf::SortedLayout one = plain_layout(f, align, &fine);
            const double slots_per_window = fine ? static_cast<double>(one.slots) / align : 1.0;
            over = std::max(over, slots_per_window * static_cast<double>(per) / 2147483648.0);
        }
        chunks = std::max<uint64_t>(chunks + 1,
                                    static_cast<uint64_t>(std::ceil(static_cast<double>(chunks) * over * 1.02)));
    }
    if (forced) return fail(CUSBF_ERR_RESOURCE, "sorted insert workspace does not fit its limit");
    return CUSBF_OK;
}

// Windows [first, first + n) of `in` as an input of their own: n + k - 1
// symbols from symbol `first` (a multiple of the symbols of a dense word).
sbf::SequenceInput window_range(const cusbf_filter* f, const sbf::SequenceInput& in, uint64_t first,
                                uint64_t n) {
    sbf::SequenceInput part = in;
    part.num_symbols = n + f->cfg.k - 1;
    if (in.dense) {
        const uint64_t per_word = 64u / f->d.symbol_bits;
        const uint64_t w0 = first / per_word;
        part.words = in.words + w0;
        part.num_words = std::min(in.num_words - w0, (part.num_symbols + per_word - 1) / per_word);
    } else {
        part.seq = in.seq + first * f->d.symbol_width;
    }
    return part;
}

// Makes the filter's workspace hold `bytes` and orders this stream after the
// previous sorted insert (which may have run on another stream).
int begin_sorted(cusbf_filter* f, uint64_t bytes, hipStream_t stream) {
    if (!f->sorted_done) CUSBF_HIP_TRY(hipEventCreateWithFlags(&f->sorted_done, hipEventDisableTiming));
    if (bytes > f->sorted_bytes) {
        if (f->d_sorted) {
            CUSBF_HIP_TRY(hipEventSynchronize(f->sorted_done));
            CUSBF_HIP_TRY(hipFree(f->d_sorted));
        }
        f->d_sorted = nullptr;
        f->sorted_bytes = 0;
        CUSBF_HIP_TRY(hipMalloc(&f->d_sorted, bytes));
        f->sorted_bytes = bytes;
    } else if (f->sorted_used) {
        CUSBF_HIP_TRY(hipStreamWaitEvent(stream, f->sorted_done, 0));
    }
    CUSBF_HIP_TRY(sbf::sorted::prepare());
    return CUSBF_OK;
}

int end_sorted(cusbf_filter* f, hipStream_t stream) {
    CUSBF_HIP_TRY(hipEventRecord(f->sorted_done, stream));
    f->sorted_used = true;
    return CUSBF_OK;
}

// Zeroes the region counts and the overflow flag (taking the filter's zero
// flag), then runs `emit(sink)` (the insert kernel writing records) and the
// partition and apply kernels, in stream order.
template <class Emit>
hipError_t enqueue_sorted(const cusbf_filter* f, const sbf::SortedLayout& l, void* base,
                          hipStream_t stream, Emit&& emit) {
    const sbf::RecordSink sink = sbf::sorted::record_sink(l, base);
    hipError_t e = sbf::sorted::sorted_begin(l, base, view(f), stream);
    if (e == hipSuccess) e = emit(sink);
    if (e == hipSuccess) e = sbf::sorted::sorted_apply(l, base, view(f), stream);
    return e;
}

// Inserts into one filter run one after another whatever their streams: the
// sorted apply rewrites whole slices, so an insert overlapping it could lose
// updates. Inside a capture the graph's order is its caller's.
int order_insert(cusbf_filter* f, hipStream_t stream) {
    if (capturing(stream)) return CUSBF_OK;
    if (!f->insert_done) {
        CUSBF_HIP_TRY(hipEventCreateWithFlags(&f->insert_done, hipEventDisableTiming));
    } else if (f->insert_recorded) {
        CUSBF_HIP_TRY(hipStreamWaitEvent(stream, f->insert_done, 0));
    }
    return CUSBF_OK;
}

int mark_insert(cusbf_filter* f, hipStream_t stream) {
    if (capturing(stream) || !f->insert_done) return CUSBF_OK;
    CUSBF_HIP_TRY(hipEventRecord(f->insert_done, stream));
    f->insert_recorded = true;
    return CUSBF_OK;
}

int launch_insert_ordered(cusbf_filter* f, const sbf::SequenceInput& in, hipStream_t stream);

int launch_insert(cusbf_filter* f, const sbf::SequenceInput& in, hipStream_t stream) {
    int rc = order_insert(f, stream);
    if (rc == CUSBF_OK) rc = launch_insert_ordered(f, in, stream);
    if (rc == CUSBF_OK) rc = mark_insert(f, stream);
    return rc;
}

// A generic insert on the sorted path: one sorted insert (no window ranges) when
// forced, or under AUTO when the cost model prefers it and the atomic insert
// would use direct tiles too, and the workspace fits; else the atomic path. The
// generic kernels' waves take 64 windows of a block tile at a time, so a region
// reserves for its wave's share of 64-window tiles, each cut into runs at its
// edges as well.
// Decides the path of a generic-kernel insert over `windows` windows (plain or
// batched: the records of a batch lie back to back), as one sorted insert;
// *layout is filled when the sorted path is taken.
int choose_generic_sorted(const cusbf_filter* f, uint64_t windows, hipStream_t stream,
                          sbf::SortedLayout* layout, bool* use_out) {
    bool use = false;
    sbf::SortedLayout l;
    if (windows != 0 && f->insert_path != CUSBF_INSERT_ATOMIC && !capturing(stream)) {
        const bool forced = f->insert_path == CUSBF_INSERT_SORTED;
        const uint64_t tiles = (windows + sbf::kGenericWaveWindows - 1) / sbf::kGenericWaveWindows;
        uint64_t blocks = (tiles + 3) / 4;
        blocks = std::min<uint64_t>(blocks, 16ull * f->compute_units);
        blocks = std::min<uint64_t>(blocks, std::max<uint64_t>(4ull * f->compute_units, (tiles + 63) / 64));
        blocks = std::max<uint64_t>(1, std::min<uint64_t>(blocks, sbf::kStripMaxBlocks));
        const uint64_t regions = 4 * blocks;
        const double w = static_cast<double>(f->cfg.k - f->cfg.m + 1);
        const uint64_t reserve = std::min<uint64_t>(
            sbf::kGenericWaveWindows,
            static_cast<uint64_t>(std::ceil(sbf::kGenericWaveWindows * 2.5 / (w + 1.0))) + 2);
        const uint64_t cap = std::max<uint64_t>(1, (tiles + regions - 1) / regions * reserve);
        if (regions * cap < (1ull << 31)) {
            l = sbf::sorted::sorted_layout(static_cast<uint32_t>(regions), static_cast<uint32_t>(cap),
                                           f->num_shards, f->slice_bits);
            use = (forced || generic_sorted_pays(f, windows, l)) && workspace_fits(f, l.bytes);
        }
        if (forced && !use) {
            return fail(CUSBF_ERR_RESOURCE, "sorted insert workspace does not fit its limit");
        }
    }
    *layout = l;
    *use_out = use;
    return CUSBF_OK;
}

int launch_generic_insert(cusbf_filter* f, const sbf::SequenceInput& in, hipStream_t stream) {
    const uint64_t windows = in.num_symbols < f->cfg.k ? 0 : in.num_symbols - f->cfg.k + 1;
    bool use = false;
    sbf::SortedLayout l;
    const int chosen = choose_generic_sorted(f, windows, stream, &l, &use);
    if (chosen != CUSBF_OK) return chosen;
    if (!use) {
        const hipError_t e = sbf::generic_insert(f->kc, in, view(f), nullptr, stream);
        if (e != hipSuccess) return fail_hip(e, __FILE__, __LINE__);
        return CUSBF_OK;
    }
    int rc = begin_sorted(f, l.bytes, stream);
    if (rc != CUSBF_OK) return rc;
    const hipError_t e = enqueue_sorted(f, l, f->d_sorted, stream, [&](const sbf::RecordSink& sink) {
        return sbf::generic_insert(f->kc, in, view(f), &sink, stream);
    });
    if (e != hipSuccess) return fail_hip(e, __FILE__, __LINE__);
    f->last_path = CUSBF_INSERT_SORTED;
    f->last_chunks = 1;
    return end_sorted(f, stream);
}

int launch_insert_ordered(cusbf_filter* f, const sbf::SequenceInput& in, hipStream_t stream) {
    f->last_path = CUSBF_INSERT_ATOMIC;
    f->last_chunks = 0;
    if (!strip_reads(f, in)) return launch_generic_insert(f, in, stream);
    const uint64_t windows = in.num_symbols < f->cfg.k ? 0 : in.num_symbols - f->cfg.k + 1;
    SortedChunks c;
    int rc = plan_plain_sorted(f, windows, stream, &c);
    if (rc != CUSBF_OK) return rc;
    if (!c.use) {
        const hipError_t e = f->strip->insert(f->kc, in, view(f), nullptr, stream);
        if (e != hipSuccess) return fail_hip(e, __FILE__, __LINE__);
        return CUSBF_OK;
    }
    rc = begin_sorted(f, c.bytes, stream);
    if (rc != CUSBF_OK) return rc;
    // Ranges one after another in stream order: each one's first kernel finds
    // the zero-words flag reset by the previous one's insert kernel, so only
    // the first range may skip loading its slices.
    for (uint64_t first = 0; first < windows; first += c.chunk_windows) {
        const uint64_t n = std::min(c.chunk_windows, windows - first);
        const sbf::SequenceInput part = c.chunks == 1 ? in : window_range(f, in, first, n);
        bool ok = false;
        const sbf::SortedLayout l = plain_layout(f, n, &ok);  // no larger than the first range's
        if (!ok || l.bytes > c.bytes) return fail(CUSBF_ERR_RESOURCE, "sorted insert range exceeds its workspace");
        const hipError_t e = enqueue_sorted(f, l, f->d_sorted, stream, [&](const sbf::RecordSink& sink) {
            return f->strip->insert(f->kc, part, view(f), &sink, stream);
        });
        if (e != hipSuccess) return fail_hip(e, __FILE__, __LINE__);
    }
    f->last_path = CUSBF_INSERT_SORTED;
    f->last_chunks = static_cast<uint32_t>(std::min<uint64_t>(c.chunks, UINT32_MAX));
    return end_sorted(f, stream);
}

int launch_contains(const cusbf_filter* f, const sbf::SequenceInput& in, uint8_t* out,
                    hipStream_t stream) {
    const hipError_t e = strip_reads(f, in) ? f->strip->contains(f->kc, in, view(f), out, stream)
                                            : sbf::generic_contains(f->kc, in, view(f), out, stream);
    if (e != hipSuccess) return fail_hip(e, __FILE__, __LINE__);
    return CUSBF_OK;
}

sbf::SequenceInput byte_input(const cusbf_filter* f, const char* seq, uint64_t bytes) {
    sbf::SequenceInput in;
    in.seq = seq;
    in.num_symbols = symbols_of(f, bytes);
    return in;
}

// filter.cuh:429-457 (validate_device_record_batch) / normalized_record_batch.hpp:123-147.
int validate_ranges(const cusbf_filter* f, uint64_t bytes, const cusbf_record_range* ranges,
                    uint64_t n, const char* exceeds_message) {
    if (n != 0 && ranges == nullptr) return fail(CUSBF_ERR_INVALID_ARGUMENT, "record ranges pointer is null");
    uint64_t next_offset = 0;
    for (uint64_t i = 0; i < n; ++i) {
        const cusbf_record_range& r = ranges[i];
        if (r.sequence_offset < next_offset) {
            return fail(CUSBF_ERR_INVALID_ARGUMENT, "record batch ranges must be ordered and non-overlapping");
        }
        if (r.sequence_offset > bytes || r.sequence_bytes > bytes - r.sequence_offset) {
            return fail(CUSBF_ERR_INVALID_ARGUMENT, exceeds_message);
        }
        if (r.sequence_offset % f->d.symbol_width != 0 || r.sequence_bytes % f->d.symbol_width != 0) {
            return fail(CUSBF_ERR_INVALID_ARGUMENT,
                        "record batch ranges must align to the configured alphabet symbol width");
        }
        next_offset = r.sequence_offset + r.sequence_bytes;
    }
    return CUSBF_OK;
}

uint64_t record_kmers(const cusbf_filter* f, uint64_t bytes) { return windows_of(f, bytes); }

// normalized_record_batch.hpp:87-110 (valid_record_kmer_count): windows of
// the record with no invalid symbol, in one pass tracking the last invalid symbol.
uint64_t valid_record_kmers(const cusbf_filter* f, const char* seq, uint64_t bytes) {
    const uint64_t sym = symbols_of(f, bytes);
    const uint64_t k = f->cfg.k;
    if (sym < k) return 0;
    const uint8_t inv = f->d.invalid_symbol;
    const uint32_t w = f->d.symbol_width;
    uint64_t valid = 0;
    int64_t last_invalid = -1;
    if (f->cfg.alphabet == CUSBF_ALPHABET_DNA) {
        // Vectorisable pre-check: records without invalid bytes have every window valid.
        const uint8_t* p = reinterpret_cast<const uint8_t*>(seq);
        uint8_t any_bad = 0;
        for (uint64_t i = 0; i < sym; ++i) {
            const uint8_t u = p[i] & 0xDFu;
            any_bad |= static_cast<uint8_t>((u != 'A') & (u != 'C') & (u != 'G') & (u != 'T'));
        }
        if (!any_bad) return sym - k + 1;
    }
    if (w == 1) {
        bool bad[256];
        for (int b = 0; b < 256; ++b) {
            const char c = static_cast<char>(b);
            bad[b] = sbf::encode_symbol(f->kc, &c) == inv;
        }
        const uint8_t* p = reinterpret_cast<const uint8_t*>(seq);
        for (uint64_t i = 0; i < sym; ++i) {
            if (bad[p[i]]) last_invalid = static_cast<int64_t>(i);
            if (i + 1 >= k) valid += last_invalid < static_cast<int64_t>(i + 1 - k);
        }
        return valid;
    }
    for (uint64_t i = 0; i < sym; ++i) {
        if (sbf::encode_symbol(f->kc, seq + i * w) == inv) last_invalid = static_cast<int64_t>(i);
        if (i + 1 >= k) valid += last_invalid < static_cast<int64_t>(i + 1 - k);
    }
    return valid;
}

// normalized_record_batch.hpp:208-276: records joined by the alphabet
// separator, padded to the symbol width.
// A host record batch as the buffer the reference normalises it into
// (normalized_record_batch.hpp:208-276): records joined by the separator, each
// boundary padded to the symbol width. The bytes are produced chunk by chunk
// straight into the pinned ring by the host pool, never as a buffer of their own.
struct NormPlan {
    std::vector<uint64_t> at;  // per record: first byte in the normalised buffer
    uint64_t total = 0;        // bytes of the normalised buffer
    uint64_t bases = 0;        // record bytes
};

void norm_plan(const cusbf_filter* f, const cusbf_record_range* r, uint64_t n, NormPlan* p) {
    const uint32_t w = f->d.symbol_width;
    p->at.resize(n);
    p->total = 0;
    p->bases = 0;
    for (uint64_t i = 0; i < n; ++i) {
        if (p->total) p->total += (p->total % w ? w - p->total % w : 0) + w;
        p->at[i] = p->total;
        p->total += r[i].sequence_bytes;
        p->bases += r[i].sequence_bytes;
    }
}

// Bytes [a, a + len) of the normalised buffer into dst.
void norm_fill(const cusbf_filter* f, const char* seq, const cusbf_record_range* r, const NormPlan& p,
               uint64_t a, uint64_t len, char* dst) {
    const char sep = static_cast<char>(f->d.separator);
    const uint64_t n = p.at.size();
    // the first record that ends past a
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) / 2;
        if (p.at[mid] + r[mid].sequence_bytes <= a) lo = mid + 1; else hi = mid;
    }
    uint64_t i = lo, pos = a;
    const uint64_t end = a + len;
    while (pos < end) {
        if (i < n && pos >= p.at[i]) {
            const uint64_t rec_end = p.at[i] + r[i].sequence_bytes;
            const uint64_t m = std::min(end, rec_end) - pos;
            std::memcpy(dst + (pos - a), seq + r[i].sequence_offset + (pos - p.at[i]), m);
            pos += m;
            if (pos == rec_end) ++i;
        } else {
            const uint64_t stop = i < n ? std::min(end, p.at[i]) : end;
            std::memset(dst + (pos - a), sep, stop - pos);
            pos = stop;
        }
    }
}

// The normalised batch into device memory, in stream order (does not wait).
int copy_in_normalized(cusbf_filter* f, void* d_dst, const char* seq, const cusbf_record_range* r,
                       const NormPlan& p, hipStream_t stream) {
    constexpr uint64_t kPart = 4ull << 20;
    for (uint64_t at = 0; at < p.total; at += kRingBytes) {
        const uint64_t n = std::min(kRingBytes, p.total - at);
        uint32_t s = 0;
        const int rc = ring_slot(f, &s);
        if (rc != CUSBF_OK) return rc;
        char* buf = f->h_ring[s];
        cusbf_internal::host_pool().run((n + kPart - 1) / kPart, [&](size_t i) {
            const uint64_t off = i * kPart;
            norm_fill(f, seq, r, p, at + off, std::min(kPart, n - off), buf + off);
        });
        CUSBF_HIP_TRY(hipMemcpyAsync(static_cast<char*>(d_dst) + at, buf, n, hipMemcpyHostToDevice, stream));
        CUSBF_HIP_TRY(hipEventRecord(f->ring_done[s], stream));
        f->ring_pending[s] = true;
    }
    return CUSBF_OK;
}

// Per record of a staged normalised batch, on the device: windows with no
// invalid symbol (normalized_record_batch.hpp:87-110) and, with d_hits, answers
// equal to 1; summed into *valid and *positive. Waits for the stream.
int record_counts(cusbf_filter* f, const cusbf_record_range* r, const NormPlan& p, const uint8_t* d_hits,
                  hipStream_t stream, uint64_t* valid, uint64_t* positive) {
    const uint64_t n = p.at.size();
    *valid = 0;
    if (positive) *positive = 0;
    if (n == 0 || p.total == 0) return CUSBF_OK;
    const uint32_t w = f->d.symbol_width;
    std::vector<uint64_t> req(2 * n), counts(2 * n);
    for (uint64_t i = 0; i < n; ++i) {
        req[2 * i] = p.at[i] / w;
        req[2 * i + 1] = record_kmers(f, r[i].sequence_bytes);
    }
    if (f->rec_cap < 4 * n) {
        if (f->d_rec) CUSBF_HIP_TRY(hipFree(f->d_rec));
        f->d_rec = nullptr;
        f->rec_cap = 0;
        CUSBF_HIP_TRY(hipMalloc(&f->d_rec, 4 * n * sizeof(uint64_t)));
        f->rec_cap = 4 * n;
    }
    uint64_t* d_req = f->d_rec;
    uint64_t* d_counts = f->d_rec + 2 * n;
    int rc = copy_in(f, d_req, req.data(), 2 * n * sizeof(uint64_t), stream);
    if (rc == CUSBF_OK) {
        rc = cusbf_internal::launch_request_counts(f, f->d_stage, p.total, d_hits, d_req, n, d_counts, stream);
    }
    if (rc == CUSBF_OK) rc = copy_out(f, counts.data(), d_counts, 2 * n * sizeof(uint64_t), stream);
    if (rc != CUSBF_OK) return rc;
    for (uint64_t i = 0; i < n; ++i) {
        *valid += counts[2 * i];
        if (positive) *positive += counts[2 * i + 1];
    }
    return CUSBF_OK;
}

// --- batched record form ------------------------------------------------------

// Launch shape used until a tuner sets the table (the generic kernels ignore
// it): the shapes the tuner picked on MI355X for 1M log-normal records, DNA
// k=31 and protein k=10 -- atomic insert: three waves per SIMD, two
// persistent blocks per compute unit, strided tiles; query: 32 blocks per
// compute unit, strided tiles, four waves per SIMD where that variant exists;
// sorted insert: as many blocks as stay resident, strided tiles.
void default_tuning(cusbf_filter* f) {
    const bool wide = f->strip && f->strip->batched_insert[1] != nullptr;
    for (auto& bucket : f->tuning) {
        bucket[CUSBF_BATCH_INSERT] = cusbf_batch_tuning{3u, 2u, 0u, 0u};
        bucket[CUSBF_BATCH_CONTAINS] = cusbf_batch_tuning{wide ? 4u : 3u, 32u, 0u, 0u};
        // sorted path: the insert kernel only computes and stores, so fill the SIMDs
        bucket[CUSBF_BATCH_INSERT_SORTED] = cusbf_batch_tuning{wide ? 4u : 3u, wide ? 4u : 3u, 0u, 0u};
    }
}

int check_tuning(const cusbf_filter* f, const cusbf_batch_tuning& t) {
    if (t.waves_per_simd != 3 && t.waves_per_simd != 4) {
        return fail(CUSBF_ERR_INVALID_ARGUMENT, "waves_per_simd must be 3 or 4");
    }
    if (t.waves_per_simd == 4 && f->strip && f->strip->batched_insert[1] == nullptr) {
        return fail(CUSBF_ERR_INVALID_ARGUMENT, "this configuration's batched kernels are built for 3 waves per SIMD");
    }
    if (t.blocks_per_cu < 1 || t.blocks_per_cu > 32) {
        return fail(CUSBF_ERR_INVALID_ARGUMENT, "blocks_per_cu must be in [1, 32]");
    }
    if (t.claim > 64) return fail(CUSBF_ERR_INVALID_ARGUMENT, "claim must be in [0, 64]");
    return CUSBF_OK;
}

bool batched_strip(const cusbf_filter* f) { return f->strip && f->strip->batched_insert[0]; }

uint64_t batch_tiles(const cusbf_filter* f, uint64_t num_symbols) {
    if (!batched_strip(f)) return 0;
    const uint64_t windows = num_symbols < f->cfg.k ? 0 : num_symbols - f->cfg.k + 1;
    return (windows + f->strip->wave_windows - 1) / f->strip->wave_windows;
}

int validate_batched(const cusbf_filter* f, const char* d_symbols, uint64_t num_symbols,
                     const uint64_t* d_offsets, uint64_t num_records) {
    if (!f) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null filter");
    if (!d_offsets) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null device offsets");
    if (num_symbols && !d_symbols) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null device sequence");
    if (num_records >= 0xFFFFFFFFull) {
        return fail(CUSBF_ERR_INVALID_ARGUMENT, "a batched call holds fewer than 2^32 - 1 records");
    }
    if (batch_tiles(f, num_symbols) >= (1ull << 31)) {
        return fail(CUSBF_ERR_INVALID_ARGUMENT, "batched sequence is too long for one launch");
    }
    return CUSBF_OK;
}

uint64_t batched_windows(const cusbf_filter* f, uint64_t num_symbols) {
    return num_symbols < f->cfg.k ? 0 : num_symbols - f->cfg.k + 1;
}

int ensure_starts(cusbf_filter* f, uint64_t num_symbols) {
    const uint64_t words = sbf::start_words(num_symbols);
    if (words <= f->starts_cap) return CUSBF_OK;
    if (f->d_starts) {
        CUSBF_HIP_TRY(hipDeviceSynchronize());  // an earlier batched launch may still read it
        CUSBF_HIP_TRY(hipFree(f->d_starts));
    }
    f->d_starts = nullptr;
    f->starts_cap = 0;
    CUSBF_HIP_TRY(hipMalloc(&f->d_starts, words * sizeof(uint64_t)));
    f->starts_cap = words;
    return CUSBF_OK;
}

const cusbf_batch_tuning& tuning_for(const cusbf_filter* f, uint32_t op, uint64_t num_symbols,
                                     uint64_t num_records) {
    return f->tuning[cusbf_batch_bucket(num_symbols, num_records)][op];
}

// Reset, prepass and kernel of one batched call, in stream order; allocates
// nothing, so it can be captured into a graph. `sorted` (insert only, or
// null): the layout of the sorted insert's workspace at `sorted_base`.
int enqueue_batched(const cusbf_filter* f, uint32_t op, const char* d_symbols,
                    uint64_t num_symbols, const uint64_t* d_offsets, uint64_t num_records,
                    uint32_t* d_state, uint64_t* d_starts, const cusbf_batch_tuning& t,
                    uint8_t* d_output, hipStream_t stream,
                    const sbf::SortedLayout* sorted = nullptr, void* sorted_base = nullptr) {
    sbf::SequenceInput in = byte_input(f, d_symbols, num_symbols * f->d.symbol_width);
    in.offsets = d_offsets;
    in.num_records = num_records;
    in.starts = d_starts;
    in.state = d_state;
    CUSBF_HIP_TRY(sbf::batch_prepass(in, d_starts, d_state, stream));
    hipError_t e;
    if (batched_strip(f)) {
        const int variant = t.waves_per_simd == 4 && f->strip->batched_insert[1] ? 1 : 0;
        const sbf::BatchLaunch shape{t.blocks_per_cu, f->compute_units, t.claim};
        if (op == CUSBF_BATCH_INSERT && sorted) {
            e = enqueue_sorted(f, *sorted, sorted_base, stream, [&](const sbf::RecordSink& sink) {
                return f->strip->batched_insert[variant](f->kc, in, view(f), shape, &sink, stream);
            });
        } else {
            e = op == CUSBF_BATCH_INSERT
                    ? f->strip->batched_insert[variant](f->kc, in, view(f), shape, nullptr, stream)
                    : f->strip->batched_contains[variant](f->kc, in, view(f), shape, d_output, stream);
        }
    } else if (op == CUSBF_BATCH_INSERT && sorted) {
        e = enqueue_sorted(f, *sorted, sorted_base, stream, [&](const sbf::RecordSink& sink) {
            return sbf::generic_insert(f->kc, in, view(f), &sink, stream);
        });
    } else {
        e = op == CUSBF_BATCH_INSERT ? sbf::generic_insert(f->kc, in, view(f), nullptr, stream)
                                     : sbf::generic_contains(f->kc, in, view(f), d_output, stream);
    }
    if (e != hipSuccess) return fail_hip(e, __FILE__, __LINE__);
    return CUSBF_OK;
}

// Sorted path of a batched insert: the regions are the persistent grid's
// waves; a work queue (claim > 0) hands a wave a varying share of the tiles,
// so its regions reserve for half as many tiles again.
int choose_batched_sorted(const cusbf_filter* f, uint64_t num_symbols, const cusbf_batch_tuning& t,
                          hipStream_t stream, sbf::SortedLayout* layout, bool* use) {
    const uint64_t tiles = batch_tiles(f, num_symbols);
    const sbf::BatchLaunch shape{t.blocks_per_cu, f->compute_units, t.claim};
    const uint32_t blocks = sbf::batched_grid(shape, static_cast<uint32_t>(tiles));
    return choose_sorted(f, batched_windows(f, num_symbols), tiles, blocks, t.claim ? 1.5 : 1.0,
                         stream, layout, use);
}

int read_batch_flag(const uint32_t* d_state, hipStream_t stream) {
    CUSBF_HIP_TRY(hipStreamSynchronize(stream));
    uint32_t flag = 0;
    CUSBF_HIP_TRY(hipMemcpy(&flag, d_state + 1, sizeof(flag), hipMemcpyDeviceToHost));
    if (flag != 0) {
        return fail(CUSBF_ERR_INVALID_ARGUMENT,
                    "batched offsets must be non-decreasing from 0 to num_symbols");
    }
    return CUSBF_OK;
}

}  // namespace

namespace cusbf_internal {

WorkerPool& host_pool() {
    // never destroyed: its threads wait for work until the process ends
    static WorkerPool* pool = new WorkerPool(std::clamp(std::thread::hardware_concurrency() / 2, 2u, 8u));
    return *pool;
}

int set_error(int status, const char* message) { return fail(status, message); }
int set_hip_error(hipError_t err, const char* file, int line) { return fail_hip(err, file, line); }
int device(const cusbf_filter* f) { return f->device; }
uint32_t symbol_width(const cusbf_filter* f) { return f->d.symbol_width; }
uint8_t separator(const cusbf_filter* f) { return f->d.separator; }
uint64_t window_count(const cusbf_filter* f, uint64_t bytes) { return windows_of(f, bytes); }
uint64_t valid_kmers(const cusbf_filter* f, const char* seq, uint64_t bytes) {
    return valid_record_kmers(f, seq, bytes);
}
int validate_host_ranges(const cusbf_filter* f, uint64_t bytes, const cusbf_record_range* ranges,
                         uint64_t n) {
    return validate_ranges(f, bytes, ranges, n, "record batch range exceeds the source sequence");
}
int launch_contains_bytes(const cusbf_filter* f, const char* d_seq, uint64_t bytes, uint8_t* d_out,
                          hipStream_t stream) {
    return launch_contains(f, byte_input(f, d_seq, bytes), d_out, stream);
}
uint32_t kmer_length(const cusbf_filter* f) { return f->cfg.k; }
const cusbf_config& config(const cusbf_filter* f) { return f->cfg; }
uint64_t fastx_chunk_bytes(const cusbf_filter* f) { return f->fastx_chunk_bytes; }
int set_fastx_error(const char* source, uint32_t line, uint32_t column, const char* detail) {
    char buf[64];
    std::snprintf(buf, sizeof(buf), ":%u:%u: ", line, column);
    const int st = fail(CUSBF_ERR_FASTX_PARSE, std::string(source) + buf + detail);
    g_last_file = source;
    g_last_line = line;
    g_last_column = column;
    return st;
}
int contains_batched(const cusbf_filter* f_const, const char* d_symbols, uint64_t num_symbols,
                     const uint64_t* d_offsets, uint64_t num_records, uint8_t* d_output,
                     uint64_t d_output_len, uint32_t* d_state, uint64_t* d_starts,
                     hipStream_t stream) {
    int rc = validate_batched(f_const, d_symbols, num_symbols, d_offsets, num_records);
    if (rc != CUSBF_OK) return rc;
    // the batch scratch is internal state, like the reference's mutable staging
    cusbf_filter* f = const_cast<cusbf_filter*>(f_const);
    f->batch_check_pending = false;
    const uint64_t windows = batched_windows(f, num_symbols);
    if (windows == 0) return CUSBF_OK;
    if (!d_output) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null device output");
    if (d_output_len < windows) return fail(CUSBF_ERR_INVALID_ARGUMENT, "batched query output span is too small");
    if (d_state == nullptr) {  // the filter's own scratch, grown on demand
        rc = ensure_starts(f, num_symbols);
        if (rc != CUSBF_OK) return rc;
        d_state = f->d_batch_state;
        d_starts = f->d_starts;
    }
    rc = enqueue_batched(f, CUSBF_BATCH_CONTAINS, d_symbols, num_symbols, d_offsets, num_records,
                         d_state, d_starts,
                         tuning_for(f, CUSBF_BATCH_CONTAINS, num_symbols, num_records), d_output,
                         stream);
    f->batch_check_pending = rc == CUSBF_OK;
    f->batch_check_state = d_state;
    return rc;
}

int launch_request_counts(const cusbf_filter* f, const char* d_seq, uint64_t bytes,
                          const uint8_t* d_hits, const uint64_t* d_req, uint64_t n_req,
                          uint64_t* d_counts, hipStream_t stream) {
    CUSBF_HIP_TRY(sbf::fill_u64(d_counts, 0, 2 * n_req, stream));
    CUSBF_HIP_TRY(sbf::request_counts(f->kc, d_seq, symbols_of(f, bytes), d_hits, d_req, n_req,
                                      d_counts, stream));
    return CUSBF_OK;
}

}  // namespace cusbf_internal

extern "C" {

int cusbf_abi_version(void) { return CUSBF_ABI_VERSION; }

int cusbf_last_error_location(char* file, size_t file_len, uint32_t* line, uint32_t* column) {
    if (file && file_len) {
        const size_t n = std::min(file_len - 1, g_last_file.size());
        std::memcpy(file, g_last_file.data(), n);
        file[n] = '\0';
    }
    if (line) *line = g_last_line;
    if (column) *column = g_last_column;
    return g_last_status;
}

int cusbf_fastx_set_chunk_bytes(cusbf_filter* f, uint64_t bytes) {
    if (!f) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null filter");
    f->fastx_chunk_bytes = bytes;
    return CUSBF_OK;
}

int cusbf_last_error(char* buf, size_t buf_len) {
    if (buf && buf_len > 0) {
        const size_t n = std::min(buf_len - 1, g_last_message.size());
        std::memcpy(buf, g_last_message.data(), n);
        buf[n] = '\0';
    }
    return g_last_status;
}

int cusbf_config_init(cusbf_config* cfg, uint32_t k, uint32_t s, uint32_t m, uint32_t hash_count,
                      uint32_t alphabet) {
    if (!cfg) return fail(CUSBF_ERR_INVALID_ARGUMENT, "config pointer is null");
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->k = k;
    cfg->s = s;
    cfg->m = m;
    cfg->hash_count = hash_count;
    cfg->alphabet = alphabet;
    if (alphabet == CUSBF_ALPHABET_CUSTOM) return fail(CUSBF_ERR_INVALID_ARGUMENT, "custom alphabets need a table");
    Derived d;
    int rc = derive(cfg, &d);
    if (rc != CUSBF_OK) return rc;
    cfg->symbol_count = d.symbol_count;
    cfg->invalid_symbol = d.invalid_symbol;
    cfg->separator = d.separator;
    return validate(cfg, d);
}

int cusbf_config_symbol_width(const cusbf_config* cfg, uint32_t* out) {
    Derived d;
    const int rc = derive(cfg, &d);
    if (rc == CUSBF_OK) *out = d.symbol_width;
    return rc;
}

int cusbf_config_symbol_bits(const cusbf_config* cfg, uint32_t* out) {
    Derived d;
    const int rc = derive(cfg, &d);
    if (rc == CUSBF_OK) *out = d.symbol_bits;
    return rc;
}

int cusbf_create(const cusbf_config* cfg, uint64_t requested_filter_bits, cusbf_filter** out) {
    if (!cfg || !out) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    Derived d;
    int rc = derive(cfg, &d);
    if (rc != CUSBF_OK) return rc;
    rc = validate(cfg, d);
    if (rc != CUSBF_OK) return rc;

    cusbf_filter* f = new cusbf_filter();
    f->cfg = *cfg;
    f->d = d;
    f->kc.k = cfg->k;
    f->kc.s = cfg->s;
    f->kc.m = cfg->m;
    f->kc.h = cfg->hash_count;
    f->kc.alphabet = cfg->alphabet;
    f->kc.symbol_width = d.symbol_width;
    f->kc.symbol_bits = d.symbol_bits;
    f->kc.invalid_symbol = d.invalid_symbol;
    std::memcpy(f->kc.lut, cfg->lut, 256);
    // test hooks: CUSBF_FORCE_GENERIC=1 skips the lane-strip kernels;
    // CUSBF_FORCE_WIDE=1 takes the runtime-span lane strips where they apply;
    // CUSBF_GENERIC_FORM=direct|hashed fixes the generic kernels' form
    const char* form = std::getenv("CUSBF_GENERIC_FORM");
    f->kc.form = form && std::strcmp(form, "direct") == 0   ? sbf::kFormDirect
                 : form && std::strcmp(form, "hashed") == 0 ? sbf::kFormHashed
                                                            : sbf::kFormAuto;
    const char* force_generic = std::getenv("CUSBF_FORCE_GENERIC");
    const char* force_wide = std::getenv("CUSBF_FORCE_WIDE");
    if (force_generic && force_generic[0] == '1') {
        f->strip = nullptr;
    } else if (force_wide && force_wide[0] == '1') {
        f->strip = sbf::find_wide_kernels(cfg->k, cfg->s, cfg->m, cfg->hash_count, cfg->alphabet);
    } else {
        f->strip = sbf::find_strip_kernels(cfg->k, cfg->s, cfg->m, cfg->hash_count, cfg->alphabet);
    }
    default_tuning(f);
    // filter.cuh:108-117
    uint64_t shards = requested_filter_bits / 256 + (requested_filter_bits % 256 != 0);
    if (shards < 1) shards = 1;
    uint64_t p = 1;
    while (p < shards) p <<= 1;
    f->num_shards = p;
    f->filter_bits = p * 256;
    hipError_t e = hipGetDevice(&f->device);
    if (e == hipSuccess) e = hipMalloc(&f->d_words, f->num_shards * 32);
    if (e == hipSuccess) e = hipMalloc(&f->d_scalar, 4 * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMalloc(&f->d_batch_state, 64);
    if (e == hipSuccess) e = hipMalloc(&f->d_zero, 256);
    if (e == hipSuccess && cusbf_internal::noalloc_setup(f->device) != CUSBF_OK) {
        const int st = g_last_status;
        cusbf_destroy(f);
        return st;
    }
    if (e == hipSuccess) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, f->device) == hipSuccess &&
            cus > 0) {
            f->compute_units = static_cast<uint32_t>(cus);
        }
    }
    if (e != hipSuccess) {
        const int st = fail_hip(e, __FILE__, __LINE__);
        cusbf_destroy(f);
        return st;
    }
    rc = cusbf_clear(f, nullptr);
    if (rc != CUSBF_OK) {
        cusbf_destroy(f);
        return rc;
    }
    *out = f;
    return CUSBF_OK;
}

void cusbf_destroy(cusbf_filter* f) {
    if (!f) return;
    release_staging(f);
    release_ring(f);
    if (f->d_words) (void)hipFree(f->d_words);
    if (f->d_scalar) (void)hipFree(f->d_scalar);
    if (f->d_rec) (void)hipFree(f->d_rec);
    if (f->d_batch_state) (void)hipFree(f->d_batch_state);
    if (f->d_zero) (void)hipFree(f->d_zero);
    if (f->d_starts) (void)hipFree(f->d_starts);
    if (f->d_sorted) {
        (void)hipEventSynchronize(f->sorted_done);
        (void)hipFree(f->d_sorted);
    }
    if (f->sorted_done) (void)hipEventDestroy(f->sorted_done);
    if (f->insert_done) (void)hipEventDestroy(f->insert_done);
    delete f;
}

// filter.cuh:793-803: clears the words and synchronizes the stream. The words
// are cleared by a fill kernel, not a memset node, so a clear may also be
// captured into a graph; under capture it skips the synchronize (which would
// end the capture) and never marks the words as known to be zero, because a
// replay may follow writes through the words pointer that the library
// cannot see.
int cusbf_clear(cusbf_filter* f, void* stream) {
    if (!f) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null filter");
    const bool captured = capturing(as_stream(stream));
    CUSBF_HIP_TRY(sbf::fill_u64(f->d_words, 0, f->num_shards * 4, as_stream(stream)));
    CUSBF_HIP_TRY(sbf::fill_u64(f->d_zero, f->words_shared || captured ? 0 : 1, 1, as_stream(stream)));
    if (!captured) CUSBF_HIP_TRY(hipStreamSynchronize(as_stream(stream)));
    return CUSBF_OK;
}

// filter.cuh:810-823
int cusbf_load_factor(const cusbf_filter* f, float* out) {
    if (!f || !out) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null argument");
    CUSBF_HIP_TRY(sbf::fill_u64(f->d_scalar, 0, 1, nullptr));
    CUSBF_HIP_TRY(sbf::popcount_words(f->d_words, f->num_shards * 4, f->d_scalar, nullptr));
    uint64_t set_bits = 0;
    CUSBF_HIP_TRY(hipMemcpy(&set_bits, f->d_scalar, sizeof(set_bits), hipMemcpyDeviceToHost));
    *out = static_cast<float>(set_bits) / static_cast<float>(f->filter_bits);
    return CUSBF_OK;
}

uint64_t cusbf_filter_bits(const cusbf_filter* f) { return f ? f->filter_bits : 0; }
uint64_t cusbf_num_shards(const cusbf_filter* f) { return f ? f->num_shards : 0; }

int cusbf_words_device_ptr(const cusbf_filter* f, uint64_t** out) {
    if (!f || !out) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null argument");
    if (!f->words_shared) {  // the caller may write the words: never assume them zero again
        f->words_shared = true;
        CUSBF_HIP_TRY(sbf::fill_u64(f->d_zero, 0, 1, nullptr));
        CUSBF_HIP_TRY(hipStreamSynchronize(nullptr));
    }
    *out = f->d_words;
    return CUSBF_OK;
}

int cusbf_copy_words_to_host(const cusbf_filter* f, uint64_t* host_words, uint64_t num_words) {
    if (!f || (!host_words && num_words)) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null argument");
    if (num_words < f->num_shards * 4) return fail(CUSBF_ERR_INVALID_ARGUMENT, "host word buffer is too small");
    CUSBF_HIP_TRY(hipDeviceSynchronize());
    // the pinned ring is mutable state of the filter, like its staging
    return copy_out(const_cast<cusbf_filter*>(f), host_words, f->d_words, f->num_shards * 32, nullptr);
}

const char* cusbf_kernel_family(const cusbf_filter* f) {
    if (!f) return "";
    if (!f->strip) return "generic";
    return f->strip->k == 0 ? "lane_strip_wide" : "lane_strip";
}

int cusbf_set_insert_path(cusbf_filter* f, uint32_t path, uint32_t slice_bits,
                          uint32_t reserve_per_tile) {
    if (!f) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null filter");
    if (path > CUSBF_INSERT_SORTED) return fail(CUSBF_ERR_INVALID_ARGUMENT, "unknown insert path");
    if (slice_bits > 12) return fail(CUSBF_ERR_INVALID_ARGUMENT, "slice_bits must be in [1, 12]");
    if (reserve_per_tile > 1024) {
        return fail(CUSBF_ERR_INVALID_ARGUMENT, "reserve_per_tile must be in [1, 1024]");
    }
    f->insert_path = path;
    f->slice_bits = slice_bits ? slice_bits : 12;
    f->reserve_per_tile = reserve_per_tile;
    return CUSBF_OK;
}

int cusbf_get_insert_path(const cusbf_filter* f, uint32_t* path, uint32_t* slice_bits,
                          uint32_t* reserve_per_tile, uint32_t* last_path) {
    if (!f) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null filter");
    if (path) *path = f->insert_path;
    if (slice_bits) *slice_bits = f->slice_bits;
    if (reserve_per_tile) *reserve_per_tile = f->reserve_per_tile;
    if (last_path) *last_path = f->last_path;
    return CUSBF_OK;
}

int cusbf_set_sorted_workspace_limit(cusbf_filter* f, uint64_t bytes) {
    if (!f) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null filter");
    f->sorted_limit = bytes;
    if (bytes != 0 && f->d_sorted != nullptr && f->sorted_bytes > bytes) {
        CUSBF_HIP_TRY(hipEventSynchronize(f->sorted_done));
        CUSBF_HIP_TRY(hipFree(f->d_sorted));
        f->d_sorted = nullptr;
        f->sorted_bytes = 0;
        f->sorted_used = false;
    }
    return CUSBF_OK;
}

int cusbf_get_sorted_workspace(const cusbf_filter* f, uint64_t* limit, uint64_t* allocated,
                               uint32_t* last_chunks) {
    if (!f) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null filter");
    if (limit) *limit = f->sorted_limit;
    if (allocated) *allocated = f->sorted_bytes;
    if (last_chunks) *last_chunks = f->last_chunks;
    return CUSBF_OK;
}

// filter.cuh:179-190
int cusbf_insert_sequence_async(cusbf_filter* f, const char* d_sequence, uint64_t bytes,
                                void* stream, uint64_t* inserted) {
    if (!f) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null filter");
    const uint64_t total = windows_of(f, bytes);
    if (inserted) *inserted = total;
    if (total == 0) return CUSBF_OK;
    if (!d_sequence) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null device sequence");
    return launch_insert(f, byte_input(f, d_sequence, bytes), as_stream(stream));
}

// filter.cuh:154-167
int cusbf_insert_sequence(cusbf_filter* f, const char* h_sequence, uint64_t bytes, void* stream,
                          uint64_t* inserted) {
    if (!f) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null filter");
    if (inserted) *inserted = 0;
    if (symbols_of(f, bytes) < f->cfg.k) return CUSBF_OK;
    int rc = ensure_stage(f, bytes);
    if (rc != CUSBF_OK) return rc;
    rc = copy_in(f, f->d_stage, h_sequence, bytes, as_stream(stream));
    if (rc != CUSBF_OK) return rc;
    rc = launch_insert(f, byte_input(f, f->d_stage, bytes), as_stream(stream));
    if (rc != CUSBF_OK) return rc;
    CUSBF_HIP_TRY(hipStreamSynchronize(as_stream(stream)));
    if (inserted) *inserted = windows_of(f, bytes);
    return CUSBF_OK;
}

// filter.cuh:464-474 (no output-size check, as in the reference)
int cusbf_contains_sequence_async(const cusbf_filter* f, const char* d_sequence, uint64_t bytes,
                                  uint8_t* d_output, uint64_t d_output_len, void* stream) {
    (void)d_output_len;
    if (!f) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null filter");
    if (windows_of(f, bytes) == 0) return CUSBF_OK;
    if (!d_sequence || !d_output) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null device buffer");
    return launch_contains(f, byte_input(f, d_sequence, bytes), d_output, as_stream(stream));
}

// filter.cuh:514-533
int cusbf_contains_sequence(const cusbf_filter* f_const, const char* h_sequence, uint64_t bytes,
                            uint8_t* h_output, uint64_t h_output_len, void* stream) {
    cusbf_filter* f = const_cast<cusbf_filter*>(f_const);  // staging is mutable state
    if (!f) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null filter");
    const uint64_t n = windows_of(f, bytes);
    if (n == 0) return CUSBF_OK;
    if (h_output_len < n) return fail(CUSBF_ERR_INVALID_ARGUMENT, "host output buffer is too small");
    int rc = ensure_stage(f, bytes);
    if (rc == CUSBF_OK) rc = ensure_result(f, n);
    if (rc != CUSBF_OK) return rc;
    rc = copy_in(f, f->d_stage, h_sequence, bytes, as_stream(stream));
    if (rc != CUSBF_OK) return rc;
    rc = launch_contains(f, byte_input(f, f->d_stage, bytes), f->d_result, as_stream(stream));
    if (rc != CUSBF_OK) return rc;
    return copy_out(f, h_output, f->d_result, n, as_stream(stream));
}

// filter.cuh:486-501
int cusbf_contains_sequence_host_to_device(const cusbf_filter* f_const, const char* h_sequence,
                                           uint64_t bytes, uint8_t* d_output,
                                           uint64_t d_output_len, void* stream) {
    (void)d_output_len;
    cusbf_filter* f = const_cast<cusbf_filter*>(f_const);
    if (!f) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null filter");
    if (windows_of(f, bytes) == 0) return CUSBF_OK;
    if (!d_output) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null device buffer");
    int rc = ensure_stage(f, bytes);
    if (rc != CUSBF_OK) return rc;
    rc = copy_in(f, f->d_stage, h_sequence, bytes, as_stream(stream));
    if (rc != CUSBF_OK) return rc;
    rc = launch_contains(f, byte_input(f, f->d_stage, bytes), d_output, as_stream(stream));
    if (rc != CUSBF_OK) return rc;
    CUSBF_HIP_TRY(hipStreamSynchronize(as_stream(stream)));
    return CUSBF_OK;
}

uint64_t cusbf_dense_packed_word_count(const cusbf_config* cfg, uint64_t num_symbols) {
    Derived d;
    if (!cfg || derive(cfg, &d) != CUSBF_OK || d.symbol_bits == 0) return 0;
    const uint64_t per_word = 64u / d.symbol_bits;
    return (num_symbols + per_word - 1) / per_word;
}

uint64_t cusbf_dense_packed_kmer_count(const cusbf_config* cfg, uint64_t num_symbols) {
    if (!cfg) return 0;
    return num_symbols < cfg->k ? 0 : num_symbols - cfg->k + 1;
}

// dense_packed.hpp:57-73
int cusbf_pack_dense_sequence_async(const cusbf_config* cfg, const char* d_sequence,
                                    uint64_t num_symbols, uint64_t* d_words, uint64_t num_words,
                                    void* stream) {
    if (!cfg) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null config");
    Derived d;
    int rc = derive(cfg, &d);
    if (rc != CUSBF_OK) return rc;
    if (num_symbols == 0) return CUSBF_OK;
    if (num_words < cusbf_dense_packed_word_count(cfg, num_symbols)) {
        return fail(CUSBF_ERR_INVALID_ARGUMENT, "dense packed span is too small for num_symbols");
    }
    sbf::KernelConfig kc{};
    kc.k = cfg->k;
    kc.alphabet = cfg->alphabet;
    kc.symbol_width = d.symbol_width;
    kc.symbol_bits = d.symbol_bits;
    kc.invalid_symbol = d.invalid_symbol;
    std::memcpy(kc.lut, cfg->lut, 256);
    CUSBF_HIP_TRY(sbf::pack_dense(kc, d_sequence, num_symbols, d_words, as_stream(stream)));
    return CUSBF_OK;
}

// filter.cuh:207-219 + launch checks of :1742-1764
int cusbf_insert_dense_packed_async(cusbf_filter* f, const uint64_t* d_words, uint64_t num_words,
                                    uint64_t num_symbols, void* stream, uint64_t* inserted) {
    if (!f) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null filter");
    const uint64_t total = cusbf_dense_packed_kmer_count(&f->cfg, num_symbols);
    if (inserted) *inserted = total;
    if (total == 0) return CUSBF_OK;
    if (num_words < cusbf_dense_packed_word_count(&f->cfg, num_symbols)) {
        if (inserted) *inserted = 0;
        return fail(CUSBF_ERR_INVALID_ARGUMENT, "dense packed span is too small for num_symbols");
    }
    sbf::SequenceInput in;
    in.words = d_words;
    in.num_words = num_words;
    in.num_symbols = num_symbols;
    in.dense = true;
    return launch_insert(f, in, as_stream(stream));
}

// filter.cuh:248-259 + launch checks of :1766-1796
int cusbf_contains_dense_packed_async(const cusbf_filter* f, const uint64_t* d_words,
                                      uint64_t num_words, uint64_t num_symbols, uint8_t* d_output,
                                      uint64_t d_output_len, void* stream) {
    if (!f) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null filter");
    const uint64_t n = cusbf_dense_packed_kmer_count(&f->cfg, num_symbols);
    if (n == 0) return CUSBF_OK;
    if (num_words < cusbf_dense_packed_word_count(&f->cfg, num_symbols)) {
        return fail(CUSBF_ERR_INVALID_ARGUMENT, "dense packed span is too small for num_symbols");
    }
    if (d_output_len < n) return fail(CUSBF_ERR_INVALID_ARGUMENT, "dense packed query output span is too small");
    sbf::SequenceInput in;
    in.words = d_words;
    in.num_words = num_words;
    in.num_symbols = num_symbols;
    in.dense = true;
    return launch_contains(f, in, d_output, as_stream(stream));
}

// Host dense words staged into library-owned device memory (word-aligned part of the stage).
static int stage_words(cusbf_filter* f, const uint64_t* h_words, uint64_t num_words,
                       hipStream_t stream) {
    const int rc = ensure_stage(f, num_words * sizeof(uint64_t));
    if (rc != CUSBF_OK) return rc;
    return copy_in(f, f->d_stage, h_words, num_words * sizeof(uint64_t), stream);
}

// filter.cuh:226-240
int cusbf_insert_dense_packed(cusbf_filter* f, const uint64_t* h_words, uint64_t num_words,
                              uint64_t num_symbols, void* stream, uint64_t* inserted) {
    if (!f) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null filter");
    if (inserted) *inserted = 0;
    if (cusbf_dense_packed_kmer_count(&f->cfg, num_symbols) == 0) return CUSBF_OK;
    if (num_words < cusbf_dense_packed_word_count(&f->cfg, num_symbols)) {
        return fail(CUSBF_ERR_INVALID_ARGUMENT, "dense packed span is too small for num_symbols");
    }
    int rc = stage_words(f, h_words, num_words, as_stream(stream));
    if (rc != CUSBF_OK) return rc;
    rc = cusbf_insert_dense_packed_async(f, reinterpret_cast<const uint64_t*>(f->d_stage), num_words,
                                         num_symbols, stream, inserted);
    if (rc != CUSBF_OK) return rc;
    CUSBF_HIP_TRY(hipStreamSynchronize(as_stream(stream)));
    return CUSBF_OK;
}

// filter.cuh:267-282
int cusbf_contains_dense_packed_host_to_device(const cusbf_filter* f_const, const uint64_t* h_words,
                                               uint64_t num_words, uint64_t num_symbols,
                                               uint8_t* d_output, uint64_t d_output_len,
                                               void* stream) {
    cusbf_filter* f = const_cast<cusbf_filter*>(f_const);
    if (!f) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null filter");
    if (cusbf_dense_packed_kmer_count(&f->cfg, num_symbols) == 0) return CUSBF_OK;
    if (num_words < cusbf_dense_packed_word_count(&f->cfg, num_symbols)) {
        return fail(CUSBF_ERR_INVALID_ARGUMENT, "dense packed span is too small for num_symbols");
    }
    int rc = stage_words(f, h_words, num_words, as_stream(stream));
    if (rc != CUSBF_OK) return rc;
    rc = cusbf_contains_dense_packed_async(f, reinterpret_cast<const uint64_t*>(f->d_stage),
                                           num_words, num_symbols, d_output, d_output_len, stream);
    if (rc != CUSBF_OK) return rc;
    CUSBF_HIP_TRY(hipStreamSynchronize(as_stream(stream)));
    return CUSBF_OK;
}

// filter.cuh:289-312
int cusbf_contains_dense_packed(const cusbf_filter* f_const, const uint64_t* h_words,
                                uint64_t num_words, uint64_t num_symbols, uint8_t* h_output,
                                uint64_t h_output_len, void* stream) {
    cusbf_filter* f = const_cast<cusbf_filter*>(f_const);
    if (!f) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null filter");
    const uint64_t n = cusbf_dense_packed_kmer_count(&f->cfg, num_symbols);
    if (n == 0) return CUSBF_OK;
    if (h_output_len < n) return fail(CUSBF_ERR_INVALID_ARGUMENT, "host output buffer is too small");
    if (num_words < cusbf_dense_packed_word_count(&f->cfg, num_symbols)) {
        return fail(CUSBF_ERR_INVALID_ARGUMENT, "dense packed span is too small for num_symbols");
    }
    int rc = ensure_result(f, n);
    if (rc == CUSBF_OK) rc = stage_words(f, h_words, num_words, as_stream(stream));
    if (rc != CUSBF_OK) return rc;
    // words in, answers out and the kernel in one stream order, one wait
    rc = cusbf_contains_dense_packed_async(f, reinterpret_cast<const uint64_t*>(f->d_stage), num_words,
                                           num_symbols, f->d_result, n, stream);
    if (rc != CUSBF_OK) return rc;
    return copy_out(f, h_output, f->d_result, n, as_stream(stream));
}

// filter.cuh:385-400
int cusbf_insert_record_batch_async(cusbf_filter* f, const char* d_sequence, uint64_t bytes,
                                    const cusbf_record_range* ranges, uint64_t num_ranges,
                                    void* stream, uint64_t* inserted) {
    if (!f) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null filter");
    if (inserted) *inserted = 0;
    int rc = validate_ranges(f, bytes, ranges, num_ranges, "record batch range exceeds device sequence");
    if (rc != CUSBF_OK) return rc;
    uint64_t total = 0;
    for (uint64_t i = 0; i < num_ranges; ++i) total += record_kmers(f, ranges[i].sequence_bytes);
    if (total == 0) return CUSBF_OK;
    rc = launch_insert(f, byte_input(f, d_sequence, bytes), as_stream(stream));
    if (rc != CUSBF_OK) return rc;
    if (inserted) *inserted = total;
    return CUSBF_OK;
}

// filter.cuh:603-618
int cusbf_contains_record_batch_async(const cusbf_filter* f, const char* d_sequence,
                                      uint64_t bytes, const cusbf_record_range* ranges,
                                      uint64_t num_ranges, uint8_t* d_output,
                                      uint64_t d_output_len, void* stream) {
    if (!f) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null filter");
    const int rc = validate_ranges(f, bytes, ranges, num_ranges, "record batch range exceeds device sequence");
    if (rc != CUSBF_OK) return rc;
    const uint64_t total = windows_of(f, bytes);
    if (total == 0) return CUSBF_OK;
    if (d_output_len < total) return fail(CUSBF_ERR_INVALID_ARGUMENT, "record batch query output span is too small");
    return launch_contains(f, byte_input(f, d_sequence, bytes), d_output, as_stream(stream));
}

// filter.cuh:338-364
int cusbf_insert_record_batch(cusbf_filter* f, const char* h_sequence, uint64_t bytes,
                              const cusbf_record_range* ranges, uint64_t num_ranges, void* stream,
                              cusbf_insert_report* report) {
    if (!f) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null filter");
    int rc = validate_ranges(f, bytes, ranges, num_ranges, "record batch range exceeds the source sequence");
    if (rc != CUSBF_OK) return rc;
    NormPlan p;
    norm_plan(f, ranges, num_ranges, &p);
    cusbf_insert_report rep{num_ranges, p.bases, 0};
    if (p.total) {
        rc = ensure_stage(f, p.total);
        if (rc == CUSBF_OK) rc = copy_in_normalized(f, f->d_stage, h_sequence, ranges, p, as_stream(stream));
        if (rc == CUSBF_OK) rc = launch_insert(f, byte_input(f, f->d_stage, p.total), as_stream(stream));
        if (rc == CUSBF_OK) rc = record_counts(f, ranges, p, nullptr, as_stream(stream), &rep.inserted_kmers, nullptr);
        if (rc != CUSBF_OK) return rc;
    }
    release_staging(f);
    if (report) *report = rep;
    return CUSBF_OK;
}

// filter.cuh:930-990 (aggregate query of a host record batch).
int cusbf_query_record_batch(const cusbf_filter* f_const, const char* h_sequence, uint64_t bytes,
                             const cusbf_record_range* ranges, uint64_t num_ranges, void* stream,
                             cusbf_query_report* report) {
    cusbf_filter* f = const_cast<cusbf_filter*>(f_const);
    if (!f) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null filter");
    int rc = validate_ranges(f, bytes, ranges, num_ranges, "record batch range exceeds the source sequence");
    if (rc != CUSBF_OK) return rc;
    NormPlan p;
    norm_plan(f, ranges, num_ranges, &p);
    cusbf_query_report rep{num_ranges, p.bases, 0, 0};
    // query_layout.cuh:66-67: total hits = windows of the normalised buffer.
    const uint64_t hits = windows_of(f, p.total);
    if (hits) {
        rc = ensure_stage(f, p.total);
        if (rc == CUSBF_OK) rc = ensure_result(f, hits);
        if (rc == CUSBF_OK) rc = copy_in_normalized(f, f->d_stage, h_sequence, ranges, p, as_stream(stream));
        if (rc == CUSBF_OK) rc = launch_contains(f, byte_input(f, f->d_stage, p.total), f->d_result, as_stream(stream));
        if (rc == CUSBF_OK) {
            rc = record_counts(f, ranges, p, f->d_result, as_stream(stream), &rep.queried_kmers, &rep.positive_kmers);
        }
        if (rc != CUSBF_OK) return rc;
    }
    release_staging(f);
    if (report) *report = rep;
    return CUSBF_OK;
}

// filter.cuh:566-582, 1007-1131: host record batch in, per-window hits of the
// normalised batch out (the consumer form reports them per record).
int cusbf_query_record_batch_hits(const cusbf_filter* f_const, const char* h_sequence, uint64_t bytes,
                                  const cusbf_record_range* ranges, uint64_t num_ranges, void* stream,
                                  uint8_t* h_hits, uint64_t hits_len, cusbf_query_report* report) {
    cusbf_filter* f = const_cast<cusbf_filter*>(f_const);
    if (!f) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null filter");
    int rc = validate_ranges(f, bytes, ranges, num_ranges, "record batch range exceeds the source sequence");
    if (rc != CUSBF_OK) return rc;
    NormPlan p;
    norm_plan(f, ranges, num_ranges, &p);
    const uint64_t hits = windows_of(f, p.total);
    if (hits_len < hits) return fail(CUSBF_ERR_INVALID_ARGUMENT, "record batch hit buffer is too small");
    cusbf_query_report rep{num_ranges, p.bases, 0, 0};
    if (hits) {
        rc = ensure_stage(f, p.total);
        if (rc == CUSBF_OK) rc = ensure_result(f, hits);
        if (rc == CUSBF_OK) rc = copy_in_normalized(f, f->d_stage, h_sequence, ranges, p, as_stream(stream));
        if (rc == CUSBF_OK) rc = launch_contains(f, byte_input(f, f->d_stage, p.total), f->d_result, as_stream(stream));
        if (rc == CUSBF_OK) {
            rc = record_counts(f, ranges, p, f->d_result, as_stream(stream), &rep.queried_kmers, &rep.positive_kmers);
        }
        if (rc == CUSBF_OK) rc = copy_out(f, h_hits, f->d_result, hits, as_stream(stream));
        if (rc != CUSBF_OK) return rc;
    }
    if (report) *report = rep;
    return CUSBF_OK;
}

int cusbf_count_positive_async(const uint8_t* d_hits, uint64_t num_hits, uint64_t* d_count,
                               void* stream) {
    if (!d_count) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null count pointer");
    CUSBF_HIP_TRY(sbf::fill_u64(d_count, 0, 1, as_stream(stream)));
    if (num_hits == 0) return CUSBF_OK;
    if (!d_hits) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null hits pointer");
    CUSBF_HIP_TRY(sbf::count_ones(d_hits, num_hits, d_count, as_stream(stream)));
    return CUSBF_OK;
}

int cusbf_count_positive_per_record_async(const uint8_t* d_hits, const uint64_t* d_offsets,
                                          const uint64_t* d_counts, uint64_t num_records,
                                          uint64_t* d_positive, void* stream) {
    if (num_records == 0) return CUSBF_OK;
    if (!d_hits || !d_offsets || !d_counts || !d_positive) {
        return fail(CUSBF_ERR_INVALID_ARGUMENT, "null device buffer");
    }
    CUSBF_HIP_TRY(sbf::count_ones_per_record(d_hits, d_offsets, d_counts, num_records, d_positive, as_stream(stream)));
    return CUSBF_OK;
}

int cusbf_count_windows_runs_async(const cusbf_filter* f, const char* d_sequence, uint64_t bytes,
                                   uint64_t* d_counts, void* stream) {
    if (!f || !d_counts) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null argument");
    CUSBF_HIP_TRY(sbf::fill_u64(d_counts, 0, 2, as_stream(stream)));
    if (windows_of(f, bytes) == 0) return CUSBF_OK;
    CUSBF_HIP_TRY(sbf::generic_count(f->kc, byte_input(f, d_sequence, bytes), f->num_shards, d_counts, as_stream(stream)));
    return CUSBF_OK;
}

// --- batched record form --------------------------------------------------------

uint32_t cusbf_batch_bucket(uint64_t num_symbols, uint64_t num_records) {
    const uint64_t mean = num_records ? num_symbols / num_records : num_symbols;
    if (mean < 256) return 0;
    if (mean < 4096) return 1;
    if (mean < 65536) return 2;
    return 3;
}

uint32_t cusbf_batch_variants(const cusbf_filter* f) {
    if (!f || !batched_strip(f)) return 0;
    return 1u | (f->strip->batched_insert[1] ? 2u : 0u);
}

int cusbf_batch_tuning_set(cusbf_filter* f, uint32_t bucket, uint32_t op,
                           const cusbf_batch_tuning* tuning) {
    if (!f || !tuning) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null argument");
    if (bucket >= CUSBF_BATCH_BUCKETS || op > CUSBF_BATCH_INSERT_SORTED) {
        return fail(CUSBF_ERR_INVALID_ARGUMENT, "batch bucket or op out of range");
    }
    const int rc = check_tuning(f, *tuning);
    if (rc != CUSBF_OK) return rc;
    f->tuning[bucket][op] = *tuning;
    return CUSBF_OK;
}

int cusbf_batch_tuning_get(const cusbf_filter* f, uint32_t bucket, uint32_t op,
                           cusbf_batch_tuning* out) {
    if (!f || !out) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null argument");
    if (bucket >= CUSBF_BATCH_BUCKETS || op > CUSBF_BATCH_INSERT_SORTED) {
        return fail(CUSBF_ERR_INVALID_ARGUMENT, "batch bucket or op out of range");
    }
    *out = f->tuning[bucket][op];
    return CUSBF_OK;
}

int cusbf_insert_batched_async(cusbf_filter* f, const char* d_symbols, uint64_t num_symbols,
                               const uint64_t* d_offsets, uint64_t num_records, void* stream) {
    int rc = validate_batched(f, d_symbols, num_symbols, d_offsets, num_records);
    if (rc != CUSBF_OK) return rc;
    f->batch_check_pending = false;
    if (batched_windows(f, num_symbols) == 0) return CUSBF_OK;
    rc = ensure_starts(f, num_symbols);
    if (rc != CUSBF_OK) return rc;
    const cusbf_batch_tuning& ts = tuning_for(f, CUSBF_BATCH_INSERT_SORTED, num_symbols, num_records);
    sbf::SortedLayout l;
    bool sorted = false;
    f->last_path = CUSBF_INSERT_ATOMIC;
    rc = batched_strip(f)
             ? choose_batched_sorted(f, num_symbols, ts, as_stream(stream), &l, &sorted)
             : choose_generic_sorted(f, batched_windows(f, num_symbols), as_stream(stream), &l, &sorted);
    if (rc == CUSBF_OK && sorted) rc = begin_sorted(f, l.bytes, as_stream(stream));
    if (rc != CUSBF_OK) return rc;
    const cusbf_batch_tuning& t =
        sorted ? ts : tuning_for(f, CUSBF_BATCH_INSERT, num_symbols, num_records);
    rc = order_insert(f, as_stream(stream));
    if (rc != CUSBF_OK) return rc;
    rc = enqueue_batched(f, CUSBF_BATCH_INSERT, d_symbols, num_symbols, d_offsets, num_records,
                         f->d_batch_state, f->d_starts, t, nullptr, as_stream(stream),
                         sorted ? &l : nullptr, f->d_sorted);
    if (rc == CUSBF_OK) rc = mark_insert(f, as_stream(stream));
    if (rc == CUSBF_OK && sorted) {
        f->last_path = CUSBF_INSERT_SORTED;
        rc = end_sorted(f, as_stream(stream));
    }
    f->batch_check_pending = rc == CUSBF_OK;
    f->batch_check_state = f->d_batch_state;
    return rc;
}

int cusbf_contains_batched_async(const cusbf_filter* f, const char* d_symbols,
                                 uint64_t num_symbols, const uint64_t* d_offsets,
                                 uint64_t num_records, uint8_t* d_output, uint64_t d_output_len,
                                 void* stream) {
    return cusbf_internal::contains_batched(f, d_symbols, num_symbols, d_offsets, num_records,
                                            d_output, d_output_len, nullptr, nullptr,
                                            as_stream(stream));
}

int cusbf_join_records_async(const cusbf_filter* f, const char* d_symbols, uint64_t num_symbols,
                             const uint64_t* d_offsets, uint64_t num_records, char* d_out,
                             uint64_t d_out_len, void* stream) {
    if (!f) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null filter");
    if (num_records == 0) return CUSBF_OK;
    if (!d_offsets) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null device offsets");
    if (num_symbols && !d_symbols) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null device sequence");
    const uint64_t w = f->d.symbol_width;
    const uint64_t need = (num_symbols + num_records - 1) * w;
    if (need && !d_out) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null device output");
    if (d_out_len < need) return fail(CUSBF_ERR_INVALID_ARGUMENT, "joined record batch output is too small");
    CUSBF_HIP_TRY(sbf::join_records(d_symbols, num_symbols, d_offsets, num_records, f->d.symbol_width,
                                    f->d.separator, d_out, as_stream(stream)));
    return CUSBF_OK;
}

int cusbf_batched_status(const cusbf_filter* f, void* stream) {
    if (!f) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null filter");
    CUSBF_HIP_TRY(hipStreamSynchronize(as_stream(stream)));
    if (!f->batch_check_pending) return CUSBF_OK;
    return read_batch_flag(f->batch_check_state ? f->batch_check_state : f->d_batch_state,
                           as_stream(stream));
}

int cusbf_batch_plan_create(cusbf_filter* f, uint32_t op, const char* d_symbols,
                            uint64_t num_symbols, const uint64_t* d_offsets, uint64_t num_records,
                            uint8_t* d_output, uint64_t d_output_len,
                            const cusbf_batch_tuning* tuning, cusbf_batch_plan** out) {
    if (!out) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    int rc = validate_batched(f, d_symbols, num_symbols, d_offsets, num_records);
    if (rc != CUSBF_OK) return rc;
    if (op > CUSBF_BATCH_CONTAINS) return fail(CUSBF_ERR_INVALID_ARGUMENT, "batch op out of range");
    const uint64_t windows = batched_windows(f, num_symbols);
    if (op == CUSBF_BATCH_CONTAINS && windows != 0) {
        if (!d_output) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null device output");
        if (d_output_len < windows) return fail(CUSBF_ERR_INVALID_ARGUMENT, "batched query output span is too small");
    }
    // an explicit tuning is used on whichever insert path the plan takes
    cusbf_batch_tuning t = tuning ? *tuning : tuning_for(f, op, num_symbols, num_records);
    const cusbf_batch_tuning ts =
        tuning ? *tuning : tuning_for(f, CUSBF_BATCH_INSERT_SORTED, num_symbols, num_records);
    rc = check_tuning(f, t);
    if (rc == CUSBF_OK) rc = check_tuning(f, ts);
    if (rc != CUSBF_OK) return rc;

    cusbf_batch_plan* p = new cusbf_batch_plan();
    p->f = f;
    p->op = op;
    p->empty = windows == 0;
    if (p->empty) {
        *out = p;
        return CUSBF_OK;
    }
    hipStream_t capture = nullptr;
    sbf::SortedLayout l;
    bool sorted = false;
    if (op == CUSBF_BATCH_INSERT && batched_strip(f)) {
        rc = choose_batched_sorted(f, num_symbols, ts, nullptr, &l, &sorted);
        if (rc != CUSBF_OK) {
            cusbf_batch_plan_destroy(p);
            return rc;
        }
        if (sorted) t = ts;
    }
    hipError_t e = hipMalloc(&p->d_state, 64);
    if (e == hipSuccess) e = hipMalloc(&p->d_starts, sbf::start_words(num_symbols) * sizeof(uint64_t));
    if (e == hipSuccess && sorted) e = hipMalloc(&p->d_sorted, l.bytes);
    if (e == hipSuccess && sorted) e = sbf::sorted::prepare();
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&capture, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamBeginCapture(capture, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        rc = fail_hip(e, __FILE__, __LINE__);
    } else {
        rc = enqueue_batched(f, op, d_symbols, num_symbols, d_offsets, num_records, p->d_state,
                             p->d_starts, t, d_output, capture, sorted ? &l : nullptr, p->d_sorted);
        const hipError_t end = hipStreamEndCapture(capture, &p->graph);
        if (rc == CUSBF_OK && end != hipSuccess) rc = fail_hip(end, __FILE__, __LINE__);
        if (rc == CUSBF_OK) {
            e = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0);
            if (e != hipSuccess) rc = fail_hip(e, __FILE__, __LINE__);
        }
    }
    if (capture) (void)hipStreamDestroy(capture);
    if (rc != CUSBF_OK) {
        cusbf_batch_plan_destroy(p);
        return rc;
    }
    *out = p;
    return CUSBF_OK;
}

// An insert plan is ordered like every other insert into its filter: after
// the previous insert on any stream and before the next (its sorted form
// rewrites whole slices, so an insert overlapping it could lose updates).
int cusbf_batch_plan_launch(cusbf_batch_plan* p, void* stream) {
    if (!p) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null plan");
    if (p->empty) return CUSBF_OK;
    const bool insert = p->op == CUSBF_BATCH_INSERT;
    if (insert) {
        const int rc = order_insert(p->f, as_stream(stream));
        if (rc != CUSBF_OK) return rc;
    }
    CUSBF_HIP_TRY(hipGraphLaunch(p->exec, as_stream(stream)));
    return insert ? mark_insert(p->f, as_stream(stream)) : CUSBF_OK;
}

int cusbf_batch_plan_status(cusbf_batch_plan* p, void* stream) {
    if (!p) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null plan");
    CUSBF_HIP_TRY(hipStreamSynchronize(as_stream(stream)));
    if (p->empty) return CUSBF_OK;
    return read_batch_flag(p->d_state, as_stream(stream));
}

void cusbf_batch_plan_destroy(cusbf_batch_plan* p) {
    if (!p) return;
    if (p->exec || p->d_state) (void)hipDeviceSynchronize();  // a replay may still run
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->d_state) (void)hipFree(p->d_state);
    if (p->d_starts) (void)hipFree(p->d_starts);
    if (p->d_sorted) (void)hipFree(p->d_sorted);
    delete p;
}

// No reference counterpart: valid windows and shard runs of a batched input
// (windows across a record start are not valid), for the byte model.
int cusbf_count_batched_windows_runs_async(const cusbf_filter* f, const char* d_symbols,
                                           uint64_t num_symbols, const uint64_t* d_offsets,
                                           uint64_t num_records, uint64_t* d_counts,
                                           void* stream) {
    int rc = validate_batched(f, d_symbols, num_symbols, d_offsets, num_records);
    if (rc != CUSBF_OK) return rc;
    if (!d_counts) return fail(CUSBF_ERR_INVALID_ARGUMENT, "null argument");
    CUSBF_HIP_TRY(sbf::fill_u64(d_counts, 0, 2, as_stream(stream)));
    if (batched_windows(f, num_symbols) == 0) return CUSBF_OK;
    sbf::SequenceInput in = byte_input(f, d_symbols, num_symbols * f->d.symbol_width);
    in.offsets = d_offsets;
    in.num_records = num_records;
    CUSBF_HIP_TRY(sbf::generic_count(f->kc, in, f->num_shards, d_counts, as_stream(stream)));
    return CUSBF_OK;
}

}  // extern "C"
