// serve.hip -- request server for the query path (continuous batching), libcusbf_hip.so.
//
// Requests are host record batches. A scheduler thread takes queued requests
// in arrival order into a free batch slot until the slot's staging is full,
// normalises them into the slot's pinned buffer (records and requests joined by
// the alphabet separator), and issues H2D copy, query kernel and D2H copy on the
// slot's stream. A completion thread waits for each slot's event, hands every
// request its windows and report, and returns the slot; the scheduler admits
// new requests as slots come back. With several slots the copies and kernels
// of different batches overlap on the device. Each request's report (valid and
// positive windows) is counted on the device in the same batch, so no host
// thread scans request bytes or answers. The filter is read-only here.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/cusbf_hip.h"
#include "filter_internal.hpp"
#include "host_pool.hpp"
#include "launch.hpp"

namespace {

using Clock = std::chrono::steady_clock;
namespace ci = cusbf_internal;
using cusbf_internal::WorkerPool;

struct Request {
    uint64_t ticket = 0;
    const char* seq = nullptr;
    std::vector<cusbf_record_range> ranges;
    uint8_t* hits = nullptr;
    uint64_t norm_bytes = 0;   // normalised size
    uint64_t windows = 0;      // windows of the normalised batch
    uint64_t batch_offset = 0; // byte offset inside the slot's batch
    cusbf_query_report report{};
    Clock::time_point submitted;
    double latency = 0.0;
    bool done = false;
    int status = CUSBF_OK;
    std::string message;
};

// Requests per batch (bounds the per-slot request descriptors and counters).
constexpr uint64_t kMaxBatchRequests = 4096;

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t start = nullptr;  // recorded before the batch's host-to-device copy
    hipEvent_t done = nullptr;
    char* d_seq = nullptr;
    uint8_t* d_hits = nullptr;
    char* h_seq = nullptr;
    uint64_t* d_bits = nullptr;  // the batch's answers packed into bits (PCIe moves these)
    uint64_t* h_bits = nullptr;
    uint64_t* d_req = nullptr;     // per request: first window in the batch, window count
    uint64_t* h_req = nullptr;
    uint64_t* d_counts = nullptr;  // per request: valid windows, positive windows
    uint64_t* h_counts = nullptr;
    uint64_t batch_bytes = 0;
    int status = CUSBF_OK;
    std::string message;
    std::vector<Request*> batch;
};

constexpr uint64_t kBatchTarget = 256ull << 20;

// Answers cross PCIe packed into bits (one eighth of the bytes; the device packs
// them after the query) and are expanded into the callers' byte buffers on the
// host pool: kExpand[b] holds bit j of b in byte j.
uint64_t bits_bytes(uint64_t windows) { return (windows + 63) / 64 * sizeof(uint64_t); }

struct ExpandTable {
    uint64_t v[256];
    constexpr ExpandTable() : v{} {
        for (int b = 0; b < 256; ++b) {
            uint64_t x = 0;
            for (int j = 0; j < 8; ++j) x |= static_cast<uint64_t>((b >> j) & 1) << (8 * j);
            v[b] = x;
        }
    }
};
constexpr ExpandTable kExpand{};

// Answers [first, first + n) of a batch as bytes 0 / 1 at dst.
void expand_bits(const uint64_t* bits, uint64_t first, uint64_t n, uint8_t* dst) {
    uint64_t i = 0;
    for (; i < n && ((first + i) & 7u) != 0; ++i) dst[i] = static_cast<uint8_t>((bits[(first + i) >> 6] >> ((first + i) & 63u)) & 1u);
    const uint8_t* bytes = reinterpret_cast<const uint8_t*>(bits);  // little-endian: byte k holds bits 8k..8k+7
    for (; i + 8 <= n; i += 8) {
        const uint64_t v = kExpand.v[bytes[(first + i) >> 3]];
        std::memcpy(dst + i, &v, sizeof(v));
    }
    for (; i < n; ++i) dst[i] = static_cast<uint8_t>((bits[(first + i) >> 6] >> ((first + i) & 63u)) & 1u);
}

// Copy parts of at most 4 MiB.
constexpr uint64_t kCopyPart = 4ull << 20;

// Separator padding of normalized_record_batch.hpp:208-232.
uint64_t boundary_bytes(uint64_t current, uint32_t width) {
    if (current == 0) return 0;
    const uint64_t rem = current % width;
    return (rem ? width - rem : 0) + width;
}

uint64_t normalized_size(const cusbf_record_range* r, uint64_t n, uint32_t width) {
    uint64_t size = 0;
    for (uint64_t i = 0; i < n; ++i) size += boundary_bytes(size, width) + r[i].sequence_bytes;
    return size;
}

}  // namespace

struct cusbf_server {
    const cusbf_filter* filter = nullptr;
    cusbf_server_options options{};
    uint32_t width = 1;
    char sep = 'N';
    uint64_t seq_cap = 0;       // per slot
    uint64_t batch_target = 0;  // a batch closes once it holds this many bytes
    std::vector<Slot> slots;

    std::mutex mu;
    std::condition_variable cv_queue, cv_slots, cv_complete, cv_done;
    std::deque<Request*> queue;
    std::vector<int> free_slots;
    std::deque<int> completing;
    std::unordered_map<uint64_t, std::unique_ptr<Request>> requests;
    uint64_t next_ticket = 1;
    bool stop = false;
    bool scheduler_done = false;
    cusbf_server_stats stats{};
    std::unique_ptr<WorkerPool> pool;  // host copies of the batch fill and the answers
    std::thread scheduler, completer;
};

namespace {

// True when the request's records are laid out in its buffer exactly as in its
// normalised batch apart from the separator bytes: each record starts one symbol
// after the previous one ends. Such a request is copied as one span, and only the
// symbol between records is overwritten with the separator.
// (Empty records are excluded: a leading empty record writes no separator.)
bool span_layout(const Request& r, uint32_t width) {
    for (size_t i = 0; i < r.ranges.size(); ++i) {
        if (r.ranges[i].sequence_bytes == 0) return false;
        if (i > 0 && r.ranges[i].sequence_offset !=
                         r.ranges[i - 1].sequence_offset + r.ranges[i - 1].sequence_bytes + width) {
            return false;
        }
    }
    return !r.ranges.empty();
}

// Writes the slot's batch: requests one after another, each request's records
// joined by the separator (padded to the symbol width), a separator between
// requests. Offsets are laid out first; the copies then run on up to eight
// threads over balanced groups of pieces. A request in span layout is one piece
// (cut into 16 MiB parts so that large requests spread over the threads) whose
// copy is followed by separator stores between its records.
void fill_batch(cusbf_server* s, Slot& slot) {
    struct Piece {
        uint64_t pad_at;  // separator bytes written at [pad_at, pad_at + pad)
        uint64_t pad;
        const char* src;
        uint64_t len;
        const Request* span;  // span piece: patch the separators between its records
        uint64_t span_base;   // batch offset of the request's first record
    };
    constexpr uint64_t kPart = kCopyPart;
    std::vector<Piece> pieces;
    uint64_t pos = 0;
    for (Request* r : slot.batch) {
        const uint64_t b = boundary_bytes(pos, s->width);
        pieces.push_back(Piece{pos, b, nullptr, 0, nullptr, 0});
        pos += b;
        r->batch_offset = pos;
        if (span_layout(*r, s->width)) {
            const char* src = r->seq + r->ranges.front().sequence_offset;
            for (uint64_t done = 0; done < r->norm_bytes; done += kPart) {
                const uint64_t len = std::min(kPart, r->norm_bytes - done);
                pieces.push_back(Piece{pos + done, 0, src + done, len, r, pos});
            }
            pos += r->norm_bytes;
            continue;
        }
        uint64_t local = 0;
        for (const cusbf_record_range& rr : r->ranges) {
            const uint64_t lb = boundary_bytes(local, s->width);
            pieces.push_back(Piece{pos, lb, r->seq + rr.sequence_offset, rr.sequence_bytes, nullptr, 0});
            pos += lb + rr.sequence_bytes;
            local += lb + rr.sequence_bytes;
        }
    }
    slot.batch_bytes = pos;
    char* out = slot.h_seq;
    const uint32_t width = s->width;
    const char sep = s->sep;
    auto patch = [&](const Piece& p) {
        // separator symbols of the request that fall inside this part
        const Request& r = *p.span;
        const uint64_t first = r.ranges.front().sequence_offset;
        const uint64_t lo = p.pad_at - p.span_base;
        const uint64_t hi = lo + p.len;
        // record i ends at ranges[i].offset + bytes - first (relative); its separator follows
        auto it = std::lower_bound(r.ranges.begin(), r.ranges.end(), lo,
                                   [&](const cusbf_record_range& x, uint64_t v) {
                                       return x.sequence_offset + x.sequence_bytes + width - first <= v;
                                   });
        for (; it != r.ranges.end() && it + 1 != r.ranges.end(); ++it) {
            const uint64_t at = it->sequence_offset + it->sequence_bytes - first;
            if (at >= hi) break;
            for (uint32_t j = 0; j < width; ++j) {
                if (at + j >= lo && at + j < hi) out[p.span_base + at + j] = sep;
            }
        }
    };
    auto run = [&](size_t first, size_t last) {
        for (size_t i = first; i < last; ++i) {
            const Piece& p = pieces[i];
            if (p.pad) std::memset(out + p.pad_at, sep, p.pad);
            if (p.len) std::memcpy(out + p.pad_at + p.pad, p.src, p.len);
            if (p.span) patch(p);
        }
    };
    // units of consecutive pieces holding about kCopyPart bytes each, run on the pool
    std::vector<size_t> unit_start{0};
    uint64_t unit_bytes = 0;
    for (size_t i = 0; i < pieces.size(); ++i) {
        unit_bytes += pieces[i].pad + pieces[i].len;
        if (unit_bytes >= kCopyPart && i + 1 < pieces.size()) {
            unit_start.push_back(i + 1);
            unit_bytes = 0;
        }
    }
    unit_start.push_back(pieces.size());
    s->pool->run(unit_start.size() - 1, [&](size_t u) { run(unit_start[u], unit_start[u + 1]); });
}

void launch_batch(cusbf_server* s, Slot& slot) {
    slot.status = CUSBF_OK;
    const uint64_t windows = ci::window_count(s->filter, slot.batch_bytes);
    const uint64_t n_req = slot.batch.size();
    for (uint64_t i = 0; i < n_req; ++i) {
        slot.h_req[2 * i] = slot.batch[i]->batch_offset / s->width;
        slot.h_req[2 * i + 1] = slot.batch[i]->windows;
    }
    hipError_t e = hipEventRecord(slot.start, slot.stream);
    if (e == hipSuccess) {
        e = hipMemcpyAsync(slot.d_seq, slot.h_seq, slot.batch_bytes, hipMemcpyHostToDevice, slot.stream);
    }
    if (e == hipSuccess) {
        e = hipMemcpyAsync(slot.d_req, slot.h_req, 2 * n_req * sizeof(uint64_t), hipMemcpyHostToDevice,
                           slot.stream);
    }
    if (e == hipSuccess && windows) {
        int rc = ci::launch_contains_bytes(s->filter, slot.d_seq, slot.batch_bytes, slot.d_hits, slot.stream);
        if (rc == CUSBF_OK) {
            rc = ci::launch_request_counts(s->filter, slot.d_seq, slot.batch_bytes, slot.d_hits, slot.d_req,
                                           n_req, slot.d_counts, slot.stream);
        }
        if (rc != CUSBF_OK) {
            char buf[512];
            cusbf_last_error(buf, sizeof(buf));
            slot.status = rc;
            slot.message = buf;
        }
        e = sbf::pack_hits(slot.d_hits, windows, slot.d_bits, slot.stream);
        if (e == hipSuccess) {
            e = hipMemcpyAsync(slot.h_bits, slot.d_bits, (windows + 63) / 64 * sizeof(uint64_t),
                               hipMemcpyDeviceToHost, slot.stream);
        }
        if (e == hipSuccess) {
            e = hipMemcpyAsync(slot.h_counts, slot.d_counts, 2 * n_req * sizeof(uint64_t),
                               hipMemcpyDeviceToHost, slot.stream);
        }
    }
    if (e == hipSuccess) e = hipEventRecord(slot.done, slot.stream);
    if (e != hipSuccess) {
        slot.status = CUSBF_ERR_DEVICE;
        slot.message = hipGetErrorString(e);
    }
}

void scheduler_loop(cusbf_server* s) {
    (void)hipSetDevice(ci::device(s->filter));
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        s->cv_queue.wait(lk, [&] { return s->stop || !s->queue.empty(); });
        if (s->queue.empty()) break;  // stop requested and drained
        if (s->options.max_wait_us) {
            const auto until = s->queue.front()->submitted + std::chrono::microseconds(s->options.max_wait_us);
            s->cv_queue.wait_until(lk, until, [&] {
                uint64_t bytes = 0;
                for (Request* r : s->queue) bytes += r->norm_bytes + s->width * 2;
                return s->stop || bytes >= s->batch_target;
            });
        }
        s->cv_slots.wait(lk, [&] { return !s->free_slots.empty(); });
        const int idx = s->free_slots.back();
        s->free_slots.pop_back();
        Slot& slot = s->slots[idx];
        slot.batch.clear();
        uint64_t bytes = 0;
        while (!s->queue.empty() && slot.batch.size() < kMaxBatchRequests) {
            Request* r = s->queue.front();
            const uint64_t need = boundary_bytes(bytes, s->width) + r->norm_bytes;
            if (!slot.batch.empty() && (bytes + need > s->seq_cap || bytes >= s->batch_target)) break;
            s->queue.pop_front();
            slot.batch.push_back(r);
            bytes += need;
        }
        s->stats.batches += 1;
        s->stats.requests += slot.batch.size();
        s->stats.max_requests_per_batch = std::max<uint64_t>(s->stats.max_requests_per_batch, slot.batch.size());
        s->stats.bytes += bytes;
        lk.unlock();
        const auto t0 = Clock::now();
        fill_batch(s, slot);
        const auto t1 = Clock::now();
        launch_batch(s, slot);
        const auto t2 = Clock::now();
        lk.lock();
        s->stats.fill_seconds += std::chrono::duration<double>(t1 - t0).count();
        s->stats.launch_seconds += std::chrono::duration<double>(t2 - t1).count();
        s->completing.push_back(idx);
        s->cv_complete.notify_one();
    }
    s->scheduler_done = true;
    s->cv_complete.notify_one();
}

void completion_loop(cusbf_server* s) {
    (void)hipSetDevice(ci::device(s->filter));
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        s->cv_complete.wait(lk, [&] { return !s->completing.empty() || s->scheduler_done; });
        if (s->completing.empty()) break;
        const int idx = s->completing.front();
        s->completing.pop_front();
        Slot& slot = s->slots[idx];
        lk.unlock();
        float device_ms = 0.0f;
        if (slot.status == CUSBF_OK) {
            hipError_t e = hipEventSynchronize(slot.done);
            if (e == hipSuccess) e = hipEventElapsedTime(&device_ms, slot.start, slot.done);
            if (e != hipSuccess) {
                slot.status = CUSBF_ERR_DEVICE;
                slot.message = hipGetErrorString(e);
            }
        }
        const auto copy_start = Clock::now();
        uint64_t windows = 0;
        struct Part {
            uint8_t* dst;
            uint64_t first;  // the part's first window in the batch
            uint64_t len;
        };
        std::vector<Part> parts;
        for (size_t i = 0; i < slot.batch.size(); ++i) {
            Request* r = slot.batch[i];
            r->status = slot.status;
            r->message = slot.message;
            if (slot.status == CUSBF_OK && r->windows) {
                const uint64_t first = r->batch_offset / s->width;
                for (uint64_t done = 0; done < r->windows; done += kCopyPart) {
                    parts.push_back(Part{r->hits + done, first + done, std::min(kCopyPart, r->windows - done)});
                }
                r->report.queried_kmers = slot.h_counts[2 * i];
                r->report.positive_kmers = slot.h_counts[2 * i + 1];
            }
            windows += r->windows;
        }
        // small parts are grouped so that one task copies about kCopyPart bytes
        std::vector<size_t> task_start{0};
        uint64_t task_bytes = 0;
        for (size_t j = 0; j < parts.size(); ++j) {
            task_bytes += parts[j].len;
            if (task_bytes >= kCopyPart && j + 1 < parts.size()) {
                task_start.push_back(j + 1);
                task_bytes = 0;
            }
        }
        task_start.push_back(parts.size());
        s->pool->run(task_start.size() - 1, [&](size_t t) {
            for (size_t j = task_start[t]; j < task_start[t + 1]; ++j) {
                expand_bits(slot.h_bits, parts[j].first, parts[j].len, parts[j].dst);
            }
        });
        const auto now = Clock::now();
        for (Request* r : slot.batch) r->latency = std::chrono::duration<double>(now - r->submitted).count();
        lk.lock();
        s->stats.windows += windows;
        s->stats.device_seconds += device_ms * 1e-3;
        s->stats.copy_seconds += std::chrono::duration<double>(now - copy_start).count();
        for (Request* r : slot.batch) r->done = true;
        slot.batch.clear();
        s->free_slots.push_back(idx);
        s->cv_slots.notify_one();
        s->cv_done.notify_all();
    }
}

void free_slots(cusbf_server* s) {
    for (Slot& slot : s->slots) {
        if (slot.stream) (void)hipStreamDestroy(slot.stream);
        if (slot.done) (void)hipEventDestroy(slot.done);
        if (slot.start) (void)hipEventDestroy(slot.start);
        if (slot.d_seq) (void)hipFree(slot.d_seq);
        if (slot.d_hits) (void)hipFree(slot.d_hits);
        if (slot.h_seq) (void)hipHostFree(slot.h_seq);
        if (slot.d_bits) (void)hipFree(slot.d_bits);
        if (slot.h_bits) (void)hipHostFree(slot.h_bits);
        if (slot.d_req) (void)hipFree(slot.d_req);
        if (slot.h_req) (void)hipHostFree(slot.h_req);
        if (slot.d_counts) (void)hipFree(slot.d_counts);
        if (slot.h_counts) (void)hipHostFree(slot.h_counts);
    }
}

}  // namespace

extern "C" {

int cusbf_server_create(const cusbf_filter* f, const cusbf_server_options* options,
                        cusbf_server** out) {
    if (!f || !options || !out) return ci::set_error(CUSBF_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (options->slots == 0 || options->slots > 64) {
        return ci::set_error(CUSBF_ERR_INVALID_ARGUMENT, "server slots must be in [1, 64]");
    }
    auto s = std::make_unique<cusbf_server>();
    s->filter = f;
    s->options = *options;
    s->width = ci::symbol_width(f);
    s->sep = static_cast<char>(ci::separator(f));
    s->seq_cap = options->staging_bytes / options->slots / 2;
    if (s->seq_cap < 4096) return ci::set_error(CUSBF_ERR_INVALID_ARGUMENT, "staging budget is too small");
    // Batches of at most kBatchTarget bytes (a larger request still goes alone, up to
    // the slot's staging): under a backlog the slots then take turns, one batch's
    // answers coming back while the next one's bytes go out, instead of every slot
    // filling its whole staging at once.
    s->batch_target = std::min(s->seq_cap, kBatchTarget);
    if (const char* env = std::getenv("CUSBF_SERVER_BATCH_BYTES")) {
        const uint64_t v = std::strtoull(env, nullptr, 10);
        if (v) s->batch_target = std::min(s->seq_cap, v);
    }
    s->slots.resize(options->slots);
    for (Slot& slot : s->slots) {
        hipError_t e = hipStreamCreateWithFlags(&slot.stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreate(&slot.start);
        if (e == hipSuccess) e = hipEventCreate(&slot.done);
        if (e == hipSuccess) e = hipMalloc(&slot.d_seq, s->seq_cap);
        if (e == hipSuccess) e = hipMalloc(&slot.d_hits, s->seq_cap);
        if (e == hipSuccess) e = hipHostMalloc(&slot.h_seq, s->seq_cap, hipHostMallocDefault);
        if (e == hipSuccess) e = hipMalloc(&slot.d_bits, bits_bytes(s->seq_cap));
        if (e == hipSuccess) e = hipHostMalloc(&slot.h_bits, bits_bytes(s->seq_cap), hipHostMallocDefault);
        const uint64_t desc = 2 * kMaxBatchRequests * sizeof(uint64_t);
        if (e == hipSuccess) e = hipMalloc(&slot.d_req, desc);
        if (e == hipSuccess) e = hipHostMalloc(&slot.h_req, desc, hipHostMallocDefault);
        if (e == hipSuccess) e = hipMalloc(&slot.d_counts, desc);
        if (e == hipSuccess) e = hipHostMalloc(&slot.h_counts, desc, hipHostMallocDefault);
        if (e != hipSuccess) {
            free_slots(s.get());
            return ci::set_hip_error(e, __FILE__, __LINE__);
        }
    }
    for (int i = static_cast<int>(options->slots) - 1; i >= 0; --i) s->free_slots.push_back(i);
    const unsigned hw = std::thread::hardware_concurrency();
    s->pool = std::make_unique<WorkerPool>(std::min(12u, std::max(2u, hw > 2 ? hw - 2 : 2u)));
    // the slots' pinned buffers touched once now, on the pool, rather than page by
    // page (first-touch faults) inside the batch fills and answer copies
    for (Slot& slot : s->slots) {
        ci::touch_pages(*s->pool, slot.h_seq, s->seq_cap);
        ci::touch_pages(*s->pool, slot.h_bits, bits_bytes(s->seq_cap));
    }
    cusbf_server* raw = s.release();
    raw->scheduler = std::thread(scheduler_loop, raw);
    raw->completer = std::thread(completion_loop, raw);
    *out = raw;
    return CUSBF_OK;
}

void cusbf_server_destroy(cusbf_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_queue.notify_all();
    s->scheduler.join();
    s->completer.join();
    free_slots(s);
    delete s;
}

uint64_t cusbf_server_request_windows(const cusbf_server* s, const cusbf_record_range* ranges,
                                      uint64_t num_ranges) {
    if (!s || (!ranges && num_ranges)) return 0;
    return ci::window_count(s->filter, normalized_size(ranges, num_ranges, s->width));
}

int cusbf_server_submit(cusbf_server* s, const char* h_sequence, uint64_t bytes,
                        const cusbf_record_range* ranges, uint64_t num_ranges, uint8_t* h_hits,
                        uint64_t hits_len, uint64_t* ticket) {
    if (!s || !ticket) return ci::set_error(CUSBF_ERR_INVALID_ARGUMENT, "null argument");
    const int rc = ci::validate_host_ranges(s->filter, bytes, ranges, num_ranges);
    if (rc != CUSBF_OK) return rc;
    auto r = std::make_unique<Request>();
    r->seq = h_sequence;
    r->ranges.assign(ranges, ranges + num_ranges);
    r->hits = h_hits;
    r->norm_bytes = normalized_size(ranges, num_ranges, s->width);
    r->report.records_queried = num_ranges;
    for (uint64_t i = 0; i < num_ranges; ++i) r->report.queried_bases += ranges[i].sequence_bytes;
    r->windows = ci::window_count(s->filter, r->norm_bytes);
    if (hits_len < r->windows) return ci::set_error(CUSBF_ERR_INVALID_ARGUMENT, "request hit buffer is too small");
    if (r->norm_bytes > s->seq_cap) {
        return ci::set_error(CUSBF_ERR_RESOURCE, "request exceeds the per-batch staging budget");
    }
    r->submitted = Clock::now();
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return ci::set_error(CUSBF_ERR_RESOURCE, "server is shutting down");
    r->ticket = s->next_ticket++;
    *ticket = r->ticket;
    Request* raw = r.get();
    s->requests.emplace(raw->ticket, std::move(r));
    if (raw->norm_bytes == 0) {
        raw->done = true;  // nothing to query
    } else {
        s->queue.push_back(raw);
        s->cv_queue.notify_one();
    }
    return CUSBF_OK;
}

int cusbf_server_wait(cusbf_server* s, uint64_t ticket, cusbf_query_report* report,
                      double* latency_seconds) {
    if (!s) return ci::set_error(CUSBF_ERR_INVALID_ARGUMENT, "null server");
    std::unique_ptr<Request> r;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->requests.find(ticket);
        if (it == s->requests.end()) return ci::set_error(CUSBF_ERR_INVALID_ARGUMENT, "unknown ticket");
        Request* raw = it->second.get();
        s->cv_done.wait(lk, [&] { return raw->done; });
        r = std::move(it->second);
        s->requests.erase(it);
    }
    if (latency_seconds) *latency_seconds = r->latency;
    if (r->status != CUSBF_OK) return ci::set_error(r->status, r->message.c_str());
    // FastxQueryReport of the request (filter.cuh:951-990): records and bases
    // from the submit, valid and positive windows counted on the device.
    if (report) *report = r->report;
    return CUSBF_OK;
}

int cusbf_server_stats_get(const cusbf_server* s, cusbf_server_stats* out) {
    if (!s || !out) return ci::set_error(CUSBF_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard<std::mutex> lk(const_cast<cusbf_server*>(s)->mu);
    *out = s->stats;
    return CUSBF_OK;
}

}  // extern "C"
