// cusbf/filter.hpp -- cusbf::filter<Config> for MI355X, over the C ABI of libcusbf_hip.so.
//
// Host C++ drop-in for the hot-path part of the reference's include/cusbf/filter.cuh
// (insert/query of byte sequences, dense packed sequences and record batches,
// clear/load_factor/filter_bits/num_shards) with the same method names,
// argument order, counting conventions and Result<T> errors. Every call
// forwards to an extern "C" entry point of include/cusbf_hip.h; the kernels
// are hand-written HIP for gfx950. Streams are hipStream_t (nullptr = default).
//
// Also here, as in the reference: RecordRange / RecordBatchView (Fastx.hpp:23-37),
// the insert/query reports (Fastx.hpp:135-170), record normalisation
// (normalized_record_batch.hpp), the query layout (detail/query_layout.cuh) and
// the host dense packer (dense_packed.hpp:35-49).
#pragma once

#include <cstdint>
#include <cstring>
#include <exception>
#include <filesystem>
#include <istream>
#include <span>
#include <string>
#include <string_view>
#include <type_traits>
#include <utility>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "../cusbf_hip.h"
#include "config.hpp"
#include "error.hpp"
#include "filter_ref.hpp"

namespace cusbf {

/// Non-owning view of device memory (device_span.cuh:20-33).
template <typename T>
class device_span {
   public:
    constexpr device_span() noexcept = default;
    constexpr device_span(T* data, size_t size) noexcept : data_(data), size_(size) {}
    template <typename U>
        requires std::is_convertible_v<U (*)[], T (*)[]>
    constexpr device_span(device_span<U> other) noexcept : data_(other.data()), size_(other.size()) {}
    [[nodiscard]] constexpr T* data() const noexcept { return data_; }
    [[nodiscard]] constexpr size_t size() const noexcept { return size_; }
    [[nodiscard]] constexpr bool empty() const noexcept { return size_ == 0; }

   private:
    T* data_ = nullptr;
    size_t size_ = 0;
};

struct RecordRange {
    uint64_t sequenceOffset{};
    uint64_t sequenceBytes{};
};
static_assert(sizeof(RecordRange) == sizeof(cusbf_record_range));

struct RecordBatchView {
    std::string_view sequence{};
    std::span<const RecordRange> records{};
};

struct FastxInsertReport {
    uint64_t recordsIndexed{};
    uint64_t indexedBases{};
    uint64_t insertedKmers{};
};

struct FastxQueryReport {
    uint64_t recordsQueried{};
    uint64_t queriedBases{};
    uint64_t queriedKmers{};
    uint64_t positive_kmers{};
};

/// Fastx.hpp:172-190: one record of a FASTX query (hits valid during the call only).
struct FastxRecordView {
    uint64_t record_index{};
    std::string_view header{};
    std::string_view sequence{};
    uint64_t queriedBases{};
    uint64_t queriedKmers{};
    uint64_t positive_kmers{};
    std::span<const uint8_t> hits{};
};

/// Fastx.hpp:193-208
struct FastxDetailedQueryRecord {
    uint64_t record_index{};
    std::string header;
    std::string sequence;
    uint64_t queriedBases{};
    uint64_t queriedKmers{};
    uint64_t positive_kmers{};
    std::vector<uint8_t> hits;
};

/// Fastx.hpp:211-216
struct FastxDetailedQueryReport {
    FastxQueryReport summary{};
    std::vector<FastxDetailedQueryRecord> records;
};

/// Fastx.hpp:230-238
template <typename Functor>
concept FastxRecordConsumer = requires(Functor f, const FastxRecordView& record) {
    { f(record) } -> std::same_as<void>;
};

struct RecordQueryView {
    uint64_t record_index{};
    std::string_view sequence{};
    uint64_t queriedBases{};
    uint64_t queriedKmers{};
    uint64_t positive_kmers{};
    std::span<const uint8_t> hits{};
};

template <typename Functor>
concept RecordQueryConsumer = requires(Functor f, const RecordQueryView& view) {
    { f(view) } -> std::same_as<void>;
};

struct NormalizedRecord {
    uint64_t record_index{};
    uint64_t input_offset{};
    uint64_t size{};
    uint64_t valid_kmers{};
};

namespace detail {

template <typename Config>
[[nodiscard]] constexpr uint64_t record_symbol_count(uint64_t bytes) noexcept {
    return bytes / Config::symbolWidth;
}

template <typename Config>
[[nodiscard]] constexpr uint64_t record_kmer_count(uint64_t bytes) noexcept {
    const uint64_t symbols = record_symbol_count<Config>(bytes);
    return symbols < Config::k ? 0 : symbols - Config::k + 1;
}

/// normalized_record_batch.hpp:87-110
template <typename Config>
[[nodiscard]] uint64_t valid_record_kmer_count(std::string_view seq) {
    const uint64_t n = record_kmer_count<Config>(seq.size());
    if (n == 0) return 0;
    using A = typename Config::Alphabet;
    auto bad = [&](uint64_t i) -> uint64_t {
        return A::encode(seq.data() + i * Config::symbolWidth) == A::invalidSymbol;
    };
    uint64_t invalid = 0;
    for (uint64_t i = 0; i < Config::k; ++i) invalid += bad(i);
    uint64_t valid = invalid == 0;
    for (uint64_t start = 1; start < n; ++start) {
        invalid -= bad(start - 1);
        invalid += bad(start + Config::k - 1);
        valid += invalid == 0;
    }
    return valid;
}

/// normalized_record_batch.hpp:123-147
template <typename Config>
[[nodiscard]] Result<void> validate_record_batch(uint64_t bytes, std::span<const RecordRange> records,
                                                 const char* exceeds) {
    uint64_t next = 0;
    for (const RecordRange& r : records) {
        if (r.sequenceOffset < next) {
            return Err(Error::invalid_argument("record batch ranges must be ordered and non-overlapping"));
        }
        if (r.sequenceOffset > bytes || r.sequenceBytes > bytes - r.sequenceOffset) {
            return Err(Error::invalid_argument(exceeds));
        }
        if (r.sequenceOffset % Config::symbolWidth != 0 || r.sequenceBytes % Config::symbolWidth != 0) {
            return Err(Error::invalid_argument(
                "record batch ranges must align to the configured alphabet symbol width"));
        }
        next = r.sequenceOffset + r.sequenceBytes;
    }
    return {};
}

struct QueryLayoutRecord {
    uint64_t record_index{};
    uint64_t input_offset{};
    uint64_t size{};
    uint64_t valid_kmers{};
    uint64_t hit_offset{};
    uint64_t hit_count{};
};

/// detail/query_layout.cuh:41-72: hit spans per record in the normalised buffer.
class QueryLayout {
   public:
    [[nodiscard]] std::span<const QueryLayoutRecord> records() const noexcept { return records_; }
    [[nodiscard]] uint64_t total_hit_count() const noexcept { return total_; }
    [[nodiscard]] std::span<const uint8_t> hits_for_record(std::span<const uint8_t> hits,
                                                           size_t i) const noexcept {
        return hits.subspan(records_[i].hit_offset, records_[i].hit_count);
    }

    template <typename Config>
    [[nodiscard]] static QueryLayout build(std::span<const NormalizedRecord> records) {
        QueryLayout layout;
        layout.records_.reserve(records.size());
        uint64_t logical = 0;
        for (const NormalizedRecord& r : records) {
            if (logical != 0) {
                const uint64_t rem = logical % Config::symbolWidth;
                if (rem != 0) logical += Config::symbolWidth - rem;
                logical += Config::symbolWidth;
            }
            layout.records_.push_back(QueryLayoutRecord{r.record_index, r.input_offset, r.size,
                                                        r.valid_kmers, logical / Config::symbolWidth,
                                                        record_kmer_count<Config>(r.size)});
            logical += r.size;
        }
        layout.total_ = record_kmer_count<Config>(logical);
        return layout;
    }

   private:
    std::vector<QueryLayoutRecord> records_{};
    uint64_t total_{};
};

inline Unexpected status_error(int status) {
    char buf[4096];
    cusbf_last_error(buf, sizeof(buf));
    const auto category = static_cast<ErrorCategory>(status - 1);
    if (category == ErrorCategory::fastx_parse) {
        char file[4096];
        uint32_t line = 0, column = 0;
        cusbf_last_error_location(file, sizeof(file), &line, &column);
        return Err(Error::fastx_parse(SourceLocation{file, line, column}, buf));
    }
    return Err(Error(category, buf));
}

inline Result<void> check(int status) {
    if (status == CUSBF_OK) return {};
    return status_error(status);
}

// Reader over a std::istream for cusbf_*_fastx_reader.
inline int64_t istream_read(void* user, char* dst, uint64_t capacity) {
    auto* in = static_cast<std::istream*>(user);
    in->read(dst, static_cast<std::streamsize>(capacity));
    const std::streamsize got = in->gcount();
    if (got > 0) return got;
    return in->bad() ? -1 : 0;
}

// A C++ consumer behind the C callback: an exception it throws stops the
// query and is rethrown to the caller once the library has returned.
template <typename Consumer>
struct ConsumerBox {
    Consumer* consume;
    std::exception_ptr error;

    static int call(const cusbf_fastx_record_view* v, void* user) {
        auto* box = static_cast<ConsumerBox*>(user);
        try {
            (*box->consume)(FastxRecordView{
                v->record_index, std::string_view(v->header, v->header_len),
                std::string_view(v->sequence, v->sequence_len), v->queried_bases, v->queried_kmers,
                v->positive_kmers, std::span<const uint8_t>(v->hits, v->hits_len)});
            return 0;
        } catch (...) {
            box->error = std::current_exception();
            return 1;
        }
    }
};

inline FastxInsertReport insert_report(const cusbf_insert_report& r) {
    return FastxInsertReport{r.records_indexed, r.indexed_bases, r.inserted_kmers};
}

inline FastxQueryReport query_report(const cusbf_query_report& r) {
    return FastxQueryReport{r.records_queried, r.queried_bases, r.queried_kmers, r.positive_kmers};
}

template <typename Config>
cusbf_config c_config() {
    using A = typename Config::Alphabet;
    cusbf_config c{};
    uint32_t code = CUSBF_ALPHABET_CUSTOM;
    if constexpr (std::is_same_v<A, DnaAlphabet>) {
        code = CUSBF_ALPHABET_DNA;
    } else if constexpr (std::is_same_v<A, DnaTripletAlphabet>) {
        code = CUSBF_ALPHABET_DNA_TRIPLET;
    } else if constexpr (std::is_same_v<A, ProteinAlphabet>) {
        code = CUSBF_ALPHABET_PROTEIN;
    }
    if (code != CUSBF_ALPHABET_CUSTOM) {
        cusbf_config_init(&c, Config::k, Config::s, Config::m, Config::hashCount, code);
        return c;
    }
    static_assert(std::is_same_v<A, DnaAlphabet> || std::is_same_v<A, DnaTripletAlphabet> ||
                      std::is_same_v<A, ProteinAlphabet> || A::symbolWidth == 1,
                  "custom alphabets must read one byte per symbol");
    c.k = Config::k;
    c.s = Config::s;
    c.m = Config::m;
    c.hash_count = Config::hashCount;
    c.alphabet = CUSBF_ALPHABET_CUSTOM;
    c.symbol_count = static_cast<uint32_t>(A::symbolCount);
    c.invalid_symbol = A::invalidSymbol;
    c.separator = A::separator;
    for (int b = 0; b < 256; ++b) {
        const char byte = static_cast<char>(b);
        c.lut[b] = A::encode(&byte);
    }
    return c;
}

}  // namespace detail

/// normalized_record_batch.hpp:288-304: separators between records, per-record metadata.
template <typename Config>
[[nodiscard]] Result<void> normalize_record_batch_into(RecordBatchView batch, std::string& sequence_out,
                                                       std::vector<NormalizedRecord>& records_out) {
    CUSBF_TRY(detail::validate_record_batch<Config>(batch.sequence.size(), batch.records,
                                                    "record batch range exceeds the source sequence"));
    sequence_out.clear();
    records_out.clear();
    const char sep = static_cast<char>(Config::Alphabet::separator);
    for (uint64_t i = 0; i < batch.records.size(); ++i) {
        const RecordRange& r = batch.records[i];
        if (!sequence_out.empty()) {
            const uint64_t rem = sequence_out.size() % Config::symbolWidth;
            if (rem != 0) sequence_out.append(Config::symbolWidth - rem, sep);
            sequence_out.append(Config::symbolWidth, sep);
        }
        const std::string_view rec = batch.sequence.substr(r.sequenceOffset, r.sequenceBytes);
        sequence_out.append(rec);
        records_out.push_back(NormalizedRecord{i, r.sequenceOffset, rec.size(),
                                               detail::valid_record_kmer_count<Config>(rec)});
    }
    return {};
}

class NormalizedRecordBatch {
   public:
    [[nodiscard]] const std::string& sequence() const noexcept { return sequence_; }
    [[nodiscard]] std::span<const NormalizedRecord> records() const noexcept { return records_; }
    [[nodiscard]] uint64_t total_valid_kmers() const noexcept {
        uint64_t t = 0;
        for (const NormalizedRecord& r : records_) t += r.valid_kmers;
        return t;
    }

    std::string sequence_;
    std::vector<NormalizedRecord> records_;
};

template <typename Config>
[[nodiscard]] Result<NormalizedRecordBatch> normalize_record_batch(RecordBatchView batch) {
    NormalizedRecordBatch out;
    CUSBF_TRY(normalize_record_batch_into<Config>(batch, out.sequence_, out.records_));
    return out;
}

/// dense_packed.cuh:19-28
template <typename Config>
[[nodiscard]] constexpr uint64_t dense_packed_word_count(uint64_t num_symbols) {
    constexpr uint64_t per = 64 / Config::symbolBits;
    return (num_symbols + per - 1) / per;
}

template <typename Config>
[[nodiscard]] constexpr uint64_t dense_packed_kmer_count(uint64_t num_symbols) {
    return num_symbols < Config::k ? 0 : num_symbols - Config::k + 1;
}

/// dense_packed.hpp:35-49: LSB-first host packing.
template <typename Config>
[[nodiscard]] std::vector<uint64_t> pack_dense_sequence(std::string_view sequence) {
    const uint64_t n = sequence.size() / Config::symbolWidth;
    constexpr uint64_t per = 64 / Config::symbolBits;
    std::vector<uint64_t> words(dense_packed_word_count<Config>(n), 0);
    for (uint64_t i = 0; i < n; ++i) {
        const uint64_t sym = Config::Alphabet::encode(sequence.data() + i * Config::symbolWidth) &
                             Config::symbolMask;
        words[i / per] |= sym << ((i % per) * Config::symbolBits);
    }
    return words;
}

/// dense_packed.hpp:57-73
template <typename Config>
[[nodiscard]] Result<void> pack_dense_sequence_device(const char* d_sequence, uint64_t num_symbols,
                                                      uint64_t* d_words, hipStream_t stream = nullptr) {
    const cusbf_config c = detail::c_config<Config>();
    return detail::check(cusbf_pack_dense_sequence_async(
        &c, d_sequence, num_symbols, d_words, dense_packed_word_count<Config>(num_symbols), stream));
}

template <typename Config>
class filter {
   public:
    static_assert(Config::blockWordCount == 4, "Filter only supports the fused 256-bit shard path");

    /// filter.cuh:108-120; throws std::runtime_error when the device allocation fails.
    explicit filter(uint64_t requestedFilterBits) {
        const cusbf_config c = detail::c_config<Config>();
        const int st = cusbf_create(&c, requestedFilterBits, &handle_);
        if (st != CUSBF_OK) throw std::runtime_error(detail::status_error(st).error.message());
    }
    filter(const filter&) = delete;
    filter& operator=(const filter&) = delete;
    filter(filter&& o) noexcept : handle_(std::exchange(o.handle_, nullptr)) {}
    filter& operator=(filter&& o) noexcept {
        if (this != &o) {
            cusbf_destroy(handle_);
            handle_ = std::exchange(o.handle_, nullptr);
        }
        return *this;
    }
    ~filter() { cusbf_destroy(handle_); }

    /// filter.cuh:137-142: view of the shard storage for device code (device members in
    /// cusbf/filter_ref_device.hpp). The library stops assuming the words are zero once
    /// the view has been handed out. If the device refuses that step the view is empty
    /// (blocks() == nullptr, num_blocks() == 0); checked_ref() returns the error instead.
    [[nodiscard]] filter_ref<Config> ref() const noexcept {
        uint64_t* w = nullptr;
        if (cusbf_words_device_ptr(handle_, &w) != CUSBF_OK) return filter_ref<Config>{};
        return filter_ref<Config>{reinterpret_cast<filter_block<Config>*>(w), num_shards()};
    }
    [[nodiscard]] Result<filter_ref<Config>> checked_ref() const {
        uint64_t* w = nullptr;
        const int st = cusbf_words_device_ptr(handle_, &w);
        if (st != CUSBF_OK) return detail::status_error(st);
        return filter_ref<Config>{reinterpret_cast<filter_block<Config>*>(w), num_shards()};
    }

    // --- byte sequences (filter.cuh:154-190, 464-533) ---
    [[nodiscard]] Result<uint64_t> insert_sequence(std::string_view sequence, hipStream_t stream = nullptr) {
        uint64_t n = 0;
        CUSBF_TRY(detail::check(cusbf_insert_sequence(handle_, sequence.data(), sequence.size(), stream, &n)));
        return n;
    }
    [[nodiscard]] Result<uint64_t> insert_sequence_async(device_span<const char> d_sequence,
                                                         hipStream_t stream = nullptr) {
        uint64_t n = 0;
        CUSBF_TRY(detail::check(
            cusbf_insert_sequence_async(handle_, d_sequence.data(), d_sequence.size(), stream, &n)));
        return n;
    }
    [[nodiscard]] Result<void> contains_sequence_async(device_span<const char> d_sequence,
                                                       device_span<uint8_t> d_output,
                                                       hipStream_t stream = nullptr) const {
        return detail::check(cusbf_contains_sequence_async(handle_, d_sequence.data(), d_sequence.size(),
                                                           d_output.data(), d_output.size(), stream));
    }
    [[nodiscard]] Result<void> contains_sequence(std::string_view sequence, device_span<uint8_t> d_output,
                                                 hipStream_t stream = nullptr) const {
        return detail::check(cusbf_contains_sequence_host_to_device(
            handle_, sequence.data(), sequence.size(), d_output.data(), d_output.size(), stream));
    }
    [[nodiscard]] Result<std::vector<uint8_t>> contains_sequence(std::string_view sequence,
                                                                 hipStream_t stream = nullptr) const {
        std::vector<uint8_t> out(detail::record_kmer_count<Config>(sequence.size()));
        if (out.empty()) return out;
        CUSBF_TRY(detail::check(cusbf_contains_sequence(handle_, sequence.data(), sequence.size(),
                                                        out.data(), out.size(), stream)));
        return out;
    }

    // --- dense packed (filter.cuh:207-322) ---
    [[nodiscard]] Result<uint64_t> insert_dense_packed_async(device_span<const uint64_t> d_words,
                                                             uint64_t num_symbols,
                                                             hipStream_t stream = nullptr) {
        uint64_t n = 0;
        CUSBF_TRY(detail::check(cusbf_insert_dense_packed_async(handle_, d_words.data(), d_words.size(),
                                                                num_symbols, stream, &n)));
        return n;
    }
    [[nodiscard]] Result<uint64_t> insert_dense_packed(std::span<const uint64_t> words, uint64_t num_symbols,
                                                       hipStream_t stream = nullptr) {
        uint64_t n = 0;
        CUSBF_TRY(detail::check(
            cusbf_insert_dense_packed(handle_, words.data(), words.size(), num_symbols, stream, &n)));
        return n;
    }
    [[nodiscard]] Result<void> contains_dense_packed_async(device_span<const uint64_t> d_words,
                                                           uint64_t num_symbols,
                                                           device_span<uint8_t> d_output,
                                                           hipStream_t stream = nullptr) const {
        return detail::check(cusbf_contains_dense_packed_async(handle_, d_words.data(), d_words.size(),
                                                               num_symbols, d_output.data(),
                                                               d_output.size(), stream));
    }
    [[nodiscard]] Result<void> contains_dense_packed(std::span<const uint64_t> words, uint64_t num_symbols,
                                                     device_span<uint8_t> d_output,
                                                     hipStream_t stream = nullptr) const {
        return detail::check(cusbf_contains_dense_packed_host_to_device(
            handle_, words.data(), words.size(), num_symbols, d_output.data(), d_output.size(), stream));
    }
    [[nodiscard]] Result<std::vector<uint8_t>> contains_dense_packed(std::span<const uint64_t> words,
                                                                     uint64_t num_symbols,
                                                                     hipStream_t stream = nullptr) const {
        std::vector<uint8_t> out(cusbf::dense_packed_kmer_count<Config>(num_symbols));
        if (out.empty()) return out;
        CUSBF_TRY(detail::check(cusbf_contains_dense_packed(handle_, words.data(), words.size(), num_symbols,
                                                            out.data(), out.size(), stream)));
        return out;
    }
    [[nodiscard]] static constexpr uint64_t dense_packed_word_count(uint64_t num_symbols) {
        return cusbf::dense_packed_word_count<Config>(num_symbols);
    }
    [[nodiscard]] constexpr uint64_t dense_packed_kmer_count(uint64_t num_symbols) const {
        return cusbf::dense_packed_kmer_count<Config>(num_symbols);
    }

    // --- record batches (filter.cuh:338-400, 549-618) ---
    [[nodiscard]] Result<FastxInsertReport> insert_record_batch(RecordBatchView batch,
                                                                hipStream_t stream = nullptr) {
        cusbf_insert_report r{};
        CUSBF_TRY(detail::check(cusbf_insert_record_batch(
            handle_, batch.sequence.data(), batch.sequence.size(), ranges_of(batch.records),
            batch.records.size(), stream, &r)));
        return FastxInsertReport{r.records_indexed, r.indexed_bases, r.inserted_kmers};
    }
    [[nodiscard]] Result<uint64_t> insert_record_batch_async(device_span<const char> d_sequence,
                                                             std::span<const RecordRange> records,
                                                             hipStream_t stream = nullptr) {
        uint64_t n = 0;
        CUSBF_TRY(detail::check(cusbf_insert_record_batch_async(handle_, d_sequence.data(), d_sequence.size(),
                                                                ranges_of(records), records.size(), stream,
                                                                &n)));
        return n;
    }
    [[nodiscard]] Result<FastxQueryReport> query_record_batch(RecordBatchView batch,
                                                              hipStream_t stream = nullptr) const {
        cusbf_query_report r{};
        CUSBF_TRY(detail::check(cusbf_query_record_batch(
            handle_, batch.sequence.data(), batch.sequence.size(), ranges_of(batch.records),
            batch.records.size(), stream, &r)));
        return FastxQueryReport{r.records_queried, r.queried_bases, r.queried_kmers, r.positive_kmers};
    }
    /// filter.cuh:566-582, 1007-1131: per-record hits delivered to `consume` in source order.
    template <RecordQueryConsumer Consumer>
    [[nodiscard]] Result<FastxQueryReport> query_record_batch(RecordBatchView batch, Consumer&& consume,
                                                              hipStream_t stream = nullptr) const {
        std::string norm;
        std::vector<NormalizedRecord> records;
        CUSBF_TRY(normalize_record_batch_into<Config>(batch, norm, records));
        const detail::QueryLayout layout = detail::QueryLayout::build<Config>(records);
        FastxQueryReport report;
        report.recordsQueried = records.size();
        for (const NormalizedRecord& r : records) {
            report.queriedBases += r.size;
            report.queriedKmers += r.valid_kmers;
        }
        if (norm.empty()) return report;
        std::vector<uint8_t> hits(layout.total_hit_count());
        if (!hits.empty()) {
            CUSBF_TRY(detail::check(cusbf_contains_sequence(handle_, norm.data(), norm.size(), hits.data(),
                                                            hits.size(), stream)));
        }
        const std::span<const uint8_t> all(hits);
        for (size_t i = 0; i < layout.records().size(); ++i) {
            const detail::QueryLayoutRecord& r = layout.records()[i];
            const std::string_view seq = batch.sequence.substr(r.input_offset, r.size);
            std::span<const uint8_t> h{};
            uint64_t positive = 0;
            if (r.hit_count != 0) {
                h = layout.hits_for_record(all, i);
                for (uint8_t x : h) positive += x;
            }
            report.positive_kmers += positive;
            consume(RecordQueryView{r.record_index, seq, r.size, r.valid_kmers, positive, h});
        }
        return report;
    }
    [[nodiscard]] Result<void> contains_record_batch_async(device_span<const char> d_sequence,
                                                           std::span<const RecordRange> records,
                                                           device_span<uint8_t> d_output,
                                                           hipStream_t stream = nullptr) const {
        return detail::check(cusbf_contains_record_batch_async(handle_, d_sequence.data(), d_sequence.size(),
                                                               ranges_of(records), records.size(),
                                                               d_output.data(), d_output.size(), stream));
    }

    // --- FASTA / FASTQ (filter.cuh:402-453, 620-800) ---
    [[nodiscard]] Result<FastxInsertReport> insert_fastx(std::istream& input, double fill_fraction = 0.7,
                                                         hipStream_t stream = nullptr) {
        cusbf_insert_report r{};
        CUSBF_TRY(detail::check(cusbf_insert_fastx_reader(handle_, &detail::istream_read, &input,
                                                          "<stream>", fill_fraction, stream, &r)));
        return detail::insert_report(r);
    }
    [[nodiscard]] Result<FastxInsertReport> insert_fastx_file(const std::filesystem::path& path,
                                                              double fill_fraction = 0.7,
                                                              hipStream_t stream = nullptr) {
        cusbf_insert_report r{};
        CUSBF_TRY(detail::check(cusbf_insert_fastx_file(handle_, path.c_str(), fill_fraction, stream, &r)));
        return detail::insert_report(r);
    }
    [[nodiscard]] Result<FastxQueryReport> query_fastx(std::istream& input, double fill_fraction = 0.7,
                                                       hipStream_t stream = nullptr) const {
        cusbf_query_report r{};
        CUSBF_TRY(detail::check(cusbf_query_fastx_reader(handle_, &detail::istream_read, &input, "<stream>",
                                                         fill_fraction, stream, nullptr, nullptr, &r)));
        return detail::query_report(r);
    }
    [[nodiscard]] Result<FastxQueryReport> query_fastx_file(const std::filesystem::path& path,
                                                            double fill_fraction = 0.7,
                                                            hipStream_t stream = nullptr) const {
        cusbf_query_report r{};
        CUSBF_TRY(detail::check(cusbf_query_fastx_file(handle_, path.c_str(), fill_fraction, stream, nullptr,
                                                       nullptr, &r)));
        return detail::query_report(r);
    }
    template <FastxRecordConsumer Consumer>
    [[nodiscard]] Result<FastxQueryReport> query_fastx_records(std::istream& input, Consumer&& consume,
                                                               double fill_fraction = 0.7,
                                                               hipStream_t stream = nullptr) const {
        using C = std::remove_reference_t<Consumer>;
        detail::ConsumerBox<C> box{&consume, nullptr};
        cusbf_query_report r{};
        const int st = cusbf_query_fastx_reader(handle_, &detail::istream_read, &input, "<stream>",
                                                fill_fraction, stream, &detail::ConsumerBox<C>::call, &box, &r);
        if (box.error) std::rethrow_exception(box.error);
        CUSBF_TRY(detail::check(st));
        return detail::query_report(r);
    }
    template <FastxRecordConsumer Consumer>
    [[nodiscard]] Result<FastxQueryReport> query_fastx_file_records(const std::filesystem::path& path,
                                                                    Consumer&& consume,
                                                                    double fill_fraction = 0.7,
                                                                    hipStream_t stream = nullptr) const {
        using C = std::remove_reference_t<Consumer>;
        detail::ConsumerBox<C> box{&consume, nullptr};
        cusbf_query_report r{};
        const int st = cusbf_query_fastx_file(handle_, path.c_str(), fill_fraction, stream,
                                              &detail::ConsumerBox<C>::call, &box, &r);
        if (box.error) std::rethrow_exception(box.error);
        CUSBF_TRY(detail::check(st));
        return detail::query_report(r);
    }
    [[nodiscard]] Result<FastxDetailedQueryReport> query_fastx_detailed(std::istream& input,
                                                                        double fill_fraction = 0.7,
                                                                        hipStream_t stream = nullptr) const {
        FastxDetailedQueryReport out;
        out.summary = CUSBF_TRY(query_fastx_records(
            input, [&](const FastxRecordView& v) { out.records.push_back(detailed(v)); }, fill_fraction, stream));
        return out;
    }
    [[nodiscard]] Result<FastxDetailedQueryReport> query_fastx_file_detailed(const std::filesystem::path& path,
                                                                             double fill_fraction = 0.7,
                                                                             hipStream_t stream = nullptr) const {
        FastxDetailedQueryReport out;
        out.summary = CUSBF_TRY(query_fastx_file_records(
            path, [&](const FastxRecordView& v) { out.records.push_back(detailed(v)); }, fill_fraction, stream));
        return out;
    }

    // --- batched record form (no single-call counterpart in the reference) ---
    // Records back to back in d_symbols, d_offsets = num_records + 1 symbol
    // offsets (device memory, non-decreasing from 0 to d_symbols' symbol
    // count). Windows inside a record behave as insert_sequence /
    // contains_sequence on that record alone; windows across a record start
    // are skipped / answered 0. Query output: one byte per window start.
    [[nodiscard]] Result<void> insert_batched_async(device_span<const char> d_symbols,
                                                    device_span<const uint64_t> d_offsets,
                                                    hipStream_t stream = nullptr) {
        return detail::check(cusbf_insert_batched_async(handle_, d_symbols.data(), symbols_of(d_symbols),
                                                        d_offsets.data(), records_of(d_offsets), stream));
    }
    [[nodiscard]] Result<void> contains_batched_async(device_span<const char> d_symbols,
                                                      device_span<const uint64_t> d_offsets,
                                                      device_span<uint8_t> d_output,
                                                      hipStream_t stream = nullptr) const {
        return detail::check(cusbf_contains_batched_async(handle_, d_symbols.data(), symbols_of(d_symbols),
                                                          d_offsets.data(), records_of(d_offsets),
                                                          d_output.data(), d_output.size(), stream));
    }
    /// Synchronises; invalid_argument when the last batched call's offsets were malformed.
    [[nodiscard]] Result<void> batched_status(hipStream_t stream = nullptr) const {
        return detail::check(cusbf_batched_status(handle_, stream));
    }
    [[nodiscard]] Result<void> set_batch_tuning(uint32_t bucket, cusbf_batch_op op,
                                                const cusbf_batch_tuning& tuning) {
        return detail::check(cusbf_batch_tuning_set(handle_, bucket, op, &tuning));
    }
    [[nodiscard]] Result<cusbf_batch_tuning> batch_tuning(uint32_t bucket, cusbf_batch_op op) const {
        cusbf_batch_tuning t{};
        CUSBF_TRY(detail::check(cusbf_batch_tuning_get(handle_, bucket, op, &t)));
        return t;
    }
    [[nodiscard]] static uint64_t symbols_of(device_span<const char> d) {
        return d.size() / Config::Alphabet::symbolWidth;
    }
    [[nodiscard]] static uint64_t records_of(device_span<const uint64_t> off) {
        return off.empty() ? 0 : off.size() - 1;
    }

    // --- state (filter.cuh:793-833) ---
    [[nodiscard]] Result<void> clear(hipStream_t stream = nullptr) {
        return detail::check(cusbf_clear(handle_, stream));
    }
    [[nodiscard]] float load_factor() const {
        float lf = 0.0f;
        if (cusbf_load_factor(handle_, &lf) != CUSBF_OK) {
            throw std::runtime_error(detail::status_error(CUSBF_ERR_DEVICE).error.message());
        }
        return lf;
    }
    [[nodiscard]] uint64_t filter_bits() const { return cusbf_filter_bits(handle_); }
    [[nodiscard]] uint64_t num_shards() const { return cusbf_num_shards(handle_); }
    /// All filter words on the host (parity checks).
    [[nodiscard]] Result<std::vector<uint64_t>> words() const {
        std::vector<uint64_t> w(num_shards() * 4);
        CUSBF_TRY(detail::check(cusbf_copy_words_to_host(handle_, w.data(), w.size())));
        return w;
    }
    [[nodiscard]] cusbf_filter* native_handle() const noexcept { return handle_; }

    // --- insert path (no reference counterpart; cusbf_set_insert_path) ---
    enum class InsertPath : uint32_t {
        automatic = CUSBF_INSERT_AUTO,
        atomic = CUSBF_INSERT_ATOMIC,
        sorted = CUSBF_INSERT_SORTED
    };
    /// Atomic ORs per shard word, or records partitioned by filter slice and
    /// applied in LDS; both give identical words. slice_bits 1..12 (0 = 12),
    /// reserve_per_tile 1..1024 (0 = from the minimizer span).
    [[nodiscard]] Result<void> set_insert_path(InsertPath path, uint32_t slice_bits = 0,
                                               uint32_t reserve_per_tile = 0) {
        return detail::check(cusbf_set_insert_path(handle_, static_cast<uint32_t>(path), slice_bits,
                                                   reserve_per_tile));
    }
    /// The path the last insert call took (atomic or sorted).
    [[nodiscard]] InsertPath last_insert_path() const {
        uint32_t last = CUSBF_INSERT_ATOMIC;
        (void)cusbf_get_insert_path(handle_, nullptr, nullptr, nullptr, &last);
        return static_cast<InsertPath>(last);
    }
    /// Bounds the sorted insert's workspace in bytes (0 = half of the free device
    /// memory); a plain insert whose records do not fit runs as several sorted
    /// inserts over consecutive window ranges.
    [[nodiscard]] Result<void> set_sorted_workspace_limit(uint64_t bytes) {
        return detail::check(cusbf_set_sorted_workspace_limit(handle_, bytes));
    }
    /// Sorted inserts the last plain insert ran as (0 when it took the atomic path).
    [[nodiscard]] uint32_t last_sorted_chunks() const {
        uint32_t chunks = 0;
        (void)cusbf_get_sorted_workspace(handle_, nullptr, nullptr, &chunks);
        return chunks;
    }

    // --- filter images (no reference counterpart) ---
    /// Writes header + words to `path` (cusbf_save).
    [[nodiscard]] Result<void> save(const std::filesystem::path& path) const {
        return detail::check(cusbf_save(handle_, path.c_str()));
    }
    /// Recreates a filter from an image; invalid_argument when the image's
    /// configuration is not Config's.
    [[nodiscard]] static Result<filter> load(const std::filesystem::path& path) {
        cusbf_config stored{};
        uint64_t bits = 0;
        CUSBF_TRY(detail::check(cusbf_image_config(path.c_str(), &stored, &bits)));
        const cusbf_config want = detail::c_config<Config>();
        if (stored.k != want.k || stored.s != want.s || stored.m != want.m ||
            stored.hash_count != want.hash_count || stored.alphabet != want.alphabet ||
            std::memcmp(stored.lut, want.lut, sizeof(want.lut)) != 0) {
            return Err(Error::invalid_argument("filter image configuration does not match Config"));
        }
        cusbf_filter* h = nullptr;
        CUSBF_TRY(detail::check(cusbf_load(path.c_str(), &h)));
        return filter(h);
    }

   private:
    explicit filter(cusbf_filter* adopted) noexcept : handle_(adopted) {}

    static FastxDetailedQueryRecord detailed(const FastxRecordView& v) {
        return FastxDetailedQueryRecord{v.record_index, std::string(v.header), std::string(v.sequence),
                                        v.queriedBases, v.queriedKmers, v.positive_kmers,
                                        std::vector<uint8_t>(v.hits.begin(), v.hits.end())};
    }
    static const cusbf_record_range* ranges_of(std::span<const RecordRange> r) {
        return reinterpret_cast<const cusbf_record_range*>(r.data());
    }

    cusbf_filter* handle_ = nullptr;
};

/// A batched insert or query captured once into a hipGraph over fixed buffers
/// and replayed with one launch (cusbf_batch_plan_*). Move-only; the buffers
/// and the filter must outlive it.
class batch_plan {
   public:
    template <typename Config>
    [[nodiscard]] static Result<batch_plan> create(filter<Config>& f, cusbf_batch_op op,
                                                   device_span<const char> d_symbols,
                                                   device_span<const uint64_t> d_offsets,
                                                   device_span<uint8_t> d_output = {},
                                                   const cusbf_batch_tuning* tuning = nullptr) {
        batch_plan p;
        CUSBF_TRY(detail::check(cusbf_batch_plan_create(
            f.native_handle(), op, d_symbols.data(), filter<Config>::symbols_of(d_symbols),
            d_offsets.data(), filter<Config>::records_of(d_offsets), d_output.data(), d_output.size(),
            tuning, &p.handle_)));
        return p;
    }
    batch_plan(const batch_plan&) = delete;
    batch_plan& operator=(const batch_plan&) = delete;
    batch_plan(batch_plan&& o) noexcept : handle_(std::exchange(o.handle_, nullptr)) {}
    batch_plan& operator=(batch_plan&& o) noexcept {
        if (this != &o) {
            cusbf_batch_plan_destroy(handle_);
            handle_ = std::exchange(o.handle_, nullptr);
        }
        return *this;
    }
    ~batch_plan() { cusbf_batch_plan_destroy(handle_); }

    [[nodiscard]] Result<void> launch(hipStream_t stream = nullptr) {
        return detail::check(cusbf_batch_plan_launch(handle_, stream));
    }
    /// Synchronises; invalid_argument when the offsets were malformed.
    [[nodiscard]] Result<void> status(hipStream_t stream = nullptr) {
        return detail::check(cusbf_batch_plan_status(handle_, stream));
    }

   private:
    batch_plan() = default;
    cusbf_batch_plan* handle_ = nullptr;
};

}  // namespace cusbf
