// test_filter_api.cpp -- the reference's filter tests, written against the C++ mirror
// (include/cusbf/filter.hpp) over libcusbf_hip.so.
//
// Mirrors tests/test_bloom_filter.cu (GPU), tests/test_query_layout.cpp,
// tests/test_normalized_record_batch.cpp (host) of the reference, with a tiny
// runner instead of GoogleTest. `--host-only` runs the tests that need no GPU.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <array>
#include <cstdio>
#include <cstring>
#include <functional>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unistd.h>
#include <vector>

#include "cusbf/filter.hpp"

// filter_ref / filter_block (filter_ref.cuh:20-113, filter_impl.cuh:18-77) in host
// constant expressions: the view and the block arithmetic compile without hipcc.
using RefCfg = cusbf::Config<31, 28, 16, 4>;
static_assert(cusbf::filter_ref<RefCfg>::shard_index(0xABCDEFull, 1u << 8) == 0xEF);
static_assert(cusbf::filter_ref<RefCfg>{nullptr, 64}.num_blocks() == 64);
static_assert(cusbf::filter_ref<RefCfg>{nullptr, 64}.shard_index(131) == 3);
static_assert(cusbf::filter_block<RefCfg>::sectorizedBitAddress<1>(0x3Full << 16) == 0x3F);
static_assert(cusbf::filter_block<cusbf::Config<31, 28, 16, 16>>::sectorizedBitAddress<0>(1) ==
              (0x9E3779B97F4A7C15ull >> 58));
static_assert(sizeof(cusbf::filter_block<RefCfg>) == 32 && alignof(cusbf::filter_block<RefCfg>) == 32);

namespace {

struct Case {
    const char* name;
    bool gpu;
    std::function<void()> fn;
};
std::vector<Case>& registry() {
    static std::vector<Case> r;
    return r;
}
struct Reg {
    Reg(const char* n, bool gpu, std::function<void()> f) { registry().push_back({n, gpu, std::move(f)}); }
};
int g_failures = 0;

#define TEST_GPU(name) \
    void name();       \
    Reg reg_##name(#name, true, name); \
    void name()
#define TEST_HOST(name) \
    void name();        \
    Reg reg_##name(#name, false, name); \
    void name()
#define EXPECT_TRUE(c)                                                          \
    do {                                                                        \
        if (!(c)) {                                                             \
            std::fprintf(stderr, "  %s:%d: EXPECT_TRUE(%s)\n", __FILE__, __LINE__, #c); \
            ++g_failures;                                                       \
        }                                                                       \
    } while (0)
#define EXPECT_EQ(a, b) EXPECT_TRUE((a) == (b))

using TestConfig = cusbf::Config<5, 4, 3, 4>;
using ProteinTestConfig = cusbf::Config<5, 4, 3, 4, 256, cusbf::ProteinAlphabet>;
using TripletTestConfig = cusbf::Config<3, 2, 2, 4, 256, cusbf::DnaTripletAlphabet>;
using LayoutConfig = cusbf::Config<3, 2, 2, 4>;

// test_support.cuh:24-44
struct CustomAlphabet {
    static constexpr uint64_t symbolWidth = 1;
    static constexpr uint64_t symbolCount = 3;
    static constexpr uint8_t invalidSymbol = 7;
    static constexpr uint8_t separator = '!';
    static constexpr char validBytes[] = "xyz";
    [[nodiscard]] static constexpr uint8_t encode(const char* input) {
        switch (static_cast<uint8_t>(input[0])) {
            case 'x': return 0;
            case 'y': return 1;
            case 'z': return 2;
            default: return invalidSymbol;
        }
    }
};
using CustomAlphabetTestConfig = cusbf::Config<3, 2, 2, 4, 256, CustomAlphabet>;

static_assert(TestConfig::symbolBits == 2);
static_assert(ProteinTestConfig::symbolBits == 5);
static_assert(TripletTestConfig::symbolBits == 6);
static_assert(TripletTestConfig::symbolWidth == 3);
static_assert(CustomAlphabetTestConfig::symbolBits == 2);

bool all_ones(const std::vector<uint8_t>& v) {
    return std::all_of(v.begin(), v.end(), [](uint8_t x) { return x == 1; });
}

template <typename T>
struct DeviceBuffer {
    T* ptr = nullptr;
    size_t n = 0;
    explicit DeviceBuffer(size_t count) : n(count) {
        if (hipMalloc(&ptr, std::max<size_t>(1, count) * sizeof(T)) != hipSuccess) throw std::runtime_error("hipMalloc");
    }
    DeviceBuffer(const T* host, size_t count) : DeviceBuffer(count) {
        if (count) (void)hipMemcpy(ptr, host, count * sizeof(T), hipMemcpyHostToDevice);
    }
    ~DeviceBuffer() { (void)hipFree(ptr); }
    std::vector<T> host() const {
        std::vector<T> v(n);
        if (n) (void)hipMemcpy(v.data(), ptr, n * sizeof(T), hipMemcpyDeviceToHost);
        return v;
    }
};

// ---- tests/test_bloom_filter.cu ------------------------------------------------------

TEST_GPU(InsertAndQuerySameSequenceHasNoFalseNegatives) {
    cusbf::filter<TestConfig> filter(1 << 12);
    const std::string sequence = "ACGTACGTACGTACGT";
    const uint64_t inserted = CUSBF_UNWRAP(filter.insert_sequence(sequence));
    const auto hits = CUSBF_UNWRAP(filter.contains_sequence(sequence));
    EXPECT_EQ(inserted, sequence.size() - TestConfig::k + 1);
    EXPECT_EQ(hits.size(), inserted);
    EXPECT_TRUE(all_ones(hits));
}

TEST_GPU(DensePackedMatchesByteSequenceInsertAndQuery) {
    cusbf::filter<TestConfig> filter(1 << 12);
    const std::string sequence = "ACGTACGTACGTACGTACGTACGT";
    const std::vector<uint64_t> packed = cusbf::pack_dense_sequence<TestConfig>(sequence);
    const uint64_t inserted = CUSBF_UNWRAP(filter.insert_dense_packed(packed, sequence.size()));
    const auto hits = CUSBF_UNWRAP(filter.contains_dense_packed(packed, sequence.size()));
    EXPECT_EQ(inserted, sequence.size() - TestConfig::k + 1);
    EXPECT_EQ(hits.size(), inserted);
    EXPECT_TRUE(all_ones(hits));
    EXPECT_TRUE(hits == CUSBF_UNWRAP(filter.contains_sequence(sequence)));
}

TEST_GPU(DensePackedDeviceAsyncMatchesHostPath) {
    cusbf::filter<TestConfig> filter(1 << 12);
    const std::string sequence = "ACGTACGTACGTACGT";
    const std::vector<uint64_t> packed = cusbf::pack_dense_sequence<TestConfig>(sequence);
    DeviceBuffer<uint64_t> d_packed(packed.data(), packed.size());
    const uint64_t inserted = CUSBF_UNWRAP(filter.insert_dense_packed_async(
        cusbf::device_span<const uint64_t>{d_packed.ptr, d_packed.n}, sequence.size()));
    EXPECT_EQ(inserted, sequence.size() - TestConfig::k + 1);
    (void)hipDeviceSynchronize();
    DeviceBuffer<uint8_t> d_out(inserted);
    cusbf::require_void(filter.contains_dense_packed_async(
        cusbf::device_span<const uint64_t>{d_packed.ptr, d_packed.n}, sequence.size(),
        cusbf::device_span<uint8_t>{d_out.ptr, d_out.n}));
    (void)hipDeviceSynchronize();
    EXPECT_TRUE(all_ones(d_out.host()));
}

TEST_GPU(DensePackedProteinMatchesByteSequenceInsertAndQuery) {
    cusbf::filter<ProteinTestConfig> filter(1 << 14);
    const std::string sequence = "ACDEFGHIKLMNPQRSTVWYACDEF";
    const auto packed = cusbf::pack_dense_sequence<ProteinTestConfig>(sequence);
    const uint64_t inserted = CUSBF_UNWRAP(filter.insert_dense_packed(packed, sequence.size()));
    const auto hits = CUSBF_UNWRAP(filter.contains_dense_packed(packed, sequence.size()));
    EXPECT_EQ(inserted, sequence.size() - ProteinTestConfig::k + 1);
    EXPECT_TRUE(all_ones(hits));
    EXPECT_TRUE(hits == CUSBF_UNWRAP(filter.contains_sequence(sequence)));
}

TEST_GPU(InvalidBasesResetForwardWindows) {
    cusbf::filter<TestConfig> filter(1 << 12);
    const std::string sequence = "ACGTNACGTACGTA";
    const auto inserted = CUSBF_UNWRAP(filter.insert_sequence(sequence));
    const auto hits = CUSBF_UNWRAP(filter.contains_sequence(sequence));
    EXPECT_EQ(inserted, hits.size());
    EXPECT_TRUE(hits == (std::vector<uint8_t>{0, 0, 0, 0, 0, 1, 1, 1, 1, 1}));
}

TEST_GPU(RepeatedInsertionIsIdempotent) {
    cusbf::filter<TestConfig> filter(1 << 12);
    const std::string sequence = "ACGTACGTACGTACGT";
    const auto first = CUSBF_UNWRAP(filter.insert_sequence(sequence));
    const float lf1 = filter.load_factor();
    const auto second = CUSBF_UNWRAP(filter.insert_sequence(sequence));
    EXPECT_EQ(first, second);
    EXPECT_EQ(lf1, filter.load_factor());
}

TEST_GPU(ShortSequenceInsertAndQueryReturnEmpty) {
    cusbf::filter<TestConfig> filter(1 << 12);
    EXPECT_EQ(CUSBF_UNWRAP(filter.insert_sequence("ACGT")), 0u);
    EXPECT_TRUE(CUSBF_UNWRAP(filter.contains_sequence("ACGT")).empty());
}

TEST_GPU(ShortSequenceDeviceOutputBufferRemainsUnchanged) {
    cusbf::filter<TestConfig> filter(1 << 12);
    const char seq[4] = {'A', 'C', 'G', 'T'};
    DeviceBuffer<char> d_seq(seq, 4);
    const uint8_t sentinel = 0xAB;
    DeviceBuffer<uint8_t> d_out(&sentinel, 1);
    cusbf::require_void(filter.contains_sequence_async(cusbf::device_span<const char>{d_seq.ptr, 4},
                                                       cusbf::device_span<uint8_t>{d_out.ptr, 1}));
    EXPECT_EQ(d_out.host()[0], 0xAB);
}

TEST_GPU(ClearResetsMembership) {
    cusbf::filter<TestConfig> filter(1 << 12);
    const std::string sequence = "ACGTACGTACGTACGT";
    (void)CUSBF_UNWRAP(filter.insert_sequence(sequence));
    (void)CUSBF_UNWRAP(filter.clear());
    const auto hits = CUSBF_UNWRAP(filter.contains_sequence(sequence));
    EXPECT_TRUE(std::all_of(hits.begin(), hits.end(), [](uint8_t v) { return v == 0; }));
}

TEST_GPU(DeviceOutputMatchesHostContainsResults) {
    cusbf::filter<TestConfig> filter(1 << 13);
    (void)CUSBF_UNWRAP(filter.insert_sequence("ACGTACGTACGTACGTACGTACGT"));
    const std::string query = "TACGTACGTACGTACGTACGTACG";
    const auto host_hits = CUSBF_UNWRAP(filter.contains_sequence(query));
    EXPECT_TRUE(!host_hits.empty());
    DeviceBuffer<char> d_q(query.data(), query.size());
    DeviceBuffer<uint8_t> d_out(host_hits.size());
    cusbf::require_void(filter.contains_sequence_async(cusbf::device_span<const char>{d_q.ptr, d_q.n},
                                                       cusbf::device_span<uint8_t>{d_out.ptr, d_out.n}));
    (void)hipDeviceSynchronize();
    EXPECT_TRUE(d_out.host() == host_hits);
}

TEST_GPU(HostSequenceDeviceOutputMatchesHostContainsResults) {
    cusbf::filter<TestConfig> filter(1 << 13);
    (void)CUSBF_UNWRAP(filter.insert_sequence("ACGTACGTACGTACGTACGTACGT"));
    const std::string query = "TACGTACGTACGTACGTACGTACG";
    const auto host_hits = CUSBF_UNWRAP(filter.contains_sequence(query));
    DeviceBuffer<uint8_t> d_out(host_hits.size());
    cusbf::require_void(filter.contains_sequence(query, cusbf::device_span<uint8_t>{d_out.ptr, d_out.n}));
    EXPECT_TRUE(d_out.host() == host_hits);
}

TEST_GPU(MultipleInsertionsRemainQueryable) {
    cusbf::filter<TestConfig> filter(1 << 14);
    (void)CUSBF_UNWRAP(filter.insert_sequence("ACGTACGTACGTACGT"));
    (void)CUSBF_UNWRAP(filter.insert_sequence("TGCATGCATGCATGCA"));
    EXPECT_TRUE(all_ones(CUSBF_UNWRAP(filter.contains_sequence("ACGTACGTACGTACGT"))));
    EXPECT_TRUE(all_ones(CUSBF_UNWRAP(filter.contains_sequence("TGCATGCATGCATGCA"))));
}

TEST_GPU(LowercaseInsertionMatchesUppercaseQuery) {
    cusbf::filter<TestConfig> filter(1 << 12);
    (void)CUSBF_UNWRAP(filter.insert_sequence("acgtacgtacgtacgt"));
    EXPECT_TRUE(all_ones(CUSBF_UNWRAP(filter.contains_sequence("ACGTACGTACGTACGT"))));
    EXPECT_TRUE(all_ones(CUSBF_UNWRAP(filter.contains_sequence("acgtacgtacgtacgt"))));
}

TEST_GPU(ProteinAlphabetInsertAndQuerySameSequenceHasNoFalseNegatives) {
    cusbf::filter<ProteinTestConfig> filter(1 << 12);
    const std::string sequence = "ABCDEFGHIJKLMNOPQRSTUVWXYZ";
    const uint64_t inserted = CUSBF_UNWRAP(filter.insert_sequence(sequence));
    const auto hits = CUSBF_UNWRAP(filter.contains_sequence(sequence));
    EXPECT_EQ(inserted, sequence.size() - ProteinTestConfig::k + 1);
    EXPECT_TRUE(all_ones(hits));
}

TEST_GPU(ProteinAlphabetLowercaseMatchesUppercaseQuery) {
    cusbf::filter<ProteinTestConfig> filter(1 << 12);
    (void)CUSBF_UNWRAP(filter.insert_sequence("abcdefghijklmnopqrstuvwxyz"));
    EXPECT_TRUE(all_ones(CUSBF_UNWRAP(filter.contains_sequence("ABCDEFGHIJKLMNOPQRSTUVWXYZ"))));
    EXPECT_TRUE(all_ones(CUSBF_UNWRAP(filter.contains_sequence("abcdefghijklmnopqrstuvwxyz"))));
}

TEST_GPU(ProteinAlphabetInvalidSymbolsResetForwardWindows) {
    cusbf::filter<ProteinTestConfig> filter(1 << 12);
    const std::string sequence = "ACDE*ACDEFGHI";
    (void)CUSBF_UNWRAP(filter.insert_sequence(sequence));
    EXPECT_TRUE(CUSBF_UNWRAP(filter.contains_sequence(sequence)) ==
                (std::vector<uint8_t>{0, 0, 0, 0, 0, 1, 1, 1, 1}));
}

TEST_GPU(CustomAlphabetUsesExplicitInvalidSentinel) {
    cusbf::filter<CustomAlphabetTestConfig> filter(1 << 12);
    const std::string sequence = "xyz!xyz";
    (void)CUSBF_UNWRAP(filter.insert_sequence(sequence));
    EXPECT_TRUE(CUSBF_UNWRAP(filter.contains_sequence(sequence)) == (std::vector<uint8_t>{1, 0, 0, 0, 1}));
}

TEST_GPU(DnaTripletAlphabetInsertAndQuerySameSequenceHasNoFalseNegatives) {
    cusbf::filter<TripletTestConfig> filter(1 << 12);
    const std::string sequence = "ACGTACGTTAAACCCGGGTTT";
    const uint64_t inserted = CUSBF_UNWRAP(filter.insert_sequence(sequence));
    EXPECT_EQ(inserted, sequence.size() / 3 - TripletTestConfig::k + 1);
    EXPECT_TRUE(all_ones(CUSBF_UNWRAP(filter.contains_sequence(sequence))));
}

TEST_GPU(DnaTripletAlphabetInvalidTripletsResetForwardWindows) {
    cusbf::filter<TripletTestConfig> filter(1 << 12);
    const std::string sequence = "ACGTACNNNGGGTTTAAA";
    (void)CUSBF_UNWRAP(filter.insert_sequence(sequence));
    EXPECT_TRUE(CUSBF_UNWRAP(filter.contains_sequence(sequence)) == (std::vector<uint8_t>{0, 0, 0, 1}));
}

TEST_GPU(DnaTripletAlphabetIgnoresTrailingIncompleteTriplet) {
    cusbf::filter<TripletTestConfig> filter(1 << 12);
    const std::string sequence = "ACGTACGTTAAACCCGGGTTTAA";
    EXPECT_EQ(CUSBF_UNWRAP(filter.insert_sequence(sequence)), 5u);
    const auto hits = CUSBF_UNWRAP(filter.contains_sequence(sequence));
    EXPECT_EQ(hits.size(), 5u);
    EXPECT_TRUE(all_ones(hits));
}

TEST_GPU(DeviceRecordBatchInsertMatchesHostSequencePath) {
    cusbf::filter<TestConfig> filter(1 << 14);
    const std::string r1 = "ACGTACGTACGT", r2 = "TTGGAACCTTGG";
    const std::string dev = r1 + 'N' + r2;
    DeviceBuffer<char> d_seq(dev.data(), dev.size());
    const std::array<cusbf::RecordRange, 2> ranges{{{0, 12}, {13, 12}}};
    EXPECT_EQ(CUSBF_UNWRAP(filter.insert_record_batch_async(cusbf::device_span<const char>{d_seq.ptr, d_seq.n},
                                                            ranges)),
              16u);
    (void)hipDeviceSynchronize();
    EXPECT_TRUE(all_ones(CUSBF_UNWRAP(filter.contains_sequence(r1))));
    EXPECT_TRUE(all_ones(CUSBF_UNWRAP(filter.contains_sequence(r2))));
}

TEST_GPU(DeviceRecordBatchQueryMatchesHostContains) {
    cusbf::filter<TestConfig> filter(1 << 14);
    const std::string r1 = "ACGTACGTACGT", r2 = "TTGGAACCTTGG";
    (void)CUSBF_UNWRAP(filter.insert_sequence(r1));
    (void)CUSBF_UNWRAP(filter.insert_sequence(r2));
    const std::string q = r1 + 'N' + r2;
    DeviceBuffer<char> d_seq(q.data(), q.size());
    const std::array<cusbf::RecordRange, 2> ranges{{{0, 12}, {13, 12}}};
    DeviceBuffer<uint8_t> d_out(21);
    cusbf::require_void(filter.contains_record_batch_async(cusbf::device_span<const char>{d_seq.ptr, d_seq.n},
                                                           ranges, cusbf::device_span<uint8_t>{d_out.ptr, 21}));
    (void)hipDeviceSynchronize();
    const auto hits = d_out.host();
    for (int i = 0; i < 8; ++i) EXPECT_EQ(hits[i], 1);
    for (int i = 8; i < 13; ++i) EXPECT_EQ(hits[i], 0);
    for (int i = 13; i < 21; ++i) EXPECT_EQ(hits[i], 1);
    // output span too small (filter.cuh:614-616)
    const auto r = filter.contains_record_batch_async(cusbf::device_span<const char>{d_seq.ptr, d_seq.n}, ranges,
                                                      cusbf::device_span<uint8_t>{d_out.ptr, 20});
    EXPECT_TRUE(!r && r.error().category() == cusbf::ErrorCategory::invalid_argument);
    EXPECT_EQ(r.error().message(), std::string("record batch query output span is too small"));
}

// Batched record form: the two records of the device record-batch test above,
// back to back without the separator. Windows inside a record answer as in the
// record batch; the 4 windows across the record start answer 0.
TEST_GPU(BatchedRecordsMatchRecordBatchAnswers) {
    cusbf::filter<TestConfig> filter(1 << 14);
    const std::string r1 = "ACGTACGTACGT", r2 = "TTGGAACCTTGG";
    const std::string all = r1 + r2;
    DeviceBuffer<char> d_seq(all.data(), all.size());
    const std::array<uint64_t, 3> off{0, 12, 24};
    DeviceBuffer<uint64_t> d_off(off.data(), off.size());
    const cusbf::device_span<const char> syms{d_seq.ptr, d_seq.n};
    const cusbf::device_span<const uint64_t> offs{d_off.ptr, d_off.n};
    cusbf::require_void(filter.insert_batched_async(syms, offs));
    cusbf::require_void(filter.batched_status());
    EXPECT_TRUE(all_ones(CUSBF_UNWRAP(filter.contains_sequence(r1))));
    EXPECT_TRUE(all_ones(CUSBF_UNWRAP(filter.contains_sequence(r2))));
    DeviceBuffer<uint8_t> d_out(20);
    cusbf::require_void(filter.contains_batched_async(syms, offs, cusbf::device_span<uint8_t>{d_out.ptr, 20}));
    cusbf::require_void(filter.batched_status());
    const auto hits = d_out.host();
    for (int i = 0; i < 8; ++i) EXPECT_EQ(hits[i], 1);
    for (int i = 8; i < 12; ++i) EXPECT_EQ(hits[i], 0);
    for (int i = 12; i < 20; ++i) EXPECT_EQ(hits[i], 1);
    const auto small = filter.contains_batched_async(syms, offs, cusbf::device_span<uint8_t>{d_out.ptr, 19});
    EXPECT_TRUE(!small && small.error().category() == cusbf::ErrorCategory::invalid_argument);
    // malformed offsets are reported by batched_status
    const std::array<uint64_t, 3> bad{0, 13, 12};
    DeviceBuffer<uint64_t> d_bad(bad.data(), bad.size());
    cusbf::require_void(filter.insert_batched_async(syms, cusbf::device_span<const uint64_t>{d_bad.ptr, 3}));
    const auto st = filter.batched_status();
    EXPECT_TRUE(!st && st.error().category() == cusbf::ErrorCategory::invalid_argument);
}

TEST_GPU(BatchPlanReplaysMatchEagerCall) {
    cusbf::filter<TestConfig> filter(1 << 14);
    const std::string all = "ACGTACGTACGTTTGGAACCTTGGACNTTGACCA";
    DeviceBuffer<char> d_seq(all.data(), all.size());
    const std::array<uint64_t, 4> off{0, 12, 24, 34};
    DeviceBuffer<uint64_t> d_off(off.data(), off.size());
    const cusbf::device_span<const char> syms{d_seq.ptr, d_seq.n};
    const cusbf::device_span<const uint64_t> offs{d_off.ptr, d_off.n};
    cusbf::require_void(filter.insert_batched_async(syms, offs));
    cusbf::require_void(filter.batched_status());
    const auto words = CUSBF_UNWRAP(filter.words());
    cusbf::require_void(filter.clear());
    auto ins = CUSBF_UNWRAP(cusbf::batch_plan::create(filter, CUSBF_BATCH_INSERT, syms, offs));
    DeviceBuffer<uint8_t> d_out(30);
    auto qry = CUSBF_UNWRAP(cusbf::batch_plan::create(filter, CUSBF_BATCH_CONTAINS, syms, offs,
                                                      cusbf::device_span<uint8_t>{d_out.ptr, 30}));
    for (int rep = 0; rep < 3; ++rep) {
        cusbf::require_void(ins.launch());
        cusbf::require_void(ins.status());
        EXPECT_TRUE(CUSBF_UNWRAP(filter.words()) == words);
        cusbf::require_void(qry.launch());
        cusbf::require_void(qry.status());
        const auto hits = d_out.host();
        uint64_t ones = 0;
        for (uint8_t h : hits) ones += h;
        // records of 12, 12 and 10 symbols; the N at position 2 of the third
        // leaves its windows 3..5: 8 + 8 + 3 valid windows, all inserted
        EXPECT_EQ(ones, 8u + 8u + 3u);
    }
    cusbf_batch_tuning t{3, 2, 1, 0};
    cusbf::require_void(filter.set_batch_tuning(1, CUSBF_BATCH_CONTAINS, t));
    EXPECT_EQ(CUSBF_UNWRAP(filter.batch_tuning(1, CUSBF_BATCH_CONTAINS)).blocks_per_cu, 2u);
}

TEST_GPU(HostRecordBatchReportsAndPerRecordHits) {
    cusbf::filter<TestConfig> filter(1 << 14);
    const std::string seq = "ACGTACGTACGTTTGGAACCTTGGACNTT";
    const std::array<cusbf::RecordRange, 3> ranges{{{0, 12}, {12, 12}, {24, 5}}};
    const auto rep = CUSBF_UNWRAP(filter.insert_record_batch(cusbf::RecordBatchView{seq, ranges}));
    EXPECT_EQ(rep.recordsIndexed, 3u);
    EXPECT_EQ(rep.indexedBases, 29u);
    EXPECT_EQ(rep.insertedKmers, 16u);  // 8 + 8 + 0 (the third record holds an N)
    std::vector<uint64_t> positives;
    const auto q = CUSBF_UNWRAP(filter.query_record_batch(
        cusbf::RecordBatchView{seq, ranges},
        [&](const cusbf::RecordQueryView& v) { positives.push_back(v.positive_kmers); }));
    EXPECT_TRUE(positives == (std::vector<uint64_t>{8, 8, 0}));
    EXPECT_EQ(q.positive_kmers, 16u);
    EXPECT_EQ(q.queriedKmers, 16u);
    const auto agg = CUSBF_UNWRAP(filter.query_record_batch(cusbf::RecordBatchView{seq, ranges}));
    EXPECT_EQ(agg.positive_kmers, 16u);
}

TEST_GPU(DenseSpanTooSmallIsInvalidArgument) {
    cusbf::filter<TestConfig> filter(1 << 12);
    const std::vector<uint64_t> words(1);
    const auto r = filter.insert_dense_packed(words, 100);
    EXPECT_TRUE(!r && r.error().category() == cusbf::ErrorCategory::invalid_argument);
    EXPECT_EQ(r.error().message(), std::string("dense packed span is too small for num_symbols"));
}

// ---- host-only: tests/test_query_layout.cpp, tests/test_normalized_record_batch.cpp -----

// ---- tests/test_fastx.cu ------------------------------------------------------------

// A temporary file holding `text`, removed when it goes out of scope.
struct TempFile {
    std::string path;
    explicit TempFile(const std::string& text) {
        char name[] = "/tmp/cusbf_fastx_XXXXXX";
        const int fd = mkstemp(name);
        if (fd < 0) throw std::runtime_error("mkstemp");
        if (write(fd, text.data(), text.size()) != static_cast<ssize_t>(text.size())) {
            close(fd);
            throw std::runtime_error("write");
        }
        close(fd);
        path = name;
    }
    ~TempFile() { unlink(path.c_str()); }
};

TEST_GPU(InsertFastxFileParsesWrappedFastaRecords) {
    cusbf::filter<TestConfig> filter(1 << 12);
    const TempFile file(">wrapped\nACGT\nACGT\nACGT\n");
    const auto report = CUSBF_UNWRAP(filter.insert_fastx_file(file.path));
    EXPECT_EQ(report.recordsIndexed, 1u);
    EXPECT_EQ(report.indexedBases, 12u);
    EXPECT_EQ(report.insertedKmers, 12u - TestConfig::k + 1);
    EXPECT_TRUE(all_ones(CUSBF_UNWRAP(filter.contains_sequence("ACGTACGTACGT"))));
}

TEST_GPU(QueryFastxFileRecordsParsesWrappedFastqWithCrLf) {
    cusbf::filter<TestConfig> filter(1 << 12);
    (void)CUSBF_UNWRAP(filter.insert_sequence("ACGTACGTACGT"));
    const TempFile file("@wrapped\r\nACGTAC\r\nGTACGT\r\n+\r\nIIIIII\r\nIIIIII\r\n");
    const auto report = CUSBF_UNWRAP(filter.query_fastx_file(file.path));
    EXPECT_EQ(report.recordsQueried, 1u);
    EXPECT_EQ(report.queriedBases, 12u);
    EXPECT_EQ(report.queriedKmers, 8u);
    EXPECT_EQ(report.positive_kmers, 8u);
    std::vector<cusbf::FastxDetailedQueryRecord> seen;
    const auto summary = CUSBF_UNWRAP(filter.query_fastx_file_records(
        file.path, [&](const cusbf::FastxRecordView& r) {
            seen.push_back({r.record_index, std::string(r.header), std::string(r.sequence), r.queriedBases,
                            r.queriedKmers, r.positive_kmers, std::vector<uint8_t>(r.hits.begin(), r.hits.end())});
        }));
    EXPECT_EQ(summary.queriedKmers, 8u);
    EXPECT_EQ(seen.size(), 1u);
    EXPECT_EQ(seen[0].header, std::string("wrapped"));
    EXPECT_EQ(seen[0].sequence, std::string("ACGTACGTACGT"));
    EXPECT_EQ(seen[0].hits.size(), 8u);
    EXPECT_TRUE(all_ones(seen[0].hits));
}

TEST_GPU(QueryFastxDoesNotCreateCrossRecordKmers) {
    cusbf::filter<TestConfig> filter(1 << 12);
    (void)CUSBF_UNWRAP(filter.insert_sequence("ACGTACGT"));
    (void)CUSBF_UNWRAP(filter.insert_sequence("TGCATGCA"));
    std::istringstream input(">first\nACGT\nACGT\n>second\nTGCA\nTGCA\n");
    const auto report = CUSBF_UNWRAP(filter.query_fastx_detailed(input));
    EXPECT_EQ(report.summary.recordsQueried, 2u);
    EXPECT_EQ(report.summary.queriedBases, 16u);
    EXPECT_EQ(report.summary.queriedKmers, 8u);
    EXPECT_EQ(report.summary.positive_kmers, 8u);
    EXPECT_EQ(report.records.size(), 2u);
    EXPECT_EQ(report.records[1].header, std::string("second"));
    EXPECT_EQ(report.records[1].record_index, 1u);
}

TEST_GPU(TripletQueryFastxRecordsDoNotCreateCrossRecordKmers) {
    cusbf::filter<TripletTestConfig> filter(1 << 12);
    (void)CUSBF_UNWRAP(filter.insert_sequence("ACGTACGTT"));
    (void)CUSBF_UNWRAP(filter.insert_sequence("GGGTTTAAA"));
    const TempFile file(">first\nACGTACGTT\n>second\nGGGTTTAAA\n");
    const auto report = CUSBF_UNWRAP(filter.query_fastx_file_detailed(file.path));
    EXPECT_EQ(report.summary.queriedKmers, 2u);
    EXPECT_EQ(report.summary.positive_kmers, 2u);
    EXPECT_EQ(report.records.size(), 2u);
    EXPECT_TRUE(report.records[0].hits == (std::vector<uint8_t>{1}));
    EXPECT_TRUE(report.records[1].hits == (std::vector<uint8_t>{1}));
}

TEST_GPU(FastxDetailedQueryReportsInvalidWindowsAsMisses) {
    cusbf::filter<TestConfig> filter(1 << 12);
    const TempFile file(">with-invalid\nACGTNACGTACGTA\n");
    const auto ins = CUSBF_UNWRAP(filter.insert_fastx_file(file.path));
    EXPECT_EQ(ins.insertedKmers, 5u);
    const auto report = CUSBF_UNWRAP(filter.query_fastx_file_detailed(file.path));
    EXPECT_EQ(report.summary.queriedKmers, 5u);
    EXPECT_EQ(report.summary.positive_kmers, 5u);
    EXPECT_TRUE(report.records[0].hits == (std::vector<uint8_t>{0, 0, 0, 0, 0, 1, 1, 1, 1, 1}));
}

TEST_GPU(ForcedFastxChunkFlushPreservesRecordOrderAndCounts) {
    cusbf::filter<TestConfig> filter(1 << 12);
    const TempFile file(">first\nACGTACGT\n>second\nACGTNACGTACGTA\n>third\nTGCATGCA\n");
    const auto ins = CUSBF_UNWRAP(filter.insert_fastx_file(file.path, 0.0));
    EXPECT_EQ(ins.recordsIndexed, 3u);
    EXPECT_EQ(ins.indexedBases, 30u);
    EXPECT_EQ(ins.insertedKmers, 13u);
    const auto report = CUSBF_UNWRAP(filter.query_fastx_file_detailed(file.path, 0.0));
    EXPECT_EQ(report.records.size(), 3u);
    EXPECT_EQ(report.records[2].header, std::string("third"));
    EXPECT_TRUE(report.records[1].hits == (std::vector<uint8_t>{0, 0, 0, 0, 0, 1, 1, 1, 1, 1}));
    EXPECT_EQ(report.summary.positive_kmers, 13u);
}

TEST_GPU(MalformedFastqReportsParseLocation) {
    cusbf::filter<TestConfig> filter(1 << 12);
    const TempFile file("@broken\nACGTACGT\n+\nIIIIIII\n");
    const auto result = filter.query_fastx_file(file.path);
    EXPECT_TRUE(!result);
    EXPECT_TRUE(result.error().category() == cusbf::ErrorCategory::fastx_parse);
    const cusbf::FastxParseError* e = result.error().as_fastx_parse();
    EXPECT_TRUE(e != nullptr);
    EXPECT_EQ(e->location.line, 4u);
    EXPECT_EQ(e->location.column, 8u);
    EXPECT_EQ(e->location.file, file.path);
    EXPECT_EQ(result.error().message(),
              file.path + ":4:8: FASTQ quality length does not match sequence length");
}

TEST_GPU(CallbackAbortStopsProcessingPreservingFilterReuse) {
    cusbf::filter<TestConfig> filter(1 << 12);
    (void)CUSBF_UNWRAP(filter.insert_sequence("ACGTACGTACGT"));
    const std::string text = ">seqA\nACGTACGTACGT\n>seqB\nTGCATGCA\n>seqC\nGGGAAAACCC\n";
    size_t calls = 0;
    bool thrown = false;
    try {
        std::istringstream in(text);
        (void)filter.query_fastx_records(in, [&](const cusbf::FastxRecordView&) {
            ++calls;
            throw std::runtime_error("abort");
        });
    } catch (const std::runtime_error& e) {
        thrown = std::string(e.what()) == "abort";
    }
    EXPECT_TRUE(thrown);
    EXPECT_EQ(calls, 1u);
    std::istringstream again(text);
    size_t n = 0;
    const auto summary = CUSBF_UNWRAP(filter.query_fastx_records(again, [&](const cusbf::FastxRecordView&) { ++n; }));
    EXPECT_EQ(n, 3u);
    EXPECT_EQ(summary.recordsQueried, 3u);
}

// ---- filter images (no reference counterpart) -------------------------------------------

TEST_GPU(FilterImageRoundTripAndConfigCheck) {
    cusbf::filter<TestConfig> filter(1 << 14);
    (void)CUSBF_UNWRAP(filter.insert_sequence("ACGTACGTACGTTTGGAACC"));
    char name[] = "/tmp/cusbf_image_XXXXXX";
    const int fd = mkstemp(name);
    close(fd);
    cusbf::require_void(filter.save(name));
    auto loaded = CUSBF_UNWRAP(cusbf::filter<TestConfig>::load(name));
    EXPECT_TRUE(CUSBF_UNWRAP(loaded.words()) == CUSBF_UNWRAP(filter.words()));
    EXPECT_TRUE(all_ones(CUSBF_UNWRAP(loaded.contains_sequence("ACGTACGTACGTTTGGAACC"))));
    const auto other = cusbf::filter<ProteinTestConfig>::load(name);
    EXPECT_TRUE(!other && other.error().category() == cusbf::ErrorCategory::invalid_argument);
    unlink(name);
}

TEST_GPU(SortedInsertPathMatchesAtomicPath) {
    std::string sequence(300000, 'A');
    uint64_t x = 0x9E3779B97F4A7C15ull;
    for (char& c : sequence) {
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        c = "ACGTN"[(x >> 20) % 64 == 0 ? 4 : (x >> 30) & 3];
    }
    using Filter = cusbf::filter<TestConfig>;
    Filter sorted(1 << 20);
    Filter atomic(1 << 20);
    cusbf::require_void(sorted.set_insert_path(Filter::InsertPath::sorted, 3));
    cusbf::require_void(atomic.set_insert_path(Filter::InsertPath::atomic));
    EXPECT_EQ(CUSBF_UNWRAP(sorted.insert_sequence(sequence)), CUSBF_UNWRAP(atomic.insert_sequence(sequence)));
    EXPECT_TRUE(sorted.last_insert_path() == Filter::InsertPath::sorted);
    EXPECT_TRUE(atomic.last_insert_path() == Filter::InsertPath::atomic);
    EXPECT_TRUE(CUSBF_UNWRAP(sorted.words()) == CUSBF_UNWRAP(atomic.words()));
    const auto bad = sorted.set_insert_path(Filter::InsertPath::sorted, 13);
    EXPECT_TRUE(!bad && bad.error().category() == cusbf::ErrorCategory::invalid_argument);
}

TEST_HOST(QueryLayoutDerivesPerRecordHitSpans) {
    const std::array<cusbf::NormalizedRecord, 2> records{{{0, 0, 6, 1}, {1, 6, 8, 2}}};
    const auto layout = cusbf::detail::QueryLayout::build<LayoutConfig>(records);
    EXPECT_EQ(layout.total_hit_count(), 13u);
    EXPECT_EQ(layout.records()[0].hit_offset, 0u);
    EXPECT_EQ(layout.records()[0].hit_count, 4u);
    EXPECT_EQ(layout.records()[1].hit_offset, 7u);
    EXPECT_EQ(layout.records()[1].hit_count, 6u);
    const std::array<uint8_t, 13> hits{{0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12}};
    const auto second = layout.hits_for_record(hits, 1);
    EXPECT_EQ(second.size(), 6u);
    EXPECT_EQ(second[0], 7u);
}

TEST_HOST(QueryLayoutTotalHitCountIncludesBoundaryWindows) {
    const std::array<cusbf::NormalizedRecord, 2> records{{{0, 0, 4, 0}, {1, 4, 4, 0}}};
    const auto layout = cusbf::detail::QueryLayout::build<LayoutConfig>(records);
    EXPECT_EQ(layout.records()[0].hit_count, 2u);
    EXPECT_EQ(layout.records()[1].hit_count, 2u);
    EXPECT_EQ(layout.total_hit_count(), 7u);
}

TEST_HOST(NormalizedRecordBatchInjectsSeparators) {
    const std::array<cusbf::RecordRange, 2> ranges{{{0, 4}, {4, 4}}};
    const auto n = CUSBF_UNWRAP(cusbf::normalize_record_batch<LayoutConfig>(cusbf::RecordBatchView{"ACGTACGTACGT", ranges}));
    EXPECT_EQ(n.records().size(), 2u);
    EXPECT_TRUE(n.sequence().size() > 8u);
    EXPECT_EQ(n.total_valid_kmers(), 4u);
    EXPECT_EQ(n.records()[1].input_offset, 4u);
}

TEST_HOST(NormalizedRecordBatchRejectsOverlappingRanges) {
    const std::array<cusbf::RecordRange, 2> ranges{{{4, 4}, {0, 4}}};
    const auto n = cusbf::normalize_record_batch<LayoutConfig>(cusbf::RecordBatchView{"ACGTACGT", ranges});
    EXPECT_TRUE(!n);
    EXPECT_TRUE(n.error().category() == cusbf::ErrorCategory::invalid_argument);
}

TEST_HOST(DensePackingIsLsbFirst) {
    const auto w = cusbf::pack_dense_sequence<TestConfig>("ACGTN");
    // A0 C1 G3 T2, N -> 0xFF & 3 = 3 (dense_packed.hpp:183-185)
    EXPECT_EQ(w.size(), 1u);
    EXPECT_EQ(w[0], (0ull) | (1ull << 2) | (3ull << 4) | (2ull << 6) | (3ull << 8));
}

TEST_HOST(ResultPropagatesErrors) {
    auto inner = []() -> cusbf::Result<uint64_t> { return cusbf::Err(cusbf::Error::resource("budget")); };
    auto outer = [&]() -> cusbf::Result<uint64_t> {
        const uint64_t v = CUSBF_TRY(inner());
        return v + 1;
    };
    const auto r = outer();
    EXPECT_TRUE(!r && r.error().category() == cusbf::ErrorCategory::resource);
    bool threw = false;
    try {
        (void)CUSBF_UNWRAP(outer());
    } catch (const std::runtime_error&) {
        threw = true;
    }
    EXPECT_TRUE(threw);
}

}  // namespace

int main(int argc, char** argv) {
    const bool host_only = argc > 1 && std::strcmp(argv[1], "--host-only") == 0;
    int run = 0;
    for (const Case& c : registry()) {
        if (host_only && c.gpu) continue;
        const int before = g_failures;
        try {
            c.fn();
        } catch (const std::exception& e) {
            std::fprintf(stderr, "  exception: %s\n", e.what());
            ++g_failures;
        }
        std::printf("[%s] %s\n", g_failures == before ? "PASS" : "FAIL", c.name);
        ++run;
    }
    std::printf("%d tests, %d failures\n", run, g_failures);
    return g_failures == 0 ? 0 : 1;
}
