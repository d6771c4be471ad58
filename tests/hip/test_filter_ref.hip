// test_filter_ref.hip -- a caller's own HIP kernels on cusbf::filter<Config>::ref().
//
// The reference hands device code a filter_ref<Config> (filter.cuh:137-142,
// filter_ref.cuh:20-113). This program does what such a caller does: one thread
// per k-mer window packs the k-mer, takes its minimizer shard with
// ref.shard_index(), builds the masks with filter_block::sectorizedHashToMasks()
// and ORs them in with ref.apply_word_masks(); a second kernel answers windows
// with ref.sectorized_contains_packed_kmer(). The same constexpr functions are
// also run on the host over the copied words.
//
// Usage: test_filter_ref <insert sequence file> <query sequence file> <out prefix>
// Writes <prefix>.words (uint64 filter words) and <prefix>.hits (one byte per
// query window); the Python test compares both with the oracle. Prints "ok" and
// returns 0 when its own checks pass.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <fstream>
#include <iterator>
#include <string>
#include <vector>

#include <cusbf/filter.hpp>
#include <cusbf/filter_ref_device.hpp>

using Cfg = cusbf::Config<31, 28, 16, 4>;
using Ref = cusbf::filter_ref<Cfg>;

// Packed k-mer of window j (first symbol in the highest bits); false when the
// window holds an invalid symbol.
__host__ __device__ inline bool pack_window(const char* seq, uint64_t j, uint64_t& packed) {
    packed = 0;
    for (uint64_t i = 0; i < Cfg::k; ++i) {
        const uint8_t sym = Cfg::Alphabet::encode(seq + j + i);
        if (sym == Cfg::Alphabet::invalidSymbol) return false;
        packed = (packed << Cfg::symbolBits) | (sym & Cfg::symbolMask);
    }
    return true;
}

__global__ void user_insert(Ref ref, const char* seq, uint64_t windows) {
    const uint64_t j = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (j >= windows) return;
    uint64_t packed;
    if (!pack_window(seq, j, packed)) return;
    const uint64_t shard = ref.shard_index(cusbf::detail::packed_kmer_minimizer_hash<Cfg>(packed));
    uint64_t m0 = 0, m1 = 0, m2 = 0, m3 = 0;
    for (uint64_t o = 0; o < Cfg::findereSpan; ++o) {
        Ref::block_type::sectorizedHashToMasks(cusbf::detail::packed_kmer_smer_hash<Cfg>(packed, o),
                                               m0, m1, m2, m3);
    }
    ref.apply_word_masks(ref.blocks()[shard], m0, m1, m2, m3);
}

__global__ void user_contains(Ref ref, const char* seq, uint64_t windows, uint8_t* out) {
    const uint64_t j = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (j >= windows) return;
    uint64_t packed;
    bool hit = false;
    if (pack_window(seq, j, packed)) {
        const uint64_t shard = ref.shard_index(cusbf::detail::packed_kmer_minimizer_hash<Cfg>(packed));
        hit = Ref::sectorized_contains_packed_kmer(packed, ref.blocks()[shard].words);
    }
    out[j] = hit ? 1 : 0;
}

static std::string read_file(const char* path) {
    std::ifstream in(path, std::ios::binary);
    return std::string(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
}

static bool write_file(const std::string& path, const void* data, size_t bytes) {
    std::ofstream out(path, std::ios::binary);
    out.write(static_cast<const char*>(data), static_cast<std::streamsize>(bytes));
    return static_cast<bool>(out);
}

#define CHECK(cond, msg)                                   \
    do {                                                   \
        if (!(cond)) {                                     \
            std::fprintf(stderr, "FAILED: %s\n", msg);     \
            return 1;                                      \
        }                                                  \
    } while (0)

int main(int argc, char** argv) {
    if (argc != 4) {
        std::fprintf(stderr, "usage: %s <insert file> <query file> <out prefix>\n", argv[0]);
        return 2;
    }
    const std::string ins = read_file(argv[1]);
    const std::string qry = read_file(argv[2]);
    CHECK(ins.size() >= Cfg::k && qry.size() >= Cfg::k, "inputs shorter than k");
    const uint64_t ins_windows = ins.size() - Cfg::k + 1;
    const uint64_t q_windows = qry.size() - Cfg::k + 1;

    cusbf::filter<Cfg> f(16 * ins_windows);
    const Ref ref = f.ref();
    CHECK(ref.blocks() != nullptr, "ref() returned an empty view");
    CHECK(ref.num_blocks() == f.num_shards(), "ref().num_blocks() != num_shards()");
    auto checked = f.checked_ref();
    CHECK(checked && checked->blocks() == ref.blocks(), "checked_ref() differs from ref()");

    char *d_ins = nullptr, *d_qry = nullptr;
    uint8_t* d_out = nullptr;
    CHECK(hipMalloc(&d_ins, ins.size()) == hipSuccess, "hipMalloc");
    CHECK(hipMalloc(&d_qry, qry.size()) == hipSuccess, "hipMalloc");
    CHECK(hipMalloc(&d_out, q_windows) == hipSuccess, "hipMalloc");
    CHECK(hipMemcpy(d_ins, ins.data(), ins.size(), hipMemcpyHostToDevice) == hipSuccess, "copy");
    CHECK(hipMemcpy(d_qry, qry.data(), qry.size(), hipMemcpyHostToDevice) == hipSuccess, "copy");

    const unsigned block = 256;
    hipLaunchKernelGGL(user_insert, dim3((ins_windows + block - 1) / block), dim3(block), 0, nullptr,
                       ref, d_ins, ins_windows);
    CHECK(hipGetLastError() == hipSuccess, "user_insert launch");
    hipLaunchKernelGGL(user_contains, dim3((q_windows + block - 1) / block), dim3(block), 0, nullptr,
                       ref, d_qry, q_windows, d_out);
    CHECK(hipGetLastError() == hipSuccess, "user_contains launch");
    CHECK(hipDeviceSynchronize() == hipSuccess, "kernels");

    std::vector<uint8_t> hits(q_windows);
    CHECK(hipMemcpy(hits.data(), d_out, q_windows, hipMemcpyDeviceToHost) == hipSuccess, "copy back");
    auto words = f.words();
    CHECK(static_cast<bool>(words), "words()");

    // the library's own insert and query on the same input give the same words and answers
    cusbf::filter<Cfg> g(16 * ins_windows);
    CHECK(static_cast<bool>(g.insert_sequence(ins)), "library insert");
    auto g_words = g.words();
    CHECK(static_cast<bool>(g_words) && *g_words == *words, "user kernel words differ from the library's");
    auto g_hits = g.contains_sequence(qry);
    CHECK(static_cast<bool>(g_hits) && *g_hits == hits, "user kernel answers differ from the library's");

    // the constexpr members on the host, over the copied words
    const auto* blocks = reinterpret_cast<const Ref::block_type*>(words->data());
    for (uint64_t j = 0; j < q_windows && j < 20000; ++j) {
        uint64_t packed;
        bool hit = false;
        if (pack_window(qry.data(), j, packed)) {
            const uint64_t shard = Ref::shard_index(cusbf::detail::packed_kmer_minimizer_hash<Cfg>(packed),
                                                    ref.num_blocks());
            hit = Ref::sectorized_contains_packed_kmer(packed, blocks[shard].words);
        }
        CHECK(hit == (hits[j] == 1), "host evaluation differs from the user kernel");
    }

    const std::string prefix = argv[3];
    CHECK(write_file(prefix + ".words", words->data(), words->size() * 8), "write words");
    CHECK(write_file(prefix + ".hits", hits.data(), hits.size()), "write hits");
    (void)hipFree(d_ins);
    (void)hipFree(d_qry);
    (void)hipFree(d_out);
    std::printf("ok %llu windows inserted, %llu queried\n", static_cast<unsigned long long>(ins_windows),
                static_cast<unsigned long long>(q_windows));
    return 0;
}
