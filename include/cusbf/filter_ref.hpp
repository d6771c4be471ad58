// cusbf/filter_ref.hpp -- filter_block<Config> and filter_ref<Config>: the shard layout
// and the trivially copyable view of a filter's shard storage, as the reference's
// include/cusbf/detail/filter_impl.cuh:18-77 and include/cusbf/filter_ref.cuh:20-113.
//
// Everything here is arithmetic on the 256-bit block layout: constexpr functions,
// usable from host code and (compiled by hipcc) from device code. The one member
// that needs device atomics, filter_ref::apply_word_masks, is defined in the
// HIP-only header cusbf/filter_ref_device.hpp, which device code that calls it
// includes.
//
// Layout: shard i of a filter is filter_block words[0..3] at blocks()[i]; hash index
// j of an s-mer sets or tests bit sectorizedBitAddress<j>(hash) of word j % 4.
#pragma once

#include <hip/hip_runtime_api.h>  // __device__ (empty under a host compiler)

#include <bit>
#include <cstdint>
#include <limits>
#include <type_traits>
#include <utility>

#include "config.hpp"

namespace cusbf {
namespace detail {

/// filter_common.cuh:25: no minimizer seen yet.
inline constexpr uint64_t kInvalidHash = std::numeric_limits<uint64_t>::max();

/// hashutil.cuh:192-196: the s-mer hash behind the Bloom bit positions.
[[nodiscard]] constexpr uint64_t hash64(uint64_t key) {
    key *= 0x9e3779b97f4a7c15ULL;
    key ^= key >> 33;
    return key;
}

/// hashutil.cuh:207-209: the m-mer hash whose minimum selects the shard.
[[nodiscard]] constexpr uint64_t minimizer_hash64(uint64_t key) { return key * 0x9E3779B97F4A7C15ULL; }

/// filter_common.cuh:27-100: multiplicative salt of hash index `Index` (used when
/// 64 / hashCount < 6). The values and their order must be the reference's.
template <uint64_t Index>
[[nodiscard]] constexpr uint64_t multiplicativeSaltLiteral() {
    static_assert(Index < 16, "Salt index out of range");
    constexpr uint64_t salts[16] = {
        0x9E3779B97F4A7C15ULL, 0xC2B2AE3D27D4EB4FULL, 0x165667B19E3779F9ULL, 0x85EBCA77C2B2AE63ULL,
        0x27D4EB2F165667C5ULL, 0x94D049BB133111EFULL, 0xBF58476D1CE4E5B9ULL, 0xD6E8FEB86659FD93ULL,
        0xA0761D6478BD642FULL, 0xE7037ED1A0B428DBULL, 0x8EBC6AF09C88C6E3ULL, 0x589965CC75374CC3ULL,
        0x1D8E4E27C47D124FULL, 0xEB449C93FBBEA6B5ULL, 0xDB4F0B9175AE2165ULL, 0xBBE056FDADE14B91ULL};
    return salts[Index];
}

/// filter_common.cuh:103-115: fn(integral_constant<uint64_t, i>) for i in [0, hashCount).
template <typename Config, typename Fn, uint64_t... I>
constexpr void forEachHashIndexImpl(Fn&& fn, std::integer_sequence<uint64_t, I...>) {
    (fn(std::integral_constant<uint64_t, I>{}), ...);
}
template <typename Config, typename Fn>
constexpr void forEachHashIndex(Fn&& fn) {
    forEachHashIndexImpl<Config>(static_cast<Fn&&>(fn),
                                 std::make_integer_sequence<uint64_t, Config::hashCount>{});
}

/// filter_common.cuh:118-125: the low `Length` symbols of a packed window.
template <typename Config, uint64_t Length>
[[nodiscard]] constexpr uint64_t packedWindowMask() {
    if constexpr (Length * Config::symbolBits >= 64) {
        return std::numeric_limits<uint64_t>::max();
    } else {
        return (uint64_t{1} << (Config::symbolBits * Length)) - 1;
    }
}

/// filter_common.cuh:127-134: the sub-window of `WindowLength` symbols at `start` of a
/// packed K-mer (first symbol in the highest bits).
template <typename Config, uint64_t WindowLength, uint64_t K>
[[nodiscard]] constexpr uint64_t extractPackedSubwindow(uint64_t packed_kmer, uint64_t start) {
    static_assert(WindowLength <= K, "WindowLength must not exceed K");
    return (packed_kmer >> (Config::symbolBits * (K - (start + WindowLength)))) &
           packedWindowMask<Config, WindowLength>();
}

/// sequence_kmer.cuh:40-50: minimum minimizer hash over the k - m + 1 m-mers of a k-mer.
template <typename Config>
[[nodiscard]] constexpr uint64_t packed_kmer_minimizer_hash(uint64_t packed_kmer) {
    uint64_t best = kInvalidHash;
    for (uint64_t offset = 0; offset < Config::minimizerSpan; ++offset) {
        const uint64_t h = minimizer_hash64(
            extractPackedSubwindow<Config, Config::m, Config::k>(packed_kmer, offset));
        best = h < best ? h : best;
    }
    return best;
}

/// sequence_kmer.cuh:53-59: Bloom hash of the s-mer at `start` of a packed k-mer.
template <typename Config>
[[nodiscard]] constexpr uint64_t packed_kmer_smer_hash(uint64_t packed_kmer, uint64_t start) {
    return hash64(extractPackedSubwindow<Config, Config::s, Config::k>(packed_kmer, start));
}

}  // namespace detail

/// filter_impl.cuh:18-77: one 256-bit shard, hash index i in word i % 4.
template <typename Config>
struct alignas(32) filter_block {
    static constexpr uint64_t wordCount = Config::blockWordCount;
    static constexpr uint64_t wordBits = Config::wordBits;
    static constexpr int wordBitsLog2 = std::bit_width(wordBits) - 1;
    static constexpr uint64_t wordMask = (1ULL << wordBitsLog2) - 1;
    static constexpr int hashShift = 64 - wordBitsLog2;
    static constexpr uint64_t sliceWidth = 64 / Config::hashCount;
    static constexpr bool useBitSlicing = sliceWidth >= wordBitsLog2;

    uint64_t words[wordCount];

    /// filter_impl.cuh:36-47: bit of word HashIndex % 4 for s-mer hash `baseHash`:
    /// a 64/hashCount-bit slice when that is at least 6 bits, else a salted multiply-shift.
    template <uint64_t HashIndex>
    [[nodiscard]] static constexpr uint64_t sectorizedBitAddress(uint64_t baseHash) {
        static_assert(HashIndex < Config::hashCount, "Hash index out of range");
        if constexpr (useBitSlicing) {
            return (baseHash >> (sliceWidth * HashIndex)) & wordMask;
        } else {
            return (baseHash * detail::multiplicativeSaltLiteral<HashIndex>()) >> hashShift;
        }
    }

    /// filter_impl.cuh:54-77: ORs the bits of every hash index into the four word masks.
    static constexpr void sectorizedHashToMasks(uint64_t baseHash, uint64_t& mask0, uint64_t& mask1,
                                                uint64_t& mask2, uint64_t& mask3) {
        uint64_t* masks[4] = {&mask0, &mask1, &mask2, &mask3};
        detail::forEachHashIndex<Config>([&](auto index) {
            constexpr uint64_t i = decltype(index)::value;
            *masks[i % 4] |= uint64_t{1} << sectorizedBitAddress<i>(baseHash);
        });
    }
};

/// filter_ref.cuh:20-113: non-owning, trivially copyable view of the shard storage,
/// passed by value into device code (filter<Config>::ref()).
template <typename Config>
class filter_ref {
   public:
    using block_type = filter_block<Config>;

    constexpr filter_ref() noexcept = default;
    /// `num_blocks` must be a power of two (the shard index takes the hash's low bits).
    constexpr filter_ref(block_type* blocks, uint64_t num_blocks) noexcept
        : blocks_(blocks), num_blocks_(num_blocks) {}

    [[nodiscard]] constexpr block_type* blocks() const noexcept { return blocks_; }
    [[nodiscard]] constexpr uint64_t num_blocks() const noexcept { return num_blocks_; }

    /// filter_ref.cuh:52-55: shard of a minimizer hash (its low bits).
    [[nodiscard]] static constexpr uint64_t shard_index(uint64_t minimizer_hash, uint64_t num_blocks) noexcept {
        return minimizer_hash & (num_blocks - 1);
    }
    [[nodiscard]] constexpr uint64_t shard_index(uint64_t minimizer_hash) const noexcept {
        return shard_index(minimizer_hash, num_blocks_);
    }

    /// filter_ref.cuh:70-87: true when every bit of every findere s-mer of the packed
    /// k-mer is set in the shard's four words.
    [[nodiscard]] static constexpr bool sectorized_contains_packed_kmer(uint64_t packed_kmer,
                                                                        const uint64_t* shard_words) {
        bool present = true;
        for (uint64_t offset = 0; offset < Config::findereSpan; ++offset) {
            const uint64_t h = detail::packed_kmer_smer_hash<Config>(packed_kmer, offset);
            detail::forEachHashIndex<Config>([&](auto index) {
                constexpr uint64_t i = decltype(index)::value;
                present &= ((shard_words[i % 4] >> block_type::template sectorizedBitAddress<i>(h)) & 1) != 0;
            });
        }
        return present;
    }

    /// filter_ref.cuh:94-108: one 64-bit device-scope atomic OR per non-zero mask
    /// (defined in cusbf/filter_ref_device.hpp).
    __device__ void apply_word_masks(block_type& block, uint64_t m0, uint64_t m1, uint64_t m2,
                                     uint64_t m3) const;

   private:
    block_type* blocks_ = nullptr;
    uint64_t num_blocks_ = 0;
};

static_assert(std::is_trivially_copyable_v<filter_ref<Config<31, 28, 16, 4>>>);
static_assert(sizeof(filter_block<Config<31, 28, 16, 4>>) == 32);

}  // namespace cusbf
