// cusbf/config.hpp -- Config<K,S,M,HashCount,CudaBlockSize,Alphabet> and alphabets (host).
//
// Same template parameters, derived constants and static checks as the
// reference's include/cusbf/config.cuh:28-118 and the alphabets of
// include/cusbf/Alphabet.cuh:113-218. CudaBlockSize is accepted for source
// compatibility; the wave64 kernels choose their own launch shapes.
#pragma once

#include <bit>
#include <concepts>
#include <cstdint>

namespace cusbf {

template <typename T>
concept Alphabet = requires(const char* input) {
    { T::symbolCount } -> std::convertible_to<uint64_t>;
    { T::symbolWidth } -> std::convertible_to<uint64_t>;
    { T::invalidSymbol } -> std::convertible_to<uint8_t>;
    { T::separator } -> std::convertible_to<uint8_t>;
    { T::validBytes } -> std::convertible_to<const char*>;
    { T::encode(input) } -> std::same_as<uint8_t>;
} && T::symbolCount > 0 && T::symbolCount <= 255 && T::symbolWidth > 0;

/// A0 C1 T2 G3, either case (Alphabet.cuh:113-142).
struct DnaAlphabet {
    static constexpr uint64_t symbolWidth = 1;
    static constexpr uint64_t symbolCount = 4;
    static constexpr uint8_t invalidSymbol = 0xFFu;
    static constexpr uint8_t separator = 'N';
    static constexpr char validBytes[] = "ACGT";

    [[nodiscard]] static constexpr uint8_t encode(const char* input) {
        const auto byte = static_cast<uint8_t>(input[0]);
        const uint8_t upper = byte & 0xDFu;
        const bool valid = upper == 'A' || upper == 'C' || upper == 'G' || upper == 'T';
        return valid ? static_cast<uint8_t>((byte >> 1u) & 3u) : invalidSymbol;
    }
};

/// Non-overlapping DNA triplets as 6-bit symbols (Alphabet.cuh:150-178).
struct DnaTripletAlphabet {
    static constexpr uint64_t symbolWidth = 3;
    static constexpr uint64_t symbolCount = 64;
    static constexpr uint8_t invalidSymbol = 0xFFu;
    static constexpr uint8_t separator = 'N';
    static constexpr char validBytes[] = "ACGT";

    [[nodiscard]] static constexpr uint8_t encode(const char* input) {
        const uint8_t a = DnaAlphabet::encode(input + 0);
        const uint8_t b = DnaAlphabet::encode(input + 1);
        const uint8_t c = DnaAlphabet::encode(input + 2);
        if (a == invalidSymbol || b == invalidSymbol || c == invalidSymbol) return invalidSymbol;
        return static_cast<uint8_t>((a << 4u) | (b << 2u) | c);
    }
};

/// A-Z (either case) as 0..25 (Alphabet.cuh:189-218).
struct ProteinAlphabet {
    static constexpr uint64_t symbolWidth = 1;
    static constexpr uint64_t symbolCount = 26;
    static constexpr uint8_t invalidSymbol = 0xFFu;
    static constexpr uint8_t separator = '*';
    static constexpr char validBytes[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZ";

    [[nodiscard]] static constexpr uint8_t encode(const char* input) {
        const auto byte = static_cast<uint8_t>(input[0]);
        const auto idx = static_cast<uint8_t>((byte & 0xDFu) - 'A');
        return idx < 26 ? idx : invalidSymbol;
    }
};

template <uint16_t K_, uint16_t S_, uint16_t M_, uint64_t HashCount_ = 4,
          uint64_t CudaBlockSize_ = 256, Alphabet Alphabet_ = DnaAlphabet>
struct Config {
    using Alphabet = Alphabet_;
    static constexpr uint16_t k = K_;
    static constexpr uint16_t m = M_;
    static constexpr uint16_t s = S_;
    static constexpr uint64_t hashCount = HashCount_;
    static constexpr uint64_t alphabetSize = Alphabet::symbolCount;
    static constexpr uint64_t symbolWidth = Alphabet::symbolWidth;
    static constexpr uint64_t symbolBits = std::bit_width(alphabetSize - 1);
    static constexpr uint64_t symbolMask = (uint64_t{1} << symbolBits) - 1;
    static constexpr uint64_t filterBlockBits = 256;
    static constexpr uint64_t cudaBlockSize = CudaBlockSize_;
    static constexpr uint64_t wordBits = 64;
    static constexpr uint64_t blockWordCount = filterBlockBits / wordBits;
    static constexpr uint64_t minimizerSpan = k - m + 1;
    static constexpr uint64_t findereSpan = k - s + 1;

    static_assert(k > 0, "k must be positive");
    static_assert(m > 0 && m <= k, "m must satisfy 0 < m <= k");
    static_assert(s > 0 && s <= k, "s must satisfy 0 < s <= k");
    static_assert(k * symbolBits <= 64, "k-mer must fit in one packed uint64_t");
    static_assert(hashCount > 0, "At least one Bloom hash is required");
    static_assert(hashCount <= 16, "This implementation provides 16 multiplicative salts");
    static_assert(hashCount >= blockWordCount,
                  "Sectorized layout requires hashCount >= blockWordCount");
    static_assert(hashCount % blockWordCount == 0,
                  "Hash count must distribute evenly across shard words");
    static_assert(cudaBlockSize % 32 == 0, "CUDA block size must be a multiple of one warp");
};

}  // namespace cusbf
