// sbf_common.hpp -- arithmetic shared by every Super Bloom kernel (host + device).
//
// The hash, minimizer, bit-address and shard formulas must be bit-identical to
// the reference; each cites the reference file:line it follows (paths relative
// to the reference checkout, include/cusbf/...).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sbf {

// hashutil.cuh:192-196 and :207-209 share this multiplier (round(2^64 / phi)).
constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;

// hashutil.cuh:192-196: one multiply then an xorshift by 33.
__host__ __device__ __forceinline__ constexpr uint64_t hash64(uint64_t key) {
    key *= kGolden;
    key ^= key >> 33;
    return key;
}

// hashutil.cuh:207-209: the m-mer hash whose minimum picks the shard.
__host__ __device__ __forceinline__ constexpr uint64_t minimizer_hash64(uint64_t key) {
    return key * kGolden;
}

// filter_common.cuh:27-93 (SaltLiteral<0..15>): salts of the multiply-shift
// bit address used when 64/hashCount < 6 (hash counts 12 and 16). The table
// has to be the reference's exactly, so it is kept in the same order.
__host__ __device__ __forceinline__ constexpr uint64_t salt(uint32_t index) {
    switch (index & 15u) {
        case 0: return 0x9E3779B97F4A7C15ull;
        case 1: return 0xC2B2AE3D27D4EB4Full;
        case 2: return 0x165667B19E3779F9ull;
        case 3: return 0x85EBCA77C2B2AE63ull;
        case 4: return 0x27D4EB2F165667C5ull;
        case 5: return 0x94D049BB133111EFull;
        case 6: return 0xBF58476D1CE4E5B9ull;
        case 7: return 0xD6E8FEB86659FD93ull;
        case 8: return 0xA0761D6478BD642Full;
        case 9: return 0xE7037ED1A0B428DBull;
        case 10: return 0x8EBC6AF09C88C6E3ull;
        case 11: return 0x589965CC75374CC3ull;
        case 12: return 0x1D8E4E27C47D124Full;
        case 13: return 0xEB449C93FBBEA6B5ull;
        case 14: return 0xDB4F0B9175AE2165ull;
        default: return 0xBBE056FDADE14B91ull;
    }
}

// filter_impl.cuh:36-47: bit position (0..63) inside word (index % 4) of the
// shard. Bit slicing when the 64/H-bit slices are at least 6 bits wide,
// otherwise the salted multiply-shift.
__host__ __device__ __forceinline__ uint32_t bit_address(uint32_t hash_count, uint32_t index,
                                                         uint64_t base_hash) {
    const uint32_t slice = 64u / hash_count;
    if (slice >= 6u) {
        return static_cast<uint32_t>((base_hash >> (slice * index)) & 63u);
    }
    return static_cast<uint32_t>((base_hash * salt(index)) >> 58);
}

template <uint32_t H, uint32_t I>
__host__ __device__ __forceinline__ uint32_t bit_address_ct(uint64_t base_hash) {
    constexpr uint32_t slice = 64u / H;
    if constexpr (slice >= 6u) {
        return static_cast<uint32_t>((base_hash >> (slice * I)) & 63u);
    } else {
        return static_cast<uint32_t>((base_hash * salt(I)) >> 58);
    }
}

// filter_common.cuh:118-125: mask of the low (len * bits) bits, all ones at 64.
__host__ __device__ __forceinline__ constexpr uint64_t window_mask(uint32_t bits, uint32_t len) {
    return bits * len >= 64u ? ~0ull : ((1ull << (bits * len)) - 1ull);
}

// Alphabets (Alphabet.cuh:113-218). Codes used by the kernels.
enum AlphabetKind : uint32_t { kDna = 0, kDnaTriplet = 1, kProtein = 2, kCustom = 3 };

// Generic-kernel form: chosen by the window spans, or fixed (tests compare both).
enum GenericForm : uint32_t { kFormAuto = 0, kFormDirect = 1, kFormHashed = 2 };

// Runtime configuration handed to the generic kernels by value.
struct KernelConfig {
    uint32_t k, s, m, h;
    uint32_t alphabet;
    uint32_t symbol_width;
    uint32_t symbol_bits;
    uint32_t invalid_symbol;
    uint32_t form;  // GenericForm
    uint8_t lut[256];
};

// Alphabet.cuh:132-141 (DnaAlphabet::encode).
__host__ __device__ __forceinline__ uint8_t encode_dna(uint8_t byte) {
    const uint8_t upper = byte & 0xDFu;
    const bool valid = (upper == 'A') | (upper == 'C') | (upper == 'G') | (upper == 'T');
    return valid ? static_cast<uint8_t>((byte >> 1u) & 3u) : static_cast<uint8_t>(0xFFu);
}

// Alphabet.cuh:132-141, :167-177, :208-217 and custom tables.
__host__ __device__ __forceinline__ uint8_t encode_symbol(const KernelConfig& c, const char* p) {
    const uint8_t b0 = static_cast<uint8_t>(p[0]);
    switch (c.alphabet) {
        case kDna:
            return encode_dna(b0);
        case kDnaTriplet: {
            const uint8_t a = encode_dna(b0);
            const uint8_t b = encode_dna(static_cast<uint8_t>(p[1]));
            const uint8_t d = encode_dna(static_cast<uint8_t>(p[2]));
            if ((a == 0xFFu) | (b == 0xFFu) | (d == 0xFFu)) return 0xFFu;
            return static_cast<uint8_t>((a << 4u) | (b << 2u) | d);
        }
        case kProtein: {
            const uint8_t idx = static_cast<uint8_t>((b0 & 0xDFu) - 'A');
            return idx < 26u ? idx : static_cast<uint8_t>(0xFFu);
        }
        default:
            return c.lut[b0];
    }
}

}  // namespace sbf
