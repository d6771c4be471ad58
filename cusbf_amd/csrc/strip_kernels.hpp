// strip_kernels.hpp -- wave64 "lane-strip" Super Bloom insert / query kernels for gfx950
// (device code and launcher templates; instantiated by kernels_strip.hip and
// kernels_strip_batched.hip).
//
// Same results as the reference kernels (include/cusbf/detail/kernels.cuh:40-172
// and :179-312), computed differently:
//
//  * Stage 1 (per wave): each lane reads 16 input symbols with one 128-bit load
//    (or one 32-bit half word of the dense packed DNA input) and writes them to
//    LDS in the front end's layout, plus a 16-bit invalid-symbol mask:
//      - Dna2: SWAR form of DnaAlphabet::encode (Alphabet.cuh:132-141) into one
//        32-bit word of 2-bit symbols, first symbol in the top bits;
//      - Protein5: ProteinAlphabet::encode (Alphabet.cuh:208-217) into 16 code
//        bytes (5-bit symbols).
//  * Stage 2 (per lane): lane l owns the 16 windows starting at symbols
//    16l..16l+15 of the wave tile. It forms every m-mer and s-mer starting in
//    its 16 positions, hashes each of them ONCE (the reference hashes k-m+1
//    m-mers and k-s+1 s-mers per window, sequence_kmer.cuh:40-59), and takes
//    the hashes of the next lane's first positions through lane shuffles. The
//    window minimum uses the van Herk / Gil-Werman split when the minimizer
//    span is 16 (one block per lane), and a direct sliding minimum otherwise.
//  * Window validity is a bit smear of the invalid-symbol mask (the reference
//    scans the k symbols, sequence_kmer.cuh:92-101).
//  * Runs of consecutive windows with the same shard (super-k-mer pieces) get
//    LDS slots; the query loads each run's 32-byte shard once with dense
//    all-lane loads and tests every window against it (filter_ref.cuh:70-87);
//    the insert ORs each run's masks in LDS and issues one 64-bit atomic OR per
//    non-zero word (filter_ref.cuh:94-108 issues them per warp run).
//
// Two drivers share the tile code:
//  * plain: one sequence (the reference's insert_sequence / contains_sequence
//    and the normalised record batch with separators); waves stride over tiles;
//  * batched: many records concatenated WITHOUT separators plus an offsets
//    array; a prepass turns the offsets into a record-start bitmap, each lane
//    loads its 16 bits of it beside its 16 symbols, and windows that cross a
//    record start are invalid (the answer the reference gives for a window
//    across a separator, normalized_record_batch.hpp:208-232). Waves take
//    tiles from a work queue (one atomic counter) or stride over them.
//
// The last HALO lanes of each wave own no windows; they only produce the
// symbols and hashes their left neighbours read.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "launch.hpp"
#include "sbf_common.hpp"

// Minimum waves per SIMD the register allocation of the plain DNA kernels must allow.
#ifndef SBF_DNA_WAVES_PER_EU
#define SBF_DNA_WAVES_PER_EU 3
#endif

namespace sbf::strip {

constexpr int kWave = 64;
constexpr int kLaneWindows = 16;
constexpr int kWavesPerBlock = 4;
constexpr int kBlock = kWave * kWavesPerBlock;
constexpr int kLdsWords = kWave + 4;

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// --- front ends -------------------------------------------------------------

// DNA, 2-bit symbols packed in 32-bit words, first symbol in the top bits.
struct Dna2 {
    static constexpr int kBits = 2;
    static constexpr int kSlots = 256;  // LDS run slots per wave tile (~100-120 runs used)
    static constexpr int kMaxK = 32;
    static constexpr uint32_t kAlphabet = kDna;
    struct Lds {
        uint32_t p[kLdsWords];
    };
    struct Ctx {
        uint64_t top;  // symbols 0..31 of the lane's span
        uint32_t w2;   // symbols 32..47
    };
};

// Protein (or any 1-byte alphabet of at most 32 symbols): 16 code bytes per lane.
struct Protein5 {
    static constexpr int kBits = 5;
    // ~330 runs per tile with a 5-m-mer span: 384 slots take most tiles in one
    // chunk (one pass of the front end); the LDS then holds two blocks per CU
    static constexpr int kSlots = 384;
    static constexpr int kMaxK = 12;
    static constexpr uint32_t kAlphabet = kProtein;
    struct Lds {
        u32x4 c[kLdsWords];
    };
    struct Ctx {
        uint32_t d[8];  // code bytes of symbols 0..31 of the lane's span, first in the low byte
    };
};

template <int K, int S, int M, int H, class FE>
struct Shape {
    static constexpr int kMinSpan = K - M + 1;      // m-mers per window (config.cuh:64)
    static constexpr int kFindereSpan = K - S + 1;  // s-mers per window (config.cuh:66)
    static constexpr int kHaloLanes = (K - 1 + 15) / 16 > 1 ? (K - 1 + 15) / 16 : 1;
    static constexpr int kWindowLanes = kWave - kHaloLanes;
    static constexpr int kWaveWindows = kWindowLanes * kLaneWindows;
    static_assert(K >= 1 && K <= FE::kMaxK, "k-mer does not fit the front end's packed word");
    static_assert(K * FE::kBits <= 64, "k * symbolBits must fit 64 bits (config.cuh:74-76)");
    static_assert(M >= 1 && M <= K && S >= 1 && S <= K, "m and s must be in [1, k]");
    static_assert(kMinSpan <= kLaneWindows + 1, "minimizer span must fit two lanes");
    static_assert(kFindereSpan <= 8, "findere span of the lane-strip kernels is at most 8");
    static_assert(H == 4 || H == 8 || H == 12 || H == 16, "hash count must be 4, 8, 12 or 16");
    static_assert(FE::kBits != 5 || K <= 17, "byte-code front end reads one neighbour lane");
};

// --- stage 1: Dna2 ------------------------------------------------------------

// Four input bytes (first symbol in the low byte) -> 8 bits of 2-bit symbols
// (first symbol in the top bits) and a 4-bit invalid mask (bit i = byte i).
__device__ __forceinline__ void encode4(uint32_t x, uint32_t& packed, uint32_t& invalid) {
    const uint32_t sym = (x >> 1) & 0x03030303u;    // (byte >> 1) & 3: A0 C1 T2 G3
    const uint32_t upper = x & 0xDFDFDFDFu;         // case fold for validation only
    // Letter each code must come from: table bytes 'A','C','T','G' picked by code.
    const uint32_t expect = __builtin_amdgcn_perm(0u, 0x47544341u, sym);
    const uint32_t diff = upper ^ expect;  // zero byte <=> valid symbol
    const uint32_t nz = ((((diff & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | diff) & 0x80808080u) >> 7;
    uint32_t inv = nz | (nz >> 7);  // bits 0, 8, 16, 24 -> 0, 1 (+ copies)
    inv |= inv >> 14;               // bits 16, 17 -> 2, 3
    invalid = inv & 0xFu;
    uint32_t r = __builtin_amdgcn_perm(0u, sym, 0x00010203u);  // byte reverse
    r |= r >> 6;
    r |= r >> 12;
    packed = r & 0xFFu;
}

__device__ __forceinline__ void encode16(u32x4 v, uint32_t& packed, uint32_t& invalid) {
    uint32_t p0, p1, p2, p3, i0, i1, i2, i3;
    encode4(v.x, p0, i0);
    encode4(v.y, p1, i1);
    encode4(v.z, p2, i2);
    encode4(v.w, p3, i3);
    packed = (p0 << 24) | (p1 << 16) | (p2 << 8) | p3;
    invalid = i0 | (i1 << 4) | (i2 << 8) | (i3 << 12);
}

// The 16 input bytes of a lane's chunk when they lie in bounds and 16-byte
// aligned (one 128-bit load, issued ahead of its use); otherwise `fast` is
// false and the chunk is read byte by byte where it is used.
struct RawChunk {
    u32x4 v;
    bool fast;
};

__device__ __forceinline__ RawChunk load_raw(const SequenceInput& in, uint64_t first) {
    RawChunk r;
    r.v = u32x4{0u, 0u, 0u, 0u};
    const char* src = in.seq + first;
    r.fast = first + 16 <= in.num_symbols && (reinterpret_cast<uintptr_t>(src) & 15u) == 0;
    if (r.fast) r.v = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(src));
    return r;
}

// Byte input: chunk of 16 symbols starting at symbol `first`.
__device__ __forceinline__ void load_chunk_bytes(const char* seq, uint64_t num_symbols,
                                                 uint64_t first, uint32_t& packed,
                                                 uint32_t& invalid) {
    const char* src = seq + first;
    if (first + 16 <= num_symbols && (reinterpret_cast<uintptr_t>(src) & 15u) == 0) {
        const u32x4 v = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(src));
        encode16(v, packed, invalid);
        return;
    }
    packed = 0;
    invalid = 0;
    for (int i = 0; i < 16; ++i) {
        const uint64_t p = first + static_cast<uint64_t>(i);
        const uint8_t code = p < num_symbols ? encode_dna(static_cast<uint8_t>(seq[p])) : 0xFFu;
        if (code == 0xFFu) {
            invalid |= 1u << i;
        } else {
            packed |= static_cast<uint32_t>(code) << (30 - 2 * i);
        }
    }
}

// Dense packed input (dense_packed.cuh:33-39: LSB-first, 32 symbols per word):
// the 16 symbols starting at `first` (a multiple of 16) are one 32-bit half
// word; reversing its bit order and swapping each bit pair puts the first
// symbol in the top bits. A decoded DNA symbol never equals the 0xFF sentinel,
// so only symbols past the end are marked (dense_packed.cuh:110-117).
__device__ __forceinline__ void load_chunk_dense(const uint64_t* words, uint64_t num_words,
                                                 uint64_t num_symbols, uint64_t first,
                                                 uint32_t& packed, uint32_t& invalid) {
    const uint64_t half = first >> 4;
    uint32_t x = 0;
    if (half < 2 * num_words) {
        x = __builtin_nontemporal_load(reinterpret_cast<const uint32_t*>(words) + half);
    }
    uint32_t r = __builtin_bitreverse32(x);
    r = ((r >> 1) & 0x55555555u) | ((r & 0x55555555u) << 1);
    packed = r;
    if (first + 16 <= num_symbols) {
        invalid = 0;
    } else if (first >= num_symbols) {
        invalid = 0xFFFFu;
    } else {
        invalid = (0xFFFFu << static_cast<uint32_t>(num_symbols - first)) & 0xFFFFu;
    }
}

// Byte input from a chunk loaded ahead (load_raw).
__device__ __forceinline__ void stage_lane_raw(const RawChunk& r, const SequenceInput& in, uint64_t first,
                                               int lane, Dna2::Lds& st, uint32_t* inv) {
    uint32_t packed, invalid;
    if (r.fast) {
        encode16(r.v, packed, invalid);
    } else {
        load_chunk_bytes(in.seq, in.num_symbols, first, packed, invalid);
    }
    st.p[lane] = packed;
    inv[lane] = invalid;
    if (lane < kLdsWords - kWave) {
        st.p[kWave + lane] = 0;
        inv[kWave + lane] = 0xFFFFu;
    }
}

template <bool Dense>
__device__ __forceinline__ void stage_lane(const SequenceInput& in, uint64_t first, int lane,
                                           Dna2::Lds& st, uint32_t* inv) {
    uint32_t packed, invalid;
    if constexpr (Dense) {
        load_chunk_dense(in.words, in.num_words, in.num_symbols, first, packed, invalid);
    } else {
        load_chunk_bytes(in.seq, in.num_symbols, first, packed, invalid);
    }
    st.p[lane] = packed;
    inv[lane] = invalid;
    if (lane < kLdsWords - kWave) {
        st.p[kWave + lane] = 0;
        inv[kWave + lane] = 0xFFFFu;
    }
}

__device__ __forceinline__ Dna2::Ctx lane_context(const Dna2::Lds& st, int lane) {
    Dna2::Ctx c;
    c.top = (static_cast<uint64_t>(st.p[lane]) << 32) | st.p[lane + 1];
    c.w2 = st.p[lane + 2];
    return c;
}

// Values of the L-symbol substrings starting at positions 0..15 (MSB-first,
// filter_common.cuh:128-134); position p is bits [2p, 2p + 2L) of top:w2
// counted from the top.
// Value of the L-symbol substring starting at position p (0 <= p, p + L <=
// 48; MSB-first, filter_common.cuh:128-134): top:w2 as one 96-bit value
// B = t1:t0:w, the substring is (B >> r) masked to 2L bits, r = 96 - 2p - 2L,
// taken with funnel shifts.
template <int L>
__device__ __forceinline__ uint64_t substring_at(const Dna2::Ctx& c, int p) {
    const uint32_t t1 = static_cast<uint32_t>(c.top >> 32);
    const uint32_t t0 = static_cast<uint32_t>(c.top);
    const uint32_t w = c.w2;
    const int r = 96 - 2 * p - 2 * L;
    uint32_t lo, hi;
    if (r >= 64) {
        lo = t1 >> (r - 64);
        hi = 0;
    } else if (r >= 32) {
        lo = __builtin_amdgcn_alignbit(t1, t0, static_cast<uint32_t>(r - 32));
        hi = t1 >> (r - 32);
    } else {
        lo = __builtin_amdgcn_alignbit(t0, w, static_cast<uint32_t>(r));
        hi = __builtin_amdgcn_alignbit(t1, t0, static_cast<uint32_t>(r));
    }
    const uint64_t x = (static_cast<uint64_t>(hi) << 32) | lo;
    return L == 32 ? x : (x & ((1ull << (2 * L)) - 1ull));
}

// The substrings starting at positions 0..15.
template <int L>
__device__ __forceinline__ void substring_values(const Dna2::Ctx& c, uint64_t (&v)[kLaneWindows]) {
#pragma unroll
    for (int p = 0; p < kLaneWindows; ++p) v[p] = substring_at<L>(c, p);
}

// --- stage 1: Protein5 --------------------------------------------------------

// ProteinAlphabet::encode (Alphabet.cuh:208-217) of four bytes at once: the
// code bytes (0..25, or 0xFF) and a 4-bit invalid mask.
__device__ __forceinline__ uint32_t encode_protein4(uint32_t x, uint32_t& invalid) {
    uint32_t codes = 0;
    invalid = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const uint32_t idx = (((x >> (8 * i)) & 0xDFu) - 0x41u) & 0xFFu;
        const bool ok = idx < 26u;
        codes |= (ok ? idx : 0xFFu) << (8 * i);
        invalid |= (ok ? 0u : 1u) << i;
    }
    return codes;
}

// The 16 code bytes and the invalid mask of a lane's chunk (loaded ahead in `r`,
// or read byte by byte when it is not a fast chunk).
__device__ __forceinline__ void encode_protein_chunk(const RawChunk& r, const SequenceInput& in, uint64_t first,
                                                     u32x4& codes, uint32_t& invalid) {
    if (r.fast) {
        const u32x4 v = r.v;
        uint32_t i0, i1, i2, i3;
        codes.x = encode_protein4(v.x, i0);
        codes.y = encode_protein4(v.y, i1);
        codes.z = encode_protein4(v.z, i2);
        codes.w = encode_protein4(v.w, i3);
        invalid = i0 | (i1 << 4) | (i2 << 8) | (i3 << 12);
    } else {
        uint32_t w[4] = {0, 0, 0, 0};
        invalid = 0;
        for (int i = 0; i < 16; ++i) {
            const uint64_t p = first + static_cast<uint64_t>(i);
            uint32_t code = 0xFFu;
            if (p < in.num_symbols) {
                const uint32_t idx = ((static_cast<uint8_t>(in.seq[p]) & 0xDFu) - 0x41u) & 0xFFu;
                code = idx < 26u ? idx : 0xFFu;
            }
            w[i >> 2] |= code << (8 * (i & 3));
            if (code == 0xFFu) invalid |= 1u << i;
        }
        codes = u32x4{w[0], w[1], w[2], w[3]};
    }
}

__device__ __forceinline__ void stage_lane_raw(const RawChunk& r, const SequenceInput& in, uint64_t first,
                                               int lane, Protein5::Lds& st, uint32_t* inv) {
    u32x4 codes;
    uint32_t invalid;
    encode_protein_chunk(r, in, first, codes, invalid);
    st.c[lane] = codes;
    inv[lane] = invalid;
    if (lane < kLdsWords - kWave) {
        st.c[kWave + lane] = u32x4{0u, 0u, 0u, 0u};
        inv[kWave + lane] = 0xFFFFu;
    }
}

template <bool Dense>
__device__ __forceinline__ void stage_lane(const SequenceInput& in, uint64_t first, int lane,
                                           Protein5::Lds& st, uint32_t* inv) {
    static_assert(!Dense, "the byte-code front end reads byte sequences only");
    stage_lane_raw(load_raw(in, first), in, first, lane, st, inv);
}

__device__ __forceinline__ Protein5::Ctx lane_context(const Protein5::Lds& st, int lane) {
    Protein5::Ctx c;
    const u32x4 a = st.c[lane];
    const u32x4 b = st.c[lane + 1];
    c.d[0] = a.x; c.d[1] = a.y; c.d[2] = a.z; c.d[3] = a.w;
    c.d[4] = b.x; c.d[5] = b.y; c.d[6] = b.z; c.d[7] = b.w;
    return c;
}

// MSB-first packing of the 5-bit symbols p..p+L-1 (sequence_kmer.cuh:74-89
// packs `symbol & symbolMask`; p + L <= 32).
template <int L>
__device__ __forceinline__ uint64_t substring_at(const Protein5::Ctx& c, int p) {
    constexpr uint64_t mask = L * 5 >= 64 ? ~0ull : ((1ull << (5 * L)) - 1ull);
    uint64_t acc = 0;
#pragma unroll
    for (int i = 0; i < L; ++i) {
        const int j = p + i;
        acc = (acc << 5) | ((c.d[j >> 2] >> (8 * (j & 3))) & 0x1Fu);
    }
    return acc & mask;
}

// Rolling MSB-first packing of 5-bit symbols (sequence_kmer.cuh:74-89 packs
// `symbol & symbolMask`); windows holding an invalid symbol are never used.
template <int L>
__device__ __forceinline__ void substring_values(const Protein5::Ctx& c,
                                                 uint64_t (&v)[kLaneWindows]) {
    constexpr uint64_t mask = L * 5 >= 64 ? ~0ull : ((1ull << (5 * L)) - 1ull);
    uint64_t acc = 0;
#pragma unroll
    for (int i = 0; i < kLaneWindows + L - 1; ++i) {
        const uint64_t sym = (c.d[i >> 2] >> (8 * (i & 3))) & 0x1Fu;
        acc = (acc << 5) | sym;
        if (i >= L - 1) v[i - (L - 1)] = acc & mask;
    }
}

// --- stage 2 helpers ---------------------------------------------------------

// y_j = OR of x_{j..j+N-1}: bit j of the result is set when window j holds an
// invalid symbol (N = k) or a record start after its first symbol (N = k-1 on
// the mask shifted by one).
template <int N>
__device__ __forceinline__ uint64_t smear(uint64_t x) {
    if constexpr (N <= 0) {
        return 0;
    } else {
        uint64_t y = x;
        int len = 1;
#pragma unroll
        for (int step = 0; step < 6; ++step) {
            if (2 * len <= N) {
                y |= y >> len;
                len *= 2;
            }
        }
        if (len < N) y |= y >> (N - len);
        return y;
    }
}

// Lane exchanges as DPP wave shifts (one VALU move per 32 bits, no LDS access):
// wave_shl:1 (DPP control 0x130) gives lane i the value of lane i + 1, wave_shr:1
// (0x138) the value of lane i - 1; the lane at the wave's edge keeps its own value.
__device__ __forceinline__ uint32_t from_next_lane(uint32_t v) {
    return static_cast<uint32_t>(
        __builtin_amdgcn_update_dpp(static_cast<int>(v), static_cast<int>(v), 0x130, 0xF, 0xF, false));
}
__device__ __forceinline__ uint32_t from_prev_lane(uint32_t v) {
    return static_cast<uint32_t>(
        __builtin_amdgcn_update_dpp(static_cast<int>(v), static_cast<int>(v), 0x138, 0xF, 0xF, false));
}
__device__ __forceinline__ uint64_t next_lane(uint64_t v) {
    return static_cast<uint64_t>(from_next_lane(static_cast<uint32_t>(v))) |
           (static_cast<uint64_t>(from_next_lane(static_cast<uint32_t>(v >> 32))) << 32);
}

__device__ __forceinline__ uint64_t umin64(uint64_t a, uint64_t b) { return a < b ? a : b; }

// Bit positions of one s-mer: 6 bits per hash index, hash i at bits [6i, 6i+6).
template <int H>
struct PosCode {
    static constexpr int kWords = (6 * H + 31) / 32;
    uint32_t w[kWords];
};

// filter_impl.cuh:36-47: the H bit positions of an s-mer hash.
template <int H>
__device__ __forceinline__ PosCode<H> pos_code(uint64_t h) {
    PosCode<H> c;
#pragma unroll
    for (int j = 0; j < PosCode<H>::kWords; ++j) c.w[j] = 0;
#pragma unroll
    for (int i = 0; i < H; ++i) {
        uint32_t pos;
        constexpr uint32_t slice = 64u / H;
        if constexpr (slice >= 6u) {
            pos = static_cast<uint32_t>((h >> (slice * i)) & 63u);
        } else {
            pos = static_cast<uint32_t>((h * salt(static_cast<uint32_t>(i))) >> 58);
        }
        const int bit = 6 * i;
        c.w[bit / 32] |= pos << (bit % 32);
        if (bit % 32 > 26) c.w[bit / 32 + 1] |= pos >> (32 - bit % 32);
    }
    return c;
}

template <int H>
__device__ __forceinline__ uint32_t code_pos(const PosCode<H>& c, int i) {
    const int bit = 6 * i;
    uint32_t v = c.w[bit / 32] >> (bit % 32);
    if (bit % 32 > 26) v |= c.w[bit / 32 + 1] << (32 - bit % 32);
    return v & 63u;
}

// filter_impl.cuh:54-77: hash i sets bit pos_i of word i % 4.
template <int H>
__device__ __forceinline__ void code_masks(const PosCode<H>& c, uint64_t (&m)[4]) {
    m[0] = m[1] = m[2] = m[3] = 0;
#pragma unroll
    for (int i = 0; i < H; ++i) m[i & 3] |= 1ull << code_pos<H>(c, i);
}

// What a lane keeps per s-mer: the bit positions when they come from slices of the
// hash (64 / H >= 6: H = 4 or 8); the s-mer hash itself when they come from salted
// multiplies (H = 12 or 16), whose 6-bit positions would take more registers than
// the hash and are computed when the masks are built.
template <int H>
using SmerCode = typename std::conditional<(64 / H >= 6), PosCode<H>, uint64_t>::type;

template <int H>
__device__ __forceinline__ SmerCode<H> smer_code(uint64_t h) {
    if constexpr (64 / H >= 6) {
        return pos_code<H>(h);
    } else {
        return h;
    }
}

// filter_impl.cuh:36-47 salted multiply-shift and :54-77 masks, from the s-mer hash.
template <int H>
__device__ __forceinline__ void code_masks(const uint64_t& h, uint64_t (&m)[4]) {
    m[0] = m[1] = m[2] = m[3] = 0;
#pragma unroll
    for (int i = 0; i < H; ++i) m[i & 3] |= 1ull << ((h * salt(static_cast<uint32_t>(i))) >> 58);
}


// --- per-wave LDS ---------------------------------------------------------------
//
// A run is a maximal sequence of consecutive valid windows of the wave tile
// with the same shard (a super-k-mer cut at the tile edges). The first FE::kSlots
// runs of a tile get a slot in the wave's LDS: the query loads each run's
// 32-byte shard once into its slot with dense, all-lane loads; the insert ORs
// each run's masks into its slot and then issues one 64-bit atomic OR per
// non-zero word. A tile with more runs than slots (sequences whose minimizer
// changes at nearly every position) handles the excess runs with direct
// global loads / atomics in the same pass.

template <class FE, bool Batched>
struct WaveLds {
    static constexpr int kSlots = FE::kSlots;
    static constexpr bool kBatched = Batched;
    typename FE::Lds st;
    uint32_t inv[kLdsWords];
    uint32_t bnd[Batched ? kLdsWords : 1];  // bit i of word l: symbol 16l+i starts a record
    uint32_t slot_shard[FE::kSlots];
    alignas(16) uint64_t slot_w[FE::kSlots][4];
};

// Orders this wave's LDS writes before later LDS reads by its other lanes. LDS
// operations of one wave execute in order, so it is enough to wait for them
// and keep the compiler from moving accesses across this point; outstanding
// global loads and atomics are not waited for.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0), vmcnt and expcnt unconstrained
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
}

// Exclusive prefix sum over the wave of c (< 64) and the wave total.
__device__ __forceinline__ uint32_t wave_exclusive_sum(uint32_t c, int lane, uint32_t& total) {
    const uint64_t below = (1ull << lane) - 1ull;
    uint32_t base = 0;
    total = 0;
#pragma unroll
    for (int b = 0; b < 6; ++b) {
        const uint64_t bal = __ballot((c >> b) & 1u);
        base += static_cast<uint32_t>(__popcll(bal & below)) << b;
        total += static_cast<uint32_t>(__popcll(bal)) << b;
    }
    return base;
}

// Batched form: the record-start bits of the 16 symbols starting at `first`
// (a multiple of 16); bit i set when symbol first+i starts a record. Chunks
// past the bitmap read as no starts (their symbols are past the end anyway).
__device__ __forceinline__ uint32_t record_starts16(const SequenceInput& in, uint64_t first) {
    const uint64_t chunk = first >> 4;
    if (chunk >= 4 * start_words(in.num_symbols)) return 0;
    return reinterpret_cast<const uint16_t*>(in.starts)[chunk];
}

// --- stage 2 ---------------------------------------------------------------------

// Per-lane shard indices, window validity and s-mer bit positions.
template <int K, int S, int M, int H, class FE>
struct LaneWindows {
    uint32_t shard[kLaneWindows];
    uint32_t valid;  // bit o: window o valid and in range
    SmerCode<H> code[kLaneWindows + Shape<K, S, M, H, FE>::kFindereSpan - 1];
};

template <int K, int S, int M, int H, class FE, bool Batched>
__device__ __forceinline__ void lane_windows(const WaveLds<FE, Batched>& L, int lane,
                                             bool owns_windows, uint64_t first_window,
                                             uint64_t num_windows, uint32_t shard_mask,
                                             LaneWindows<K, S, M, H, FE>& lw) {
    using Sh = Shape<K, S, M, H, FE>;
    constexpr int W = Sh::kMinSpan;
    constexpr int F = Sh::kFindereSpan;

    const uint64_t inv = static_cast<uint64_t>(L.inv[lane]) |
                         (static_cast<uint64_t>(L.inv[lane + 1]) << 16) |
                         (static_cast<uint64_t>(L.inv[lane + 2]) << 32);
    uint64_t bad = smear<K>(inv);
    if constexpr (Batched) {
        const uint64_t bnd = static_cast<uint64_t>(L.bnd[lane]) |
                             (static_cast<uint64_t>(L.bnd[lane + 1]) << 16) |
                             (static_cast<uint64_t>(L.bnd[lane + 2]) << 32);
        bad |= smear<K - 1>(bnd >> 1);
    }
    uint32_t valid = ~static_cast<uint32_t>(bad) & 0xFFFFu;
    if (!owns_windows) valid = 0;
    if (first_window >= num_windows) {
        valid = 0;
    } else if (num_windows - first_window < kLaneWindows) {
        valid &= (1u << static_cast<uint32_t>(num_windows - first_window)) - 1u;
    }
    lw.valid = valid;

    const typename FE::Ctx ctx = lane_context(L.st, lane);
    uint64_t mh[kLaneWindows];
    substring_values<M>(ctx, mh);
#pragma unroll
    for (int p = 0; p < kLaneWindows; ++p) mh[p] = minimizer_hash64(mh[p]);

    uint64_t mn[kLaneWindows];
    if constexpr (W == kLaneWindows) {
        // van Herk / Gil-Werman: window o = suffix-min of own positions [o, 16)
        // joined with prefix-min of the next lane's positions [0, o).
        uint64_t pre[kLaneWindows - 1];
        pre[0] = mh[0];
#pragma unroll
        for (int i = 1; i < kLaneWindows - 1; ++i) pre[i] = umin64(pre[i - 1], mh[i]);
        uint64_t pre_next[kLaneWindows - 1];
#pragma unroll
        for (int i = 0; i < kLaneWindows - 1; ++i) pre_next[i] = next_lane(pre[i]);
        uint64_t suf = mh[kLaneWindows - 1];
        mn[kLaneWindows - 1] = umin64(suf, pre_next[kLaneWindows - 2]);
#pragma unroll
        for (int o = kLaneWindows - 2; o >= 0; --o) {
            suf = umin64(mh[o], suf);
            mn[o] = o == 0 ? suf : umin64(suf, pre_next[o - 1]);
        }
    } else {
        uint64_t v[kLaneWindows + W - 1];
#pragma unroll
        for (int i = 0; i < kLaneWindows; ++i) v[i] = mh[i];
#pragma unroll
        for (int i = 0; i < W - 1; ++i) v[kLaneWindows + i] = next_lane(mh[i]);
#pragma unroll
        for (int o = 0; o < kLaneWindows; ++o) {
            uint64_t best = v[o];
#pragma unroll
            for (int d = 1; d < W; ++d) best = umin64(best, v[o + d]);
            mn[o] = best;
        }
    }
    // filter_ref.cuh:52-55 with the uint32_t narrowing of kernels.cuh:240-241.
#pragma unroll
    for (int o = 0; o < kLaneWindows; ++o) lw.shard[o] = static_cast<uint32_t>(mn[o]) & shard_mask;

    // s-mer bit positions (second pass keeps the minimizer temporaries dead)
    uint64_t sv[kLaneWindows];
    substring_values<S>(ctx, sv);
#pragma unroll
    for (int p = 0; p < kLaneWindows; ++p) lw.code[p] = smer_code<H>(hash64(sv[p]));
    // s-mers 16 .. 16 + F - 2 (the first ones of the next lane's span) lie in
    // this lane's context too: computed here, no lane exchange
#pragma unroll
    for (int i = 0; i < F - 1; ++i) {
        lw.code[kLaneWindows + i] = smer_code<H>(hash64(substring_at<S>(ctx, kLaneWindows + i)));
    }
}

// Bit o of the result: window o starts a run. Bit o of `cont`: window o
// continues the run of the window before it (for o = 0, the previous lane's
// last window).
template <class LW>
__device__ __forceinline__ uint32_t run_heads(const LW& lw, int lane, uint32_t& cont) {
    const uint32_t prev_shard = from_prev_lane(lw.shard[kLaneWindows - 1]);
    const uint32_t prev_valid = from_prev_lane((lw.valid >> (kLaneWindows - 1)) & 1u);
    uint32_t same = 0;
#pragma unroll
    for (int o = 1; o < kLaneWindows; ++o) same |= (lw.shard[o] == lw.shard[o - 1] ? 1u : 0u) << o;
    same &= lw.valid & (lw.valid << 1);
    if (lane > 0 && prev_valid && (lw.valid & 1u) && lw.shard[0] == prev_shard) same |= 1u;
    cont = same;
    return lw.valid & ~same;
}

// Publishes the shard of every run head that has an LDS slot; `base` is the
// slot of the lane's first head (wrapping below zero for heads of an earlier
// chunk, which are skipped like those past the slots).
template <class LW, class Lds>
__device__ __forceinline__ void publish_heads(const LW& lw, uint32_t heads, uint32_t base, Lds& L) {
#pragma unroll
    for (int o = 0; o < kLaneWindows; ++o) {
        if ((heads >> o) & 1u) {
            const uint32_t slot = base + static_cast<uint32_t>(__popc(heads & ((1u << o) - 1u)));
            if (slot < static_cast<uint32_t>(Lds::kSlots)) L.slot_shard[slot] = lw.shard[o];
        }
    }
}

// Zeroes the slots' mask words (each lane its share).
template <class Lds>
__device__ __forceinline__ void zero_slots(Lds& L, int lane) {
#pragma unroll
    for (int j = 0; j < Lds::kSlots / kWave; ++j) {
        const u32x4 z = {0u, 0u, 0u, 0u};
        u32x4* dst = reinterpret_cast<u32x4*>(L.slot_w[lane + j * kWave]);
        dst[0] = z;
        dst[1] = z;
    }
}

// First window of tile `tile` (wave-uniform), and the first window of the
// lane's 16 within it. The empty asm statement pins the tile's offset as a
// scalar register value of its own: otherwise the compiler fuses
// tile * windows + 16 * lane into one multiply-add whose 64-bit lane operand
// stays live in a vector register pair for the whole kernel.
template <int WaveWindows>
__device__ __forceinline__ uint64_t tile_start(uint64_t tile) {
    uint64_t t = tile * static_cast<uint64_t>(WaveWindows);
    asm volatile("" : "+s"(t));
    return t;
}

__device__ __forceinline__ uint64_t lane_start(uint64_t tile_first, int lane) {
    return tile_first + (static_cast<uint32_t>(lane) << 4);
}

template <int K, int S, int M, int H, class FE, bool Dense, bool Batched>
__device__ __forceinline__ void stage_tile(const SequenceInput& in, uint64_t tile, int lane,
                                           WaveLds<FE, Batched>& L) {
    using Sh = Shape<K, S, M, H, FE>;
    const uint64_t first = lane_start(tile_start<Sh::kWaveWindows>(tile), lane);
    uint32_t starts = 0;
    if constexpr (Batched) starts = record_starts16(in, first);  // issued beside the symbol load
    stage_lane<Dense>(in, first, lane, L.st, L.inv);
    if constexpr (Batched) L.bnd[lane] = starts;
}

// The halo record-start words stay zero (the halo symbols are all invalid).
template <class Lds>
__device__ __forceinline__ void init_wave_lds(Lds& L, int lane) {
    if constexpr (Lds::kBatched) {
        if (lane < kLdsWords - kWave) L.bnd[kWave + lane] = 0;
        wave_lds_sync();
    }
}

// --- tile bodies -----------------------------------------------------------------

__device__ __forceinline__ uint32_t lo32(uint64_t x) { return static_cast<uint32_t>(x); }
__device__ __forceinline__ uint32_t hi32(uint64_t x) { return static_cast<uint32_t>(x >> 32); }

// (a | c) & ~b in one v_bitop3_b32 (truth table 0x32: symmetric in a and c,
// so it does not depend on which operand indexes the table's high bit).
__device__ __forceinline__ uint32_t or_and_not(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x32);
}

// Shard words of window o's run: its LDS slot, or (a run past the slots) the
// filter itself.
template <int K, int S, int M, int H, class FE, bool Batched, bool AllSlotted>
__device__ __forceinline__ void run_words(const LaneWindows<K, S, M, H, FE>& lw, int o,
                                          uint32_t heads, uint32_t base,
                                          const WaveLds<FE, Batched>& L, const FilterView& filter,
                                          uint64_t (&w)[4]) {
    const uint32_t slot = base + static_cast<uint32_t>(__popc(heads & ((2u << o) - 1u))) - 1u;
    if (AllSlotted || slot < FE::kSlots) {
        const u32x4* src = reinterpret_cast<const u32x4*>(L.slot_w[AllSlotted ? slot : (slot < FE::kSlots ? slot : 0)]);
        const u32x4 a = src[0];
        const u32x4 b = src[1];
        w[0] = a.x | (static_cast<uint64_t>(a.y) << 32);
        w[1] = a.z | (static_cast<uint64_t>(a.w) << 32);
        w[2] = b.x | (static_cast<uint64_t>(b.y) << 32);
        w[3] = b.z | (static_cast<uint64_t>(b.w) << 32);
    } else {
        const uint64_t* src = filter.words + 4ull * lw.shard[o];
#pragma unroll
        for (int q = 0; q < 4; ++q) w[q] = src[q];
    }
}

// filter_ref.cuh:70-87 for the 16 windows of the lane: window o is present
// when every bit of its k-s+1 s-mers' masks is set in its run's shard words.
// The s-mer masks slide through a ring; with four s-mers per window the ring
// holds pair masks (s-mers p and p+1), so a window's mask is the OR of two
// pairs, folded with the and-not of the test into one op per 32-bit half.
template <int K, int S, int M, int H, class FE, bool Batched, bool AllSlotted>
__device__ __forceinline__ void findere_test(const LaneWindows<K, S, M, H, FE>& lw, uint32_t heads,
                                             uint32_t base, const WaveLds<FE, Batched>& L,
                                             const FilterView& filter, uint32_t (&result)[4]) {
    constexpr int F = Shape<K, S, M, H, FE>::kFindereSpan;
    if constexpr (F == 4 && AllSlotted) {
        // Every run has a slot: no branches. Window o reads slot (run heads up
        // to o) - 1 at a running LDS offset; slot -1 (windows before the lane's
        // first head that continue no run: all invalid) lies in this wave's LDS
        // just below the slots. Invalid windows start as misses.
        uint64_t last[4];
        uint64_t pair[3][4];
        uint64_t m1[4], m2[4];
        code_masks<H>(lw.code[0], last);
        code_masks<H>(lw.code[1], m1);
        code_masks<H>(lw.code[2], m2);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            pair[0][q] = last[q] | m1[q];
            pair[1][q] = m1[q] | m2[q];
            last[q] = m2[q];
        }
        uint32_t misses[4];  // byte o & 3 of word o >> 2: window o missed
        const uint32_t invalid = ~lw.valid;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            // bits 4j..4j+3 of `invalid` to bytes (the shifted copies do not overlap)
            misses[j] = (((invalid >> (4 * j)) & 0xFu) * 0x00204081u) & 0x01010101u;
        }
        const char* slot0 = reinterpret_cast<const char*>(&L.slot_w[0][0]) - 32 + 32 * static_cast<int>(base);
#pragma unroll
        for (int o = 0; o < kLaneWindows; ++o) {
            uint64_t nm[4];
            code_masks<H>(lw.code[o + 3], nm);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                pair[(o + 2) % 3][q] = last[q] | nm[q];
                last[q] = nm[q];
            }
            slot0 += static_cast<int>(__builtin_amdgcn_ubfe(heads, o, 1)) << 5;
            const u32x4* src = reinterpret_cast<const u32x4*>(slot0);
            const u32x4 a = src[0];
            const u32x4 b = src[1];
            const uint32_t wv[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
            uint32_t miss = 0;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                miss |= or_and_not(lo32(pair[o % 3][q]), wv[2 * q], lo32(pair[(o + 2) % 3][q]));
                miss |= or_and_not(hi32(pair[o % 3][q]), wv[2 * q + 1], hi32(pair[(o + 2) % 3][q]));
            }
            misses[o >> 2] |= (miss < 1u ? miss : 1u) << (8 * (o & 3));
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) result[j] = ~misses[j] & 0x01010101u;
    } else if constexpr (F == 4) {
        uint64_t last[4];     // mask of the newest s-mer
        uint64_t pair[3][4];  // pair[p % 3] = mask(p) | mask(p + 1)
        uint64_t m1[4], m2[4];
        code_masks<H>(lw.code[0], last);
        code_masks<H>(lw.code[1], m1);
        code_masks<H>(lw.code[2], m2);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            pair[0][q] = last[q] | m1[q];
            pair[1][q] = m1[q] | m2[q];
            last[q] = m2[q];
        }
#pragma unroll
        for (int o = 0; o < kLaneWindows; ++o) {
            uint64_t nm[4];
            code_masks<H>(lw.code[o + 3], nm);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                pair[(o + 2) % 3][q] = last[q] | nm[q];
                last[q] = nm[q];
            }
            if ((lw.valid >> o) & 1u) {
                uint64_t w[4];
                run_words<K, S, M, H, FE, Batched, AllSlotted>(lw, o, heads, base, L, filter, w);
                uint32_t miss = 0;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    miss |= or_and_not(lo32(pair[o % 3][q]), lo32(w[q]), lo32(pair[(o + 2) % 3][q]));
                    miss |= or_and_not(hi32(pair[o % 3][q]), hi32(w[q]), hi32(pair[(o + 2) % 3][q]));
                }
                result[o >> 2] |= (miss == 0 ? 1u : 0u) << (8 * (o & 3));
            }
        }
    } else {
        uint64_t ring[F][4];
#pragma unroll
        for (int p = 0; p < F - 1; ++p) code_masks<H>(lw.code[p], ring[p]);
#pragma unroll
        for (int o = 0; o < kLaneWindows; ++o) {
            code_masks<H>(lw.code[o + F - 1], ring[(o + F - 1) % F]);
            if ((lw.valid >> o) & 1u) {
                uint64_t w[4];
                run_words<K, S, M, H, FE, Batched, AllSlotted>(lw, o, heads, base, L, filter, w);
                uint64_t miss = 0;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    uint64_t x = 0;
#pragma unroll
                    for (int d = 0; d < F - 1; ++d) x |= ring[(o + d) % F][q];
                    miss |= (x | ring[(o + F - 1) % F][q]) & ~w[q];
                }
                result[o >> 2] |= (miss == 0 ? 1u : 0u) << (8 * (o & 3));
            }
        }
    }
}

// Query tile. Prefetch (plain byte input): `raw` holds this tile's input
// chunk, loaded ahead; once the shard loads are issued the lane loads the
// chunk of tile `next` into it (none when next >= num_tiles), so that load's
// latency hides behind the findere test instead of stalling the next tile.
template <int K, int S, int M, int H, class FE, bool Dense, bool Batched, bool Prefetch = false>
__device__ __forceinline__ void contains_tile(const SequenceInput& in, const FilterView& filter,
                                              uint64_t num_windows, uint8_t* out, uint64_t tile,
                                              int lane, WaveLds<FE, Batched>& L, RawChunk* raw = nullptr,
                                              uint64_t next = 0, uint64_t num_tiles = 0) {
    using Sh = Shape<K, S, M, H, FE>;
    constexpr int F = Sh::kFindereSpan;
    const uint64_t tile_first = tile_start<Sh::kWaveWindows>(tile);

    if constexpr (Prefetch) {
        stage_lane_raw(*raw, in, lane_start(tile_first, lane), lane, L.st, L.inv);
    } else {
        stage_tile<K, S, M, H, FE, Dense, Batched>(in, tile, lane, L);
    }
    wave_lds_sync();

    const bool owns = lane < Sh::kWindowLanes;
    const uint64_t first_window = lane_start(tile_first, lane);
    LaneWindows<K, S, M, H, FE> lw;
    lane_windows<K, S, M, H, FE, Batched>(L, lane, owns, first_window, num_windows,
                                          static_cast<uint32_t>(filter.num_shards - 1), lw);
    uint32_t cont;
    const uint32_t heads = run_heads(lw, lane, cont);
    uint32_t total;
    const uint32_t base = wave_exclusive_sum(static_cast<uint32_t>(__popc(heads)), lane, total);
    publish_heads(lw, heads, base, L);
    wave_lds_sync();

    // dense shard loads: lanes 2i and 2i+1 fetch the two 16-byte halves of one
    // slot's shard in the same instruction, so each shard costs the memory
    // system one request for its line (slots 32 j + i, j = 0..7)
    const u32x4* shards = reinterpret_cast<const u32x4*>(filter.words);
    const uint32_t n_slots = total < FE::kSlots ? total : FE::kSlots;
    constexpr int kHalves = 2 * FE::kSlots / kWave;
    const uint32_t half = static_cast<uint32_t>(lane) & 1u;
    u32x4 sw[kHalves];
#pragma unroll
    for (int j = 0; j < kHalves; ++j) {
        const uint32_t q = static_cast<uint32_t>(lane + j * kWave) >> 1;
        if (q < n_slots) sw[j] = shards[2 * static_cast<uint64_t>(L.slot_shard[q]) + half];
    }
    if constexpr (Prefetch) {
        if (next < num_tiles) {
            *raw = load_raw(in, lane_start(tile_start<Sh::kWaveWindows>(next), lane));
        } else {
            raw->fast = false;
        }
        // compiler-only memory barrier: keeps the load here, ahead of the test,
        // instead of next to its use in the next tile
        asm volatile("" ::: "memory");
    }
#pragma unroll
    for (int j = 0; j < kHalves; ++j) {
        const uint32_t q = static_cast<uint32_t>(lane + j * kWave) >> 1;
        if (q < n_slots) reinterpret_cast<u32x4*>(L.slot_w[q])[half] = sw[j];
    }
    wave_lds_sync();

    // findere test of every window against its run's shard words
    uint32_t result[4] = {0, 0, 0, 0};
    if (total <= FE::kSlots) {  // wave-uniform: every run of the tile has its LDS slot
        findere_test<K, S, M, H, FE, Batched, true>(lw, heads, base, L, filter, result);
    } else {
        findere_test<K, S, M, H, FE, Batched, false>(lw, heads, base, L, filter, result);
    }
    wave_lds_sync();  // the next tile of this wave reuses the LDS

    if (!owns || first_window >= num_windows) return;
    uint8_t* dst = out + first_window;
    if (first_window + kLaneWindows <= num_windows &&
        (reinterpret_cast<uintptr_t>(dst) & 15u) == 0) {
        const u32x4 r = {result[0], result[1], result[2], result[3]};
        __builtin_nontemporal_store(r, reinterpret_cast<u32x4*>(dst));
    } else {
        const uint64_t n = num_windows - first_window;
#pragma unroll
        for (int o = 0; o < kLaneWindows; ++o) {
            if (static_cast<uint64_t>(o) < n) {
                dst[o] = static_cast<uint8_t>((result[o >> 2] >> (8 * (o & 3))) & 0xFFu);
            }
        }
    }
}

// Insert tile. Runs get LDS slots in chunks of FE::kSlots: for each chunk the
// lanes OR the masks of their pieces of the chunk's runs into the slots with
// LDS atomics, then the wave flushes the slots -- one 64-bit device-scope
// atomic OR per non-zero word of each run, or (sorted insert, `sink.pairs`
// non-null) one record per run written into the wave's region (`region`,
// `emitted` records so far) with coalesced stores; runs past the region's
// capacity keep the atomics. Tiles with more runs than slots (short minimizer
// spans, e.g. protein with m = 6) take more than one chunk and recompute the
// lane windows for each (keeping them live across a flush would spill).
template <int K, int S, int M, int H, class FE, bool Dense, bool Batched>
__device__ __forceinline__ void insert_tile(const SequenceInput& in, const FilterView& filter,
                                            uint64_t num_windows, uint64_t tile, int lane,
                                            WaveLds<FE, Batched>& L, const RecordSink& sink,
                                            uint32_t region, uint32_t& emitted) {
    using Sh = Shape<K, S, M, H, FE>;
    constexpr int F = Sh::kFindereSpan;
    constexpr uint32_t kSlotsU = static_cast<uint32_t>(FE::kSlots);
    const uint64_t tile_first = tile_start<Sh::kWaveWindows>(tile);
    const bool owns = lane < Sh::kWindowLanes;
    const uint64_t first_window = lane_start(tile_first, lane);

    stage_tile<K, S, M, H, FE, Dense, Batched>(in, tile, lane, L);
    uint32_t first = 0;  // first run of the chunk (wave-uniform)
    for (;;) {
        zero_slots(L, lane);
        wave_lds_sync();

        LaneWindows<K, S, M, H, FE> lw;
        lane_windows<K, S, M, H, FE, Batched>(L, lane, owns, first_window, num_windows,
                                              static_cast<uint32_t>(filter.num_shards - 1), lw);
        uint32_t cont;
        const uint32_t heads = run_heads(lw, lane, cont);
        uint32_t total;
        const uint32_t base = wave_exclusive_sum(static_cast<uint32_t>(__popc(heads)), lane, total) - first;
        publish_heads(lw, heads, base, L);
        // bit o: window o is the last window of its run inside this lane
        const uint32_t ends = lw.valid & ~(cont >> 1);

        // each lane ORs the masks of its piece of every run of the chunk into its slot
        uint64_t ring[F][4];
#pragma unroll
        for (int p = 0; p < F - 1; ++p) code_masks<H>(lw.code[p], ring[p]);
        uint64_t acc[4] = {0, 0, 0, 0};
#pragma unroll
        for (int o = 0; o < kLaneWindows; ++o) {
            code_masks<H>(lw.code[o + F - 1], ring[(o + F - 1) % F]);
            if ((cont >> o) & 1u) {
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[q] |= ring[(o + F - 1) % F][q];
            } else {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    uint64_t x = 0;
#pragma unroll
                    for (int d = 0; d < F; ++d) x |= ring[(o + d) % F][q];
                    acc[q] = x;
                }
            }
            if ((ends >> o) & 1u) {
                const uint32_t slot = base + static_cast<uint32_t>(__popc(heads & ((2u << o) - 1u))) - 1u;
                if (slot < kSlotsU) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        __hip_atomic_fetch_or(&L.slot_w[slot][q], acc[q], __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_WORKGROUP);
                    }
                }
            }
        }
        wave_lds_sync();

        // flush the chunk's slots
        const uint32_t n_slots = total - first < kSlotsU ? total - first : kSlotsU;
        const uint32_t room = sink.pairs != nullptr ? sink.cap - emitted : 0u;
#pragma unroll
        for (int j = 0; j < FE::kSlots / kWave; ++j) {
            const uint32_t q = static_cast<uint32_t>(lane + j * kWave);
            if (q < n_slots) {
                if (q < room) {
                    const uint64_t at = static_cast<uint64_t>(region) * sink.cap + emitted + q;
                    sink.pairs[at] = (static_cast<uint64_t>(L.slot_shard[q]) << 32) | at;
                    const u32x4* src = reinterpret_cast<const u32x4*>(L.slot_w[q]);
                    u32x4* dst = reinterpret_cast<u32x4*>(sink.masks + 4 * at);
                    dst[0] = src[0];
                    dst[1] = src[1];
                } else {
                    uint64_t* dst = filter.words + 4ull * L.slot_shard[q];
#pragma unroll
                    for (int w = 0; w < 4; ++w) {
                        const uint64_t m = L.slot_w[q][w];
                        if (m != 0) __hip_atomic_fetch_or(dst + w, m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
                }
            }
        }
        emitted += n_slots < room ? n_slots : room;
        wave_lds_sync();  // the next chunk or tile of this wave reuses the LDS
        first += kSlotsU;
        if (first >= total) break;
    }
}

// --- plain kernels: waves stride over the tiles ------------------------------------

// The wave's index in its block, as a wave-uniform (scalar) value: tile indices,
// region numbers and the loop bounds derived from it then live in scalar
// registers instead of taking vector registers for the whole kernel.
__device__ __forceinline__ int wave_in_block() {
    return __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / kWave);
}

template <int K, int S, int M, int H, class FE, bool Dense, int WPE>
__global__ __launch_bounds__(kBlock, WPE) void contains_kernel(SequenceInput in, FilterView filter,
                                                               uint64_t num_windows,
                                                               uint64_t num_tiles, uint8_t* out) {
    __shared__ WaveLds<FE, false> lds[kWavesPerBlock];
    const int lane = static_cast<int>(threadIdx.x) & (kWave - 1);
    const int wave = wave_in_block();
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kWavesPerBlock;
    const uint64_t first = static_cast<uint64_t>(blockIdx.x) * kWavesPerBlock + wave;
    if constexpr (Dense) {
        for (uint64_t tile = first; tile < num_tiles; tile += stride) {
            contains_tile<K, S, M, H, FE, Dense, false>(in, filter, num_windows, out, tile, lane, lds[wave]);
        }
    } else {
        RawChunk raw;
        raw.fast = false;
        if (first < num_tiles) {
            raw = load_raw(in, lane_start(tile_start<Shape<K, S, M, H, FE>::kWaveWindows>(first), lane));
        }
        for (uint64_t tile = first; tile < num_tiles; tile += stride) {
            contains_tile<K, S, M, H, FE, false, false, true>(in, filter, num_windows, out, tile, lane,
                                                             lds[wave], &raw, tile + stride, num_tiles);
        }
    }
}

template <int K, int S, int M, int H, class FE, bool Dense, int WPE>
__global__ __launch_bounds__(kBlock, WPE) void insert_kernel(SequenceInput in, FilterView filter,
                                                             uint64_t num_windows,
                                                             uint64_t num_tiles, RecordSink sink) {
    __shared__ WaveLds<FE, false> lds[kWavesPerBlock];
    mark_words_written(filter);
    const int lane = static_cast<int>(threadIdx.x) & (kWave - 1);
    const int wave = wave_in_block();
    const uint32_t region = blockIdx.x * kWavesPerBlock + static_cast<uint32_t>(wave);
    uint32_t emitted = 0;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kWavesPerBlock;
    for (uint64_t tile = region; tile < num_tiles; tile += stride) {
        insert_tile<K, S, M, H, FE, Dense, false>(in, filter, num_windows, tile, lane, lds[wave], sink,
                                                  region, emitted);
    }
    if (sink.pairs != nullptr && lane == 0) {
        sink.counts[region] = emitted;
        if (emitted == sink.cap) *sink.overflow = 1u;  // later runs may have used atomics
    }
}

// --- batched kernels: persistent waves over the tiles ---------------------------------

// Lane 0 takes `claim` tiles from the queue; every lane gets the first one.
__device__ __forceinline__ uint32_t claim_tiles(uint32_t* counter, uint32_t claim, int lane) {
    uint32_t v = 0;
    if (lane == 0) v = __hip_atomic_fetch_add(counter, claim, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return __builtin_amdgcn_readfirstlane(v);
}

// Runs body(tile) for this wave's tiles. Queue: each wave takes `claim` tiles
// at a time from the work queue, asking for its next claim before it works
// through the current one; otherwise the waves stride over the tiles. Every
// wave leaves once the tiles are used up. (Two kernels rather than a runtime
// switch: the queue's live state costs the strided loop registers.)
template <bool Queue, class Body>
__device__ __forceinline__ void for_each_tile(uint32_t* counter, uint32_t num_tiles,
                                              uint32_t claim, int lane, int wave, Body&& body) {
    if constexpr (!Queue) {
        const uint32_t stride = gridDim.x * kWavesPerBlock;
        for (uint32_t t = blockIdx.x * kWavesPerBlock + static_cast<uint32_t>(wave); t < num_tiles;
             t += stride) {
            body(t);
        }
    } else {
        uint32_t t = claim_tiles(counter, claim, lane);
        while (t < num_tiles) {
            const uint32_t next = claim_tiles(counter, claim, lane);
            const uint32_t end = num_tiles - t < claim ? num_tiles : t + claim;
            for (uint32_t x = t; x < end; ++x) body(x);
            t = next;
        }
    }
}

template <int K, int S, int M, int H, class FE, int WPE, bool Queue>
__global__ __launch_bounds__(kBlock, Queue ? 2 : WPE) void batched_contains_kernel(SequenceInput in,
                                                                       FilterView filter,
                                                                       uint64_t num_windows,
                                                                       uint32_t num_tiles,
                                                                       uint32_t claim, uint8_t* out) {
    __shared__ WaveLds<FE, true> lds[kWavesPerBlock];
    const int lane = static_cast<int>(threadIdx.x) & (kWave - 1);
    const int wave = wave_in_block();
    if (__builtin_amdgcn_readfirstlane(in.state[1]) != 0) return;  // malformed offsets
    init_wave_lds(lds[wave], lane);
    for_each_tile<Queue>(in.state, num_tiles, claim, lane, wave, [&](uint32_t tile) {
        contains_tile<K, S, M, H, FE, false, true>(in, filter, num_windows, out, tile, lane,
                                                   lds[wave]);
    });
}

template <int K, int S, int M, int H, class FE, int WPE, bool Queue>
__global__ __launch_bounds__(kBlock, Queue ? 2 : WPE) void batched_insert_kernel(SequenceInput in,
                                                                     FilterView filter,
                                                                     uint64_t num_windows,
                                                                     uint32_t num_tiles,
                                                                     uint32_t claim, RecordSink sink) {
    __shared__ WaveLds<FE, true> lds[kWavesPerBlock];
    mark_words_written(filter);
    const int lane = static_cast<int>(threadIdx.x) & (kWave - 1);
    const int wave = wave_in_block();
    const uint32_t region = blockIdx.x * kWavesPerBlock + static_cast<uint32_t>(wave);
    uint32_t emitted = 0;
    if (__builtin_amdgcn_readfirstlane(in.state[1]) == 0) {  // else malformed offsets: no work
        init_wave_lds(lds[wave], lane);
        for_each_tile<Queue>(in.state, num_tiles, claim, lane, wave, [&](uint32_t tile) {
            insert_tile<K, S, M, H, FE, false, true>(in, filter, num_windows, tile, lane, lds[wave],
                                                     sink, region, emitted);
        });
    }
    if (sink.pairs != nullptr && lane == 0) {
        sink.counts[region] = emitted;
        if (emitted == sink.cap) *sink.overflow = 1u;  // later runs may have used atomics
    }
}

// --- host launchers ----------------------------------------------------------

// Grid of at most kStripMaxBlocks blocks (a dispatch must stay below 2^32
// work-items); waves stride over the tiles.
static_assert(kWavesPerBlock == kStripWavesPerBlock, "launch.hpp block shape");

template <int K, int S, int M, int H, class FE>
uint32_t tile_grid(uint64_t num_windows, uint64_t& num_tiles) {
    using Sh = Shape<K, S, M, H, FE>;
    num_tiles = (num_windows + Sh::kWaveWindows - 1) / Sh::kWaveWindows;
    uint64_t blocks = (num_tiles + kWavesPerBlock - 1) / kWavesPerBlock;
    if (blocks > kStripMaxBlocks) blocks = kStripMaxBlocks;
    return static_cast<uint32_t>(blocks);
}

// Sorted insert: never more blocks than the sink has regions for.
inline uint32_t sink_grid(uint32_t grid, const RecordSink* sink) {
    if (sink == nullptr) return grid;
    const uint32_t most = sink->regions / kWavesPerBlock;
    return grid < most ? grid : most;
}

template <int K>
uint64_t windows_of(uint64_t num_symbols) {
    return num_symbols < static_cast<uint64_t>(K) ? 0 : num_symbols - K + 1;
}

// Waves per SIMD the plain kernels are built for: a register budget under
// which they do not spill (spilled values reloaded while other vector memory
// operations are in flight have read stale data on this GPU). Eight and more
// hashes per s-mer need the larger budget of two waves per SIMD.
template <class FE, int H>
constexpr int plain_waves() {
    return FE::kBits == 2 ? (H == 4 ? SBF_DNA_WAVES_PER_EU : 2) : 3;
}

template <int K, int S, int M, int H, class FE>
hipError_t launch_insert(const KernelConfig&, const SequenceInput& in, FilterView f,
                         const RecordSink* sink, hipStream_t stream) {
    const uint64_t n = windows_of<K>(in.num_symbols);
    const RecordSink s = sink ? *sink : RecordSink{};
    uint64_t tiles = 0;
    const uint32_t grid = n == 0 ? 0u : sink_grid(tile_grid<K, S, M, H, FE>(n, tiles), sink);
    if (grid == 0) return hipSuccess;
    constexpr int WPE = plain_waves<FE, H>();
    (void)hipGetLastError();  // report only this launch
    if constexpr (FE::kBits == 2) {
        if (in.dense) {
            hipLaunchKernelGGL((insert_kernel<K, S, M, H, FE, true, WPE>), dim3(grid), dim3(kBlock), 0,
                               stream, in, f, n, tiles, s);
            return hipGetLastError();
        }
    }
    hipLaunchKernelGGL((insert_kernel<K, S, M, H, FE, false, WPE>), dim3(grid), dim3(kBlock), 0,
                       stream, in, f, n, tiles, s);
    return hipGetLastError();
}

template <int K, int S, int M, int H, class FE>
hipError_t launch_contains(const KernelConfig&, const SequenceInput& in, FilterView f, uint8_t* out,
                           hipStream_t stream) {
    const uint64_t n = windows_of<K>(in.num_symbols);
    if (n == 0) return hipSuccess;
    uint64_t tiles;
    const uint32_t grid = tile_grid<K, S, M, H, FE>(n, tiles);
    constexpr int WPE = plain_waves<FE, H>();
    (void)hipGetLastError();  // report only this launch
    if constexpr (FE::kBits == 2) {
        if (in.dense) {
            hipLaunchKernelGGL((contains_kernel<K, S, M, H, FE, true, WPE>), dim3(grid), dim3(kBlock),
                               0, stream, in, f, n, tiles, out);
            return hipGetLastError();
        }
    }
    hipLaunchKernelGGL((contains_kernel<K, S, M, H, FE, false, WPE>), dim3(grid), dim3(kBlock), 0,
                       stream, in, f, n, tiles, out);
    return hipGetLastError();
}

// Persistent grid (batched_grid in launch.hpp): blocks_per_cu blocks on each
// compute unit, never more blocks than the tiles need. A sorted insert with no
// windows still launches, so that every region count is written.
template <int K, int S, int M, int H, class FE, int WPE>
hipError_t launch_batched_insert(const KernelConfig&, const SequenceInput& in, FilterView f,
                                 const BatchLaunch& b, const RecordSink* sink, hipStream_t stream) {
    using Sh = Shape<K, S, M, H, FE>;
    const uint64_t n = windows_of<K>(in.num_symbols);
    if (n == 0 && sink == nullptr) return hipSuccess;
    const RecordSink s = sink ? *sink : RecordSink{};
    const uint32_t tiles = static_cast<uint32_t>((n + Sh::kWaveWindows - 1) / Sh::kWaveWindows);
    const uint32_t blocks = sink_grid(batched_grid(b, tiles), sink);
    if (blocks == 0) return hipSuccess;
    const dim3 grid(blocks);
    (void)hipGetLastError();  // report only this launch
    if (b.claim == 0) {
        hipLaunchKernelGGL((batched_insert_kernel<K, S, M, H, FE, WPE, false>), grid, dim3(kBlock), 0,
                           stream, in, f, n, tiles, 0u, s);
    } else {
        hipLaunchKernelGGL((batched_insert_kernel<K, S, M, H, FE, WPE, true>), grid, dim3(kBlock), 0,
                           stream, in, f, n, tiles, b.claim, s);
    }
    return hipGetLastError();
}

template <int K, int S, int M, int H, class FE, int WPE>
hipError_t launch_batched_contains(const KernelConfig&, const SequenceInput& in, FilterView f,
                                   const BatchLaunch& b, uint8_t* out, hipStream_t stream) {
    using Sh = Shape<K, S, M, H, FE>;
    const uint64_t n = windows_of<K>(in.num_symbols);
    if (n == 0) return hipSuccess;
    const uint32_t tiles = static_cast<uint32_t>((n + Sh::kWaveWindows - 1) / Sh::kWaveWindows);
    const dim3 grid(batched_grid(b, tiles));
    (void)hipGetLastError();  // report only this launch
    if (b.claim == 0) {
        hipLaunchKernelGGL((batched_contains_kernel<K, S, M, H, FE, WPE, false>), grid, dim3(kBlock),
                           0, stream, in, f, n, tiles, 0u, out);
    } else {
        hipLaunchKernelGGL((batched_contains_kernel<K, S, M, H, FE, WPE, true>), grid, dim3(kBlock),
                           0, stream, in, f, n, tiles, b.claim, out);
    }
    return hipGetLastError();
}

template <int K, int S, int M, int H, class FE>
constexpr uint32_t wave_windows() {
    return static_cast<uint32_t>(Shape<K, S, M, H, FE>::kWaveWindows);
}

// Batched launchers compiled in kernels_strip_batched.hip: (front end, k, s,
// m, h, waves per SIMD of the strided kernel; the work-queue kernel, which
// needs more registers for its loop, is built for two so that it does not
// spill).
#define SBF_STRIP_BATCHED_INSTANCES(X)                                        \
    X(Dna2, 31, 28, 16, 4, 3) X(Dna2, 31, 30, 16, 4, 3)                       \
    X(Dna2, 31, 31, 16, 4, 3) X(Dna2, 21, 19, 11, 4, 3)                       \
    X(Dna2, 5, 4, 3, 4, 3) X(Dna2, 3, 2, 2, 4, 3)                             \
    X(Protein5, 10, 8, 6, 4, 3) X(Protein5, 12, 10, 6, 4, 3)                 \
    X(Protein5, 5, 4, 3, 4, 3)

}  // namespace sbf::strip
