// strip_wide.hpp -- lane-strip kernels whose spans are runtime values (device code
// and launchers; instantiated by kernels_wide_dna.hip and kernels_wide_protein.hip).
//
// The compile-time lane strips (strip_kernels.hpp) are built for a handful of
// (k, s, m, h). These kernels take k, s and m at run time, so that one kernel
// set per (front end, h) covers:
//  * DNA: every k <= 32 with findere span k - s + 1 <= 16 and minimizer span
//    k - m + 1 <= 24 -- the reference's whole DNA parameter sweep
//    (benchmarks/param_sweep.cu:263-530: k = 31, s = 16..31, m = 8..31);
//  * protein: every k <= 12 (its protein grid, param_sweep.cu:534-591: k = 12,
//    s = 6..12, m = 4..12).
//
// Tile, run slots, shard loads, record sink and drivers are those of the
// compile-time lane strips. What differs is how a lane gets the minima and the
// masks of its 16 windows (the reference recomputes both per window,
// sequence_kmer.cuh:40-59):
//  * the lane hashes the m-mer and the s-mer that start at each of its 16
//    positions once, and takes the next lanes' hashes and s-mer codes with DPP
//    wave shifts;
//  * window minimum over a runtime span w: in-place doubling (after level l a
//    position holds the minimum of the 2^l hashes from it on), then, min being
//    idempotent, min(T(o), T(o + w - 2^L)) with L = floor(log2 w). The runtime
//    offset w - 2^L is applied as a shift of the register array by its binary
//    digits under wave-uniform branches, so no register is indexed at run time;
//  * findere mask of a window over a runtime span f: the same doubling with OR,
//    one 64-bit shard word at a time (filter_impl.cuh:54-77 sets hash i's bit in
//    word i % 4), then the test of filter_ref.cuh:70-87 per word.
#pragma once

#include "strip_kernels.hpp"

namespace sbf::strip {

// --- limits and shapes -----------------------------------------------------------

template <class FE>
struct WideLimits;

template <>
struct WideLimits<Dna2> {
    static constexpr int kMaxK = 32;   // the lane's 48-symbol context covers k - 1 + 16
    static constexpr int kMaxF = 16;   // s-mer codes of positions 16.. come from lane + 1
    static constexpr int kMaxW = 24;   // m-mer hashes of positions 32.. come from lane + 2
    static constexpr int kHaloLanes = 2;
};

template <>
struct WideLimits<Protein5> {
    static constexpr int kMaxK = 12;
    static constexpr int kMaxF = 12;
    static constexpr int kMaxW = 12;
    static constexpr int kHaloLanes = 1;
};

// Largest w - 2^floor(log2 w) over 1 <= w <= n.
constexpr int max_span_rest(int n) {
    int best = 0;
    for (int w = 1; w <= n; ++w) {
        int p = 1;
        while (2 * p <= w) p *= 2;
        if (w - p > best) best = w - p;
    }
    return best;
}

template <class FE>
struct WideShape {
    using Lim = WideLimits<FE>;
    static constexpr int kWindowLanes = kWave - Lim::kHaloLanes;
    static constexpr int kWaveWindows = kWindowLanes * kLaneWindows;
    static constexpr int kMinPositions = kLaneWindows + Lim::kMaxW - 1;   // m-mer hashes per lane
    static constexpr int kSmerPositions = kLaneWindows + Lim::kMaxF - 1;  // s-mer codes per lane
    static_assert(kMinPositions <= 3 * kLaneWindows, "m-mer hashes come from two lanes ahead at most");
    static_assert(kSmerPositions <= 2 * kLaneWindows, "s-mer codes come from the next lane only");
};

// Runtime spans of one configuration (wide_spans below).
struct Spans {
    uint32_t k;
    uint32_t f, fl, fd;  // findere span k - s + 1 = 2^fl + fd, 0 <= fd < 2^fl
    uint32_t w, wl, wd;  // minimizer span k - m + 1 = 2^wl + wd
    uint32_t m_shift;    // m-mer = (value of the lane position's span) >> m_shift
    uint32_t s_shift;    // s-mer, likewise
};

template <class FE>
constexpr bool wide_fits(uint32_t k, uint32_t s, uint32_t m) {
    using Lim = WideLimits<FE>;
    return k >= 1 && k <= static_cast<uint32_t>(Lim::kMaxK) && s >= 1 && s <= k && m >= 1 && m <= k &&
           k - s + 1 <= static_cast<uint32_t>(Lim::kMaxF) && k - m + 1 <= static_cast<uint32_t>(Lim::kMaxW);
}

inline uint32_t floor_log2(uint32_t x) { return 31u - static_cast<uint32_t>(__builtin_clz(x)); }

template <class FE>
Spans wide_spans(const KernelConfig& c) {
    Spans sp;
    sp.k = c.k;
    sp.f = c.k - c.s + 1;
    sp.fl = floor_log2(sp.f);
    sp.fd = sp.f - (1u << sp.fl);
    sp.w = c.k - c.m + 1;
    sp.wl = floor_log2(sp.w);
    sp.wd = sp.w - (1u << sp.wl);
    if constexpr (FE::kBits == 2) {
        sp.m_shift = 64u - 2u * c.m;  // own spans are 32 symbols
        sp.s_shift = 64u - 2u * c.s;
    } else {
        sp.m_shift = 5u * (12u - c.m);  // own spans are 12 symbols
        sp.s_shift = 5u * (12u - c.s);
    }
    return sp;
}

// Per-wave LDS of the runtime-span kernels: WaveLds with 512 run slots. A
// tile has at most one run per window (992 or 1008), and short minimizer spans
// give several hundred (2 / (w + 1) per window on random input): 512 slots take
// every tile down to w = 3 in one chunk. 19-20 KB per wave, so two 4-wave blocks
// (two waves per SIMD) share a compute unit.
template <class FE, bool Batched>
struct WideLds {
    static constexpr int kSlots = 512;
    static constexpr bool kBatched = Batched;
    typename FE::Lds st;
    uint32_t inv[kLdsWords];
    uint32_t bnd[Batched ? kLdsWords : 1];  // bit i of word l: symbol 16l+i starts a record
    uint32_t slot_shard[kSlots];
    alignas(16) uint64_t slot_w[kSlots][4];
};
static_assert(4 * sizeof(WideLds<Protein5, true>) <= 80 * 1024, "two blocks per compute unit");

// --- per-lane values ---------------------------------------------------------------

// MSB-first values of the 32 symbols starting at each own position p = 0..15
// (symbols p..p+31 of top:w2), shifted right by `shift`: the substring of
// length 32 - shift / 2 starting at p (filter_common.cuh:128-134).
__device__ __forceinline__ void own_spans(const Dna2::Ctx& c, uint32_t shift, uint64_t (&v)[kLaneWindows]) {
    const uint32_t t1 = hi32(c.top);
    const uint32_t t0 = lo32(c.top);
    v[0] = c.top >> shift;
#pragma unroll
    for (int p = 1; p < kLaneWindows; ++p) {
        const uint32_t hi = __builtin_amdgcn_alignbit(t1, t0, static_cast<uint32_t>(32 - 2 * p));
        const uint32_t lo = __builtin_amdgcn_alignbit(t0, c.w2, static_cast<uint32_t>(32 - 2 * p));
        v[p] = ((static_cast<uint64_t>(hi) << 32) | lo) >> shift;
    }
}

// MSB-first packing of the 5-bit symbols p..p+11 (sequence_kmer.cuh:74-89 packs
// `symbol & symbolMask`), shifted right by `shift`: the substring of length
// 12 - shift / 5 starting at p.
__device__ __forceinline__ void own_spans(const Protein5::Ctx& c, uint32_t shift,
                                          uint64_t (&v)[kLaneWindows]) {
    constexpr uint64_t mask = (1ull << 60) - 1ull;
    uint64_t acc = 0;
#pragma unroll
    for (int i = 0; i < kLaneWindows + 11; ++i) {
        const uint64_t sym = (c.d[i >> 2] >> (8 * (i & 3))) & 0x1Fu;
        acc = ((acc << 5) | sym) & mask;
        if (i >= 11) v[i - 11] = acc >> shift;
    }
}

// y_j = OR of x_{j..j+n-1} for a runtime n (smear with a wave-uniform length).
__device__ __forceinline__ uint64_t smear_rt(uint64_t x, uint32_t n) {
    if (n == 0) return 0;
    uint64_t y = x;
    uint32_t len = 1;
#pragma unroll
    for (int step = 0; step < 6; ++step) {
        if (2 * len <= n) {
            y |= y >> len;
            len *= 2;
        }
    }
    if (len < n) y |= y >> (n - len);
    return y;
}

template <int H>
__device__ __forceinline__ SmerCode<H> next_lane_code(const SmerCode<H>& c) {
    if constexpr (64 / H >= 6) {
        SmerCode<H> r;
#pragma unroll
        for (int j = 0; j < PosCode<H>::kWords; ++j) r.w[j] = from_next_lane(c.w[j]);
        return r;
    } else {
        return next_lane(c);
    }
}

// Word q of an s-mer's mask: the bits of its hashes i = q, q + 4, ... (filter_impl.cuh:54-77).
// q may be a runtime (wave-uniform) value: the code is read as one 64-bit value
// and the salts from a constant table, so nothing is indexed in registers.
template <int H>
__device__ __forceinline__ uint64_t word_bits(const SmerCode<H>& c, int q) {
    uint64_t m = 0;
    if constexpr (64 / H >= 6) {
        static_assert(PosCode<H>::kWords <= 2, "positions of at most eight hashes");
        uint64_t x = c.w[0];
        if constexpr (PosCode<H>::kWords > 1) x |= static_cast<uint64_t>(c.w[1]) << 32;
#pragma unroll
        for (int j = 0; j < H / 4; ++j) m |= 1ull << ((x >> (6 * (q + 4 * j))) & 63u);
    } else {
        constexpr uint64_t kSalt[16] = {salt(0), salt(1), salt(2),  salt(3),  salt(4),  salt(5),
                                        salt(6), salt(7), salt(8),  salt(9),  salt(10), salt(11),
                                        salt(12), salt(13), salt(14), salt(15)};
#pragma unroll
        for (int j = 0; j < H / 4; ++j) m |= 1ull << ((c * kSalt[q + 4 * j]) >> 58);
    }
    return m;
}

struct MinOp {
    __device__ __forceinline__ uint64_t operator()(uint64_t a, uint64_t b) const { return umin64(a, b); }
};
struct OrOp {
    __device__ __forceinline__ uint64_t operator()(uint64_t a, uint64_t b) const { return a | b; }
};

// In-place doubling: after the call, v[p] = op over v[p .. p + 2^levels) for
// every p < 16 + MaxSpan - 2^levels. Level l runs when l < levels (wave-uniform);
// its positions are the ones a span of at most MaxSpan still needs.
template <int MaxSpan, int N, class Op>
__device__ __forceinline__ void doubling(uint64_t (&v)[N], uint32_t levels, Op op) {
    static_assert(N >= kLaneWindows + MaxSpan - 1, "array shorter than the spans");
#pragma unroll
    for (int l = 0; (2 << l) <= MaxSpan; ++l) {
        if (static_cast<uint32_t>(l) < levels) {
#pragma unroll
            for (int p = 0; p < kLaneWindows + MaxSpan - (2 << l); ++p) v[p] = op(v[p], v[p + (1 << l)]);
        }
    }
}

// v[o] = old v[o + d] for o < 16 (0 <= d <= MaxShift, wave-uniform): one
// shift per binary digit of d, the largest first.
template <int MaxShift, int N>
__device__ __forceinline__ void shift_down(uint64_t (&v)[N], uint32_t d) {
#pragma unroll
    for (int b = 16; b >= 1; b >>= 1) {
        if (b <= MaxShift) {
            static_assert(N >= kLaneWindows + 2 * MaxShift - 1 || MaxShift == 0, "shift reads past the array");
            if (d & static_cast<uint32_t>(b)) {
#pragma unroll
                for (int p = 0; p < kLaneWindows + b - 1; ++p) v[p] = v[p + b];
            }
        }
    }
}

// Per-lane shard indices, window validity and s-mer codes (runtime spans).
template <class FE, int H>
struct WideLanes {
    uint32_t shard[kLaneWindows];
    uint32_t valid;  // bit o: window o valid and in range
    SmerCode<H> code[WideShape<FE>::kSmerPositions];
};

template <class FE, int H, bool Batched>
__device__ __forceinline__ void wide_lane_windows(const WideLds<FE, Batched>& L, int lane, bool owns_windows,
                                                  uint64_t first_window, uint64_t num_windows,
                                                  uint32_t shard_mask, const Spans& sp,
                                                  WideLanes<FE, H>& lw) {
    using WS = WideShape<FE>;
    using Lim = WideLimits<FE>;

    const uint64_t inv = static_cast<uint64_t>(L.inv[lane]) |
                         (static_cast<uint64_t>(L.inv[lane + 1]) << 16) |
                         (static_cast<uint64_t>(L.inv[lane + 2]) << 32);
    uint64_t bad = smear_rt(inv, sp.k);
    if constexpr (Batched) {
        const uint64_t bnd = static_cast<uint64_t>(L.bnd[lane]) |
                             (static_cast<uint64_t>(L.bnd[lane + 1]) << 16) |
                             (static_cast<uint64_t>(L.bnd[lane + 2]) << 32);
        bad |= smear_rt(bnd >> 1, sp.k - 1);
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
    uint64_t v[WS::kMinPositions];
    {
        uint64_t own[kLaneWindows];
        own_spans(ctx, sp.m_shift, own);
#pragma unroll
        for (int p = 0; p < kLaneWindows; ++p) v[p] = minimizer_hash64(own[p]);
    }
    // positions 16..31 are lane + 1's own, 32.. lane + 2's
#pragma unroll
    for (int i = 0; i < kLaneWindows && kLaneWindows + i < WS::kMinPositions; ++i) {
        v[kLaneWindows + i] = next_lane(v[i]);
    }
#pragma unroll
    for (int i = 0; 2 * kLaneWindows + i < WS::kMinPositions; ++i) {
        v[2 * kLaneWindows + i] = next_lane(v[kLaneWindows + i]);
    }
    // window o: minimum of the hashes o .. o + w - 1 (filter_ref.cuh:52-55)
    doubling<Lim::kMaxW>(v, sp.wl, MinOp{});
    uint64_t mn[kLaneWindows];
#pragma unroll
    for (int o = 0; o < kLaneWindows; ++o) mn[o] = v[o];
    if (sp.wd != 0) {
        shift_down<max_span_rest(Lim::kMaxW)>(v, sp.wd);
#pragma unroll
        for (int o = 0; o < kLaneWindows; ++o) mn[o] = umin64(mn[o], v[o]);
    }
    // filter_ref.cuh:52-55 with the uint32_t narrowing of kernels.cuh:240-241.
#pragma unroll
    for (int o = 0; o < kLaneWindows; ++o) lw.shard[o] = static_cast<uint32_t>(mn[o]) & shard_mask;

    // s-mer codes (second pass keeps the minimizer temporaries dead)
    {
        uint64_t own[kLaneWindows];
        own_spans(ctx, sp.s_shift, own);
#pragma unroll
        for (int p = 0; p < kLaneWindows; ++p) lw.code[p] = smer_code<H>(hash64(own[p]));
    }
#pragma unroll
    for (int i = 0; kLaneWindows + i < WS::kSmerPositions; ++i) {
        lw.code[kLaneWindows + i] = next_lane_code<H>(lw.code[i]);
    }
}

// Word q of every window's findere mask: a[o] = OR of word q of the masks of
// s-mers o .. o + f - 1.
template <class FE, int H>
__device__ __forceinline__ void window_words(const WideLanes<FE, H>& lw, int q, const Spans& sp,
                                             uint64_t (&a)[kLaneWindows]) {
    using Lim = WideLimits<FE>;
    uint64_t s[WideShape<FE>::kSmerPositions];
#pragma unroll
    for (int p = 0; p < WideShape<FE>::kSmerPositions; ++p) s[p] = word_bits<H>(lw.code[p], q);
    doubling<Lim::kMaxF>(s, sp.fl, OrOp{});
#pragma unroll
    for (int o = 0; o < kLaneWindows; ++o) a[o] = s[o];
    if (sp.fd != 0) {
        shift_down<max_span_rest(Lim::kMaxF)>(s, sp.fd);
#pragma unroll
        for (int o = 0; o < kLaneWindows; ++o) a[o] |= s[o];
    }
}

// filter_ref.cuh:70-87 for the lane's 16 windows: window o is present when
// every bit of its mask is set in its run's shard words (the run's LDS slot,
// or the filter for a run past the slots).
template <class FE, int H, bool Batched, bool AllSlotted>
__device__ __forceinline__ void wide_findere(const WideLanes<FE, H>& lw, uint32_t heads, uint32_t base,
                                             const WideLds<FE, Batched>& L, const FilterView& filter,
                                             const Spans& sp, uint32_t (&result)[4]) {
    // window o's slot: its run's, counted from the lane's first head (slot base);
    // windows before that head continue the previous lane's last run (base - 1)
    uint32_t slot[kLaneWindows];
#pragma unroll
    for (int o = 0; o < kLaneWindows; ++o) {
        slot[o] = base + static_cast<uint32_t>(__popc(heads & ((2u << o) - 1u))) - 1u;
    }
    uint32_t miss[kLaneWindows];
#pragma unroll
    for (int o = 0; o < kLaneWindows; ++o) miss[o] = 0;
    // one shard word per iteration; kept a loop so that the words' mask arrays
    // are never live at once
#pragma unroll 1
    for (int q = 0; q < 4; ++q) {
        uint64_t a[kLaneWindows];
        window_words<FE, H>(lw, q, sp, a);
#pragma unroll
        for (int o = 0; o < kLaneWindows; ++o) {
            uint64_t w;
            if constexpr (AllSlotted) {
                // a window before the lane's first head with base 0 reads 32 bytes
                // below the slots (this wave's slot_shard array): it is invalid
                const char* p = reinterpret_cast<const char*>(&L.slot_w[0][0]) +
                                32 * static_cast<int>(slot[o]) + 8 * q;
                w = *reinterpret_cast<const uint64_t*>(p);
            } else {
                w = 0;
                if ((lw.valid >> o) & 1u) {
                    w = slot[o] < static_cast<uint32_t>(WideLds<FE, Batched>::kSlots) ? L.slot_w[slot[o]][q]
                                                                    : filter.words[4ull * lw.shard[o] + q];
                }
            }
            const uint64_t m = a[o] & ~w;
            miss[o] |= lo32(m) | hi32(m);
        }
    }
#pragma unroll
    for (int o = 0; o < kLaneWindows; ++o) {
        const uint32_t hit = ((lw.valid >> o) & 1u) & (miss[o] == 0 ? 1u : 0u);
        result[o >> 2] |= hit << (8 * (o & 3));
    }
}

// The findere test of the tiles with many runs (wide_contains_tile_chunked): the
// slots hold the runs first .. first + 511 of the tile, and only windows whose
// run is among them are answered.
template <class FE, int H, bool Batched>
__device__ __forceinline__ void wide_findere_chunk(const WideLanes<FE, H>& lw, uint32_t heads, uint32_t base,
                                             uint32_t first, const WideLds<FE, Batched>& L,
                                             const Spans& sp, uint32_t (&result)[4]) {
    constexpr uint32_t kSlotsU = static_cast<uint32_t>(WideLds<FE, Batched>::kSlots);
    // window o's slot: its run's, counted from the lane's first head (slot base);
    // windows before that head continue the previous lane's last run (base - 1)
    uint32_t in_chunk = lw.valid;
#pragma unroll
    for (int o = 0; o < kLaneWindows; ++o) {
        const uint32_t slot = base + static_cast<uint32_t>(__popc(heads & ((2u << o) - 1u))) - 1u - first;
        if (slot >= kSlotsU) in_chunk &= ~(1u << o);
    }
    // LDS address of slot base - 1 - first; windows outside the chunk do not read
    const char* slot0 = reinterpret_cast<const char*>(&L.slot_w[0][0]) +
                        32 * (static_cast<int>(base) - 1 - static_cast<int>(first));
    uint32_t miss[kLaneWindows];
#pragma unroll
    for (int o = 0; o < kLaneWindows; ++o) miss[o] = 0;
    // One shard word per iteration, kept a loop so that the words' mask arrays
    // are never live at once. Word q of window o's mask is T(o) | T(o + fd)
    // (doubling over the s-mers' bits); since (T | U) & ~w = (T & ~w) | (U & ~w),
    // T and the shifted U are tested one after the other on one array.
    using Lim = WideLimits<FE>;
#pragma unroll 1
    for (int q = 0; q < 4; ++q) {
        uint64_t s[WideShape<FE>::kSmerPositions];
#pragma unroll
        for (int p = 0; p < WideShape<FE>::kSmerPositions; ++p) s[p] = word_bits<H>(lw.code[p], q);
        doubling<Lim::kMaxF>(s, sp.fl, OrOp{});
        for (int part = 0; part < 2; ++part) {  // wave-uniform
            if (part == 1) {
                if (sp.fd == 0) break;
                shift_down<max_span_rest(Lim::kMaxF)>(s, sp.fd);
            }
            const char* at = slot0 + 8 * q;
#pragma unroll
            for (int o = 0; o < kLaneWindows; ++o) {
                at += static_cast<int>(__builtin_amdgcn_ubfe(heads, o, 1)) << 5;  // the run of window o
                uint64_t w = 0;
                if ((in_chunk >> o) & 1u) w = *reinterpret_cast<const uint64_t*>(at);
                const uint64_t m = s[o] & ~w;
                miss[o] |= lo32(m) | hi32(m);
            }
        }
    }
#pragma unroll
    for (int o = 0; o < kLaneWindows; ++o) {
        const uint32_t hit = ((in_chunk >> o) & 1u) & (miss[o] == 0 ? 1u : 0u);
        result[o >> 2] |= hit << (8 * (o & 3));
    }
}

// --- tile bodies ---------------------------------------------------------------------

// The lane's 16 symbols (and record starts) into the wave's LDS. The halo words
// past the wave's lanes (no symbols: all invalid, no record starts) are written
// once per kernel by wide_init.
template <bool Dense, bool Batched>
__device__ __forceinline__ void wide_stage(const SequenceInput& in, uint64_t first, int lane,
                                           WideLds<Dna2, Batched>& L) {
    uint32_t starts = 0;
    if constexpr (Batched) starts = record_starts16(in, first);  // issued beside the symbol load
    uint32_t packed, invalid;
    if constexpr (Dense) {
        load_chunk_dense(in.words, in.num_words, in.num_symbols, first, packed, invalid);
    } else {
        load_chunk_bytes(in.seq, in.num_symbols, first, packed, invalid);
    }
    L.st.p[lane] = packed;
    L.inv[lane] = invalid;
    if constexpr (Batched) L.bnd[lane] = starts;
}

template <bool Dense, bool Batched>
__device__ __forceinline__ void wide_stage(const SequenceInput& in, uint64_t first, int lane,
                                           WideLds<Protein5, Batched>& L) {
    static_assert(!Dense, "the byte-code front end reads byte sequences only");
    uint32_t starts = 0;
    if constexpr (Batched) starts = record_starts16(in, first);
    u32x4 codes;
    uint32_t invalid;
    encode_protein_chunk(load_raw(in, first), in, first, codes, invalid);
    L.st.c[lane] = codes;
    L.inv[lane] = invalid;
    if constexpr (Batched) L.bnd[lane] = starts;
}

__device__ __forceinline__ void wide_halo(Dna2::Lds& st, int i) { st.p[i] = 0; }
__device__ __forceinline__ void wide_halo(Protein5::Lds& st, int i) { st.c[i] = u32x4{0u, 0u, 0u, 0u}; }

template <class FE, bool Batched>
__device__ __forceinline__ void wide_init(WideLds<FE, Batched>& L, int lane) {
    if (lane < kLdsWords - kWave) {
        wide_halo(L.st, kWave + lane);
        L.inv[kWave + lane] = 0xFFFFu;
        if constexpr (Batched) L.bnd[kWave + lane] = 0;
    }
    wave_lds_sync();
}

// First window of the lane's 16 in tile `tile`. The lane's offset 16 * lane is
// computed per tile by an instruction the compiler cannot hoist, so that no copy
// of it is kept live across the tile loop (one that spilled in the kernels
// closest to the register budget).
template <int WaveWindows>
__device__ __forceinline__ uint64_t wide_lane_start(uint64_t tile, int lane) {
    uint32_t off;
    asm volatile("v_lshlrev_b32 %0, 4, %1" : "=v"(off) : "v"(lane));
    return tile_start<WaveWindows>(tile) + off;
}

// Query tile (contains_tile with runtime spans).
template <class FE, int H, bool Dense, bool Batched>
__device__ __forceinline__ void wide_contains_tile(const SequenceInput& in, const FilterView& filter,
                                                   const Spans& sp, uint64_t num_windows, uint8_t* out,
                                                   uint64_t tile, int lane, WideLds<FE, Batched>& L) {
    using WS = WideShape<FE>;
    const uint64_t first_window = wide_lane_start<WS::kWaveWindows>(tile, lane);
    wide_stage<Dense>(in, first_window, lane, L);
    wave_lds_sync();

    const bool owns = lane < WS::kWindowLanes;
    WideLanes<FE, H> lw;
    wide_lane_windows<FE, H, Batched>(L, lane, owns, first_window, num_windows,
                                      static_cast<uint32_t>(filter.num_shards - 1), sp, lw);
    uint32_t cont;
    const uint32_t heads = run_heads(lw, lane, cont);
    uint32_t total;
    const uint32_t base = wave_exclusive_sum(static_cast<uint32_t>(__popc(heads)), lane, total);
    publish_heads(lw, heads, base, L);
    wave_lds_sync();

    // dense shard loads: lanes 2i and 2i+1 fetch the two 16-byte halves of one
    // slot's shard in the same instruction (slots 32 j + i)
    const u32x4* shards = reinterpret_cast<const u32x4*>(filter.words);
    const uint32_t n_slots = total < WideLds<FE, Batched>::kSlots ? total : WideLds<FE, Batched>::kSlots;
    constexpr int kHalves = 2 * WideLds<FE, Batched>::kSlots / kWave;
    const uint32_t half = static_cast<uint32_t>(lane) & 1u;
    // rounds of 256 slots (8 loads per lane), the second only for tiles with
    // more runs: fewer load registers live at once
#pragma unroll
    for (int r = 0; r < kHalves / 8; ++r) {
        if (static_cast<uint32_t>(r) * 256u >= n_slots) break;  // wave-uniform
        u32x4 sw[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const uint32_t q = static_cast<uint32_t>(lane + (8 * r + j) * kWave) >> 1;
            if (q < n_slots) sw[j] = shards[2 * static_cast<uint64_t>(L.slot_shard[q]) + half];
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const uint32_t q = static_cast<uint32_t>(lane + (8 * r + j) * kWave) >> 1;
            if (q < n_slots) reinterpret_cast<u32x4*>(L.slot_w[q])[half] = sw[j];
        }
    }
    wave_lds_sync();

    uint32_t result[4] = {0, 0, 0, 0};
    if (total <= WideLds<FE, Batched>::kSlots) {  // wave-uniform: every run of the tile has its LDS slot
        wide_findere<FE, H, Batched, true>(lw, heads, base, L, filter, sp, result);
    } else {
        wide_findere<FE, H, Batched, false>(lw, heads, base, L, filter, sp, result);
    }
    wave_lds_sync();  // the next tile of this wave reuses the LDS

    if (!owns || first_window >= num_windows) return;
    uint8_t* dst = out + first_window;
    if (first_window + kLaneWindows <= num_windows && (reinterpret_cast<uintptr_t>(dst) & 15u) == 0) {
        const u32x4 r = {result[0], result[1], result[2], result[3]};
        __builtin_nontemporal_store(r, reinterpret_cast<u32x4*>(dst));
    } else {
        const uint64_t n = num_windows - first_window;
#pragma unroll
        for (int o = 0; o < kLaneWindows; ++o) {
            if (static_cast<uint64_t>(o) < n) dst[o] = static_cast<uint8_t>((result[o >> 2] >> (8 * (o & 3))) & 0xFFu);
        }
    }
}

// Query tile for short minimizer spans (w <= 3: up to one run per window, more
// than the 512 slots hold): the runs in chunks of 512, each chunk loading its
// shards into the slots and answering its windows.
template <class FE, int H, bool Dense, bool Batched>
__device__ __forceinline__ void wide_contains_tile_chunked(const SequenceInput& in, const FilterView& filter,
                                                   const Spans& sp, uint64_t num_windows, uint8_t* out,
                                                   uint64_t tile, int lane, WideLds<FE, Batched>& L) {
    using WS = WideShape<FE>;
    const uint64_t first_window = wide_lane_start<WS::kWaveWindows>(tile, lane);
    wide_stage<Dense>(in, first_window, lane, L);
    wave_lds_sync();

    // the runs in chunks of kSlots (one chunk unless the tile has more runs than
    // slots; each further chunk recomputes the lane windows, which kept live
    // across a chunk would spill): dense shard loads into the slots -- lanes 2i
    // and 2i+1 fetch the two 16-byte halves of one slot's shard in the same
    // instruction, slots 32 j + i -- then the findere test of the chunk's windows
    constexpr uint32_t kSlotsU = static_cast<uint32_t>(WideLds<FE, Batched>::kSlots);
    constexpr int kHalves = 2 * WideLds<FE, Batched>::kSlots / kWave;
    const bool owns = lane < WS::kWindowLanes;
    const u32x4* shards = reinterpret_cast<const u32x4*>(filter.words);
    const uint32_t half = static_cast<uint32_t>(lane) & 1u;
    uint32_t result[4] = {0, 0, 0, 0};
    for (uint32_t first = 0;; first += kSlotsU) {  // wave-uniform
        WideLanes<FE, H> lw;
        wide_lane_windows<FE, H, Batched>(L, lane, owns, first_window, num_windows,
                                          static_cast<uint32_t>(filter.num_shards - 1), sp, lw);
        uint32_t cont;
        const uint32_t heads = run_heads(lw, lane, cont);
        uint32_t total;
        const uint32_t base = wave_exclusive_sum(static_cast<uint32_t>(__popc(heads)), lane, total);
        publish_heads(lw, heads, base - first, L);
        wave_lds_sync();
        const uint32_t n_slots = total - first < kSlotsU ? total - first : kSlotsU;
        // rounds of 256 slots (8 loads per lane), the second only for tiles
        // with more runs: fewer load registers live at once
#pragma unroll
        for (int r = 0; r < kHalves / 8; ++r) {
            if (static_cast<uint32_t>(r) * 256u >= n_slots) break;  // wave-uniform
            u32x4 sw[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const uint32_t q = static_cast<uint32_t>(lane + (8 * r + j) * kWave) >> 1;
                if (q < n_slots) sw[j] = shards[2 * static_cast<uint64_t>(L.slot_shard[q]) + half];
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const uint32_t q = static_cast<uint32_t>(lane + (8 * r + j) * kWave) >> 1;
                if (q < n_slots) reinterpret_cast<u32x4*>(L.slot_w[q])[half] = sw[j];
            }
        }
        wave_lds_sync();
        // one form for every chunk: a second inlined form spills
        wide_findere_chunk<FE, H, Batched>(lw, heads, base, first, L, sp, result);
        wave_lds_sync();  // the next chunk or tile of this wave reuses the LDS
        if (first + kSlotsU >= total) break;
    }

    if (!owns || first_window >= num_windows) return;
    uint8_t* dst = out + first_window;
    if (first_window + kLaneWindows <= num_windows && (reinterpret_cast<uintptr_t>(dst) & 15u) == 0) {
        const u32x4 r = {result[0], result[1], result[2], result[3]};
        __builtin_nontemporal_store(r, reinterpret_cast<u32x4*>(dst));
    } else {
        const uint64_t n = num_windows - first_window;
#pragma unroll
        for (int o = 0; o < kLaneWindows; ++o) {
            if (static_cast<uint64_t>(o) < n) dst[o] = static_cast<uint8_t>((result[o >> 2] >> (8 * (o & 3))) & 0xFFu);
        }
    }
}

// Insert tile (insert_tile with runtime spans): runs get LDS slots in chunks
// of 512; each lane ORs the masks of its piece of every run of the
// chunk into the run's slot, one shard word at a time, then the wave flushes
// the slots as records (sorted insert) or 64-bit atomic ORs.
template <class FE, int H, bool Dense, bool Batched>
__device__ __forceinline__ void wide_insert_tile(const SequenceInput& in, const FilterView& filter,
                                                 const Spans& sp, uint64_t num_windows, uint64_t tile,
                                                 int lane, WideLds<FE, Batched>& L, const RecordSink& sink,
                                                 uint32_t region, uint32_t& emitted) {
    using WS = WideShape<FE>;
    constexpr uint32_t kSlotsU = static_cast<uint32_t>(WideLds<FE, Batched>::kSlots);
    const bool owns = lane < WS::kWindowLanes;
    const uint64_t first_window = wide_lane_start<WS::kWaveWindows>(tile, lane);

    wide_stage<Dense>(in, first_window, lane, L);
    uint32_t first = 0;  // first run of the chunk (wave-uniform)
    for (;;) {
        zero_slots(L, lane);
        wave_lds_sync();

        WideLanes<FE, H> lw;
        wide_lane_windows<FE, H, Batched>(L, lane, owns, first_window, num_windows,
                                          static_cast<uint32_t>(filter.num_shards - 1), sp, lw);
        uint32_t cont;
        const uint32_t heads = run_heads(lw, lane, cont);
        uint32_t total;
        const uint32_t base = wave_exclusive_sum(static_cast<uint32_t>(__popc(heads)), lane, total) - first;
        publish_heads(lw, heads, base, L);
        // bit o: window o is the last window of its run inside this lane
        const uint32_t ends = lw.valid & ~(cont >> 1);

#pragma unroll
        for (int q = 0; q < 4; ++q) {
            uint64_t a[kLaneWindows];
            window_words<FE, H>(lw, q, sp, a);
            uint64_t acc = 0;
#pragma unroll
            for (int o = 0; o < kLaneWindows; ++o) {
                acc = ((cont >> o) & 1u) ? (acc | a[o]) : a[o];
                if ((ends >> o) & 1u) {
                    const uint32_t slot = base + static_cast<uint32_t>(__popc(heads & ((2u << o) - 1u))) - 1u;
                    if (slot < kSlotsU) {
                        __hip_atomic_fetch_or(&L.slot_w[slot][q], acc, __ATOMIC_RELAXED,
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
        for (int j = 0; j < WideLds<FE, Batched>::kSlots / kWave; ++j) {
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

// --- kernels (two waves per SIMD: the per-word mask arrays need the registers) -------

constexpr int kWideWaves = 2;

// Chunked: the query tiles of short minimizer spans (wide_contains_tile_chunked),
// a kernel of its own so that each form gets its own register allocation.
template <class FE, int H, bool Dense, bool Chunked>
__global__ __launch_bounds__(kBlock, kWideWaves) void wide_contains_kernel(SequenceInput in, FilterView filter,
                                                                           Spans sp, uint64_t num_windows,
                                                                           uint64_t num_tiles, uint8_t* out) {
    __shared__ WideLds<FE, false> lds[kWavesPerBlock];
    const int lane = static_cast<int>(threadIdx.x) & (kWave - 1);
    const int wave = wave_in_block();
    wide_init(lds[wave], lane);
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kWavesPerBlock;
    for (uint64_t tile = static_cast<uint64_t>(blockIdx.x) * kWavesPerBlock + wave; tile < num_tiles;
         tile += stride) {
        if constexpr (Chunked) {
            wide_contains_tile_chunked<FE, H, Dense, false>(in, filter, sp, num_windows, out, tile, lane, lds[wave]);
        } else {
            wide_contains_tile<FE, H, Dense, false>(in, filter, sp, num_windows, out, tile, lane, lds[wave]);
        }
    }
}

template <class FE, int H, bool Dense>
__global__ __launch_bounds__(kBlock, kWideWaves) void wide_insert_kernel(SequenceInput in, FilterView filter,
                                                                         Spans sp, uint64_t num_windows,
                                                                         uint64_t num_tiles, RecordSink sink) {
    __shared__ WideLds<FE, false> lds[kWavesPerBlock];
    mark_words_written(filter);
    const int lane = static_cast<int>(threadIdx.x) & (kWave - 1);
    const int wave = wave_in_block();
    wide_init(lds[wave], lane);
    const uint32_t region = blockIdx.x * kWavesPerBlock + static_cast<uint32_t>(wave);
    uint32_t emitted = 0;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kWavesPerBlock;
    for (uint64_t tile = region; tile < num_tiles; tile += stride) {
        wide_insert_tile<FE, H, Dense, false>(in, filter, sp, num_windows, tile, lane, lds[wave], sink, region,
                                              emitted);
    }
    if (sink.pairs != nullptr && lane == 0) {
        sink.counts[region] = emitted;
        if (emitted == sink.cap) *sink.overflow = 1u;  // later runs may have used atomics
    }
}

template <class FE, int H, bool Queue>
__global__ __launch_bounds__(kBlock, kWideWaves) void wide_batched_contains_kernel(
    SequenceInput in, FilterView filter, Spans sp, uint64_t num_windows, uint32_t num_tiles, uint32_t claim,
    uint8_t* out) {
    __shared__ WideLds<FE, true> lds[kWavesPerBlock];
    const int lane = static_cast<int>(threadIdx.x) & (kWave - 1);
    const int wave = wave_in_block();
    if (__builtin_amdgcn_readfirstlane(in.state[1]) != 0) return;  // malformed offsets
    wide_init(lds[wave], lane);
    for_each_tile<Queue>(in.state, num_tiles, claim, lane, wave, [&](uint32_t tile) {
        wide_contains_tile<FE, H, false, true>(in, filter, sp, num_windows, out, tile, lane, lds[wave]);
    });
}

template <class FE, int H, bool Queue>
__global__ __launch_bounds__(kBlock, kWideWaves) void wide_batched_insert_kernel(
    SequenceInput in, FilterView filter, Spans sp, uint64_t num_windows, uint32_t num_tiles, uint32_t claim,
    RecordSink sink) {
    __shared__ WideLds<FE, true> lds[kWavesPerBlock];
    mark_words_written(filter);
    const int lane = static_cast<int>(threadIdx.x) & (kWave - 1);
    const int wave = wave_in_block();
    wide_init(lds[wave], lane);
    const uint32_t region = blockIdx.x * kWavesPerBlock + static_cast<uint32_t>(wave);
    uint32_t emitted = 0;
    if (__builtin_amdgcn_readfirstlane(in.state[1]) == 0) {  // else malformed offsets: no work
        wide_init(lds[wave], lane);
        for_each_tile<Queue>(in.state, num_tiles, claim, lane, wave, [&](uint32_t tile) {
            wide_insert_tile<FE, H, false, true>(in, filter, sp, num_windows, tile, lane, lds[wave], sink,
                                                 region, emitted);
        });
    }
    if (sink.pairs != nullptr && lane == 0) {
        sink.counts[region] = emitted;
        if (emitted == sink.cap) *sink.overflow = 1u;  // later runs may have used atomics
    }
}

// --- host launchers ----------------------------------------------------------------

inline uint64_t windows_rt(uint64_t num_symbols, uint32_t k) {
    return num_symbols < k ? 0 : num_symbols - k + 1;
}

template <class FE>
uint32_t wide_grid(uint64_t num_windows, uint64_t& num_tiles) {
    num_tiles = (num_windows + WideShape<FE>::kWaveWindows - 1) / WideShape<FE>::kWaveWindows;
    uint64_t blocks = (num_tiles + kWavesPerBlock - 1) / kWavesPerBlock;
    if (blocks > kStripMaxBlocks) blocks = kStripMaxBlocks;
    return static_cast<uint32_t>(blocks);
}

template <class FE, int H>
hipError_t launch_wide_insert(const KernelConfig& c, const SequenceInput& in, FilterView f,
                              const RecordSink* sink, hipStream_t stream) {
    const uint64_t n = windows_rt(in.num_symbols, c.k);
    const RecordSink s = sink ? *sink : RecordSink{};
    uint64_t tiles = 0;
    const uint32_t grid = n == 0 ? 0u : sink_grid(wide_grid<FE>(n, tiles), sink);
    if (grid == 0) return hipSuccess;
    const Spans sp = wide_spans<FE>(c);
    (void)hipGetLastError();  // report only this launch
    if constexpr (FE::kBits == 2) {
        if (in.dense) {
            hipLaunchKernelGGL((wide_insert_kernel<FE, H, true>), dim3(grid), dim3(kBlock), 0, stream, in, f, sp,
                               n, tiles, s);
            return hipGetLastError();
        }
    }
    hipLaunchKernelGGL((wide_insert_kernel<FE, H, false>), dim3(grid), dim3(kBlock), 0, stream, in, f, sp, n,
                       tiles, s);
    return hipGetLastError();
}

template <class FE, int H>
hipError_t launch_wide_contains(const KernelConfig& c, const SequenceInput& in, FilterView f, uint8_t* out,
                                hipStream_t stream) {
    const uint64_t n = windows_rt(in.num_symbols, c.k);
    if (n == 0) return hipSuccess;
    uint64_t tiles;
    const uint32_t grid = wide_grid<FE>(n, tiles);
    const Spans sp = wide_spans<FE>(c);
    // minimizer spans up to 3 give 2 / (w + 1) >= 1/2 runs per window: most tiles
    // hold more runs than the 512 slots
    const bool chunked = sp.w <= 3;
    (void)hipGetLastError();  // report only this launch
    if constexpr (FE::kBits == 2) {
        if (in.dense) {
            if (chunked) {
                hipLaunchKernelGGL((wide_contains_kernel<FE, H, true, true>), dim3(grid), dim3(kBlock), 0, stream,
                                   in, f, sp, n, tiles, out);
            } else {
                hipLaunchKernelGGL((wide_contains_kernel<FE, H, true, false>), dim3(grid), dim3(kBlock), 0, stream,
                                   in, f, sp, n, tiles, out);
            }
            return hipGetLastError();
        }
    }
    if (chunked) {
        hipLaunchKernelGGL((wide_contains_kernel<FE, H, false, true>), dim3(grid), dim3(kBlock), 0, stream, in, f,
                           sp, n, tiles, out);
    } else {
        hipLaunchKernelGGL((wide_contains_kernel<FE, H, false, false>), dim3(grid), dim3(kBlock), 0, stream, in, f,
                           sp, n, tiles, out);
    }
    return hipGetLastError();
}

// Persistent grid as for the compile-time lane strips (a sorted insert with no
// windows still launches, so that every region count is written).
template <class FE, int H>
hipError_t launch_wide_batched_insert(const KernelConfig& c, const SequenceInput& in, FilterView f,
                                      const BatchLaunch& b, const RecordSink* sink, hipStream_t stream) {
    const uint64_t n = windows_rt(in.num_symbols, c.k);
    if (n == 0 && sink == nullptr) return hipSuccess;
    const RecordSink s = sink ? *sink : RecordSink{};
    const uint32_t tiles =
        static_cast<uint32_t>((n + WideShape<FE>::kWaveWindows - 1) / WideShape<FE>::kWaveWindows);
    const uint32_t blocks = sink_grid(batched_grid(b, tiles), sink);
    if (blocks == 0) return hipSuccess;
    const Spans sp = wide_spans<FE>(c);
    (void)hipGetLastError();  // report only this launch
    if (b.claim == 0) {
        hipLaunchKernelGGL((wide_batched_insert_kernel<FE, H, false>), dim3(blocks), dim3(kBlock), 0, stream,
                           in, f, sp, n, tiles, 0u, s);
    } else {
        hipLaunchKernelGGL((wide_batched_insert_kernel<FE, H, true>), dim3(blocks), dim3(kBlock), 0, stream, in,
                           f, sp, n, tiles, b.claim, s);
    }
    return hipGetLastError();
}

template <class FE, int H>
hipError_t launch_wide_batched_contains(const KernelConfig& c, const SequenceInput& in, FilterView f,
                                        const BatchLaunch& b, uint8_t* out, hipStream_t stream) {
    const uint64_t n = windows_rt(in.num_symbols, c.k);
    if (n == 0) return hipSuccess;
    const uint32_t tiles =
        static_cast<uint32_t>((n + WideShape<FE>::kWaveWindows - 1) / WideShape<FE>::kWaveWindows);
    const Spans sp = wide_spans<FE>(c);
    (void)hipGetLastError();  // report only this launch
    if (b.claim == 0) {
        hipLaunchKernelGGL((wide_batched_contains_kernel<FE, H, false>), dim3(batched_grid(b, tiles)),
                           dim3(kBlock), 0, stream, in, f, sp, n, tiles, 0u, out);
    } else {
        hipLaunchKernelGGL((wide_batched_contains_kernel<FE, H, true>), dim3(batched_grid(b, tiles)),
                           dim3(kBlock), 0, stream, in, f, sp, n, tiles, b.claim, out);
    }
    return hipGetLastError();
}

}  // namespace sbf::strip
