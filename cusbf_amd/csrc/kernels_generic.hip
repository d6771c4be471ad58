// kernels_generic.hip -- Super Bloom kernels for any configuration and alphabet,
// plus the packing and reduction kernels.
//
// One window per thread over an LDS tile of encoded symbols. The insert and
// query kernels hash every m-mer and s-mer of the tile once and build sparse
// tables over them in LDS (hashed_tile below), so a window's minimizer and
// findere masks cost two table reads each instead of the k-m+1 m-mer and k-s+1
// s-mer hashes the reference computes per window (include/cusbf/detail/
// kernels.cuh:40-172); the diagnostic run counter recomputes each window the
// reference's way. Configurations without lane-strip kernels (triplets, custom
// alphabets, other k/s/m/h, dense protein input) run here; the lane-strip
// kernels are checked against these and against the CPU oracle.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "launch.hpp"
#include "sbf_common.hpp"

namespace sbf::generic {

constexpr int kBlock = 256;
constexpr int kMaxK = 64;
constexpr uint64_t kMaxBlocks = 1ull << 16;  // 2^24 work-items per dispatch at most

// Encoded symbol p of the input; positions outside [0, n) get the invalid
// sentinel (they only occur in windows that are never reported).
__device__ __forceinline__ uint8_t symbol_at(const KernelConfig& c, const SequenceInput& in, int64_t p) {
    if (p < 0 || static_cast<uint64_t>(p) >= in.num_symbols) return static_cast<uint8_t>(c.invalid_symbol);
    const uint64_t up = static_cast<uint64_t>(p);
    if (in.dense) {
        // dense_packed.cuh:56-65: LSB-first decode, masked to symbolBits.
        const uint64_t per_word = 64u / c.symbol_bits;
        const uint64_t sym_mask = (1ull << c.symbol_bits) - 1ull;
        return static_cast<uint8_t>((in.words[up / per_word] >> ((up % per_word) * c.symbol_bits)) & sym_mask);
    }
    return encode_symbol(c, in.seq + up * c.symbol_width);
}

// Tile of encoded symbols [first, first + count).
__device__ __forceinline__ void load_tile(const KernelConfig& c, const SequenceInput& in, int64_t first,
                                          int count, uint8_t* tile) {
    for (int i = static_cast<int>(threadIdx.x); i < count; i += kBlock) tile[i] = symbol_at(c, in, first + i);
}

struct Window {
    bool valid;
    uint64_t packed;
};

// sequence_kmer.cuh:74-82 (pack) and :92-101 (validity).
__device__ __forceinline__ Window read_window(const KernelConfig& c, const uint8_t* t) {
    const uint64_t sym_mask = (1ull << c.symbol_bits) - 1ull;
    Window w{true, 0};
    for (uint32_t i = 0; i < c.k; ++i) {
        const uint8_t sym = t[i];
        w.valid &= sym != static_cast<uint8_t>(c.invalid_symbol);
        w.packed = (w.packed << c.symbol_bits) | (sym & sym_mask);
    }
    return w;
}

// filter_common.cuh:128-134
__device__ __forceinline__ uint64_t subwindow(const KernelConfig& c, uint64_t packed, uint32_t len,
                                              uint32_t start) {
    return (packed >> (c.symbol_bits * (c.k - (start + len)))) & window_mask(c.symbol_bits, len);
}

// sequence_kmer.cuh:40-50
__device__ __forceinline__ uint64_t minimizer(const KernelConfig& c, uint64_t packed) {
    uint64_t best = ~0ull;
    for (uint32_t o = 0; o + c.m <= c.k; ++o) {
        const uint64_t h = minimizer_hash64(subwindow(c, packed, c.m, o));
        best = h < best ? h : best;
    }
    return best;
}

// filter_common.cuh:27-93: the 16 salts, as a table indexed by a loop counter.
__constant__ const uint64_t kSaltTable[16] = {salt(0),  salt(1),  salt(2),  salt(3), salt(4),  salt(5),
                                              salt(6),  salt(7),  salt(8),  salt(9), salt(10), salt(11),
                                              salt(12), salt(13), salt(14), salt(15)};

// filter_impl.cuh:36-77 for one s-mer hash: hash i sets bit bit_address(i) of word
// i % 4 (the hash count is a multiple of 4, so the word index is a constant of
// the unrolled inner loop and the words stay in registers).
__device__ __forceinline__ void or_smer_bits(const KernelConfig& c, uint64_t h, uint64_t (&m)[4]) {
    const uint32_t slice = 64u / c.h;
    for (uint32_t g = 0; g < (c.h >> 2); ++g) {
#pragma unroll
        for (uint32_t q = 0; q < 4; ++q) {
            const uint32_t i = 4 * g + q;
            const uint32_t pos = slice >= 6u ? static_cast<uint32_t>((h >> (slice * i)) & 63u)
                                             : static_cast<uint32_t>((h * kSaltTable[i]) >> 58);
            m[q] |= 1ull << pos;
        }
    }
}

// kernels.cuh:294-304 with filter_impl.cuh:54-77.
__device__ __forceinline__ void masks(const KernelConfig& c, uint64_t packed, uint64_t (&m)[4]) {
    m[0] = m[1] = m[2] = m[3] = 0;
    for (uint32_t o = 0; o + c.s <= c.k; ++o) or_smer_bits(c, hash64(subwindow(c, packed, c.s, o)), m);
}

// Batched form: true when a record starts inside window j after its first
// symbol: a set bit among symbols j+1 .. j+k-1 of the record-start bitmap, or,
// without the bitmap, a first start after j (binary search over the offsets)
// closer than k. Malformed offsets give wrong answers, never an access outside
// the offsets array or the bitmap.
__device__ __forceinline__ bool crosses_record(const SequenceInput& in, uint64_t j, uint32_t k) {
    if (in.starts != nullptr) {
        if (k < 2) return false;
        const uint64_t first = j + 1;
        const uint64_t last = j + k - 1;  // < num_symbols for a window in range
        for (uint64_t w = first >> 6; w <= (last >> 6); ++w) {
            uint64_t bits = in.starts[w];
            if (w == (first >> 6)) bits &= ~0ull << (first & 63);
            if (w == (last >> 6) && (last & 63) != 63) bits &= (2ull << (last & 63)) - 1ull;
            if (bits) return true;
        }
        return false;
    }
    uint64_t lo = 0, hi = in.num_records + 1;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) / 2;
        if (in.offsets[mid] <= j) {
            lo = mid + 1;
        } else {
            hi = mid;
        }
    }
    return lo <= in.num_records && in.offsets[lo] - j < k;
}

__device__ __forceinline__ bool window_ok(const SequenceInput& in, const Window& w, uint64_t j,
                                          uint32_t k) {
    return w.valid && (in.offsets == nullptr || !crosses_record(in, j, k));
}

// Batched form with offsets the prepass found malformed: no work.
__device__ __forceinline__ bool batch_rejected(const SequenceInput& in) {
    return in.state != nullptr && in.state[1] != 0;
}

// --- tiles ----------------------------------------------------------------------
//
// A block's tile is R * kBlock consecutive windows; thread t owns windows
// first + t + r * kBlock (r < R), so the lanes of a wave hold consecutive windows
// of each of its R slices. Fixed per-tile latencies (the symbol loads, the
// barriers, the shard loads of the query) are paid once for the R windows of a
// thread, whose loads are in flight together.
//
// Hashed tiles: the block hashes each of the tile's m-mers and s-mers once
// (sequence_kmer.cuh:40-59 hashes k-m+1 m-mers and k-s+1 s-mers per window),
// keeps the m-mer hashes and the s-mers' word masks (filter_impl.cuh:54-77) in
// LDS, and builds sparse tables over them: level l holds the minimum (for the
// masks: the OR) over 2^l consecutive positions. A window's minimizer hash is the
// minimum of two overlapping level-L entries, L = floor(log2(k - m + 1)), and its
// findere masks are the OR of two level-L' entries, L' = floor(log2(k - s + 1));
// min and OR are idempotent, so the overlap does not change either. A bit per
// position marks invalid symbols (and positions past the input), and a window
// is valid when its k bits are clear.
template <int R>
struct HashedTile {
    static constexpr int kR = R;
    static constexpr int kWindows = R * kBlock;
    static constexpr int kSymbols = kWindows + kMaxK;  // the windows and the k - 1 symbols after them
    static_assert(kSymbols % 64 == 0, "the invalid bits are ballots of whole waves");
    uint8_t sym[kSymbols];
    uint64_t invalid[kSymbols / 64];  // bit p of the tile: symbol p invalid or past the input
    uint64_t mh[2][kSymbols];         // m-mer hashes, two sparse-table levels in turn
    uint64_t mk[2][4][kSymbols];      // s-mer masks (word q of position p at [.][q][p])
};

// Per-window form: each thread packs its windows' k symbols and hashes their
// m-mers and s-mers, as the reference does (kernels.cuh:40-172). For short
// spans this is cheaper than a hashed tile's tables.
template <int R>
struct DirectTile {
    static constexpr int kR = R;
    static constexpr int kWindows = R * kBlock;
    static constexpr int kSymbols = kWindows + kMaxK;
    uint8_t sym[kSymbols];
};

// What a thread learns about its R windows.
template <int R>
struct ThreadWindows {
    bool valid[R];  // valid and in range
    uint32_t shard[R];
    uint64_t m[R][4];
};

__device__ __forceinline__ uint32_t floor_log2(uint32_t x) { return 31u - static_cast<uint32_t>(__builtin_clz(x)); }

// sequence_kmer.cuh:74-82: len symbols packed MSB-first.
__device__ __forceinline__ uint64_t pack_symbols(const KernelConfig& c, const uint8_t* p, uint32_t len) {
    const uint64_t sym_mask = (1ull << c.symbol_bits) - 1ull;
    uint64_t v = 0;
    for (uint32_t i = 0; i < len; ++i) v = (v << c.symbol_bits) | (p[i] & sym_mask);
    return v;
}

// The windows of a tile whose symbols the block has loaded into T.sym. Every
// thread of the block calls.
template <int R>
__device__ __forceinline__ void tile_windows(const KernelConfig& c, const SequenceInput& in, const FilterView& f,
                                             uint64_t first, uint64_t num_windows, HashedTile<R>& T,
                                             ThreadWindows<R>& out) {
    using Tile = HashedTile<R>;
    const uint32_t t = threadIdx.x;
    const uint32_t count = Tile::kWindows + c.k - 1;     // symbols the tile's windows cover
    const uint32_t span_m = c.k - c.m + 1;               // m-mers per window
    const uint32_t span_s = c.k - c.s + 1;               // s-mers per window
    const uint32_t pm = Tile::kWindows + span_m - 1;     // m-mer positions of the tile
    const uint32_t ps = Tile::kWindows + span_s - 1;     // s-mer positions of the tile
    const uint32_t both = pm < ps ? pm : ps;
    const uint32_t longest = c.m > c.s ? c.m : c.s;
    // level 0: one hash per m-mer, one mask per s-mer (positions p < pm or ps
    // read symbols p .. p + len - 1 <= count - 1)
    for (uint32_t p = t; p < static_cast<uint32_t>(Tile::kSymbols); p += kBlock) {
        const bool bad = p >= count || T.sym[p] == static_cast<uint8_t>(c.invalid_symbol);
        const uint64_t ballot = __ballot(bad);  // the loop is uniform per wave
        if ((p & 63u) == 0) T.invalid[p >> 6] = ballot;
        uint64_t vm = 0, vs = 0;
        if (p < both) {
            const uint64_t v = pack_symbols(c, T.sym + p, longest);
            vm = v >> (c.symbol_bits * (longest - c.m));
            vs = v >> (c.symbol_bits * (longest - c.s));
        } else if (p < pm) {
            vm = pack_symbols(c, T.sym + p, c.m);
        } else if (p < ps) {
            vs = pack_symbols(c, T.sym + p, c.s);
        }
        if (p < pm) T.mh[0][p] = minimizer_hash64(vm);
        if (p < ps) {
            // kernels.cuh:294-304 with filter_impl.cuh:54-77, one s-mer
            uint64_t w[4] = {0, 0, 0, 0};
            or_smer_bits(c, hash64(vs), w);
#pragma unroll
            for (int q = 0; q < 4; ++q) T.mk[0][q][p] = w[q];
        }
    }
    // sparse-table levels: an entry of level l covers positions p .. p + 2^l - 1
    const uint32_t lm = floor_log2(span_m), ls = floor_log2(span_s);
    const uint32_t levels = lm > ls ? lm : ls;
    uint32_t cm = 0, cs = 0;
    for (uint32_t l = 1; l <= levels; ++l) {
        __syncthreads();
        const uint32_t half = 1u << (l - 1), span = 1u << l;
        for (uint32_t p = t; p < static_cast<uint32_t>(Tile::kSymbols); p += kBlock) {
            if (l <= lm && p + span <= pm) {
                const uint64_t a = T.mh[cm][p], b = T.mh[cm][p + half];
                T.mh[cm ^ 1u][p] = a < b ? a : b;
            }
            if (l <= ls && p + span <= ps) {
#pragma unroll
                for (int q = 0; q < 4; ++q) T.mk[cs ^ 1u][q][p] = T.mk[cs][q][p] | T.mk[cs][q][p + half];
            }
        }
        if (l <= lm) cm ^= 1u;
        if (l <= ls) cs ^= 1u;
    }
    __syncthreads();

    const uint64_t kmask = c.k >= 64 ? ~0ull : ((1ull << c.k) - 1ull);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const uint32_t tr = t + static_cast<uint32_t>(r) * kBlock;
        const uint64_t j = first + tr;
        bool valid = false;
        if (j < num_windows) {
            // symbols tr .. tr + k - 1 of the tile (tr + k - 1 < kSymbols)
            const uint32_t wi = tr >> 6, sh = tr & 63u;
            uint64_t bits = T.invalid[wi] >> sh;
            if (sh != 0) bits |= T.invalid[wi + 1] << (64u - sh);
            valid = (bits & kmask) == 0 && (in.offsets == nullptr || !crosses_record(in, j, c.k));
        }
        out.valid[r] = valid;
        out.shard[r] = 0;
        out.m[r][0] = out.m[r][1] = out.m[r][2] = out.m[r][3] = 0;
        if (valid) {
            // sequence_kmer.cuh:40-50 (minimum over the window's m-mers) and
            // filter_ref.cuh:52-55
            const uint64_t a = T.mh[cm][tr], b = T.mh[cm][tr + span_m - (1u << lm)];
            out.shard[r] = static_cast<uint32_t>((a < b ? a : b) & (f.num_shards - 1));
            const uint32_t u = tr + span_s - (1u << ls);
#pragma unroll
            for (int q = 0; q < 4; ++q) out.m[r][q] = T.mk[cs][q][tr] | T.mk[cs][q][u];
        }
    }
}

template <int R>
__device__ __forceinline__ void tile_windows(const KernelConfig& c, const SequenceInput& in, const FilterView& f,
                                             uint64_t first, uint64_t num_windows, DirectTile<R>& T,
                                             ThreadWindows<R>& out) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const uint32_t tr = threadIdx.x + static_cast<uint32_t>(r) * kBlock;
        const uint64_t j = first + tr;
        out.valid[r] = false;
        out.shard[r] = 0;
        out.m[r][0] = out.m[r][1] = out.m[r][2] = out.m[r][3] = 0;
        if (j < num_windows) {
            const Window w = read_window(c, T.sym + tr);
            if (window_ok(in, w, j, c.k)) {
                out.valid[r] = true;
                out.shard[r] = static_cast<uint32_t>(minimizer(c, w.packed) & (f.num_shards - 1));
                masks(c, w.packed, out.m[r]);
            }
        }
    }
}

// Blocks stride over tiles: a dispatch must stay below 2^32 work-items, so grids
// are capped and large inputs loop.
//
// Insert: consecutive lanes hold consecutive windows, so a run (consecutive
// valid windows with one shard) is a segment of the wave. A segmented OR scan
// over the wave (kernels.cuh:143-166 does the same per warp) leaves each run's
// masks in its last lane, which issues one 64-bit atomic OR per non-zero word:
// about one update per run instead of one per window.
//
// Sorted = true (a sorted insert): the run's last lane writes the run as a record
// into its wave's region instead (RecordSink in launch.hpp); runs past the
// region's capacity still take the atomics. That build uses direct tiles only:
// with hashed tiles it would spill scalar registers to scratch.
template <class Tile, bool Sorted>
__global__ __launch_bounds__(kBlock) void insert_kernel(KernelConfig c, SequenceInput in,
                                                        FilterView f, uint64_t num_windows,
                                                        RecordSink sink) {
    constexpr int R = Tile::kR;
    __shared__ Tile T;
    mark_words_written(f);
    const int lane = static_cast<int>(threadIdx.x) & 63;
    const uint32_t region = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    uint32_t emitted = 0;  // records this wave wrote (the same in every lane)
    if (batch_rejected(in)) {
        if (Sorted && lane == 0) sink.counts[region] = 0;
        return;
    }
    for (uint64_t first = static_cast<uint64_t>(blockIdx.x) * Tile::kWindows; first < num_windows;
         first += static_cast<uint64_t>(gridDim.x) * Tile::kWindows) {
        __syncthreads();  // the previous tile's readers are done
        load_tile(c, in, static_cast<int64_t>(first), Tile::kWindows + static_cast<int>(c.k) - 1, T.sym);
        __syncthreads();
        ThreadWindows<R> tw;
        tile_windows(c, in, f, first, num_windows, T, tw);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const bool valid = tw.valid[r];
            const uint32_t shard = tw.shard[r];
            uint64_t* m = tw.m[r];
            // run heads and the lane of this lane's run head
            const uint32_t prev_shard = __shfl_up(shard, 1, 64);
            const int prev_valid = __shfl_up(valid ? 1 : 0, 1, 64);
            const bool head = valid && (lane == 0 || !prev_valid || prev_shard != shard);
            const uint64_t heads = __ballot(head);
            const uint64_t valids = __ballot(valid);
            const uint64_t upto = heads & (lane == 63 ? ~0ull : ((2ull << lane) - 1ull));
            const int start = upto ? 63 - __clzll(static_cast<long long>(upto)) : 0;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const uint64_t t = __shfl_up(m[q], d, 64);
                    if (lane - d >= start) m[q] |= t;
                }
            }
            // the run's last lane applies its masks (or writes them as a record)
            const bool tail = valid && (lane == 63 || ((heads >> (lane + 1)) & 1ull) != 0 ||
                                        ((valids >> (lane + 1)) & 1ull) == 0);
            bool atomics = tail;
            if constexpr (Sorted) {
                const uint64_t tails = __ballot(tail);
                const uint32_t rank = static_cast<uint32_t>(__popcll(tails & ((1ull << lane) - 1ull)));
                const uint32_t room = sink.cap - emitted;
                if (tail && rank < room) {
                    const uint64_t at = static_cast<uint64_t>(region) * sink.cap + emitted + rank;
                    sink.pairs[at] = (static_cast<uint64_t>(shard) << 32) | at;
#pragma unroll
                    for (int q = 0; q < 4; ++q) sink.masks[4 * at + q] = m[q];
                    atomics = false;
                }
                const uint32_t n_tails = static_cast<uint32_t>(__popcll(tails));
                emitted += n_tails < room ? n_tails : room;
            }
            if (atomics) {
                uint64_t* dst = f.words + 4ull * shard;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    if (m[q] != 0) __hip_atomic_fetch_or(dst + q, m[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
        }
    }
    if (Sorted && lane == 0) {
        sink.counts[region] = emitted;
        if (emitted == sink.cap) *sink.overflow = 1u;  // later runs may have used atomics
    }
}

template <class Tile>
__global__ __launch_bounds__(kBlock) void contains_kernel(KernelConfig c, SequenceInput in,
                                                          FilterView f, uint64_t num_windows,
                                                          uint8_t* out) {
    constexpr int R = Tile::kR;
    __shared__ Tile T;
    if (batch_rejected(in)) return;
    for (uint64_t first = static_cast<uint64_t>(blockIdx.x) * Tile::kWindows; first < num_windows;
         first += static_cast<uint64_t>(gridDim.x) * Tile::kWindows) {
        __syncthreads();  // the previous tile's readers are done
        load_tile(c, in, static_cast<int64_t>(first), Tile::kWindows + static_cast<int>(c.k) - 1, T.sym);
        __syncthreads();
        ThreadWindows<R> tw;
        tile_windows(c, in, f, first, num_windows, T, tw);
        // the R shard loads go out together (filter_ref.cuh:70-87)
        uint64_t w[R][4];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const uint64_t* src = f.words + 4ull * tw.shard[r];
#pragma unroll
            for (int q = 0; q < 4; ++q) w[r][q] = tw.valid[r] ? src[q] : 0;
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const uint64_t j = first + threadIdx.x + static_cast<uint64_t>(r) * kBlock;
            if (j >= num_windows) continue;
            uint64_t miss = 0;
#pragma unroll
            for (int q = 0; q < 4; ++q) miss |= tw.m[r][q] & ~w[r][q];
            out[j] = tw.valid[r] && miss == 0 ? 1 : 0;
        }
    }
}

// Valid windows and shard-run heads; the tile starts one window early so each
// thread can compare with its predecessor.
__global__ __launch_bounds__(kBlock) void count_kernel(KernelConfig c, SequenceInput in,
                                                       uint64_t num_shards, uint64_t num_windows,
                                                       unsigned long long* counts) {
    __shared__ uint8_t tile[kBlock + 1 + kMaxK];
    __shared__ uint32_t shard_of[kBlock + 1];
    __shared__ uint8_t valid_of[kBlock + 1];
    __shared__ unsigned long long totals[2];
    if (threadIdx.x < 2) totals[threadIdx.x] = 0;
    for (uint64_t first = static_cast<uint64_t>(blockIdx.x) * kBlock; first < num_windows;
         first += static_cast<uint64_t>(gridDim.x) * kBlock) {
        __syncthreads();
        load_tile(c, in, static_cast<int64_t>(first) - 1, kBlock + static_cast<int>(c.k), tile);
        __syncthreads();
        for (int t = static_cast<int>(threadIdx.x); t < kBlock + 1; t += kBlock) {
            const int64_t j = static_cast<int64_t>(first) - 1 + t;
            bool valid = false;
            uint32_t shard = 0;
            if (j >= 0 && static_cast<uint64_t>(j) < num_windows) {
                const Window w = read_window(c, tile + t);
                valid = window_ok(in, w, static_cast<uint64_t>(j), c.k);
                if (valid) shard = static_cast<uint32_t>(minimizer(c, w.packed) & (num_shards - 1));
            }
            shard_of[t] = shard;
            valid_of[t] = valid;
        }
        __syncthreads();
        const uint64_t j = first + threadIdx.x;
        if (j < num_windows && valid_of[threadIdx.x + 1]) {
            const bool head =
                !valid_of[threadIdx.x] || shard_of[threadIdx.x] != shard_of[threadIdx.x + 1];
            atomicAdd(&totals[0], 1ull);
            if (head) atomicAdd(&totals[1], 1ull);
        }
    }
    __syncthreads();
    if (threadIdx.x < 2 && totals[threadIdx.x] != 0) atomicAdd(&counts[threadIdx.x], totals[threadIdx.x]);
}

// Report counts of a server batch (filter.cuh:951-990 FastxQueryReport, done on
// the device): request r owns windows [req[2r], req[2r] + req[2r+1]) of the
// batch; counts[2r] += its valid windows (no invalid symbol among the k),
// counts[2r+1] += its hit bytes equal to 1 (none without hits). A thread takes 16 consecutive
// windows: the invalid-symbol mask of their 16 + k - 1 symbols (128-bit loads;
// SWAR test for DNA), the hit bytes with one 128-bit load, then the counts go
// to the owning request -- summed over the wave first when the whole wave falls
// in one request, which is the common case.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Bit i set when byte i of x is not A, C, G or T in either case
// (DnaAlphabet::encode, Alphabet.cuh:132-141).
__device__ __forceinline__ uint32_t dna_invalid4(uint32_t x) {
    const uint32_t sym = (x >> 1) & 0x03030303u;
    const uint32_t expect = __builtin_amdgcn_perm(0u, 0x47544341u, sym);  // 'A','C','T','G' by code
    const uint32_t diff = (x & 0xDFDFDFDFu) ^ expect;
    const uint32_t nz = ((((diff & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | diff) & 0x80808080u) >> 7;
    uint32_t inv = nz | (nz >> 7);
    inv |= inv >> 14;
    return inv & 0xFu;
}

// Bit i set when symbol first+i is invalid or past the end (16 symbols).
__device__ __forceinline__ uint32_t invalid16(const KernelConfig& c, const char* seq, uint64_t n,
                                              uint64_t first) {
    if (c.symbol_width == 1 && first + 16 <= n &&
        (reinterpret_cast<uintptr_t>(seq + first) & 15u) == 0) {
        const u32x4 v = *reinterpret_cast<const u32x4*>(seq + first);
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
        uint32_t m = 0;
        if (c.alphabet == kDna) {
#pragma unroll
            for (int q = 0; q < 4; ++q) m |= dna_invalid4(w[q]) << (4 * q);
        } else {
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const char b = static_cast<char>((w[i >> 2] >> (8 * (i & 3))) & 0xFFu);
                m |= (encode_symbol(c, &b) == static_cast<uint8_t>(c.invalid_symbol) ? 1u : 0u) << i;
            }
        }
        return m;
    }
    uint32_t m = 0;
    for (int i = 0; i < 16; ++i) {
        const uint64_t q = first + static_cast<uint64_t>(i);
        const bool bad = q >= n || encode_symbol(c, seq + q * c.symbol_width) ==
                                       static_cast<uint8_t>(c.invalid_symbol);
        m |= (bad ? 1u : 0u) << i;
    }
    return m;
}

__device__ __forceinline__ unsigned long long wave_sum64(unsigned long long v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__global__ __launch_bounds__(kBlock) void request_counts_kernel(KernelConfig c, const char* seq,
                                                                uint64_t num_symbols,
                                                                const uint8_t* hits,
                                                                const uint64_t* req, uint64_t n_req,
                                                                unsigned long long* counts) {
    const uint64_t num_windows = num_symbols < c.k ? 0 : num_symbols - c.k + 1;
    const uint64_t groups = (num_windows + 15) / 16;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    const uint64_t kmask = c.k >= 64 ? ~0ull : ((1ull << c.k) - 1ull);
    // every lane runs every iteration (wave sums below)
    for (uint64_t base = static_cast<uint64_t>(blockIdx.x) * kBlock; base < groups; base += stride) {
        const uint64_t g = base + threadIdx.x;
        const bool active = g < groups;
        uint32_t valid = 0, positive = 0;
        uint64_t r = 0;
        bool whole = true;  // the lane's windows lie in request r (or the lane has none)
        if (active) {
            const uint64_t p0 = 16 * g;
            // symbols p0 .. p0 + 79 cover 16 windows of k <= 64
            const uint64_t m0 = static_cast<uint64_t>(invalid16(c, seq, num_symbols, p0)) |
                                (static_cast<uint64_t>(invalid16(c, seq, num_symbols, p0 + 16)) << 16) |
                                (static_cast<uint64_t>(invalid16(c, seq, num_symbols, p0 + 32)) << 32) |
                                (static_cast<uint64_t>(invalid16(c, seq, num_symbols, p0 + 48)) << 48);
            const uint64_t m1 = invalid16(c, seq, num_symbols, p0 + 64);
            const uint64_t n = num_windows - p0 < 16 ? num_windows - p0 : 16;
            for (uint32_t o = 0; o < n; ++o) {
                const uint64_t w = o == 0 ? m0 : ((m0 >> o) | (m1 << (64 - o)));
                valid |= ((w & kmask) == 0 ? 1u : 0u) << o;
            }
            if (hits == nullptr) {
                // insert reports: valid windows only
            } else if (n == 16 && (reinterpret_cast<uintptr_t>(hits + p0) & 15u) == 0) {
                const u32x4 h = *reinterpret_cast<const u32x4*>(hits + p0);
                const uint32_t hw[4] = {h.x, h.y, h.z, h.w};
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const uint32_t x = hw[q] ^ 0x01010101u;  // zero byte <=> hit byte == 1
                    const uint32_t nz = ((((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x) & 0x80808080u) >> 7;
                    const uint32_t bits = ~(nz | (nz >> 7) | (nz >> 14) | (nz >> 21)) & 0xFu;
                    positive |= bits << (4 * q);
                }
            } else {
                for (uint32_t o = 0; o < n; ++o) positive |= (hits[p0 + o] == 1 ? 1u : 0u) << o;
            }
            // last request starting at or before p0
            uint64_t lo = 0, hi = n_req;
            while (lo < hi) {
                const uint64_t mid = (lo + hi) / 2;
                if (req[2 * mid] <= p0) {
                    lo = mid + 1;
                } else {
                    hi = mid;
                }
            }
            r = lo == 0 ? 0 : lo - 1;
            whole = lo != 0 && p0 + n <= req[2 * r] + req[2 * r + 1];
            if (!whole) {
                // windows split over requests (or outside them): per window
                uint64_t rr = lo == 0 ? 0 : lo - 1;
                for (uint32_t o = 0; o < n; ++o) {
                    const uint64_t p = p0 + o;
                    while (rr < n_req && p >= req[2 * rr] + req[2 * rr + 1]) ++rr;
                    if (rr < n_req && p >= req[2 * rr]) {
                        if ((valid >> o) & 1u) atomicAdd(&counts[2 * rr], 1ull);
                        if ((positive >> o) & 1u) atomicAdd(&counts[2 * rr + 1], 1ull);
                    }
                }
                valid = positive = 0;
            }
        }
        const uint64_t r0 = __shfl(r, 0, 64);
        const bool uniform = __all(!active || !whole || r == r0);
        if (uniform) {
            const unsigned long long sv = wave_sum64(static_cast<unsigned long long>(__popc(valid)));
            const unsigned long long sp = wave_sum64(static_cast<unsigned long long>(__popc(positive)));
            if ((threadIdx.x & 63) == 0 && (sv | sp)) {
                atomicAdd(&counts[2 * r0], sv);
                atomicAdd(&counts[2 * r0 + 1], sp);
            }
        } else if (valid | positive) {
            atomicAdd(&counts[2 * r], static_cast<unsigned long long>(__popc(valid)));
            atomicAdd(&counts[2 * r + 1], static_cast<unsigned long long>(__popc(positive)));
        }
    }
}

// dense_packed.cuh:126-149, staged through LDS: the block reads its input
// bytes with coalesced loads, then each thread packs one output word.
constexpr int kPackMaxBytes = kBlock * 32;

__global__ __launch_bounds__(kBlock) void pack_kernel(KernelConfig c, const char* seq,
                                                      uint64_t num_symbols, uint64_t* words) {
    __shared__ uint8_t codes[kBlock * 64];
    const uint32_t per_word = 64u / c.symbol_bits;
    const uint64_t sym_mask = (1ull << c.symbol_bits) - 1ull;
    const uint32_t tile_symbols = kBlock * per_word;
    const uint64_t num_words = (num_symbols + per_word - 1) / per_word;
    for (uint64_t first_word = static_cast<uint64_t>(blockIdx.x) * kBlock; first_word < num_words;
         first_word += static_cast<uint64_t>(gridDim.x) * kBlock) {
        const uint64_t first_symbol = first_word * per_word;
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < tile_symbols; i += kBlock) {
            const uint64_t p = first_symbol + i;
            uint8_t v = 0;
            if (p < num_symbols) v = static_cast<uint8_t>(encode_symbol(c, seq + p * c.symbol_width) & sym_mask);
            codes[i] = v;
        }
        __syncthreads();
        const uint64_t word = first_word + threadIdx.x;
        if (word >= num_words) continue;
        uint64_t packed = 0;
        const uint8_t* src = codes + threadIdx.x * per_word;
        for (uint32_t i = 0; i < per_word; ++i) packed |= static_cast<uint64_t>(src[i]) << (i * c.symbol_bits);
        words[word] = packed;
    }
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__global__ __launch_bounds__(kBlock) void popcount_kernel(const uint64_t* words, uint64_t n,
                                                          unsigned long long* out) {
    unsigned long long acc = 0;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride) {
        acc += static_cast<unsigned long long>(__popcll(words[i]));
    }
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0 && acc) atomicAdd(out, acc);
}

// Filter image checksum: sum over i of mix(words[i] ^ (i * golden)) modulo 2^64
// (order-free, so waves add their sums atomically); mix is the splitmix64
// finaliser.
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(kBlock) void checksum_kernel(const uint64_t* words, uint64_t n,
                                                          unsigned long long* out) {
    unsigned long long acc = 0;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride) {
        acc += mix64(words[i] ^ (i * 0x9E3779B97F4A7C15ull));
    }
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0 && acc) atomicAdd(out, acc);
}

// count_positive_kmers.cuh:47-62: number of bytes equal to 1.
__global__ __launch_bounds__(kBlock) void count_ones_kernel(const uint8_t* hits, uint64_t n,
                                                            unsigned long long* out) {
    unsigned long long acc = 0;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    const uint64_t vec = n / 16;
    const uint4* h4 = reinterpret_cast<const uint4*>(hits);
    const bool aligned = (reinterpret_cast<uintptr_t>(hits) & 15u) == 0;
    if (aligned) {
        for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i < vec; i += stride) {
            const uint4 v = h4[i];
            const uint32_t w[4] = {v.x, v.y, v.z, v.w};
            for (int q = 0; q < 4; ++q) {
                // bytes equal to 1 (hit bytes are 0 or 1, but count exactly "== 1")
                const uint32_t x = w[q] ^ 0x01010101u;
                const uint32_t nz = (((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x) & 0x80808080u;
                acc += 4u - static_cast<unsigned>(__popc(nz));
            }
        }
    }
    const uint64_t tail_start = aligned ? vec * 16 : 0;
    for (uint64_t i = tail_start + static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride) {
        acc += hits[i] == 1;
    }
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0 && acc) atomicAdd(out, acc);
}

// count_positive_kmers.cuh:21-45, one wave per record (waves stride over records).
__global__ __launch_bounds__(kBlock) void count_ones_per_record_kernel(const uint8_t* hits,
                                                                       const uint64_t* offsets,
                                                                       const uint64_t* counts,
                                                                       uint64_t num_records,
                                                                       uint64_t* out) {
    const int lane = static_cast<int>(threadIdx.x & 63);
    const uint64_t waves = static_cast<uint64_t>(gridDim.x) * (kBlock / 64);
    for (uint64_t record = (static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x) / 64;
         record < num_records; record += waves) {
        const uint64_t begin = offsets[record];
        const uint64_t n = counts[record];
        unsigned long long acc = 0;
        for (uint64_t i = lane; i < n; i += 64) acc += hits[begin + i];
        acc = wave_sum(acc);
        if (lane == 0) out[record] = acc;
    }
}

__global__ __launch_bounds__(kBlock) void fill_kernel(uint64_t* dst, uint64_t value, uint64_t n) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride) dst[i] = value;
}

// Batched form prepass (after the state and the bitmap are zeroed): one bit
// per record start below num_symbols; state[1] is set when the offsets are not
// non-decreasing from 0 to num_symbols.
__global__ __launch_bounds__(kBlock) void batch_prepass_kernel(const uint64_t* offsets,
                                                               uint64_t num_records,
                                                               uint64_t num_symbols,
                                                               unsigned long long* starts,
                                                               uint32_t* state) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i <= num_records;
         i += stride) {
        const uint64_t v = offsets[i];
        if (i < num_records && v < num_symbols) atomicOr(&starts[v >> 6], 1ull << (v & 63));
        bool bad = i < num_records && v > offsets[i + 1];
        if (i == 0) bad = bad || v != 0 || offsets[num_records] != num_symbols;
        if (bad) atomicOr(&state[1], 1u);
    }
}

// Records back to back (`in`, offsets in symbols) -> the same records joined by one
// separator symbol each (normalized_record_batch.hpp:208-232 for records whose byte
// lengths are whole symbols): record r moves to symbol off[r] + r of `out`, and the w
// bytes after it (before record r + 1) are separators. Each thread writes one 16-byte
// chunk of `out`, found by a binary search over off[r] + r; inside one record the
// chunk's bytes are 16 contiguous input bytes, read as five aligned dwords and
// funnel-shifted. Reads stay inside `in` whatever the offsets hold.
__global__ __launch_bounds__(kBlock) void join_records_kernel(const uint8_t* in, uint64_t in_bytes,
                                                              const uint64_t* __restrict__ off,
                                                              uint64_t num_records,
                                                              uint32_t w, uint8_t sep, uint8_t* out,
                                                              uint64_t out_bytes) {
    const uint64_t chunks = (out_bytes + 15) / 16;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    for (uint64_t c = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; c < chunks; c += stride) {
        const uint64_t p0 = 16 * c;
        // last record r with (off[r] + r) * w <= p0 (record 0 starts at 0): a binary
        // search for the wave's first chunk (the same in every lane: one memory request
        // per step), then a galloping search from there for the lane's own chunk
        const uint64_t pw = p0 - 16 * static_cast<uint64_t>(threadIdx.x & 63u);
        uint64_t lo = 0, hi = num_records;
        while (hi - lo > 1) {  // wave-uniform: scalar loads
            const uint64_t mid = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>((lo + hi) >> 1)) |
                                 (static_cast<uint64_t>(__builtin_amdgcn_readfirstlane(
                                      static_cast<uint32_t>(((lo + hi) >> 1) >> 32))) << 32);
            if ((off[mid] + mid) * w <= pw) {
                lo = mid;
            } else {
                hi = mid;
            }
        }
        uint64_t step = 1;
        while (lo + step < num_records && (off[lo + step] + lo + step) * w <= p0) {
            lo += step;
            step <<= 1;
        }
        hi = lo + step < num_records ? lo + step : num_records;
        while (hi - lo > 1) {
            const uint64_t mid = lo + (hi - lo) / 2;
            if ((off[mid] + mid) * w <= p0) {
                lo = mid;
            } else {
                hi = mid;
            }
        }
        uint64_t r = lo;
        uint64_t end = (off[r + 1] + r) * w;  // record r occupies [(off[r] + r) w, end)
        const uint64_t src0 = p0 - r * w;
        if (p0 + 16 <= end && p0 + 16 <= out_bytes && (src0 & ~3ull) + 20 <= in_bytes &&
            (reinterpret_cast<uintptr_t>(out + p0) & 15u) == 0) {
            const uint32_t* d = reinterpret_cast<const uint32_t*>(in + (src0 & ~3ull));
            const uint32_t sh = static_cast<uint32_t>(src0 & 3u);
            const uint32_t a0 = d[0], a1 = d[1], a2 = d[2], a3 = d[3], a4 = d[4];
            const u32x4 v = {__builtin_amdgcn_alignbyte(a1, a0, sh), __builtin_amdgcn_alignbyte(a2, a1, sh),
                             __builtin_amdgcn_alignbyte(a3, a2, sh), __builtin_amdgcn_alignbyte(a4, a3, sh)};
            *reinterpret_cast<u32x4*>(out + p0) = v;
            continue;
        }
        for (uint64_t p = p0; p < p0 + 16 && p < out_bytes; ++p) {
            while (r + 1 < num_records && p >= (off[r + 1] + r + 1) * w) {
                ++r;
                end = (off[r + 1] + r) * w;
            }
            uint8_t b = sep;
            const uint64_t start = (off[r] + r) * w;
            if (p >= start && p < end && p - r * w < in_bytes) b = in[p - r * w];
            out[p] = b;
        }
    }
}

uint32_t grid_for(uint64_t items, uint64_t per_block, uint64_t cap) {
    uint64_t g = (items + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return static_cast<uint32_t>(g);
}

}  // namespace sbf::generic

namespace sbf {

using namespace generic;

static uint64_t window_count(const KernelConfig& c, const SequenceInput& in) {
    return in.num_symbols < c.k ? 0 : in.num_symbols - c.k + 1;
}

// Hashed tiles pay a fixed LDS cost per window (the hash tables and their
// sparse-table levels); recomputing pays for each of the window's k - m + 1
// m-mers and k - s + 1 s-mers (h bit addresses each). Measured on MI355X (k = 31,
// DNA): at (k-m+1) + (k-s+1) h/4 = 40-44 hashed tiles ran 1.7-3.5x faster, at
// 2-16 about 20% slower.
// Windows per thread of each form's tiles. Measured on MI355X: four per thread in
// the per-window form halved its occupancy (98-103 VGPRs) and ran no faster (the
// one-shard-per-window query 25% slower), so both forms take one.
constexpr int kHashedR = 1;
constexpr int kDirectR = 1;

bool hashed_tiles(const KernelConfig& c) {
    if (c.form != kFormAuto) return c.form == kFormHashed;
    return (c.k - c.m + 1) + (c.k - c.s + 1) * (c.h / 4) >= 24;
}

bool generic_hashed_tiles(const KernelConfig& c) { return hashed_tiles(c); }

hipError_t generic_insert(const KernelConfig& c, const SequenceInput& in, FilterView f,
                          const RecordSink* sink, hipStream_t stream) {
    const uint64_t n = window_count(c, in);
    if (n == 0) return hipSuccess;
    (void)hipGetLastError();  // report only this launch
    if (sink != nullptr) {
        using Tile = DirectTile<kDirectR>;
        const uint64_t most = sink->regions / (kBlock / 64);
        const uint32_t grid = static_cast<uint32_t>(std::min<uint64_t>(grid_for(n, Tile::kWindows, kMaxBlocks), most));
        hipLaunchKernelGGL((insert_kernel<Tile, true>), dim3(grid), dim3(kBlock), 0, stream, c, in, f, n, *sink);
    } else if (hashed_tiles(c)) {
        using Tile = HashedTile<kHashedR>;
        hipLaunchKernelGGL((insert_kernel<Tile, false>), dim3(grid_for(n, Tile::kWindows, kMaxBlocks)), dim3(kBlock),
                           0, stream, c, in, f, n, RecordSink{});
    } else {
        using Tile = DirectTile<kDirectR>;
        hipLaunchKernelGGL((insert_kernel<Tile, false>), dim3(grid_for(n, Tile::kWindows, kMaxBlocks)), dim3(kBlock),
                           0, stream, c, in, f, n, RecordSink{});
    }
    return hipGetLastError();
}

hipError_t generic_contains(const KernelConfig& c, const SequenceInput& in, FilterView f,
                            uint8_t* out, hipStream_t stream) {
    const uint64_t n = window_count(c, in);
    if (n == 0) return hipSuccess;
    (void)hipGetLastError();  // report only this launch
    if (hashed_tiles(c)) {
        using Tile = HashedTile<kHashedR>;
        hipLaunchKernelGGL(contains_kernel<Tile>, dim3(grid_for(n, Tile::kWindows, kMaxBlocks)), dim3(kBlock), 0,
                           stream, c, in, f, n, out);
    } else {
        using Tile = DirectTile<kDirectR>;
        hipLaunchKernelGGL(contains_kernel<Tile>, dim3(grid_for(n, Tile::kWindows, kMaxBlocks)), dim3(kBlock), 0,
                           stream, c, in, f, n, out);
    }
    return hipGetLastError();
}

hipError_t generic_count(const KernelConfig& c, const SequenceInput& in, uint64_t num_shards,
                         uint64_t* d_counts, hipStream_t stream) {
    const uint64_t n = window_count(c, in);
    if (n == 0) return hipSuccess;
    (void)hipGetLastError();  // report only this launch
    hipLaunchKernelGGL(count_kernel, dim3(grid_for(n, kBlock, kMaxBlocks)), dim3(kBlock), 0,
                       stream, c, in, num_shards, n,
                       reinterpret_cast<unsigned long long*>(d_counts));
    return hipGetLastError();
}

hipError_t pack_dense(const KernelConfig& c, const char* seq, uint64_t num_symbols,
                      uint64_t* words, hipStream_t stream) {
    if (num_symbols == 0) return hipSuccess;
    const uint64_t per_word = 64u / c.symbol_bits;
    const uint64_t num_words = (num_symbols + per_word - 1) / per_word;
    (void)hipGetLastError();  // report only this launch
    hipLaunchKernelGGL(pack_kernel, dim3(grid_for(num_words, kBlock, kMaxBlocks)), dim3(kBlock),
                       0, stream, c, seq, num_symbols, words);
    return hipGetLastError();
}

hipError_t popcount_words(const uint64_t* words, uint64_t n, uint64_t* d_out, hipStream_t stream) {
    (void)hipGetLastError();  // report only this launch
    hipLaunchKernelGGL(popcount_kernel, dim3(grid_for(n, kBlock * 8, 4096)), dim3(kBlock), 0,
                       stream, words, n, reinterpret_cast<unsigned long long*>(d_out));
    return hipGetLastError();
}

hipError_t words_checksum(const uint64_t* words, uint64_t n, uint64_t* d_out, hipStream_t stream) {
    hipError_t e = fill_u64(d_out, 0, 1, stream);
    if (e != hipSuccess || n == 0) return e;
    (void)hipGetLastError();  // report only this launch
    hipLaunchKernelGGL(checksum_kernel, dim3(grid_for(n, kBlock * 8, 4096)), dim3(kBlock), 0, stream,
                       words, n, reinterpret_cast<unsigned long long*>(d_out));
    return hipGetLastError();
}

hipError_t count_ones(const uint8_t* hits, uint64_t n, uint64_t* d_out, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    (void)hipGetLastError();  // report only this launch
    hipLaunchKernelGGL(count_ones_kernel, dim3(grid_for(n, kBlock * 64, 4096)), dim3(kBlock), 0,
                       stream, hits, n, reinterpret_cast<unsigned long long*>(d_out));
    return hipGetLastError();
}

hipError_t count_ones_per_record(const uint8_t* hits, const uint64_t* offsets,
                                 const uint64_t* counts, uint64_t num_records, uint64_t* d_out,
                                 hipStream_t stream) {
    if (num_records == 0) return hipSuccess;
    (void)hipGetLastError();  // report only this launch
    hipLaunchKernelGGL(count_ones_per_record_kernel,
                       dim3(grid_for(num_records, kBlock / 64, kMaxBlocks)), dim3(kBlock), 0,
                       stream, hits, offsets, counts, num_records, d_out);
    return hipGetLastError();
}

hipError_t batch_prepass(const SequenceInput& in, uint64_t* starts, uint32_t* state,
                         hipStream_t stream) {
    hipError_t e = fill_u64(reinterpret_cast<uint64_t*>(state), 0, 1, stream);
    if (e == hipSuccess) e = fill_u64(starts, 0, start_words(in.num_symbols), stream);
    if (e != hipSuccess) return e;
    (void)hipGetLastError();  // report only this launch
    hipLaunchKernelGGL(batch_prepass_kernel, dim3(grid_for(in.num_records + 1, kBlock, kMaxBlocks)),
                       dim3(kBlock), 0, stream, in.offsets, in.num_records, in.num_symbols,
                       reinterpret_cast<unsigned long long*>(starts), state);
    return hipGetLastError();
}

hipError_t request_counts(const KernelConfig& c, const char* seq, uint64_t num_symbols,
                          const uint8_t* hits, const uint64_t* req, uint64_t n_req,
                          uint64_t* counts, hipStream_t stream) {
    const uint64_t n = num_symbols < c.k ? 0 : num_symbols - c.k + 1;
    if (n == 0 || n_req == 0) return hipSuccess;
    (void)hipGetLastError();  // report only this launch
    hipLaunchKernelGGL(request_counts_kernel, dim3(grid_for((n + 15) / 16, kBlock, kMaxBlocks)),
                       dim3(kBlock), 0, stream, c, seq, num_symbols, hits, req, n_req,
                       reinterpret_cast<unsigned long long*>(counts));
    return hipGetLastError();
}

// bits[i / 64] bit i % 64 = hits[i] (each hit byte is 0 or 1): four hit bytes
// x = b0 | b1 << 8 | b2 << 16 | b3 << 24 give b0 | b1 << 1 | b2 << 2 | b3 << 3 as
// (x * 0x10204080) >> 28 in 32-bit arithmetic (b_i lands on bit 28 + i; every
// cross term lands on a distinct bit below 28 or above 31, so nothing carries).
__global__ __launch_bounds__(kBlock) void pack_hits_kernel(const uint8_t* __restrict__ hits, uint64_t n,
                                                           uint64_t* __restrict__ bits) {
    const uint64_t words = (n + 63) / 64;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    for (uint64_t w = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; w < words; w += stride) {
        const uint64_t p0 = 64 * w;
        uint64_t out = 0;
        if (p0 + 64 <= n && (reinterpret_cast<uintptr_t>(hits + p0) & 15u) == 0) {
            const u32x4* src = reinterpret_cast<const u32x4*>(hits + p0);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const u32x4 v = src[q];
                const uint32_t x[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    out |= static_cast<uint64_t>((x[c] * 0x10204080u) >> 28) << (16 * q + 4 * c);
                }
            }
        } else {
            const uint64_t m = n - p0 < 64 ? n - p0 : 64;
            for (uint64_t i = 0; i < m; ++i) out |= static_cast<uint64_t>(hits[p0 + i] & 1u) << i;
        }
        bits[w] = out;
    }
}

hipError_t pack_hits(const uint8_t* hits, uint64_t n, uint64_t* bits, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    (void)hipGetLastError();  // report only this launch
    hipLaunchKernelGGL(pack_hits_kernel, dim3(grid_for((n + 63) / 64, kBlock, kMaxBlocks)), dim3(kBlock), 0,
                       stream, hits, n, bits);
    return hipGetLastError();
}

hipError_t join_records(const char* in, uint64_t num_symbols, const uint64_t* offsets,
                        uint64_t num_records, uint32_t symbol_width, uint8_t separator, char* out,
                        hipStream_t stream) {
    if (num_records == 0) return hipSuccess;
    const uint64_t out_bytes = (num_symbols + num_records - 1) * symbol_width;
    if (out_bytes == 0) return hipSuccess;
    (void)hipGetLastError();  // report only this launch
    hipLaunchKernelGGL(join_records_kernel, dim3(grid_for((out_bytes + 15) / 16, kBlock, kMaxBlocks)),
                       dim3(kBlock), 0, stream, reinterpret_cast<const uint8_t*>(in),
                       num_symbols * symbol_width, offsets, num_records, symbol_width, separator,
                       reinterpret_cast<uint8_t*>(out), out_bytes);
    return hipGetLastError();
}

hipError_t fill_u64(uint64_t* dst, uint64_t value, uint64_t n, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    (void)hipGetLastError();  // report only this launch
    hipLaunchKernelGGL(fill_kernel, dim3(grid_for(n, kBlock * 4, 65536)), dim3(kBlock), 0, stream,
                       dst, value, n);
    return hipGetLastError();
}

}  // namespace sbf
