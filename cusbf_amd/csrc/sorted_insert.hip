// sorted_insert.hip -- the sorted form of the lane-strip insert (partition and apply).
//
// The atomic insert issues one 64-bit device-scope atomic OR per non-zero word
// of every shard run (filter_ref.cuh:94-108 issues them per warp run). On this
// GPU each such atomic is performed at the memory side as a read-modify-write
// of DRAM, which caps the insert at about 6.4 G runs/s whatever the kernel
// does. The sorted form produces the same filter words with ordinary,
// coalesced loads and stores:
//   1. emit (strip_kernels.hpp, insert_tile): each wave writes its tile's
//      slotted runs as records into its own region of a record buffer -- an
//      8-byte pair (shard << 32 | record index) and the run's 32-byte OR mask;
//      runs beyond the region's capacity still take the atomic path, and a
//      region that fills up raises the overflow flag;
//   2. partition (here): LSD radix passes, 8 bits each, over the pairs, keyed
//      on the slice index shard >> slice_bits, until the pairs are grouped by
//      slice. Only the grouping matters (the apply ORs a slice's records in any
//      order), so the first pass ranks with LDS atomics (unstable), and the
//      second runs its blocks over one first-pass digit group each: any order
//      inside a block keeps the groups apart, so it ranks with atomics too, and
//      its scanned histogram already holds every slice's start. A third pass
//      (more than 2^16 slices) needs a stable rank over the second-pass groups
//      and uses wave ballots;
//   3. apply (here): one block per non-empty slice loads the slice's words
//      into LDS, ORs the masks of its records into them with LDS atomics and
//      stores the slice back. Into a filter whose words are all zero (its
//      zero flag: set by a clear, reset by every insert kernel) it starts each
//      slice from zeros instead of reading it, unless the overflow flag is set.
// OR is commutative and idempotent, so the words equal the atomic insert's;
// the atomics of step 1 land before step 3 reads the slice (stream order).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "launch.hpp"

namespace sbf::sorted {

namespace {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr int kRankBlock = 256;                       // threads of histogram / scatter blocks
constexpr int kRankItems = 32;                        // pairs per thread (at most 32: bit masks)
constexpr int kBlockItems = kRankBlock * kRankItems;  // pairs per block
static_assert(kRankItems <= 32, "validity and rank bit masks are 32-bit");
constexpr int kBins = 256;                            // 8-bit digits
constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;
constexpr int kScanChunk = kScanBlock * kScanItems;
constexpr int kTopBlock = 1024;
constexpr int kApplyBlock = 1024;
constexpr uint32_t kApplyGrid = 1u << 16;

// Input of one pass: slot i of `pairs` holds a record when it lies below its
// region's count (the emit buffer has `regions` regions of `cap` slots; a
// compacted buffer is one region whose count is the record total).
struct PassInput {
    const uint64_t* pairs;
    uint32_t slots;
    uint32_t cap;
    const uint32_t* counts;
};

// Walks slots first, first + step, ... of a pass input, tracking the region
// and the position inside it without a division per slot. Near: the wave's
// slots span at most two regions (cap >= kWaveItems, or one region), whose
// counts are loaded once; otherwise each slot loads its region's count.
template <bool Near>
struct SlotCursor {
    uint32_t i, region, rem, first_region, cnt0, cnt1;
    __device__ __forceinline__ SlotCursor(const PassInput& in, uint32_t first)
        : i(first), region(first / in.cap), rem(first % in.cap), first_region(region), cnt0(0),
          cnt1(0) {
        if constexpr (Near) {
            const uint32_t regions = in.slots / in.cap;
            if (region < regions) cnt0 = in.counts[region];
            if (region + 1 < regions) cnt1 = in.counts[region + 1];
        }
    }
    __device__ __forceinline__ bool valid(const PassInput& in) const {
        if (i >= in.slots) return false;
        if constexpr (Near) {
            return rem < (region == first_region ? cnt0 : cnt1);
        } else {
            return rem < in.counts[region];
        }
    }
    __device__ __forceinline__ void advance(const PassInput& in, uint32_t step) {
        i += step;
        rem += step;
        while (rem >= in.cap) {
            rem -= in.cap;
            ++region;
        }
    }
};

__device__ __forceinline__ uint32_t digit_of(uint64_t pair, uint32_t shift, uint32_t mask) {
    return static_cast<uint32_t>(pair >> shift) & mask;
}

// Exclusive prefix of v over the block (kThreads threads, a multiple of 64)
// and the block total. `tmp` holds kThreads / 64 words; every thread must call.
template <int kThreads>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* tmp, uint32_t& total) {
    const int lane = static_cast<int>(threadIdx.x) & 63;
    const int wave = static_cast<int>(threadIdx.x) >> 6;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) tmp[wave] = x;
    __syncthreads();
    uint32_t before = 0;
    total = 0;
#pragma unroll
    for (int w = 0; w < kThreads / 64; ++w) {
        const uint32_t s = tmp[w];
        before += w < wave ? s : 0u;
        total += s;
    }
    __syncthreads();  // tmp may be reused by the caller
    return before + x - v;
}

// Block b covers slots [b * kBlockItems, (b + 1) * kBlockItems); wave w of
// the block takes the contiguous quarter w of it, 64 slots per round.
constexpr int kWaveItems = kBlockItems / (kRankBlock / 64);

__device__ __forceinline__ uint32_t wave_first_slot(uint32_t b) {
    return b * kBlockItems + (threadIdx.x >> 6) * kWaveItems + (threadIdx.x & 63);
}

// hist[d * nblocks + b] = records of block b with digit d.
template <bool Near>
__global__ __launch_bounds__(kRankBlock) void hist_kernel(PassInput in, uint32_t shift,
                                                          uint32_t digit_mask, uint32_t nblocks,
                                                          uint32_t* hist) {
    __shared__ uint32_t h[kBins];
    h[threadIdx.x] = 0;
    __syncthreads();
    // all loads first (kRankItems in flight per lane), then the LDS counts
    uint64_t p[kRankItems];
    uint32_t valid = 0;
    SlotCursor<Near> c(in, wave_first_slot(blockIdx.x));
#pragma unroll
    for (int j = 0; j < kRankItems; ++j) {
        const bool v = c.valid(in);
        p[j] = v ? in.pairs[c.i] : 0ull;
        valid |= (v ? 1u : 0u) << j;
        c.advance(in, 64);
    }
#pragma unroll
    for (int j = 0; j < kRankItems; ++j) {
        if ((valid >> j) & 1u) atomicAdd(&h[digit_of(p[j], shift, digit_mask)], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// partials[c] = sum of chunk c of v (kScanChunk values per chunk).
__global__ __launch_bounds__(kScanBlock) void scan_partials_kernel(const uint32_t* v, uint32_t n,
                                                                   uint32_t* partials) {
    __shared__ uint32_t tmp[kScanBlock / 64];
    const uint32_t first = blockIdx.x * kScanChunk + threadIdx.x * kScanItems;
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        if (first + j < n) s += v[first + j];
    }
    uint32_t total;
    (void)block_exclusive_scan<kScanBlock>(s, tmp, total);
    if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

// Exclusive scan of partials[0..n) in place (one block); *total = their sum.
__global__ __launch_bounds__(kTopBlock) void scan_top_kernel(uint32_t* partials, uint32_t n,
                                                             uint32_t* total) {
    __shared__ uint32_t tmp[kTopBlock / 64];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < n; base += kTopBlock) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t x = i < n ? partials[i] : 0u;
        uint32_t sum;
        const uint32_t pre = block_exclusive_scan<kTopBlock>(x, tmp, sum);
        if (i < n) partials[i] = carry + pre;
        carry += sum;
    }
    if (threadIdx.x == 0) *total = carry;
}

// Exclusive scan of v in place, chunk c offset by partials[c].
__global__ __launch_bounds__(kScanBlock) void scan_finish_kernel(uint32_t* v, uint32_t n,
                                                                 const uint32_t* partials) {
    __shared__ uint32_t tmp[kScanBlock / 64];
    const uint32_t first = blockIdx.x * kScanChunk + threadIdx.x * kScanItems;
    uint32_t x[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        x[j] = first + j < n ? v[first + j] : 0u;
        s += x[j];
    }
    uint32_t total;
    uint32_t run = partials[blockIdx.x] + block_exclusive_scan<kScanBlock>(s, tmp, total);
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        if (first + j < n) v[first + j] = run;
        run += x[j];
    }
}

// Stable scatter of one LSD pass. Each wave ranks its contiguous quarter of
// the block in slot order, 64 slots per round, with running per-digit counts
// of its own (no block barrier between rounds); the block then orders its
// records by (digit, wave, rank) -- slot order within each digit -- in LDS and
// writes them out in that order, so each digit's records of the block go out
// as one contiguous run.
template <bool Near>
__global__ __launch_bounds__(kRankBlock) void scatter_kernel(PassInput in, uint32_t shift,
                                                             uint32_t digit_bits, uint32_t nblocks,
                                                             const uint32_t* offsets, uint64_t* out) {
    constexpr int kWaves = kRankBlock / 64;
    __shared__ uint32_t wrun[kWaves][kBins];  // per-wave running counts, then per-wave bases
    __shared__ uint32_t start[kBins];
    __shared__ uint32_t dest[kBins];
    __shared__ uint32_t tmp[kWaves];
    __shared__ uint64_t sorted[kBlockItems];
    const int t = static_cast<int>(threadIdx.x);
    const int lane = t & 63;
    const int wave = t >> 6;
    const uint32_t mask = (1u << digit_bits) - 1u;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) wrun[w][t] = 0;
    dest[t] = offsets[static_cast<uint32_t>(t) * nblocks + blockIdx.x];
    __syncthreads();

    const uint64_t below = (1ull << lane) - 1ull;
    uint32_t* run = wrun[wave];
    uint64_t item[kRankItems];
    uint32_t rank[kRankItems];
    uint32_t valid = 0;
    SlotCursor<Near> c(in, wave_first_slot(blockIdx.x));
#pragma unroll
    for (int j = 0; j < kRankItems; ++j) {  // all loads first (kRankItems in flight per lane)
        const bool v = c.valid(in);
        item[j] = v ? in.pairs[c.i] : 0ull;
        valid |= (v ? 1u : 0u) << j;
        c.advance(in, 64);
    }
#pragma unroll
    for (int j = 0; j < kRankItems; ++j) {
        const bool v = ((valid >> j) & 1u) != 0u;
        const uint64_t p = item[j];
        const uint32_t d = digit_of(p, shift, mask);
        uint64_t peers = __ballot(v);
        for (uint32_t b = 0; b < digit_bits; ++b) {
            const bool bit = ((d >> b) & 1u) != 0u;
            const uint64_t set = __ballot(bit);
            peers &= bit ? set : ~set;
        }
        const uint32_t r = static_cast<uint32_t>(__popcll(peers & below));
        // LDS operations of one wave take effect in program order: every lane
        // reads the count before the lowest lane of its digit writes it back.
        const uint32_t pre = run[d];
        if (v && r == 0) run[d] = pre + static_cast<uint32_t>(__popcll(peers));
        rank[j] = v ? pre + r : 0xFFFFFFFFu;
    }
    __syncthreads();

    // digit t: the waves' counts -> block start of the digit and each wave's base
    uint32_t wsum[kWaves];
    uint32_t count = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
        wsum[w] = count;
        count += wrun[w][t];
    }
    uint32_t total;
    const uint32_t st = block_exclusive_scan<kRankBlock>(count, tmp, total);
    start[t] = st;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) wrun[w][t] = st + wsum[w];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kRankItems; ++j) {
        if (rank[j] != 0xFFFFFFFFu) sorted[run[digit_of(item[j], shift, mask)] + rank[j]] = item[j];
    }
    __syncthreads();
    for (uint32_t x = static_cast<uint32_t>(t); x < total; x += kRankBlock) {
        const uint64_t p = sorted[x];
        const uint32_t d = digit_of(p, shift, mask);
        out[dest[d] + (x - start[d])] = p;
    }
}

// Unstable scatter of one pass: each block ranks its items with LDS atomics
// on per-digit counters, orders them by digit in LDS and writes each digit's
// items as one contiguous run at the block's offset for that digit. Items of
// a block: Group = false, the pass-0 layout (wave w of block b takes slots
// b * kBlockItems + w * kWaveItems + lane + 64 j of the emit regions, valid
// below their region's count); Group = true, the block's chunk of its
// first-pass digit group (block_range).
__device__ __forceinline__ void block_range(const uint32_t* first_block, const uint32_t* group_start,
                                            uint32_t b, uint32_t& lo, uint32_t& hi) {
    // first_block[g]: blocks before group g (256 + 1 non-decreasing entries,
    // first_block[0] = 0); g with first_block[g] <= b < first_block[g + 1] is
    // the number of entries 1..256 at most b: each lane tests four of them
    // (independent loads) and ballots count the wave's.
    if (b >= first_block[kBins]) {  // past the last group's chunks: no items
        lo = hi = 0;
        return;
    }
    const uint32_t lane = threadIdx.x & 63u;
    bool le[kBins / 64];
#pragma unroll
    for (int r = 0; r < kBins / 64; ++r) le[r] = first_block[1u + lane + 64u * r] <= b;
    uint32_t g = 0;
#pragma unroll
    for (int r = 0; r < kBins / 64; ++r) g += static_cast<uint32_t>(__popcll(__ballot(le[r])));
    lo = group_start[g] + (b - first_block[g]) * static_cast<uint32_t>(kBlockItems);
    const uint32_t end = group_start[g + 1];
    hi = lo + static_cast<uint32_t>(kBlockItems) < end ? lo + static_cast<uint32_t>(kBlockItems) : end;
}

// Items of partition block b (every lane of the wave calls it together).
template <bool Near, bool Group>
__device__ __forceinline__ void load_block_items(const PassInput& in, const uint32_t* first_block,
                                                 const uint32_t* group_start, uint32_t b,
                                                 uint64_t (&item)[kRankItems], uint32_t& valid) {
    valid = 0;
    if constexpr (Group) {
        uint32_t lo, hi;
        block_range(first_block, group_start, b, lo, hi);
        const uint32_t first = lo + (threadIdx.x >> 6) * kWaveItems + (threadIdx.x & 63);
#pragma unroll
        for (int j = 0; j < kRankItems; ++j) {
            const uint32_t i = first + 64u * static_cast<uint32_t>(j);
            const bool v = i < hi;
            item[j] = v ? in.pairs[i] : 0ull;
            valid |= (v ? 1u : 0u) << j;
        }
    } else {
        SlotCursor<Near> c(in, wave_first_slot(b));
#pragma unroll
        for (int j = 0; j < kRankItems; ++j) {
            const bool v = c.valid(in);
            item[j] = v ? in.pairs[c.i] : 0ull;
            valid |= (v ? 1u : 0u) << j;
            c.advance(in, 64);
        }
    }
}

// hist[d * nblocks + b] = items of partition block b with digit d (all 256
// counts of every block b < nblocks, zeros when it has no items). One
// histogram block counts kHistBlocks consecutive partition blocks, so one
// digit's counts leave as kHistBlocks consecutive words (a digit-major layout
// written one word per block would cost a write request per count). Small
// inserts, whose partition blocks would give too few histogram blocks to
// fill the GPU, count one partition block per histogram block.
template <bool Near, bool Group, int kHistBlocks>
__global__ __launch_bounds__(kRankBlock) void hist_any_kernel(PassInput in, const uint32_t* first_block,
                                                              const uint32_t* group_start, uint32_t shift,
                                                              uint32_t digit_mask, uint32_t nblocks,
                                                              uint32_t* hist) {
    __shared__ uint32_t h[kHistBlocks][kBins];
#pragma unroll
    for (int s = 0; s < kHistBlocks; ++s) h[s][threadIdx.x] = 0;
    __syncthreads();
    const uint32_t b0 = blockIdx.x * kHistBlocks;
    for (int s = 0; s < kHistBlocks; ++s) {
        const uint32_t b = b0 + static_cast<uint32_t>(s);
        if (b >= nblocks) break;  // uniform
        uint64_t p[kRankItems];
        uint32_t valid;
        load_block_items<Near, Group>(in, first_block, group_start, b, p, valid);
#pragma unroll
        for (int j = 0; j < kRankItems; ++j) {
            if ((valid >> j) & 1u) atomicAdd(&h[s][digit_of(p[j], shift, digit_mask)], 1u);
        }
    }
    __syncthreads();
    // round r, thread t: word w = r * 256 + t holds digit w / kHistBlocks of
    // block b0 + w % kHistBlocks
#pragma unroll
    for (int r = 0; r < kHistBlocks; ++r) {
        const uint32_t w = static_cast<uint32_t>(r) * kRankBlock + threadIdx.x;
        const uint32_t d = w / kHistBlocks;
        const uint32_t s = w % kHistBlocks;
        if (b0 + s < nblocks) hist[d * nblocks + b0 + s] = h[s][d];
    }
}

template <bool Near, bool Group>
__global__ __launch_bounds__(kRankBlock) void scatter_any_kernel(PassInput in, const uint32_t* first_block,
                                                                 const uint32_t* group_start, uint32_t shift,
                                                                 uint32_t digit_mask, uint32_t nblocks,
                                                                 const uint32_t* offsets, uint64_t* out) {
    __shared__ uint32_t count[kBins];
    __shared__ uint32_t start[kBins];
    __shared__ uint32_t dest[kBins];
    __shared__ uint32_t tmp[kRankBlock / 64];
    __shared__ uint64_t sorted[kBlockItems];
    const int t = static_cast<int>(threadIdx.x);
    uint64_t item[kRankItems];
    uint32_t valid;
    load_block_items<Near, Group>(in, first_block, group_start, blockIdx.x, item, valid);  // in flight over the set-up
    count[t] = 0;
    dest[t] = offsets[static_cast<uint32_t>(t) * nblocks + blockIdx.x];
    __syncthreads();
    uint32_t rank[kRankItems];
#pragma unroll
    for (int j = 0; j < kRankItems; ++j) {
        rank[j] = ((valid >> j) & 1u) ? atomicAdd(&count[digit_of(item[j], shift, digit_mask)], 1u) : 0u;
    }
    __syncthreads();
    uint32_t total;
    const uint32_t st = block_exclusive_scan<kRankBlock>(count[t], tmp, total);
    start[t] = st;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kRankItems; ++j) {
        if ((valid >> j) & 1u) sorted[start[digit_of(item[j], shift, digit_mask)] + rank[j]] = item[j];
    }
    __syncthreads();
    for (uint32_t x = static_cast<uint32_t>(t); x < total; x += kRankBlock) {
        const uint64_t p = sorted[x];
        const uint32_t d = digit_of(p, shift, digit_mask);
        out[dest[d] + (x - start[d])] = p;
    }
}

// Second-pass block table from the first pass's scanned histogram: group g =
// the items of first-pass digit g, at [group_start[g], group_start[g + 1]);
// first_block[g] = blocks of kBlockItems before group g. One block of 256.
__global__ __launch_bounds__(kBins) void groups_kernel(const uint32_t* scanned, uint32_t nblocks,
                                                       const uint32_t* total, uint32_t* group_start,
                                                       uint32_t* first_block) {
    __shared__ uint32_t tmp[kBins / 64];
    const uint32_t g = threadIdx.x;
    const uint32_t lo = scanned[g * nblocks];
    const uint32_t hi = g + 1 < kBins ? scanned[(g + 1) * nblocks] : *total;
    group_start[g] = lo;
    if (g + 1 == kBins) group_start[kBins] = hi;
    const uint32_t nb = (hi - lo + static_cast<uint32_t>(kBlockItems) - 1) / static_cast<uint32_t>(kBlockItems);
    uint32_t sum;
    const uint32_t before = block_exclusive_scan<kBins>(nb, tmp, sum);
    first_block[g] = before;
    if (g + 1 == kBins) first_block[kBins] = sum;
}

// Slice starts read off a scanned histogram. One pass (slice index < 2^8
// values): bounds[s] = scanned[s * nblocks] (first block's offset of digit s).
// Two passes (slice = hi << 8 | lo, second-pass blocks grouped by lo):
// bounds[s] = scanned[hi * nblocks + first_block[lo]], the offset of the
// first block of group lo for digit hi -- exclusive scans count every item of
// smaller hi, and of equal hi in earlier groups. bounds[num_slices] = total.
__global__ __launch_bounds__(256) void bounds_from_scan_kernel(const uint32_t* scanned, uint32_t nblocks,
                                                               const uint32_t* first_block,
                                                               const uint32_t* total, uint32_t num_slices,
                                                               uint32_t* bounds) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s > num_slices) return;
    if (s == num_slices) {
        bounds[s] = *total;
    } else {
        // the first block of a group past every item is block nblocks of digit hi,
        // i.e. the start of digit hi + 1 (the total for hi = 255)
        const uint64_t i = first_block == nullptr
                               ? static_cast<uint64_t>(s) * nblocks
                               : static_cast<uint64_t>(s >> 8) * nblocks + first_block[s & (kBins - 1)];
        bounds[s] = i < static_cast<uint64_t>(kBins) * nblocks ? scanned[i] : *total;
    }
}

// bounds[b] = first record of slice b in the grouped pairs, bounds[num_slices]
// = record total. Record i writes the starts of the slices after its
// predecessor's slice up to its own; slot `total` closes the table.
__global__ __launch_bounds__(256) void bounds_kernel(const uint64_t* pairs, const uint32_t* total,
                                                     uint32_t shift, uint32_t num_slices,
                                                     uint32_t* bounds) {
    const uint32_t n = *total;
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i > n) return;
    uint32_t cur = num_slices;
    if (i < n) {
        const uint64_t s = pairs[i] >> shift;
        cur = s < num_slices ? static_cast<uint32_t>(s) : num_slices;
    }
    uint32_t from = 0;
    if (i > 0) {
        const uint64_t s = pairs[i - 1] >> shift;
        from = (s < num_slices ? static_cast<uint32_t>(s) : num_slices) + 1;
    }
    for (uint32_t b = from; b <= cur; ++b) bounds[b] = static_cast<uint32_t>(i);
}

// Zeroes n words of counts (the region counts and the overflow flag) and
// moves the filter's zero flag into counts[n] (the flag becomes 0).
__global__ __launch_bounds__(256) void begin_kernel(uint32_t* counts, uint32_t n, uint64_t* zero) {
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) counts[i] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const bool was_zero = zero != nullptr && *zero != 0;
        counts[n] = was_zero ? 1u : 0u;
        if (zero != nullptr) *zero = 0;
    }
}

// One block per non-empty slice at a time (blocks stride over the slices):
// slice words -> LDS, OR of the slice's records, LDS -> slice words.
// flags[0]: the overflow flag; flags[1]: the filter words were all zero
// before this insert. A slice is read unless they were and no region
// overflowed into atomics.
__global__ __launch_bounds__(kApplyBlock) void apply_kernel(const uint64_t* pairs,
                                                            const uint32_t* bounds,
                                                            const u32x4* masks, uint64_t* words,
                                                            uint32_t slice_bits,
                                                            uint32_t num_slices,
                                                            const uint32_t* flags) {
    extern __shared__ uint64_t slice[];  // 4 << slice_bits words
    u32x4* lds4 = reinterpret_cast<u32x4*>(slice);
    const uint32_t vecs = 2u << slice_bits;  // 16-byte vectors of one slice
    const uint32_t local_mask = (1u << slice_bits) - 1u;
    const bool load = flags[1] == 0u || flags[0] != 0u;  // uniform
    constexpr int kUnroll = 4;
    for (uint32_t b = blockIdx.x; b < num_slices; b += gridDim.x) {
        const uint32_t lo = bounds[b];
        const uint32_t hi = bounds[b + 1];
        if (lo >= hi) continue;  // uniform: every thread reads the same bounds
        u32x4* g = reinterpret_cast<u32x4*>(words + (static_cast<uint64_t>(b) << slice_bits) * 4ull);
        if (load) {
            for (uint32_t x = threadIdx.x; x < vecs; x += kApplyBlock) lds4[x] = g[x];
        } else {
            for (uint32_t x = threadIdx.x; x < vecs; x += kApplyBlock) lds4[x] = u32x4{0u, 0u, 0u, 0u};
        }
        __syncthreads();
        // threads 2r and 2r+1 take the two 16-byte halves of record r's mask in
        // one instruction: one memory request per mask line
        const uint32_t half = threadIdx.x & 1u;
        constexpr uint32_t kPerRound = kApplyBlock / 2;
        for (uint32_t i0 = lo; i0 < hi; i0 += kUnroll * kPerRound) {
            uint64_t p[kUnroll];
            u32x4 a[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) {
                const uint32_t i = i0 + static_cast<uint32_t>(u) * kPerRound + (threadIdx.x >> 1);
                p[u] = i < hi ? pairs[i] : ~0ull;
            }
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) {
                if (p[u] != ~0ull) a[u] = masks[2 * static_cast<uint64_t>(static_cast<uint32_t>(p[u])) + half];
            }
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) {
                if (p[u] != ~0ull) {
                    uint64_t* s = slice + 4u * (static_cast<uint32_t>(p[u] >> 32) & local_mask) + 2u * half;
                    const uint64_t m0 = a[u].x | (static_cast<uint64_t>(a[u].y) << 32);
                    const uint64_t m1 = a[u].z | (static_cast<uint64_t>(a[u].w) << 32);
                    if (m0 != 0) __hip_atomic_fetch_or(s, m0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    if (m1 != 0) __hip_atomic_fetch_or(s + 1, m1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                }
            }
        }
        __syncthreads();
        for (uint32_t x = threadIdx.x; x < vecs; x += kApplyBlock) g[x] = lds4[x];
        __syncthreads();  // the next slice reuses the LDS
    }
}

uint32_t div_up(uint64_t a, uint64_t b) { return static_cast<uint32_t>((a + b - 1) / b); }

uint64_t align256(uint64_t x) { return (x + 255) & ~255ull; }

uint32_t log2_u64(uint64_t x) {
    uint32_t r = 0;
    while ((1ull << r) < x) ++r;
    return r;
}

}  // namespace

SortedLayout sorted_layout(uint32_t regions, uint32_t cap, uint64_t num_shards, uint32_t slice_bits) {
    SortedLayout l{};
    l.regions = regions;
    l.cap = cap;
    l.slots = static_cast<uint64_t>(regions) * cap;
    const uint32_t shard_bits = log2_u64(num_shards);
    l.slice_bits = slice_bits < shard_bits ? slice_bits : shard_bits;
    l.slice_index_bits = shard_bits - l.slice_bits;
    l.num_slices = 1u << l.slice_index_bits;
    l.passes = l.slice_index_bits == 0 ? 1u : (l.slice_index_bits + 7) / 8;
    const uint32_t nblocks = div_up(l.slots, kBlockItems);
    // second-pass blocks: a group adds at most one partial block
    l.hist_len = static_cast<uint64_t>(kBins) * (nblocks + kBins);
    l.partials_len = div_up(l.hist_len, kScanChunk);
    uint64_t at = 0;
    l.off_pairs[0] = at;
    at += align256(l.slots * 8);
    l.off_pairs[1] = at;
    at += align256(l.slots * 8);
    l.off_masks = at;
    at += align256(l.slots * 32);
    l.off_counts = at;  // counts[regions], then the overflow flag (and one spare word)
    at += align256((static_cast<uint64_t>(regions) + 2) * 4);
    l.off_hist = at;
    at += align256(l.hist_len * 4);
    l.off_partials = at;
    at += align256(static_cast<uint64_t>(l.partials_len) * 4);
    l.off_totals = at;
    at += 256;
    l.off_bounds = at;
    at += align256((static_cast<uint64_t>(l.num_slices) + 1) * 4);
    l.off_groups = at;  // group_start[257], first_block[257]
    at += align256(2 * (kBins + 1) * 4);
    l.bytes = at;
    return l;
}

RecordSink record_sink(const SortedLayout& l, void* base) {
    char* p = static_cast<char*>(base);
    RecordSink s;
    s.pairs = reinterpret_cast<uint64_t*>(p + l.off_pairs[0]);
    s.masks = reinterpret_cast<uint64_t*>(p + l.off_masks);
    s.counts = reinterpret_cast<uint32_t*>(p + l.off_counts);
    s.overflow = s.counts + l.regions;
    s.cap = l.cap;
    s.regions = l.regions;
    return s;
}

hipError_t prepare() {
    // the apply kernel's LDS slice: up to 2^12 shards of 32 bytes
    return hipFuncSetAttribute(reinterpret_cast<const void*>(&apply_kernel),
                               hipFuncAttributeMaxDynamicSharedMemorySize, 32 << 12);
}

hipError_t sorted_begin(const SortedLayout& l, void* base, FilterView f, hipStream_t stream) {
    uint32_t* counts = reinterpret_cast<uint32_t*>(static_cast<char*>(base) + l.off_counts);
    (void)hipGetLastError();
    hipLaunchKernelGGL(begin_kernel, dim3(div_up(l.regions + 1ull, 256 * 8)), dim3(256), 0, stream,
                       counts, l.regions + 1, f.zero);
    return hipGetLastError();
}

hipError_t sorted_apply(const SortedLayout& l, void* base, FilterView f, hipStream_t stream) {
    char* p = static_cast<char*>(base);
    uint64_t* pairs[2] = {reinterpret_cast<uint64_t*>(p + l.off_pairs[0]),
                          reinterpret_cast<uint64_t*>(p + l.off_pairs[1])};
    const u32x4* masks = reinterpret_cast<const u32x4*>(p + l.off_masks);
    uint32_t* hist = reinterpret_cast<uint32_t*>(p + l.off_hist);
    uint32_t* partials = reinterpret_cast<uint32_t*>(p + l.off_partials);
    uint32_t* totals = reinterpret_cast<uint32_t*>(p + l.off_totals);
    uint32_t* bounds = reinterpret_cast<uint32_t*>(p + l.off_bounds);
    uint32_t* group_start = reinterpret_cast<uint32_t*>(p + l.off_groups);
    uint32_t* first_block = group_start + (kBins + 1);

    (void)hipGetLastError();  // report only these launches
    const uint32_t nblocks0 = div_up(l.slots, kBlockItems);
    const uint32_t nblocks1 = nblocks0 + kBins;
    auto scan = [&](uint32_t len, uint32_t* total) {
        const uint32_t parts = div_up(len, kScanChunk);
        hipLaunchKernelGGL(scan_partials_kernel, dim3(parts), dim3(kScanBlock), 0, stream, hist, len, partials);
        hipLaunchKernelGGL(scan_top_kernel, dim3(1), dim3(kTopBlock), 0, stream, partials, parts, total);
        hipLaunchKernelGGL(scan_finish_kernel, dim3(parts), dim3(kScanBlock), 0, stream, hist, len, partials);
    };
    auto digit = [&](uint32_t pass, uint32_t& shift, uint32_t& mask) {
        const uint32_t done = 8 * pass;
        const uint32_t bits = l.slice_index_bits > done
                                  ? (l.slice_index_bits - done < 8 ? l.slice_index_bits - done : 8u)
                                  : 0u;
        shift = 32 + l.slice_bits + done;
        mask = (1u << bits) - 1u;
        return bits;
    };

    // histogram blocks of eight partition blocks when that still gives 1024 or more
    auto histogram = [&](auto kernel8, auto kernel1, const PassInput& pin, const uint32_t* fb,
                         const uint32_t* gs, uint32_t sh, uint32_t mk, uint32_t nb) {
        if (nb >= 8u * 1024u) {
            hipLaunchKernelGGL(kernel8, dim3(div_up(nb, 8)), dim3(kRankBlock), 0, stream, pin, fb, gs, sh, mk, nb, hist);
        } else {
            hipLaunchKernelGGL(kernel1, dim3(nb), dim3(kRankBlock), 0, stream, pin, fb, gs, sh, mk, nb, hist);
        }
    };

    // pass 0: the emit regions, unstable
    uint32_t shift, mask;
    uint32_t bits = digit(0, shift, mask);
    PassInput in{pairs[0], static_cast<uint32_t>(l.slots), l.cap,
                 reinterpret_cast<const uint32_t*>(p + l.off_counts)};
    const bool near = in.cap >= static_cast<uint32_t>(kWaveItems) || in.cap == in.slots;
    if (near) {
        histogram(hist_any_kernel<true, false, 8>, hist_any_kernel<true, false, 1>, in, nullptr, nullptr, shift,
                  mask, nblocks0);
    } else {
        histogram(hist_any_kernel<false, false, 8>, hist_any_kernel<false, false, 1>, in, nullptr, nullptr,
                  shift, mask, nblocks0);
    }
    uint32_t* total = totals;
    scan(kBins * nblocks0, total);
    if (near) {
        hipLaunchKernelGGL((scatter_any_kernel<true, false>), dim3(nblocks0), dim3(kRankBlock), 0, stream, in,
                           nullptr, nullptr, shift, mask, nblocks0, hist, pairs[1]);
    } else {
        hipLaunchKernelGGL((scatter_any_kernel<false, false>), dim3(nblocks0), dim3(kRankBlock), 0, stream, in,
                           nullptr, nullptr, shift, mask, nblocks0, hist, pairs[1]);
    }
    int cur = 1;
    const uint32_t bgrid = div_up(static_cast<uint64_t>(l.num_slices) + 1, 256);
    if (l.passes == 1) {
        hipLaunchKernelGGL(bounds_from_scan_kernel, dim3(bgrid), dim3(256), 0, stream, hist, nblocks0,
                           nullptr, total, l.num_slices, bounds);
    } else {
        // pass 1: blocks over the pass-0 digit groups, unstable within a group
        hipLaunchKernelGGL(groups_kernel, dim3(1), dim3(kBins), 0, stream, hist, nblocks0, total,
                           group_start, first_block);
        bits = digit(1, shift, mask);
        in = PassInput{pairs[1], static_cast<uint32_t>(l.slots), static_cast<uint32_t>(l.slots), total};
        histogram(hist_any_kernel<true, true, 8>, hist_any_kernel<true, true, 1>, in, first_block, group_start,
                  shift, mask, nblocks1);
        scan(kBins * nblocks1, totals + 1);
        hipLaunchKernelGGL((scatter_any_kernel<true, true>), dim3(nblocks1), dim3(kRankBlock), 0, stream, in,
                           first_block, group_start, shift, mask, nblocks1, hist, pairs[0]);
        cur = 0;
        total = totals + 1;
        if (l.passes == 2) {
            hipLaunchKernelGGL(bounds_from_scan_kernel, dim3(bgrid), dim3(256), 0, stream, hist, nblocks1,
                               first_block, total, l.num_slices, bounds);
        } else {
            // pass 2: stable (wave-ballot ranks) over the compacted pass-1 output
            bits = digit(2, shift, mask);
            in = PassInput{pairs[0], static_cast<uint32_t>(l.slots), static_cast<uint32_t>(l.slots), total};
            hipLaunchKernelGGL(hist_kernel<true>, dim3(nblocks0), dim3(kRankBlock), 0, stream, in, shift,
                               mask, nblocks0, hist);
            scan(kBins * nblocks0, totals);
            hipLaunchKernelGGL(scatter_kernel<true>, dim3(nblocks0), dim3(kRankBlock), 0, stream, in, shift,
                               bits, nblocks0, hist, pairs[1]);
            cur = 1;
            total = totals;
            hipLaunchKernelGGL(bounds_kernel, dim3(div_up(l.slots + 1, 256)), dim3(256), 0, stream,
                               pairs[cur], total, 32 + l.slice_bits, l.num_slices, bounds);
        }
    }
    (void)bits;
    const uint32_t grid = l.num_slices < kApplyGrid ? l.num_slices : kApplyGrid;
    hipLaunchKernelGGL(apply_kernel, dim3(grid), dim3(kApplyBlock), (32u << l.slice_bits), stream,
                       pairs[cur], bounds, masks, f.words, l.slice_bits, l.num_slices,
                       reinterpret_cast<const uint32_t*>(p + l.off_counts) + l.regions);
    return hipGetLastError();
}

}  // namespace sbf::sorted
