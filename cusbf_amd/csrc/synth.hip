// synth.hip -- seeded synthetic genomes and reads on the GPU (libcusbf_synth.so).
//
// Workload generator for the benchmarks and large parity tests; not part of
// the filter's boundary. Every base is a pure function of (seed, index), so
// any slice can be regenerated on the host for checking. Chromosomes are laid
// out as one record batch with a single 'N' separator between chromosomes;
// reads are 150 bp records (or any fixed length) separated by 'N'.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__host__ __device__ __forceinline__ uint64_t rnd(uint64_t seed, uint64_t stream, uint64_t i) {
    return mix64(seed ^ mix64(stream * 0x9E3779B97F4A7C15ull + i + 0x632BE59BD9B4E019ull));
}

// Draw a base with G+C fraction gc_ppm / 1e6 from 32 random bits.
__host__ __device__ __forceinline__ uint8_t draw_base(uint32_t r, uint32_t gc_ppm) {
    const uint64_t x = (static_cast<uint64_t>(r) * 1000000ull) >> 32;  // uniform in [0, 1e6)
    const uint32_t half_gc = gc_ppm / 2;
    if (x < half_gc) return 'G';
    if (x < gc_ppm) return 'C';
    if (x < gc_ppm + (1000000u - gc_ppm) / 2) return 'A';
    return 'T';
}

__host__ __device__ __forceinline__ uint8_t substitute(uint8_t b, uint64_t r) {
    const char* bases = "ACGT";
    int idx = b == 'A' ? 0 : b == 'C' ? 1 : b == 'G' ? 2 : 3;
    return static_cast<uint8_t>(bases[(idx + 1 + static_cast<int>(r % 3)) & 3]);
}

struct RepeatCopy {
    uint64_t dst;        // genome coordinate of the first copied base
    uint32_t len;
    uint32_t family;
    uint32_t offset;     // offset into the family consensus
    uint32_t period;     // > 0: tandem repeat of a `period`-base unit
    uint32_t div_ppm;    // substitution rate of this copy
    uint32_t pad;
};

struct GenomeParams {
    uint64_t seed;
    uint64_t buffer_len;       // bytes incl. separators
    const uint64_t* chrom_start;  // buffer offsets of chromosome starts (sorted)
    uint32_t num_chrom;
    uint32_t gc_ppm;
    const RepeatCopy* repeats;    // sorted by dst, non-overlapping
    uint64_t num_repeats;
};

__device__ uint8_t genome_base(const GenomeParams& p, uint64_t g, uint64_t& rep_hint) {
    // locate the last repeat copy starting at or before g (copies sorted, disjoint)
    const uint64_t n = p.num_repeats;
    uint64_t lo = rep_hint;
    const bool hint_ok = lo < n && p.repeats[lo].dst <= g && (lo + 1 == n || p.repeats[lo + 1].dst > g);
    if (!hint_ok) {
        uint64_t a = 0, b = n;
        while (a < b) {  // first copy with dst > g
            const uint64_t mid = (a + b) / 2;
            if (p.repeats[mid].dst <= g) a = mid + 1; else b = mid;
        }
        lo = a == 0 ? n : a - 1;
    }
    rep_hint = lo;
    if (lo < p.num_repeats) {
        const RepeatCopy& c = p.repeats[lo];
        if (g < c.dst + c.len) {
            uint64_t pos = g - c.dst;
            if (c.period) pos %= c.period;
            pos += c.offset;
            uint8_t b = draw_base(static_cast<uint32_t>(rnd(p.seed, 1000003ull + c.family, pos)), p.gc_ppm);
            const uint64_t r = rnd(p.seed, 7, g);
            if ((r >> 32) % 1000000u < c.div_ppm) b = substitute(b, r);
            return b;
        }
    }
    return draw_base(static_cast<uint32_t>(rnd(p.seed, 1, g)), p.gc_ppm);
}

__global__ void genome_kernel(GenomeParams p, uint8_t* out) {
    constexpr uint64_t kChunk = 64;
    const uint64_t first = (static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) * kChunk;
    if (first >= p.buffer_len) return;
    // chromosome of `first`
    uint32_t c = 0;
    while (c + 1 < p.num_chrom && p.chrom_start[c + 1] <= first) ++c;
    uint64_t hint = ~0ull;
    for (uint64_t b = first; b < first + kChunk && b < p.buffer_len; ++b) {
        while (c + 1 < p.num_chrom && p.chrom_start[c + 1] <= b) ++c;
        if (c + 1 < p.num_chrom && b == p.chrom_start[c + 1] - 1) {
            out[b] = 'N';  // separator before the next chromosome
            continue;
        }
        const uint64_t g = b - c;  // genome coordinate: one separator per earlier chromosome
        out[b] = genome_base(p, g, hint);
    }
}

struct ReadParams {
    uint64_t seed;
    const uint8_t* genome;
    uint64_t genome_len;
    uint64_t num_reads;
    uint64_t genome_reads;   // reads [0, genome_reads) are sampled from the genome
    uint32_t read_len;
    uint32_t sub_ppm;
    uint32_t gc_ppm;
};

// Read r occupies bytes [r*(L+1), r*(L+1)+L) followed by an 'N' (not after the last).
__global__ void reads_kernel(ReadParams p, uint8_t* out) {
    const uint64_t stride = static_cast<uint64_t>(p.read_len) + 1;
    const uint64_t total = p.num_reads * stride - 1;
    const uint64_t first = (static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) * 16;
    for (uint64_t b = first; b < first + 16 && b < total; ++b) {
        const uint64_t r = b / stride;
        const uint64_t j = b - r * stride;
        if (j == p.read_len) {
            out[b] = 'N';
            continue;
        }
        uint8_t base;
        if (r < p.genome_reads) {
            const uint64_t span = p.genome_len - p.read_len;
            const uint64_t start = rnd(p.seed, 21, r) % span;
            base = p.genome[start + j];
            const uint64_t x = rnd(p.seed, 22, b);
            if (base != 'N' && (x >> 32) % 1000000u < p.sub_ppm) base = substitute(base, x);
        } else {
            base = draw_base(static_cast<uint32_t>(rnd(p.seed, 23, b)), p.gc_ppm);
        }
        out[b] = base;
    }
}

// Long reads: read r is written at out[offsets[r] .. offsets[r] + len_r) with
// len_r = offsets[r + 1] - offsets[r] - 1 (an 'N' follows every read but the
// last, whose length is offsets[r + 1] - offsets[r]). Each output base is,
// with probability err/3 each, an inserted random base, a base after one
// deleted genome base, or a substituted genome base; otherwise a copy.
__global__ void long_reads_kernel(uint64_t seed, const uint8_t* genome, uint64_t genome_len,
                                  uint64_t num_reads, const uint64_t* offsets, uint32_t err_ppm,
                                  uint8_t* out) {
    const uint64_t r = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= num_reads) return;
    const bool last = r + 1 == num_reads;
    const uint64_t len = offsets[r + 1] - offsets[r] - (last ? 0 : 1);
    uint8_t* dst = out + offsets[r];
    uint64_t pos = rnd(seed, 30, r) % (genome_len - 2 * len);
    for (uint64_t i = 0; i < len; ++i) {
        const uint64_t x = rnd(seed, 31, (r << 20) + i);
        const uint32_t u = static_cast<uint32_t>((x >> 32) % 1000000u);
        uint8_t b;
        if (u < err_ppm / 3) {
            b = "ACGT"[x & 3];  // insertion
        } else if (u < 2 * (err_ppm / 3)) {
            pos += 1;  // deletion of one genome base
            b = genome[pos++];
        } else if (u < err_ppm) {
            b = genome[pos++];
            if (b != 'N') b = substitute(b, x);
        } else {
            b = genome[pos++];
        }
        dst[i] = b;
    }
    if (!last) dst[len] = 'N';
}

// Uniform letters: byte i is letters[(32 random bits of (seed, i)) * n >> 32].
struct Letters {
    uint8_t c[64];
};

__global__ void letters_kernel(uint64_t seed, Letters letters, uint32_t n_letters, uint8_t* out,
                               uint64_t n) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint32_t r = static_cast<uint32_t>(rnd(seed, 11, i));
        out[i] = letters.c[(static_cast<uint64_t>(r) * n_letters) >> 32];
    }
}

}  // namespace

extern "C" {

// n bytes drawn uniformly from `letters` (host string of 1..64 letters).
int cusbf_synth_letters(uint64_t seed, const char* letters, uint32_t n_letters, uint8_t* d_out,
                        uint64_t n, void* stream) {
    if (n_letters == 0 || n_letters > 64) return static_cast<int>(hipErrorInvalidValue);
    if (n == 0) return 0;
    Letters l{};
    for (uint32_t i = 0; i < n_letters; ++i) l.c[i] = static_cast<uint8_t>(letters[i]);
    uint64_t blocks = (n + 256 * 16 - 1) / (256 * 16);
    if (blocks > 65536) blocks = 65536;
    (void)hipGetLastError();  // report only this launch
    hipLaunchKernelGGL(letters_kernel, dim3(static_cast<uint32_t>(blocks)), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), seed, l, n_letters, d_out, n);
    return static_cast<int>(hipGetLastError());
}

int cusbf_synth_long_reads(uint64_t seed, const uint8_t* d_genome, uint64_t genome_len,
                           uint64_t num_reads, const uint64_t* d_offsets, uint32_t err_ppm,
                           uint8_t* d_out, void* stream) {
    if (num_reads == 0) return 0;
    (void)hipGetLastError();  // report only this launch
    hipLaunchKernelGGL(long_reads_kernel, dim3(static_cast<uint32_t>((num_reads + 63) / 64)), dim3(64),
                       0, reinterpret_cast<hipStream_t>(stream), seed, d_genome, genome_len, num_reads,
                       d_offsets, err_ppm, d_out);
    return static_cast<int>(hipGetLastError());
}

int cusbf_synth_genome(uint64_t seed, uint32_t gc_ppm, const uint64_t* d_chrom_start,
                       uint32_t num_chrom, const void* d_repeats, uint64_t num_repeats,
                       uint8_t* d_out, uint64_t buffer_len, void* stream) {
    GenomeParams p{seed, buffer_len, d_chrom_start, num_chrom, gc_ppm,
                   reinterpret_cast<const RepeatCopy*>(d_repeats), num_repeats};
    const uint64_t threads = (buffer_len + 63) / 64;
    const uint64_t blocks = (threads + 255) / 256;
    (void)hipGetLastError();  // report only this launch
    hipLaunchKernelGGL(genome_kernel, dim3(static_cast<uint32_t>(blocks)), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), p, d_out);
    return static_cast<int>(hipGetLastError());
}

int cusbf_synth_reads(uint64_t seed, const uint8_t* d_genome, uint64_t genome_len,
                      uint64_t num_reads, uint64_t genome_reads, uint32_t read_len,
                      uint32_t sub_ppm, uint32_t gc_ppm, uint8_t* d_out, void* stream) {
    if (num_reads == 0) return 0;
    if (genome_len <= read_len && genome_reads) return 2;
    ReadParams p{seed, d_genome, genome_len, num_reads, genome_reads, read_len, sub_ppm, gc_ppm};
    const uint64_t total = num_reads * (static_cast<uint64_t>(read_len) + 1) - 1;
    const uint64_t threads = (total + 15) / 16;
    const uint64_t blocks = (threads + 255) / 256;
    (void)hipGetLastError();  // report only this launch
    hipLaunchKernelGGL(reads_kernel, dim3(static_cast<uint32_t>(blocks)), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), p, d_out);
    return static_cast<int>(hipGetLastError());
}

uint64_t cusbf_synth_repeat_copy_size(void) { return sizeof(RepeatCopy); }

}  // extern "C"
