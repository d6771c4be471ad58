// kernels_strip.hip -- dispatch table of the lane-strip kernels (strip_kernels.hpp):
// the plain kernels are instantiated here, the batched ones in
// kernels_strip_batched.hip (a second translation unit, compiled in parallel).
#include "strip_kernels.hpp"

namespace sbf::strip {

#define SBF_STRIP_EXTERN(FE, K, S, M, H, W)                                                    \
    extern template hipError_t launch_batched_insert<K, S, M, H, FE, W>(                       \
        const KernelConfig&, const SequenceInput&, FilterView, const BatchLaunch&,             \
        const RecordSink*, hipStream_t);                                                       \
    extern template hipError_t launch_batched_contains<K, S, M, H, FE, W>(                     \
        const KernelConfig&, const SequenceInput&, FilterView, const BatchLaunch&, uint8_t*,   \
        hipStream_t);
SBF_STRIP_BATCHED_INSTANCES(SBF_STRIP_EXTERN)
#undef SBF_STRIP_EXTERN

// A configuration with plain kernels only.
#define SBF_STRIP_PLAIN(FE, K, S, M, H)                                                       \
    StripKernelSet {                                                                          \
        K, S, M, H, FE::kAlphabet, FE::kBits == 2, wave_windows<K, S, M, H, FE>(),             \
            &launch_insert<K, S, M, H, FE>, &launch_contains<K, S, M, H, FE>,                  \
            {nullptr, nullptr}, {nullptr, nullptr}                                            \
    }

// Plain and batched kernels (batched: strided kernel for three waves per SIMD).
#define SBF_STRIP_BATCHED(FE, K, S, M, H)                                                     \
    StripKernelSet {                                                                          \
        K, S, M, H, FE::kAlphabet, FE::kBits == 2, wave_windows<K, S, M, H, FE>(),             \
            &launch_insert<K, S, M, H, FE>, &launch_contains<K, S, M, H, FE>,                  \
            {&launch_batched_insert<K, S, M, H, FE, 3>, nullptr},                             \
            {&launch_batched_contains<K, S, M, H, FE, 3>, nullptr}                            \
    }

// Compiled configurations: the measured workloads (DNA k=31, s in {28, 30, 31},
// m=16, h=4; protein k=10, s=8, m=6, h=4), m=21 as in the reference's FPR
// sweep, h=8, the protein configuration of the reference's example
// (examples/main.cu:164: k=12, s=10, m=6, h=4), the reference's small test
// configurations and a mid-size one; h = 8, 12 and 16 (the salted bit
// addresses of h = 12 and 16 keep each s-mer's hash and build its masks when
// they are needed) and three short-span points of the reference's parameter
// sweep (param_sweep.cu:263-300: m = 24, 20 and 31 with s = 28, 30 and 31) with
// plain kernels. Other configurations run the generic kernels.
static const StripKernelSet kSets[] = {
    SBF_STRIP_BATCHED(Dna2, 31, 28, 16, 4),
    SBF_STRIP_BATCHED(Dna2, 31, 30, 16, 4),
    SBF_STRIP_BATCHED(Dna2, 31, 31, 16, 4),
    SBF_STRIP_PLAIN(Dna2, 31, 28, 21, 4),
    SBF_STRIP_PLAIN(Dna2, 31, 28, 16, 8),
    SBF_STRIP_PLAIN(Dna2, 31, 28, 16, 12),
    SBF_STRIP_PLAIN(Dna2, 31, 28, 16, 16),
    SBF_STRIP_PLAIN(Dna2, 31, 28, 24, 4),
    SBF_STRIP_PLAIN(Dna2, 31, 30, 20, 8),
    SBF_STRIP_PLAIN(Dna2, 31, 31, 31, 4),
    SBF_STRIP_BATCHED(Dna2, 5, 4, 3, 4),
    SBF_STRIP_BATCHED(Dna2, 3, 2, 2, 4),
    SBF_STRIP_BATCHED(Dna2, 21, 19, 11, 4),
    SBF_STRIP_BATCHED(Protein5, 10, 8, 6, 4),
    SBF_STRIP_BATCHED(Protein5, 12, 10, 6, 4),
    SBF_STRIP_BATCHED(Protein5, 5, 4, 3, 4),
};

}  // namespace sbf::strip

namespace sbf {

// kernels_wide_dna.hip, kernels_wide_protein.hip
const StripKernelSet* find_wide_dna(uint32_t k, uint32_t s, uint32_t m, uint32_t h);
const StripKernelSet* find_wide_protein(uint32_t k, uint32_t s, uint32_t m, uint32_t h);

const StripKernelSet* find_wide_kernels(uint32_t k, uint32_t s, uint32_t m, uint32_t h,
                                        uint32_t alphabet) {
    if (alphabet == kDna) return find_wide_dna(k, s, m, h);
    if (alphabet == kProtein) return find_wide_protein(k, s, m, h);
    return nullptr;
}

const StripKernelSet* find_strip_kernels(uint32_t k, uint32_t s, uint32_t m, uint32_t h,
                                         uint32_t alphabet) {
    for (const StripKernelSet& set : strip::kSets) {
        if (set.k == k && set.s == s && set.m == m && set.h == h && set.alphabet == alphabet) {
            return &set;
        }
    }
    return find_wide_kernels(k, s, m, h, alphabet);
}

}  // namespace sbf
