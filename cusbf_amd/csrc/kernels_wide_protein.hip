// kernels_wide_protein.hip -- runtime-span lane-strip kernel sets for protein, one per hash
// count (strip_wide.hpp); kernels_wide_dna.hip holds the DNA sets.
#include "strip_wide.hpp"

namespace sbf::strip {

#define SBF_WIDE_SET(FE, H)                                                                        \
    StripKernelSet {                                                                               \
        0, 0, 0, H, FE::kAlphabet, FE::kBits == 2, static_cast<uint32_t>(WideShape<FE>::kWaveWindows), \
            &launch_wide_insert<FE, H>, &launch_wide_contains<FE, H>,                              \
            {&launch_wide_batched_insert<FE, H>, nullptr}, {&launch_wide_batched_contains<FE, H>, nullptr} \
    }

static const StripKernelSet kWideProtein[] = {SBF_WIDE_SET(Protein5, 4), SBF_WIDE_SET(Protein5, 8),
                                          SBF_WIDE_SET(Protein5, 12), SBF_WIDE_SET(Protein5, 16)};

}  // namespace sbf::strip

namespace sbf {

const StripKernelSet* find_wide_protein(uint32_t k, uint32_t s, uint32_t m, uint32_t h) {
    if (!strip::wide_fits<strip::Protein5>(k, s, m)) return nullptr;
    for (const StripKernelSet& set : strip::kWideProtein) {
        if (set.h == h) return &set;
    }
    return nullptr;
}

}  // namespace sbf
