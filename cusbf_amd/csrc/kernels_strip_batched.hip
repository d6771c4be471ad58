// kernels_strip_batched.hip -- instantiates the batched lane-strip launchers and
// their kernels (strip_kernels.hpp) for the configurations the dispatch table in
// kernels_strip.hip lists.
#include "strip_kernels.hpp"

namespace sbf::strip {

#define SBF_STRIP_INSTANTIATE(FE, K, S, M, H, W)                                               \
    template hipError_t launch_batched_insert<K, S, M, H, FE, W>(                              \
        const KernelConfig&, const SequenceInput&, FilterView, const BatchLaunch&,             \
        const RecordSink*, hipStream_t);                                                       \
    template hipError_t launch_batched_contains<K, S, M, H, FE, W>(                            \
        const KernelConfig&, const SequenceInput&, FilterView, const BatchLaunch&, uint8_t*,   \
        hipStream_t);
SBF_STRIP_BATCHED_INSTANCES(SBF_STRIP_INSTANTIATE)
#undef SBF_STRIP_INSTANTIATE

}  // namespace sbf::strip
