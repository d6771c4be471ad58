// cusbf/filter_ref_device.hpp -- device-only members of cusbf::filter_ref<Config>
// (include/cusbf/filter_ref.hpp). HIP device code that inserts through a
// filter_ref includes this header; it compiles only with hipcc.
//
// filter_ref.cuh:94-108 and filter_common.cuh:137-139 of the reference: the insert
// ORs a run's masks into its shard with 64-bit device-scope atomics, skipping zero
// masks. The filter's words live in ordinary device memory, so the atomics are
// performed at the memory side and are coherent with the library's own kernels.
#pragma once

#if !defined(__HIPCC__)
#error "cusbf/filter_ref_device.hpp holds device code: compile it with hipcc"
#endif

#include <hip/hip_runtime.h>

#include "filter_ref.hpp"

namespace cusbf {
namespace detail {

/// filter_common.cuh:137-139.
__device__ __forceinline__ void atomicOrWord(uint64_t* ptr, uint64_t value) {
    __hip_atomic_fetch_or(ptr, value, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace detail

template <typename Config>
__device__ void filter_ref<Config>::apply_word_masks(block_type& block, uint64_t m0, uint64_t m1,
                                                     uint64_t m2, uint64_t m3) const {
    if (m0 != 0) detail::atomicOrWord(&block.words[0], m0);
    if (m1 != 0) detail::atomicOrWord(&block.words[1], m1);
    if (m2 != 0) detail::atomicOrWord(&block.words[2], m2);
    if (m3 != 0) detail::atomicOrWord(&block.words[3], m3);
}

}  // namespace cusbf
