// probe.hip -- memory-system ceilings the Super Bloom kernels are measured against
// (bench-support library libcusbf_synth.so; not part of the filter's boundary).
//
//  * stream copy: 16-byte vector loads and stores, grid-stride (HBM streaming ceiling);
//  * random 32-byte gather: each access reads one random 32-byte sector of a
//    buffer, lanes 2r and 2r+1 taking its two 16-byte halves in one load
//    instruction (one memory request per sector line), eight independent
//    accesses in flight per lane pair (the query's shard loads and the sorted
//    insert's mask loads have this shape);
//  * random sector update: four 64-bit atomic ORs into one random 32-byte
//    sector per access (the insert's per-run update);
//  * the gather in other load forms (nontemporal, coherent atomic loads) and on
//    other memory kinds, to see what an HBM read request costs for each.
// Sector indices come from a counter-based hash, so every run is reproducible.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

constexpr int kBlock = 256;
constexpr int kInFlight = 8;

__global__ __launch_bounds__(kBlock) void copy_kernel(const u32x4* src, u32x4* dst, uint64_t n) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride) {
        dst[i] = __builtin_nontemporal_load(src + i);
    }
}

__global__ __launch_bounds__(kBlock) void gather_kernel(const u32x4* sectors, uint64_t sector_mask,
                                                        uint64_t accesses, uint64_t seed,
                                                        uint32_t* sink) {
    const uint64_t pair = (static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x) >> 1;
    const uint64_t pairs = static_cast<uint64_t>(gridDim.x) * (kBlock / 2);
    const uint32_t half = threadIdx.x & 1u;
    uint32_t acc = 0;
    for (uint64_t base = pair * kInFlight; base < accesses; base += pairs * kInFlight) {
        u32x4 a[kInFlight];
#pragma unroll
        for (int j = 0; j < kInFlight; ++j) {
            const uint64_t s = mix64(seed + base + j) & sector_mask;
            a[j] = base + j < accesses ? sectors[2 * s + half] : u32x4{0u, 0u, 0u, 0u};
        }
#pragma unroll
        for (int j = 0; j < kInFlight; ++j) acc ^= a[j].x ^ a[j].w;
    }
    if (acc == 0x9E3779B9u) sink[0] = acc;  // keeps the loads; practically never stores
}

// The gather in four load forms: 0 plain 16-byte loads (as gather_kernel), 1 nontemporal
// 16-byte loads, 2 two 64-bit relaxed atomic loads at agent scope, 3 the same at system
// scope (the coherent forms set the sc bits on the loads).
template <int Mode>
__global__ __launch_bounds__(kBlock) void gather_form_kernel(const uint64_t* words, uint64_t sector_mask,
                                                             uint64_t accesses, uint64_t seed,
                                                             uint32_t* sink) {
    const uint64_t pair = (static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x) >> 1;
    const uint64_t pairs = static_cast<uint64_t>(gridDim.x) * (kBlock / 2);
    const uint32_t half = threadIdx.x & 1u;
    uint32_t acc = 0;
    for (uint64_t base = pair * kInFlight; base < accesses; base += pairs * kInFlight) {
        uint64_t a[kInFlight][2];
#pragma unroll
        for (int j = 0; j < kInFlight; ++j) {
            const uint64_t* p = words + 4 * (mix64(seed + base + j) & sector_mask) + 2 * half;
            if (base + j >= accesses) {
                a[j][0] = a[j][1] = 0;
            } else if constexpr (Mode == 0 || Mode == 1) {
                const u32x4 v = Mode == 0 ? *reinterpret_cast<const u32x4*>(p)
                                          : __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(p));
                a[j][0] = v.x | (static_cast<uint64_t>(v.y) << 32);
                a[j][1] = v.z | (static_cast<uint64_t>(v.w) << 32);
            } else if constexpr (Mode == 2) {
                a[j][0] = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                a[j][1] = __hip_atomic_load(p + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                a[j][0] = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                a[j][1] = __hip_atomic_load(p + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            }
        }
#pragma unroll
        for (int j = 0; j < kInFlight; ++j) acc ^= static_cast<uint32_t>(a[j][0] ^ (a[j][1] >> 7));
    }
    if (acc == 0x9E3779B9u) sink[0] = acc;  // keeps the loads; practically never stores
}

__global__ __launch_bounds__(kBlock) void atomic_kernel(uint64_t* words, uint64_t sector_mask,
                                                        uint64_t accesses, uint64_t seed) {
    const uint64_t tid = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
    const uint64_t threads = static_cast<uint64_t>(gridDim.x) * kBlock;
    for (uint64_t i = tid; i < accesses; i += threads) {
        const uint64_t r = mix64(seed + i);
        uint64_t* dst = words + 4 * (r & sector_mask);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            __hip_atomic_fetch_or(dst + q, 1ull << ((r >> (40 + 6 * q)) & 63), __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

uint32_t grid_of(uint64_t items, uint64_t per_thread) {
    uint64_t blocks = (items / per_thread + kBlock - 1) / kBlock;
    if (blocks < 1) blocks = 1;
    if (blocks > 65536) blocks = 65536;
    return static_cast<uint32_t>(blocks);
}

}  // namespace

extern "C" {

// Copies n_bytes (a multiple of 16) from src to dst.
int cusbf_probe_copy(const void* src, void* dst, uint64_t n_bytes, void* stream) {
    const uint64_t n = n_bytes / 16;
    (void)hipGetLastError();
    hipLaunchKernelGGL(copy_kernel, dim3(grid_of(n, 4)), dim3(kBlock), 0,
                       reinterpret_cast<hipStream_t>(stream), static_cast<const u32x4*>(src),
                       static_cast<u32x4*>(dst), n);
    return static_cast<int>(hipGetLastError());
}

// `accesses` random 32-byte sector reads from a buffer of num_sectors (a power of two).
int cusbf_probe_gather(const void* buffer, uint64_t num_sectors, uint64_t accesses, uint64_t seed,
                       uint32_t* d_sink, void* stream) {
    (void)hipGetLastError();
    hipLaunchKernelGGL(gather_kernel, dim3(grid_of(accesses, 2 * kInFlight)), dim3(kBlock), 0,
                       reinterpret_cast<hipStream_t>(stream), static_cast<const u32x4*>(buffer),
                       num_sectors - 1, accesses, seed, d_sink);
    return static_cast<int>(hipGetLastError());
}

// The gather in load form `mode` (0 plain, 1 nontemporal, 2 agent-scope atomic loads,
// 3 system-scope atomic loads) over a buffer of num_sectors (a power of two).
int cusbf_probe_gather_form(const void* buffer, uint64_t num_sectors, uint64_t accesses,
                            uint64_t seed, uint32_t* d_sink, int mode, void* stream) {
    (void)hipGetLastError();
    const dim3 grid(grid_of(accesses, 2 * kInFlight)), block(kBlock);
    const hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const uint64_t* w = static_cast<const uint64_t*>(buffer);
    switch (mode) {
        case 0: hipLaunchKernelGGL(gather_form_kernel<0>, grid, block, 0, s, w, num_sectors - 1, accesses, seed, d_sink); break;
        case 1: hipLaunchKernelGGL(gather_form_kernel<1>, grid, block, 0, s, w, num_sectors - 1, accesses, seed, d_sink); break;
        case 2: hipLaunchKernelGGL(gather_form_kernel<2>, grid, block, 0, s, w, num_sectors - 1, accesses, seed, d_sink); break;
        case 3: hipLaunchKernelGGL(gather_form_kernel<3>, grid, block, 0, s, w, num_sectors - 1, accesses, seed, d_sink); break;
        default: return static_cast<int>(hipErrorInvalidValue);
    }
    return static_cast<int>(hipGetLastError());
}

// Device memory of the kinds the policy probe compares: 0 hipMalloc, 1 fine-grained,
// 2 uncached, 3 contiguous. Buffers of kinds 1-3 are released only when the process ends.
int cusbf_probe_alloc(uint64_t bytes, int kind, void** out) {
    hipError_t e;
    switch (kind) {
        case 0: e = hipMalloc(out, bytes); break;
        case 1: e = hipExtMallocWithFlags(out, bytes, hipDeviceMallocFinegrained); break;
        case 2: e = hipExtMallocWithFlags(out, bytes, hipDeviceMallocUncached); break;
        case 3: e = hipExtMallocWithFlags(out, bytes, hipDeviceMallocContiguous); break;
        default: return static_cast<int>(hipErrorInvalidValue);
    }
    if (e == hipSuccess) e = hipMemset(*out, 0x5A, bytes);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    return static_cast<int>(e);
}

// `accesses` random sector updates (four 64-bit atomic ORs each).
int cusbf_probe_atomic(void* buffer, uint64_t num_sectors, uint64_t accesses, uint64_t seed,
                       void* stream) {
    (void)hipGetLastError();
    hipLaunchKernelGGL(atomic_kernel, dim3(grid_of(accesses, 4)), dim3(kBlock), 0,
                       reinterpret_cast<hipStream_t>(stream), static_cast<uint64_t*>(buffer),
                       num_sectors - 1, accesses, seed);
    return static_cast<int>(hipGetLastError());
}

}  // extern "C"
