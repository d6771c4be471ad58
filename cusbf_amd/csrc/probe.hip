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
//    other memory kinds, to see what an HBM read request costs for each;
//  * VALU issue cost: how many cycles one wave64 instruction of a given kind
//    holds its SIMD (the kernels' ALU time is priced with these);
//  * the query's memory traffic without its arithmetic: per 1 KiB tile of a
//    streamed input, a given number of random 32-byte sector reads and 1 KiB
//    written back (input, answers and shard lines of the query in the same
//    mix, with the shard loads' sectors taken from a hash).
// Sector indices come from a counter-based hash, so every run is reproducible.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <utility>

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

// Waves stride over 1 KiB tiles like the lane-strip query: each lane loads its 16
// input bytes (nontemporal), the wave reads `per_tile` random sectors (lanes 2q
// and 2q+1 the two halves of sector q, q = lane / 2 + 32 i), and each lane stores
// 16 bytes (nontemporal) that depend on the loads, so none is dropped.
// `parts`: bit 0 streams the input in, bit 1 writes the output (3: the query's mix).
__global__ __launch_bounds__(kBlock) void traffic_kernel(const u32x4* in, u32x4* out, uint64_t tiles,
                                                         uint32_t per_tile, const u32x4* sectors,
                                                         uint64_t sector_mask, uint64_t seed,
                                                         uint32_t parts, uint32_t* sink) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t half = lane & 1u;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * (kBlock / 64);
    const uint64_t rounds = (per_tile + 31) / 32;
    for (uint64_t t = static_cast<uint64_t>(blockIdx.x) * (kBlock / 64) + (threadIdx.x >> 6); t < tiles;
         t += stride) {
        u32x4 acc = {0u, 0u, 0u, 0u};
        if (parts & 1u) acc = __builtin_nontemporal_load(in + 64 * t + lane);
        for (uint64_t i = 0; i < rounds; ++i) {
            const uint64_t q = (lane >> 1) + 32 * i;
            if (q < per_tile) {
                const uint64_t sct = mix64(seed + t * per_tile + q) & sector_mask;
                acc ^= sectors[2 * sct + half];
            }
        }
        if (parts & 2u) {
            __builtin_nontemporal_store(acc, out + 64 * t + lane);
        } else if ((acc.x ^ acc.w) == 0x9E3779B9u) {
            sink[0] = acc.x;  // keeps the loads; practically never stores
        }
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

// VALU issue cost: each lane runs kValuIters rounds of eight independent chains
// of one instruction (inline asm on registers only, so nothing else is issued in
// the loop but its counter and branch); time x clock x SIMDs / wave instructions
// is the number of cycles one wave64 instruction holds its SIMD.
constexpr int kValuIters = 4096;
constexpr int kValuOps = 23;

#define SBF_CHAIN8(STMT) STMT(0) STMT(1) STMT(2) STMT(3) STMT(4) STMT(5) STMT(6) STMT(7)
#define SBF_VALU32(OP, ASM, ...)                                         \
    if constexpr (Op == OP) {                                            \
        _Pragma("unroll") for (int j = 0; j < 8; ++j) asm volatile(ASM : "+v"(a[j]) : __VA_ARGS__); \
    }
#define SBF_VALU64(OP, ASM, ...)                                         \
    if constexpr (Op == OP) {                                            \
        _Pragma("unroll") for (int j = 0; j < 8; ++j) asm volatile(ASM : "+v"(b[j]) : __VA_ARGS__); \
    }

template <int Op>
__global__ __launch_bounds__(kBlock) void valu_kernel(uint32_t x, uint32_t y, uint64_t* sink) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    const uint64_t c64 = (static_cast<uint64_t>(y) << 32) | x;
    uint32_t a[8];
    uint64_t b[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        a[j] = t * (2u * j + 3u);
        b[j] = (static_cast<uint64_t>(a[j]) << 32) | (t + j);
    }
    for (int i = 0; i < kValuIters; ++i) {
        SBF_VALU32(0, "v_add_u32 %0, %1, %0", "v"(x))
        SBF_VALU32(1, "v_or_b32 %0, %1, %0", "v"(x))
        SBF_VALU32(2, "v_and_b32 %0, %1, %0", "v"(x))
        SBF_VALU32(3, "v_xor_b32 %0, %1, %0", "v"(x))
        SBF_VALU32(4, "v_lshrrev_b32 %0, %1, %0", "v"(x))
        SBF_VALU32(5, "v_lshlrev_b32 %0, %1, %0", "v"(x))
        SBF_VALU32(6, "v_bitop3_b32 %0, %1, %0, %2 bitop3:0x32", "v"(x), "v"(y))
        SBF_VALU32(7, "v_or3_b32 %0, %1, %0, %2", "v"(x), "v"(y))
        SBF_VALU32(8, "v_add3_u32 %0, %1, %0, %2", "v"(x), "v"(y))
        SBF_VALU32(9, "v_min_u32 %0, %1, %0", "v"(x))
        SBF_VALU32(10, "v_bfe_u32 %0, %0, %1, 6", "v"(x))
        SBF_VALU32(11, "v_alignbit_b32 %0, %1, %0, 7", "v"(x))
        SBF_VALU32(12, "v_perm_b32 %0, %1, %0, %2", "v"(x), "v"(y))
        SBF_VALU32(13, "v_lshl_or_b32 %0, %0, 2, %1", "v"(x))
        SBF_VALU32(14, "v_mul_lo_u32 %0, %1, %0", "v"(x))
        SBF_VALU32(15, "v_mul_hi_u32 %0, %1, %0", "v"(x))
        SBF_VALU64(16, "v_mad_u64_u32 %0, vcc, %1, %2, %0", "v"(x), "v"(y) : "vcc")
        SBF_VALU64(17, "v_lshlrev_b64 %0, %1, %0", "v"(x))
        SBF_VALU64(18, "v_lshl_add_u64 %0, %0, 2, %1", "v"(c64))
        SBF_VALU64(19, "v_cmp_lt_u64_e64 vcc, %0, %1", "v"(c64) : "vcc")
        SBF_VALU32(20, "v_mov_b32_dpp %0, %0 wave_shl:1 row_mask:0xf bank_mask:0xf", "v"(x))
        SBF_VALU32(21, "v_bcnt_u32_b32 %0, %1, %0", "v"(x))
        SBF_VALU32(22, "v_mov_b32 %0, %1", "v"(x))
    }
    uint64_t r = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) r ^= a[j] ^ b[j];
    if (r == 0x9E3779B97F4A7C15ull) sink[0] = r;  // keeps the chains; practically never stores
}

#undef SBF_VALU64
#undef SBF_VALU32
#undef SBF_CHAIN8

template <int Op>
void launch_valu(uint32_t blocks, uint64_t* sink, hipStream_t s) {
    hipLaunchKernelGGL(valu_kernel<Op>, dim3(blocks), dim3(kBlock), 0, s, 0x9E3779B9u, 0x7F4A7C15u, sink);
}

template <int... Ops>
void launch_valu_op(int op, uint32_t blocks, uint64_t* sink, hipStream_t s,
                    std::integer_sequence<int, Ops...>) {
    ((op == Ops ? launch_valu<Ops>(blocks, sink, s) : void()), ...);
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

// The query's traffic: `tiles` tiles of 1 KiB streamed from `in` to `out` (both
// tiles * 1 KiB), with `per_tile` random 32-byte sector reads of `buffer`
// (num_sectors, a power of two) per tile. At most 2^16 blocks of four waves.
// `parts`: bit 0 streams the input in, bit 1 writes the output; with a bit clear that
// stream is left out (d_sink receives nothing in practice).
int cusbf_probe_traffic(const void* in, void* out, uint64_t tiles, uint32_t per_tile, const void* buffer,
                        uint64_t num_sectors, uint64_t seed, uint32_t parts, uint32_t* d_sink,
                        void* stream) {
    uint64_t blocks = (tiles + 3) / 4;
    if (blocks < 1) blocks = 1;
    if (blocks > 65536) blocks = 65536;
    (void)hipGetLastError();
    hipLaunchKernelGGL(traffic_kernel, dim3(static_cast<uint32_t>(blocks)), dim3(kBlock), 0,
                       reinterpret_cast<hipStream_t>(stream), static_cast<const u32x4*>(in),
                       static_cast<u32x4*>(out), tiles, per_tile, static_cast<const u32x4*>(buffer),
                       num_sectors - 1, seed, parts, d_sink);
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

// Cycles one wave64 instruction of VALU kind `op` (0 .. 22; the names are listed in
// cusbf_amd/probes.py) holds its SIMD, with `waves_per_simd` waves on every SIMD of
// the device: the second of two identical launches, timed with events (synchronous).
int cusbf_probe_valu(int op, int waves_per_simd, double* cycles) {
    if (op < 0 || op >= kValuOps || waves_per_simd < 1 || waves_per_simd > 8 || cycles == nullptr) {
        return static_cast<int>(hipErrorInvalidValue);
    }
    int device = 0, cus = 0, clock_khz = 0;
    hipError_t e = hipGetDevice(&device);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeClockRate, device);
    if (e != hipSuccess) return static_cast<int>(e);
    // 256-thread blocks are four waves: CUs x waves_per_simd blocks put that many on each SIMD
    const uint32_t blocks = static_cast<uint32_t>(cus) * static_cast<uint32_t>(waves_per_simd);
    uint64_t* sink = nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    e = hipMalloc(&sink, sizeof(uint64_t));
    if (e == hipSuccess) e = hipEventCreate(&t0);
    if (e == hipSuccess) e = hipEventCreate(&t1);
    if (e == hipSuccess) {
        const auto ops = std::make_integer_sequence<int, kValuOps>{};
        launch_valu_op(op, blocks, sink, nullptr, ops);
        e = hipEventRecord(t0, nullptr);
        if (e == hipSuccess) launch_valu_op(op, blocks, sink, nullptr, ops);
        if (e == hipSuccess) e = hipGetLastError();
        if (e == hipSuccess) e = hipEventRecord(t1, nullptr);
        if (e == hipSuccess) e = hipEventSynchronize(t1);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
        if (e == hipSuccess) {
            const double wave_instructions = static_cast<double>(blocks) * (kBlock / 64) * kValuIters * 8.0;
            const double simd_cycles = ms * 1e-3 * clock_khz * 1e3 * (4.0 * cus);
            *cycles = simd_cycles / wave_instructions;
        }
    }
    if (t0 != nullptr) (void)hipEventDestroy(t0);
    if (t1 != nullptr) (void)hipEventDestroy(t1);
    if (sink != nullptr) (void)hipFree(sink);
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
