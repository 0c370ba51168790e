// Shared helpers for the gfx950 kernels and the C-ABI host code.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include "../../include/bubbledynamics.h"

#define BD_HIP_CHECK(expr)                                                       \
    do {                                                                         \
        hipError_t _e = (expr);                                                  \
        if (_e != hipSuccess) {                                                  \
            fprintf(stderr, "bubbledynamics: %s failed: %s (%s:%d)\n", #expr,    \
                    hipGetErrorString(_e), __FILE__, __LINE__);                  \
            return BD_ERR_HIP;                                                   \
        }                                                                        \
    } while (0)

#define BD_LAUNCH_CHECK()                                                        \
    do {                                                                         \
        hipError_t _e = hipGetLastError();                                       \
        if (_e != hipSuccess) {                                                  \
            fprintf(stderr, "bubbledynamics: launch failed: %s (%s:%d)\n",       \
                    hipGetErrorString(_e), __FILE__, __LINE__);                  \
            return BD_ERR_HIP;                                                   \
        }                                                                        \
    } while (0)

namespace bd {

constexpr int kWave = 64;   // CDNA wavefront width
constexpr int kXcds = 8;    // MI355X: 8 XCDs, workgroups dispatched round-robin

// Workgroups are handed to the XCDs round-robin (block b -> XCD b % 8).  This
// remap gives every XCD a contiguous range of logical tile ids so that tiles
// that share halo data also share an L2 (the stencil kernels number their
// tiles x, then y, then z-chunk).  Bijective for any nblocks.
__device__ __forceinline__ int xcd_remap(int bid, int nblocks) {
    const int q = nblocks / kXcds, r = nblocks % kXcds;
    const int x = bid % kXcds, k = bid / kXcds;
    const int start = x * q + (x < r ? x : r);
    return start + k;
}

// Deterministic wave64 sum (fixed butterfly order).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Deterministic block sum of up to 16 waves; result valid in thread 0.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* lds /* >= NT/64 */) {
    v = wave_sum(v);
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    if (lane == 0) lds[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NT / kWave; ++i) s += lds[i];
    }
    __syncthreads();
    return s;
}

// Raw buffer access: a workgroup-uniform 128-bit resource (base, byte size)
// plus a 32-bit per-lane byte offset.  Reads past the size return 0 and
// writes past it are dropped by the hardware.
using rsrc_t = __amdgpu_buffer_rsrc_t;
constexpr int kRsrcWord3 = 0x00020000;

__device__ __forceinline__ rsrc_t make_rsrc(const void* base, size_t bytes) {
    const uint32_t n = bytes > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)bytes;
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)n, kRsrcWord3);
}

template <typename T>
__device__ __forceinline__ T bload(rsrc_t r, uint32_t voff, uint32_t soff = 0);
template <>
__device__ __forceinline__ float bload<float>(rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}
template <>
__device__ __forceinline__ double bload<double>(rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
}

// Load with the non-temporal cache policy (NT): for streams read once per
// kernel whose lines are not reused from the caches before eviction (history
// levels of the march).  Measured on the Westervelt step: 83.4 -> 76.4 us at
// 192^3 (340 MB state), but slower at 128^3, whose 100 MB state the 256 MB
// last-level cache (MALL) holds between steps.
constexpr size_t kMallBytes = 256ull << 20;
constexpr int kCachePolicyNT = 2;
template <typename T>
__device__ __forceinline__ T bload_nt(rsrc_t r, uint32_t voff) {
    if constexpr (sizeof(T) == 4)
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, 0, kCachePolicyNT));
    else
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, 0, kCachePolicyNT));
}

template <typename T>
__device__ __forceinline__ void bstore(T v, rsrc_t r, uint32_t voff);
template <>
__device__ __forceinline__ void bstore<float>(float v, rsrc_t r, uint32_t voff) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, (int)voff, 0, 0);
}
template <>
__device__ __forceinline__ void bstore<double>(double v, rsrc_t r, uint32_t voff) {
    typedef unsigned int v2u __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2u, v), r, (int)voff, 0, 0);
}

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace bd
