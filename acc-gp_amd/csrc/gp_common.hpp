// Shared definitions for the gfx950 partitioner: error handling, wave64
// helpers and the hash used by every deterministic tie-break.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace gp {

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define GP_HIP(call)                                                                   \
    do {                                                                               \
        const hipError_t gp_err__ = (call);                                            \
        if (gp_err__ != hipSuccess)                                                    \
            throw ::gp::HipError(std::string("HIP failure at ") + __FILE__ + ":" +     \
                                 std::to_string(__LINE__) + " for " #call ": " +       \
                                 hipGetErrorString(gp_err__));                         \
    } while (0)

#define GP_LAUNCH_CHECK() GP_HIP(hipGetLastError())

constexpr int kWave = 64;          // CDNA wavefront width
constexpr int kMaxParts = 128;
constexpr std::int32_t kInvalid = 0x7fffffff;

inline int ceil_div(std::int64_t a, std::int64_t b) { return static_cast<int>((a + b - 1) / b); }

// mix32: the integer hash behind every tie-break (src/coarsen.cu:173-179).
__host__ __device__ inline std::uint32_t mix32(std::uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    return x ^ (x >> 16);
}

__device__ inline std::uint64_t ew_at(const std::uint32_t* ew, std::int64_t e) {
    return ew ? static_cast<std::uint64_t>(ew[e]) : 1ULL;
}

__device__ inline int lane_id() { return static_cast<int>(__lane_id()); }

// Orders LDS traffic between lanes of one wavefront.
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename T>
__device__ inline T wave_sum(T v) {
    for (int d = kWave / 2; d; d >>= 1) v += __shfl_xor(v, d);
    return v;
}
template <typename T>
__device__ inline T wave_max(T v) {
    for (int d = kWave / 2; d; d >>= 1) {
        const T o = __shfl_xor(v, d);
        v = o > v ? o : v;
    }
    return v;
}
// Reduction within aligned groups of W lanes (W divides 64).
template <int W, typename T>
__device__ inline T group_sum(T v) {
    for (int d = W / 2; d; d >>= 1) v += __shfl_xor(v, d, W);
    return v;
}

__device__ inline unsigned long long atomic_add_u64(unsigned long long* p, unsigned long long v) {
    return atomicAdd(p, v);
}

}  // namespace gp
