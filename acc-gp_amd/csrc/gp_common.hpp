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

// Kernel launches made by the calling host thread: every launch site goes
// through GP_LAUNCH_CHECK, and a partition reads the count before and after.
std::uint64_t& thread_launches();
#define GP_LAUNCH_CHECK()              \
    do {                               \
        ++::gp::thread_launches();     \
        GP_HIP(hipGetLastError());     \
    } while (0)

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

// ------------------------------------------------ wave butterflies by DPP
// DPP controls (GFX9 encoding) of the row-local exchanges a 64-lane
// butterfly uses: quad_perm [1,0,3,2] (lane ^ 1), quad_perm [2,3,0,1]
// (lane ^ 2), row_half_mirror (lane i <-> 7 - i in each group of 8) and
// row_mirror (lane i <-> 15 - i in each row of 16).  After the two quad
// exchanges every lane holds its quad's result; the half mirror pairs each
// lane with a lane of the other quad of its 8, the mirror with a lane of the
// other 8 of its row, so four DPP moves combine a whole row.
constexpr int kDppQuadXor1 = 0xB1;
constexpr int kDppQuadXor2 = 0x4E;
constexpr int kDppRowHalfMirror = 0x141;
constexpr int kDppRowMirror = 0x140;

// Moves a trivially copyable value (a whole number of 32-bit words) across
// lanes by one DPP pattern; every lane of the wave must be active.
template <int Ctrl, typename T>
__device__ inline T dpp_move(const T& v) {
    static_assert(sizeof(T) % 4 == 0, "DPP moves 32-bit words");
    constexpr int words = sizeof(T) / 4;
    int in[words], out[words];
    __builtin_memcpy(in, &v, sizeof(T));
#pragma unroll
    for (int i = 0; i < words; ++i) out[i] = __builtin_amdgcn_update_dpp(in[i], in[i], Ctrl, 0xf, 0xf, false);
    T r;
    __builtin_memcpy(&r, out, sizeof(T));
    return r;
}

// The same across rows (lane ^ d for d = 16, 32) through the LDS crossbar.
template <typename T>
__device__ inline T xor_move(const T& v, int d) {
    static_assert(sizeof(T) % 4 == 0, "lane moves are 32-bit words");
    constexpr int words = sizeof(T) / 4;
    int in[words], out[words];
    __builtin_memcpy(in, &v, sizeof(T));
#pragma unroll
    for (int i = 0; i < words; ++i) out[i] = __shfl_xor(in[i], d);
    T r;
    __builtin_memcpy(&r, out, sizeof(T));
    return r;
}

// Every lane gets combine() over the 64 lanes' values.  combine must be
// associative and commutative (a sum, or the better of two under a total
// order), so the result does not depend on the pairing.  Called from
// converged control flow only.
template <typename T, typename Combine>
__device__ inline T wave_allreduce(T v, Combine combine) {
    v = combine(v, dpp_move<kDppQuadXor1>(v));
    v = combine(v, dpp_move<kDppQuadXor2>(v));
    v = combine(v, dpp_move<kDppRowHalfMirror>(v));
    v = combine(v, dpp_move<kDppRowMirror>(v));
    v = combine(v, xor_move(v, 16));
    v = combine(v, xor_move(v, 32));
    return v;
}

struct PlusU64 {
    __device__ unsigned long long operator()(unsigned long long x, unsigned long long y) const { return x + y; }
};
struct PlusI64 {
    __device__ std::int64_t operator()(std::int64_t x, std::int64_t y) const { return x + y; }
};

}  // namespace gp
