// Shared host/device definitions for the MI355X projector library.
//
// Conventions follow the reference (wudufan/CTProjector):
//   images       img[B][nz][ny][nx]          (cuda/projector.h:100-123)
//   projections  prj[B][nview][nv][nu]
//   geometry     rows [nview][3] as (x, y, z) with x on the fastest image axis
//   image index space has 0 at the corner of the first voxel
//                (PhysicsToImg, cuda/projector.h:63-71).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace ctp {

struct Grid {                     // cuda/projector.h:7-18
    int nx, ny, nz;
    float dx, dy, dz, cx, cy, cz;
};

struct Det {                      // cuda/projector.h:20-28
    int nu, nv;
    float du, dv, off_u, off_v;
};

// typeProjector bits of the 2D distance-driven projector
// (cuda/distanceDriven.h:33-44, cuda/distanceDriven2D.cu:12-25).
enum : int { DD_FBP = 1, DD_BRANCHLESS = 2, DD_FORCE_FBP = 4 };

// filter codes (cuda/fbp.h:5-7)
enum : int { FILTER_RL = 0, FILTER_HAMMING = 1, FILTER_HANN = 2, FILTER_COSINE = 3 };

// The reference's constants (cuda/projector.h:5, cuda/siddon.cu:11,
// cuda/distanceDriven2DBoxInt.cu:18).
constexpr float kPI = 3.141592657f;
constexpr float kEps = 1e-6f;
constexpr float kDDMargin = 0.12f;

__host__ __device__ inline int clampi(int x, int s, int e) { return x < s ? s : (x >= e ? e - 1 : x); }
__host__ __device__ inline float clampf(float x, float s, float e) { return x < s ? s : (x > e ? e : x); }

// a / b through the hardware reciprocal (v_rcp_f32, 1 ulp): the divisions of
// the per-view and per-row footprint arithmetic, where the IEEE-exact
// sequence costs a dozen instructions and two mode switches.  Operands here
// are finite and far from the denormal range; the result is within 2 ulp.
__device__ inline float fast_div(float a, float b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return a * __builtin_amdgcn_rcpf(b);
#else
    return a / b;
#endif
}
// clampf for finite x, and min / max of finite values, as one median instruction each
__device__ inline float med3_f(float a, float b, float c) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_fmed3f(a, b, c);
#else
    return a < b ? (b < c ? b : (a < c ? c : a)) : (a < c ? a : (b < c ? c : b));
#endif
}
__device__ inline float med_clamp(float x, float s, float e) { return med3_f(x, s, e); }
__device__ inline float min_f(float a, float b) { return med3_f(a, b, -3.0e38f); }
__device__ inline float max_f(float a, float b) { return med3_f(a, b, 3.0e38f); }

inline int ceil_div(long a, long b) { return (int)((a + b - 1) / b); }

// Read-only data loaded through the constant address space: wave-uniform
// loads become scalar memory instructions.  The data must not be written while
// the kernel runs.
template <typename T>
__device__ inline T const_load(const T* p, long i) {
#if defined(__HIP_DEVICE_COMPILE__)
    static_assert(sizeof(T) % 4 == 0, "const_load: size must be a multiple of 4 bytes");
    typedef const __attribute__((address_space(4))) unsigned* cptr;
    const cptr q = (cptr)(p + i);
    T out;
    unsigned* o = reinterpret_cast<unsigned*>(&out);
#pragma unroll
    for (int k = 0; k < (int)(sizeof(T) / 4); k++) o[k] = q[k];
    return out;
#else
    return p[i];
#endif
}

// Value of lane lead + t (t < LPC) of the calling lane's group of LPC
// consecutive lanes: a DPP quad permute for groups of 2 and 4 lanes.
#ifdef __HIPCC__
template <int LPC>
__device__ inline int group_bcast(int v, int t, int lead) {
    if constexpr (LPC == 4) {
        switch (t) {   // quad_perm(t, t, t, t)
            case 0: return __builtin_amdgcn_update_dpp(0, v, 0x00, 0xF, 0xF, true);
            case 1: return __builtin_amdgcn_update_dpp(0, v, 0x55, 0xF, 0xF, true);
            case 2: return __builtin_amdgcn_update_dpp(0, v, 0xAA, 0xF, 0xF, true);
            default: return __builtin_amdgcn_update_dpp(0, v, 0xFF, 0xF, 0xF, true);
        }
    } else if constexpr (LPC == 2) {   // quad_perm(t, t, 2 + t, 2 + t)
        return t == 0 ? __builtin_amdgcn_update_dpp(0, v, 0xA0, 0xF, 0xF, true)
                      : __builtin_amdgcn_update_dpp(0, v, 0xF5, 0xF, 0xF, true);
    } else {
        return __shfl(v, lead + t);
    }
}
template <int LPC>
__device__ inline float group_bcast(float v, int t, int lead) {
    return __int_as_float(group_bcast<LPC>(__float_as_int(v), t, lead));
}

// Minimum (MAX false) or maximum over the lanes l of a wavefront with the same
// l mod LPC: row rotations (DPP) within rows of 16 lanes, then two exchanges
// across rows.
template <int LPC, bool MAX>
__device__ inline int class_reduce(int v) {
    auto op = [](int a, int b) { return MAX ? max(a, b) : min(a, b); };
    if constexpr (LPC == 2) v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x122, 0xF, 0xF, false));   // row_ror:2
    v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x124, 0xF, 0xF, false));                           // row_ror:4
    v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x128, 0xF, 0xF, false));                           // row_ror:8
    v = op(v, __shfl_xor(v, 16));
    return op(v, __shfl_xor(v, 32));
}

// Work distribution of the persistent kernels: a block takes its next block
// index from a counter with an atomic add, so the blocks in flight stay a
// compact window of the index space (as with a hardware-dispatched grid) even
// when blocks take unequal times.  With 'lanes' = 8 the counters are per XCD
// (a persistent block stays on the XCD of its first dispatch, blockIdx.x mod
// 8) and hand out indices b = 8 * k + xcd, which keeps the XCD band mapping.
// The index is read back through LDS and made wave-uniform (readfirstlane), so
// that the job lookup and the job's parameters stay in scalar registers.
// Block indices fit in 31 bits (checked on the host).
__device__ inline int next_block(int* __restrict__ ctr, int lanes, int* slot) {
    const int l = lanes == 8 ? (int)(blockIdx.x & 7) : 0;
    __syncthreads();                                  // the previous index has been read by every thread
    if (threadIdx.x == 0) *slot = lanes * atomicAdd(ctr + l, 1) + l;
    __syncthreads();
    return __builtin_amdgcn_readfirstlane(*slot);
}

// Block loop of the plan kernels: persistent (work counters above) or one
// hardware-dispatched block per index (persist == 0).
__device__ inline int first_block(int* __restrict__ ctr, int lanes, int* slot, int persist) {
    return persist ? next_block(ctr, lanes, slot) : (int)blockIdx.x;
}
__device__ inline int following_block(int* __restrict__ ctr, int lanes, int* slot, int persist) {
    return persist ? next_block(ctr, lanes, slot) : 0x7fffffff;
}
#endif

}  // namespace ctp
