// Wavefront (wave64) primitives shared by the gfx950 kernels: lane id,
// DPP prefix scans, cross-lane reads and reductions.  All functions that
// exchange values between lanes must be called with the whole wave active.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace trust {

// Orders the wave's LDS accesses before and after this point (a wave runs
// in lockstep, so a fence plus the wave barrier is a wave-level barrier).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Inclusive prefix sum over the 64 lanes.
__device__ __forceinline__ int wave_scan_add(int x)
{
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);  // row_shr:1
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);  // row_shr:2
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);  // row_shr:4
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);  // row_shr:8
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);  // row_bcast:15
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return x;
}

// Inclusive prefix max over the 64 lanes of non-negative values.
__device__ __forceinline__ int wave_scan_max(int x)
{
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false));
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false));
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false));
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false));
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false));
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false));
    return x;
}

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// Value of x in lane src (any lane may name any source).
__device__ __forceinline__ int shfl_idx(int x, int src)
{
    return __builtin_amdgcn_ds_bpermute(src << 2, x);
}
__device__ __forceinline__ int64_t shfl_idx(int64_t x, int src)
{
    int lo = __builtin_amdgcn_ds_bpermute(src << 2, (int)(x & 0xffffffff));
    int hi = __builtin_amdgcn_ds_bpermute(src << 2, (int)(x >> 32));
    return ((int64_t)hi << 32) | (unsigned int)lo;
}

// Value of x in lane l (l wave-uniform), into a scalar register.
__device__ __forceinline__ int readlane(int x, int l) { return __builtin_amdgcn_readlane(x, l); }
__device__ __forceinline__ int64_t readlane(int64_t x, int l)
{
    int lo = __builtin_amdgcn_readlane((int)(x & 0xffffffff), l);
    int hi = __builtin_amdgcn_readlane((int)(x >> 32), l);
    return ((int64_t)hi << 32) | (unsigned int)lo;
}

// Sum over the 64 lanes, in every lane.
__device__ __forceinline__ unsigned long long wave_sum(unsigned long long x)
{
    for (int off = 32; off > 0; off >>= 1)
        x += __shfl_xor(x, off);
    return x;
}

// The XCD (accelerator complex die, 0..7 on MI355X) this wave runs on, from
// the hardware register XCC_ID.  Each XCD has its own L2, so work that
// shares data is steered to one XCD; callers use it only for locality, never
// for correctness.
__device__ __forceinline__ int xcc_id()
{
    return (int)(__builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 7);  // HW_REG_XCC_ID
}

}  // namespace trust
