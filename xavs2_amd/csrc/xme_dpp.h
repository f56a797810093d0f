// Cross-lane reductions with DPP (no LDS traffic) for gfx950 kernels.  They
// run with all 64 lanes active.  row_shr and the row broadcasts move data to
// higher lanes, so a sum or minimum collects in the last lane of its group.
#pragma once

#include <hip/hip_runtime.h>

template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ int dpp_add(int v)
{
    return v + __builtin_amdgcn_update_dpp(0, v, CTRL, ROW_MASK, 0xF, true);
}

template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ int dpp_min(int v)
{
    // lanes without a source keep the identity
    return min(v, __builtin_amdgcn_update_dpp(0x7fffffff, v, CTRL, ROW_MASK, 0xF, false));
}

// Sum over each aligned group of P lanes (P a power of two <= 64, uniform);
// the sum is in the group's last lane.
__device__ __forceinline__ int group_sum(int v, int P)
{
    if (P >= 2)
        v = dpp_add<0xB1>(v);          // quad_perm [1,0,3,2]
    if (P >= 4)
        v = dpp_add<0x4E>(v);          // quad_perm [2,3,0,1]
    if (P >= 8)
        v = dpp_add<0x114>(v);         // row_shr:4
    if (P >= 16)
        v = dpp_add<0x118>(v);         // row_shr:8
    if (P >= 32)
        v = dpp_add<0x142, 0xA>(v);    // row_bcast:15 into rows 1 and 3
    if (P >= 64)
        v = dpp_add<0x143, 0xC>(v);    // row_bcast:31 into rows 2 and 3
    return v;
}

// Minimum over each aligned group of P lanes (P a power of two <= 64,
// uniform); the minimum is in the group's last lane.
__device__ __forceinline__ int group_min(int v, int P)
{
    if (P >= 2)
        v = dpp_min<0xB1>(v);
    if (P >= 4)
        v = dpp_min<0x4E>(v);
    if (P >= 8)
        v = dpp_min<0x114>(v);
    if (P >= 16)
        v = dpp_min<0x118>(v);
    if (P >= 32)
        v = dpp_min<0x142, 0xA>(v);
    if (P >= 64)
        v = dpp_min<0x143, 0xC>(v);
    return v;
}

// Sum over the wave, returned to every lane.
__device__ __forceinline__ int wave_sum(int v) { return __builtin_amdgcn_readlane(group_sum(v, 64), 63); }

// Minimum of non-negative keys over the wave, returned to every lane.
__device__ __forceinline__ int wave_min_key(int v)
{
    v = dpp_min<0xB1>(v);
    v = dpp_min<0x4E>(v);
    v = dpp_min<0x114>(v);
    v = dpp_min<0x118>(v);
    v = dpp_min<0x142, 0xA>(v);
    v = dpp_min<0x143, 0xC>(v);
    return __builtin_amdgcn_readlane(v, 63);
}
