// 4x4 tile costs shared by the sub-pel kernels.  gfx950 only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

__device__ __forceinline__ int xme_bytei(uint32_t v, int i) { return (int)((v >> (8 * i)) & 0xFFu); }

// Sum of |H4 D H4^T| over one 4x4 tile (before the final halving), or the SAD,
// of current rows cr[0..3] against prediction rows rr[0..3] (4 bytes each).
// 32-bit arithmetic throughout: packed 16-bit integer forms are avoided on
// this target.
__device__ __forceinline__ int xme_tile_cost_rows(const uint32_t (&cr)[4], const uint32_t (&rr)[4], bool hadamard)
{
    if (!hadamard) {
        int s = 0;
#pragma unroll
        for (int y = 0; y < 4; y++)
            s = (int)__builtin_amdgcn_sad_u8(cr[y], rr[y], (uint32_t)s);
        return s;
    }
    int d[4][4];
#pragma unroll
    for (int y = 0; y < 4; y++)
#pragma unroll
        for (int x = 0; x < 4; x++)
            d[y][x] = xme_bytei(cr[y], x) - xme_bytei(rr[y], x);
#pragma unroll
    for (int y = 0; y < 4; y++) {
        const int a = d[y][0] + d[y][1], b = d[y][0] - d[y][1];
        const int c = d[y][2] + d[y][3], e = d[y][2] - d[y][3];
        d[y][0] = a + c;
        d[y][1] = b + e;
        d[y][2] = a - c;
        d[y][3] = b - e;
    }
    // the last butterfly level pairs (a + c, a - c) and (b + e, b - e), and
    // |p + q| + |p - q| = 2 max(|p|, |q|) for integers: the pairs' sums are
    // never formed
    int s = 0;
#pragma unroll
    for (int x = 0; x < 4; x++) {
        const int a = d[0][x] + d[1][x], b = d[0][x] - d[1][x];
        const int c = d[2][x] + d[3][x], e = d[2][x] - d[3][x];
        s += max(abs(a), abs(c)) + max(abs(b), abs(e));
    }
    return 2 * s;
}
