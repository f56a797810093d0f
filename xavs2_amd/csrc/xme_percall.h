// Launch helpers of the per-call faces (percall_kernels.hip), used by the
// g_funcs-signature wrappers in primitives.cpp.
#pragma once

#include "xme_internal.h"

// An interpolation (or pixel-average) call on staged device buffers: output
// pixel (x, y) of a w x h block reads the staged source around
// src + y * i_src + x (its margins staged too), or the int16 rows of tin.
struct PcIntpl {
    int mode;               // kPcLumaHor ... kPcAvg
    int w, h;
    int64_t i_src, i_tmp, i_dst;
    int8_t ch[8], cv[8];    // taps across / down (4 used for chroma)
};

enum {
    kPcLumaHor = 0,         // intpl_luma_hor (tout may be null: intpl_luma_block_hor)
    kPcLumaVer,             // intpl_luma_ver / intpl_luma_block_ver
    kPcLumaExt,             // intpl_luma_ext (from int16 rows)
    kPcLumaBlockExt,        // intpl_luma_block_ext
    kPcChromaHor,
    kPcChromaVer,
    kPcChromaExt,
    kPcAvg                  // pixf.avg: (a + b + 1) >> 1, src2 at the same stride
};

int xme_pc_prepare();
int xme_pc_xform(const int16_t *d_src, int16_t *d_dst, int w, int h, int fwd, int s1, int s2, int clip2, int half,
                 hipStream_t s);
int xme_pc_quant(int16_t *d_c, int n, int scale, int shift, int add, int *d_nnz, int quant, hipStream_t s);
int xme_pc_intpl(const PcIntpl &p, const uint8_t *d_src, const int16_t *d_tin, const uint8_t *d_src2, uint8_t *d_dst,
                 int16_t *d_tout, hipStream_t s);
