// Integer pattern searches (diamond, hexagon) of xavs2_me_search for every
// (PU, reference) instance of a plan.  gfx950 only.
//
// The search of one instance is a chain of dependent steps (encoder/me.c:
// 1101-1154): each step evaluates 3, 4 or 6 points around the current best
// and moves to the first strict minimum.  A group of S lanes (16, 32 or 64 by
// PU area) walks one instance; a wave holds 64 / S instances of one size
// class and steps until all of them have stopped.  Every lane of a group
// holds the instance's state, sums its share of the PU's pixel quads for all
// points of the step, and the group's DPP sums come back to every lane with
// one bpermute per point.  Reference pixels are read from global memory (a
// step reads a few PU-sized areas next to the previous step's, from L2).
#include "xme_internal.h"
#include "xme_dpp.h"

namespace {

typedef const __attribute__((address_space(1))) uint8_t *gbytes;
typedef const __attribute__((address_space(1))) uint32_t *gdwords;

// Point tables of encoder/me.c, 3 bits per entry (entry + 2 for signed
// coordinates), so that lanes index them in registers:
//  HEX2 / DIA1 with their wrap-around entries (me.c:83-92, :101-108),
//  M1MOD6 / M1MOD4 (me.c:96, :112),
//  the first hexagon of umh_step_2 in evaluation order with its direction
//  codes (me.c:1111-1112), the first diamond of umh_step_3 (me.c:1137).
constexpr uint32_t pack8(int a0, int a1, int a2, int a3, int a4, int a5, int a6, int a7, int bias)
{
    return (uint32_t)(a0 + bias) | (uint32_t)(a1 + bias) << 3 | (uint32_t)(a2 + bias) << 6 |
           (uint32_t)(a3 + bias) << 9 | (uint32_t)(a4 + bias) << 12 | (uint32_t)(a5 + bias) << 15 |
           (uint32_t)(a6 + bias) << 18 | (uint32_t)(a7 + bias) << 21;
}
constexpr uint32_t kHex2X = pack8(-1, -2, -1, 1, 2, 1, -1, -2, 2);
constexpr uint32_t kHex2Y = pack8(-2, 0, 2, 2, 0, -2, -2, 0, 2);
constexpr uint32_t kDia1X = pack8(0, -1, 0, 1, 0, -1, 0, 0, 2);
constexpr uint32_t kDia1Y = pack8(-1, 0, 1, 0, -1, 0, 0, 0, 2);
constexpr uint32_t kM1Mod6 = pack8(5, 0, 1, 2, 3, 4, 5, 0, 0);
constexpr uint32_t kM1Mod4 = pack8(3, 0, 1, 2, 3, 0, 0, 0, 0);
constexpr uint32_t kHex6X = pack8(-1, 1, -2, 2, -1, 1, 0, 0, 2);
constexpr uint32_t kHex6Y = pack8(-2, -2, 0, 0, 2, 2, 0, 0, 2);
constexpr uint32_t kHex6D = pack8(6, 5, 1, 4, 2, 3, 0, 0, 0);
constexpr uint32_t kDia4X = pack8(0, -1, 1, 0, 0, 0, 0, 0, 2);
constexpr uint32_t kDia4Y = pack8(-1, 0, 0, 1, 0, 0, 0, 0, 2);
constexpr uint32_t kDia4D = pack8(4, 1, 3, 2, 0, 0, 0, 0, 0);

__device__ __forceinline__ int tab(uint32_t t, int i, int bias) { return (int)((t >> (3 * i)) & 7u) - bias; }

// Stages: the first hexagon, half hexagons, the first diamond, half diamonds.
enum { kHexFirst = 0, kHexHalf = 1, kDiaFirst = 2, kDiaHalf = 3, kDone = 4 };

// One wave: 64 / S instances of PUs with at most S * QMAX pixel quads.
template <int S, int QMAX>
__device__ void pattern_wave(const DevParams &p, const FrameDesc *__restrict__ desc, const DevPu *__restrict__ pus,
                             DevInst *__restrict__ inst, uint64_t *__restrict__ best, int id,
                             unsigned long long *__restrict__ counts)
{
    const int lane = threadIdx.x & 63;
    const int l = lane & (S - 1);
    int phase = kDone;
    DevPu pu = {};
    DevInst in = {};
    int bcost = 0, bmx = 0, bmy = 0;
    if (id >= 0) {
        pu = pus[id / p.n_ref];
        in = inst[id];
        const uint64_t key = best[id];
        if (!in.status && key < XME_KEY_NONE) {
            const uint32_t order = (uint32_t)key;
            bcost = (int)(key >> 32);
            // the candidate-list pass leaves the best candidate as the centre
            bmx = order == 0 ? in.mvc[0][0] : in.mvc[1][0];
            bmy = order == 0 ? in.mvc[0][1] : in.mvc[1][1];
            phase = p.method == XME_ME_HEX ? kHexFirst : kDiaFirst;
        }
    }
    const int r = id >= 0 ? id - (id / p.n_ref) * p.n_ref : 0;
    const MvWindow win = xme_mv_window(p, pu.x, pu.y, pu.w, pu.h);
    const int64_t i_ref = desc->stride[r];
    const gbytes ref = (gbytes)desc->planes[r][0] + (int64_t)pu.y * i_ref + pu.x;

    // this lane's pixel quads of the PU: current dwords and offsets
    const int gw = pu.w >> 2, quads = gw * pu.h;
    uint32_t c4[QMAX];
    int32_t off[QMAX];
    int nq = 0;
    {
        const gbytes cur = (gbytes)desc->cur + (int64_t)pu.y * desc->i_cur + pu.x;
#pragma unroll
        for (int q = 0; q < QMAX; q++) {
            const int u = l + q * S;
            c4[q] = 0;
            off[q] = 0;
            if (phase != kDone && u < quads) {
                const int yy = u / gw, xx = (u - yy * gw) * 4;
                c4[q] = *(gdwords)(cur + (int64_t)yy * desc->i_cur + xx);
                off[q] = yy * (int32_t)i_ref + xx;
                nq = q + 1;
            }
        }
    }

    int idx = 0, hex_i = 0, dia_i = 0;
    uint32_t evals = 0;
    for (;;) {
        // stage ends that need no evaluation (loop heads of me.c:1120, :1136, :1145)
        const bool legal = xme_mv_legal(win, bmx, bmy);
        if (phase == kHexHalf && !(hex_i < p.R && legal))
            phase = kDiaFirst;
        if (phase == kDiaFirst && !legal)
            phase = kDone;
        if (phase == kDiaHalf && !(dia_i < p.R && legal))
            phase = kDone;
        if (__builtin_amdgcn_ballot_w64(phase != kDone) == 0)
            break;
        const int n = phase == kHexFirst ? 6 : phase == kDiaFirst ? 4 : phase == kDone ? 0 : 3;
        int dx[6], dy[6], dc[6];
#pragma unroll
        for (int k = 0; k < 6; k++) {
            if (phase == kHexFirst) {
                dx[k] = tab(kHex6X, k, 2);
                dy[k] = tab(kHex6Y, k, 2);
                dc[k] = tab(kHex6D, k, 0);
            } else if (phase == kHexHalf) {
                dx[k] = tab(kHex2X, min(idx + k, 7), 2);
                dy[k] = tab(kHex2Y, min(idx + k, 7), 2);
                dc[k] = k + 1;
            } else if (phase == kDiaFirst) {
                dx[k] = tab(kDia4X, k, 2);
                dy[k] = tab(kDia4Y, k, 2);
                dc[k] = tab(kDia4D, k, 0);
            } else {
                dx[k] = tab(kDia1X, min(idx + k, 5), 2);
                dy[k] = tab(kDia1Y, min(idx + k, 5), 2);
                dc[k] = k + 1;
            }
        }
        // SADs of the step's points (ME_COST_IPEL_X3_DIR / _X4_DIR evaluate
        // them without a range check; they lie at most 2 pixels outside the
        // MV window, inside the plane's padding)
        uint32_t sad[6];
#pragma unroll
        for (int k = 0; k < 6; k++) {
            sad[k] = 0;
            if (k < n) {
                const gbytes base = ref + (int64_t)(bmy + dy[k]) * i_ref + (bmx + dx[k]);
#pragma unroll
                for (int q = 0; q < QMAX; q++) {
                    if (q < nq) {
                        const uintptr_t a = (uintptr_t)(base + off[q]);
                        const gdwords d = (gdwords)(a & ~(uintptr_t)3);
                        sad[k] = __builtin_amdgcn_sad_u8(c4[q], __builtin_amdgcn_alignbyte(d[1], d[0], (uint32_t)(a & 3)),
                                                         sad[k]);
                    }
                }
            }
        }
        // group sums back to every lane of the group, then the points in
        // order with a strict '<' (COPY4_IF_LT, me.h:76-82)
        int dir = 0, nbx = bmx, nby = bmy;
#pragma unroll
        for (int k = 0; k < 6; k++) {
            int s = group_sum((int)sad[k], S);
            s = __builtin_amdgcn_ds_bpermute((lane | (S - 1)) << 2, s);
            if (k < n) {
                const int mx = bmx + dx[k], my = bmy + dy[k];
                const int cost = s + xme_weighted_cost(p.lambda, xme_mvbits(4 * mx - pu.mvp_x) + xme_mvbits(4 * my - pu.mvp_y));
                if (cost < bcost) {
                    bcost = cost;
                    nbx = mx;
                    nby = my;
                    dir = dc[k];
                }
            }
        }
        evals += (uint32_t)n;
        bmx = nbx;
        bmy = nby;
        if (phase == kHexFirst) {
            phase = dir ? kHexHalf : kDiaFirst;
            idx = dir - 1;
        } else if (phase == kHexHalf) {
            hex_i++;
            if (dir)
                idx = tab(kM1Mod6, dir + idx - 1, 0);
            else
                phase = kDiaFirst;
        } else if (phase == kDiaFirst) {
            phase = dir ? kDiaHalf : kDone;
            idx = dir - 1;
        } else if (phase == kDiaHalf) {
            dia_i++;
            if (dir)
                idx = tab(kM1Mod4, dir + idx - 1, 0);
            else
                phase = kDone;
        }
    }
    if (id >= 0 && l == S - 1 && evals) {
        // the sub-pel pass reads the integer best as the centre's raster
        // candidate (0, 0): order XME_ORDER_RASTER + R * 2R + R
        inst[id].omx = (int16_t)bmx;
        inst[id].omy = (int16_t)bmy;
        best[id] = xme_key((uint32_t)bcost, (uint32_t)(XME_ORDER_RASTER + p.R * 2 * p.R + p.R));
        atomicAdd(&counts[0], (unsigned long long)evals);
        atomicAdd(&counts[1], (unsigned long long)evals * (unsigned long long)(pu.w * pu.h));
    }
}

// Wave w of the launch: classes of 16-, 32-, 64- and 64-lane groups
// (PUs of <= 64 px, < 256 px, <= 1024 px, larger) at wave offsets
// wave_end[0..3]; list holds 64 / S instance ids per wave (-1: none).
__global__ void __launch_bounds__(256) pattern_kernel(DevParams p, const FrameDesc *__restrict__ desc,
                                                      const DevPu *__restrict__ pus, DevInst *__restrict__ inst,
                                                      uint64_t *__restrict__ best, const int32_t *__restrict__ list,
                                                      int4 wave_end, unsigned long long *__restrict__ counts)
{
    const int w = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4 + (threadIdx.x >> 6)));
    const int lane = threadIdx.x & 63;
    if (w < wave_end.x) {
        pattern_wave<16, 1>(p, desc, pus, inst, best, list[w * 64 / 16 + lane / 16], counts);
    } else if (w < wave_end.y) {
        const int e0 = wave_end.x * 4;
        pattern_wave<32, 2>(p, desc, pus, inst, best, list[e0 + (w - wave_end.x) * 2 + lane / 32], counts);
    } else if (w < wave_end.z) {
        const int e0 = wave_end.x * 4 + (wave_end.y - wave_end.x) * 2;
        pattern_wave<64, 4>(p, desc, pus, inst, best, list[e0 + (w - wave_end.y)], counts);
    } else if (w < wave_end.w) {
        const int e0 = wave_end.x * 4 + (wave_end.y - wave_end.x) * 2 + (wave_end.z - wave_end.y);
        pattern_wave<64, 16>(p, desc, pus, inst, best, list[e0 + (w - wave_end.z)], counts);
    }
}

}  // namespace

int xme_pattern_class(int w, int h)
{
    const int px = w * h;
    return px <= 64 ? 0 : px < 256 ? 1 : px <= 1024 ? 2 : 3;
}

int xme_launch_pattern(const DevParams &p, const FrameDesc *d_desc, const DevPu *d_pus, DevInst *d_inst,
                       uint64_t *d_best, const int32_t *d_list, const int wave_end[4], unsigned long long *d_counts,
                       hipStream_t s)
{
    const int waves = wave_end[3];
    if (waves <= 0)
        return XME_OK;
    const int4 we = make_int4(wave_end[0], wave_end[1], wave_end[2], wave_end[3]);
    hipLaunchKernelGGL(pattern_kernel, dim3((waves + 3) / 4), dim3(256), 0, s, p, d_desc, d_pus, d_inst, d_best,
                       d_list, we, d_counts);
    XME_HIP(hipGetLastError());
    return XME_OK;
}
