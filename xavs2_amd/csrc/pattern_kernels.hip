// Integer pattern searches (diamond, hexagon) of xavs2_me_search for every
// (PU, reference) instance of a plan.  gfx950 only.
//
// The search of one instance is a chain of dependent steps (encoder/me.c:
// 1101-1154): each step evaluates 3, 4 or 6 points around the current best
// and moves to the first strict minimum.  A group of S lanes (4 to 64, so that
// a lane holds at most 4 of the PU's pixel quads, 16 for the largest PUs)
// walks one instance; a wave holds 64 / S instances of one size class and
// steps until all of them have stopped.  Every lane of a group
// holds the instance's state.  A step stages the box its points read (the PU
// grown by 2 pixels a side) from the reference plane into the wave's LDS with
// row-contiguous dword loads, then sums each lane's share of the PU's pixel
// quads for all points of the step from LDS; the group's DPP sums come back
// to every lane with one bpermute per point.
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
// An entry that does not fit its 3 bits sets bit 31, which the
// static_asserts below the tables reject.
constexpr bool fits3(int v) { return v >= 0 && v < 8; }
constexpr uint32_t pack8(int a0, int a1, int a2, int a3, int a4, int a5, int a6, int a7, int bias)
{
    const bool ok = fits3(a0 + bias) && fits3(a1 + bias) && fits3(a2 + bias) && fits3(a3 + bias) &&
                    fits3(a4 + bias) && fits3(a5 + bias) && fits3(a6 + bias) && fits3(a7 + bias);
    return (uint32_t)(a0 + bias) | (uint32_t)(a1 + bias) << 3 | (uint32_t)(a2 + bias) << 6 |
           (uint32_t)(a3 + bias) << 9 | (uint32_t)(a4 + bias) << 12 | (uint32_t)(a5 + bias) << 15 |
           (uint32_t)(a6 + bias) << 18 | (uint32_t)(a7 + bias) << 21 | (ok ? 0u : 1u << 31);
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
static_assert(!((kHex2X | kHex2Y | kDia1X | kDia1Y | kM1Mod6 | kM1Mod4 | kHex6X | kHex6Y | kHex6D | kDia4X |
                 kDia4Y | kDia4D) >> 31), "a 3-bit table entry out of range");

__device__ __forceinline__ int tab(uint32_t t, int i, int bias) { return (int)((t >> (3 * i)) & 7u) - bias; }

// UMH tables (me.c:69-79, :116-126), 4 bits per entry (entry + 4): the big
// hexagons HEX4 (16 points) and FAST_HEX4 (8), and the 5x5 grid in two
// halves of 12 entries.  The grid lists (0, 2) at entries 10 and 21 and never
// (0, -2), as the reference does.
constexpr uint64_t pack16(const int (&v)[16])
{
    uint64_t r = 0;
    for (int i = 0; i < 16; i++)
        r |= (uint64_t)(v[i] + 4) << (4 * i);
    return r;
}
constexpr bool fits16(const int (&v)[16])
{
    for (int i = 0; i < 16; i++)
        if (v[i] + 4 < 0 || v[i] + 4 > 15)
            return false;
    return true;
}
constexpr int kHex4XV[16] = { 0, 0, -2, 2, -4, 4, -4, 4, -4, 4, -4, 4, -4, 4, -2, 2 };
constexpr int kHex4YV[16] = { -4, 4, -3, -3, -2, -2, -1, -1, 0, 0, 1, 1, 2, 2, 3, 3 };
constexpr int kFHex4XV[16] = { 0, 0, -2, 2, -4, 4, -2, 2, 0, 0, 0, 0, 0, 0, 0, 0 };
constexpr int kFHex4YV[16] = { -4, 4, -3, -3, 0, 0, 3, 3, 0, 0, 0, 0, 0, 0, 0, 0 };
constexpr int kGridXV[24] = { -1, 0, 1, -1, 1, -1, 0, 1, -2, -1, 0, 1, 2, -2, 2, -2, 2, -2, 2, -2, -1, 0, 1, 2 };
constexpr int kGridYV[24] = { -1, -1, -1, 0, 0, 1, 1, 1, -2, -2, 2, -2, -2, -1, -1, 0, 0, 1, 1, 2, 2, 2, 2, 2 };
constexpr int kGridXA[16] = { -1, 0, 1, -1, 1, -1, 0, 1, -2, -1, 0, 1, 0, 0, 0, 0 };
constexpr int kGridXB[16] = { 2, -2, 2, -2, 2, -2, 2, -2, -1, 0, 1, 2, 0, 0, 0, 0 };
constexpr int kGridYA[16] = { -1, -1, -1, 0, 0, 1, 1, 1, -2, -2, 2, -2, 0, 0, 0, 0 };
constexpr int kGridYB[16] = { -2, -1, -1, 0, 0, 1, 1, 2, 2, 2, 2, 2, 0, 0, 0, 0 };
constexpr uint64_t kHex4X = pack16(kHex4XV), kHex4Y = pack16(kHex4YV);
constexpr uint64_t kFHex4X = pack16(kFHex4XV), kFHex4Y = pack16(kFHex4YV);
constexpr uint64_t kGridXa = pack16(kGridXA), kGridXb = pack16(kGridXB);
constexpr uint64_t kGridYa = pack16(kGridYA), kGridYb = pack16(kGridYB);
static_assert(kGridXV[10] == 0 && kGridYV[10] == 2 && kGridXV[21] == 0 && kGridYV[21] == 2, "grid entries 10, 21");
static_assert(fits16(kHex4XV) && fits16(kHex4YV) && fits16(kFHex4XV) && fits16(kFHex4YV) && fits16(kGridXA) &&
              fits16(kGridXB) && fits16(kGridYA) && fits16(kGridYB), "a 4-bit table entry out of range");

__device__ __forceinline__ int tab4(uint64_t t, int i) { return (int)((t >> (4 * i)) & 15u) - 4; }

// Stages: the first hexagon, half hexagons, the first diamond, half diamonds.
enum { kHexFirst = 0, kHexHalf = 1, kDiaFirst = 2, kDiaHalf = 3, kDone = 4 };

// LDS dwords per wave: the step box of a 64x64 PU, 68 rows x 18 dwords.
#define XME_PAT_BOX_DW 1224

// One wave: 64 / S instances of PUs with at most S * QMAX pixel quads, whose
// step box (PU + 2 pixels on every side, dword aligned) takes at most S * NB
// dwords.
// Offsets and direction code of point k of a step (k < the step's count).
__device__ __forceinline__ void point_of(int phase, int idx, int k, int &dx, int &dy, int &dc)
{
    if (phase == kHexFirst) {
        dx = tab(kHex6X, k, 2);
        dy = tab(kHex6Y, k, 2);
        dc = tab(kHex6D, k, 0);
    } else if (phase == kHexHalf) {
        dx = tab(kHex2X, min(idx + k, 7), 2);
        dy = tab(kHex2Y, min(idx + k, 7), 2);
        dc = k + 1;
    } else if (phase == kDiaFirst) {
        dx = tab(kDia4X, k, 2);
        dy = tab(kDia4Y, k, 2);
        dc = tab(kDia4D, k, 0);
    } else {
        dx = tab(kDia1X, min(idx + k, 5), 2);
        dy = tab(kDia1Y, min(idx + k, 5), 2);
        dc = k + 1;
    }
}

// COPY4_IF_LT (me.h:76-82) of one point: SAD + MV_COST_IPEL against the
// running best with a strict '<'.
__device__ __forceinline__ void take_point(const DevParams &p, const DevPu &pu, int mx, int my, int dc, int sad,
                                           int &bcost, int &nbx, int &nby, int &dir)
{
    const int cost = sad + xme_weighted_cost(p.lambda, xme_mvbits(4 * mx - pu.mvp_x) + xme_mvbits(4 * my - pu.mvp_y));
    if (cost < bcost) {
        bcost = cost;
        nbx = mx;
        nby = my;
        dir = dc;
    }
}

// UMH stages before the hexagon (me.c:1030-1099).  Evaluating stages hold
// their points; the others are decisions.
enum {
    kUDia0 = 0,      // DIA_ITER around the MVP candidate mvc[0]
    kUDia1,          // DIA_ITER around the best, if the MVP is not (0,0) and the best moved
    kUMvp1,          // IPEL(mvp1), if mvp1 is not (0,0)
    kUMvp23,         // IPEL(mvp2) if ref_idx > 0, IPEL(mvp3) if mvp3 is not (0,0)
    kUDia2,          // DIA_ITER around the best, if it is not pmv
    kUCross,         // the unsymmetrical cross
    kUGrid,          // the 5x5 grid
    kUBig,           // one big hexagon level
    kUOff            // left for the hexagon or the diamond
};

struct UmhState {
    int st, bi;                 // stage, batch of 6 points within it
    int cx, cy;                 // stage centre
    int pmx, pmy;               // pmv (me.c:910)
    int level;                  // big hexagon level j
    int p0x, p0y;               // mvc[0]
    bool dfree;                 // the window holds the offsets (0,-1), (-1,0), (1,0), (0,1)
};

// Point t of the current UMH stage: 0 for a point the stage skips (a missing
// predictor), 1 for a point taken only inside the MV window (ME_COST_IPEL),
// 2 for a point taken without a range check.  DIA_ITER (me.c:328-334) runs
// ME_COST_IPEL_X4 on the offsets (0,-1), (-1,0), (1,0), (0,1), whose
// CHECK_MV_RANGE_X4 (me.h:51-56) tests the offsets themselves: when the
// window holds them (dfree) the four points around the centre are taken
// unchecked (me.c:244-258), otherwise the offsets as absolute MVs, each
// checked (me.c:259-264).
__device__ __forceinline__ int umh_point(const UmhState &u, const xme_umh_ctx &c, int R, int big, int t, int &mx,
                                         int &my)
{
    const int ch = (R + 1) >> 1;           // odd i <= R
    switch (u.st) {
    case kUDia0:
    case kUDia1:
    case kUDia2:
        mx = (u.dfree ? u.cx : 0) + tab(kDia4X, t, 2);
        my = (u.dfree ? u.cy : 0) + tab(kDia4Y, t, 2);
        return t < 4 ? (u.dfree ? 2 : 1) : 0;
    case kUMvp1:
        mx = (c.mvp1_x + 2) >> 2;
        my = (c.mvp1_y + 2) >> 2;
        return t == 0;
    case kUMvp23:
        mx = t == 0 ? (c.mvp2_x + 2) >> 2 : (c.mvp3_x + 2) >> 2;
        my = t == 0 ? (c.mvp2_y + 2) >> 2 : (c.mvp3_y + 2) >> 2;
        return t == 0 ? c.ref_idx > 0 : (t == 1 && (c.mvp3_x | c.mvp3_y) != 0);
    case kUCross:
        if (t < 2 * ch) {
            const int i = 2 * (t >> 1) + 1;
            mx = u.cx + ((t & 1) ? -i : i);
            my = u.cy;
        } else {
            const int j = 2 * ((t - 2 * ch) >> 1) + 1;
            mx = u.cx;
            my = u.cy + (((t - 2 * ch) & 1) ? -j : j);
        }
        return t < 2 * ch + 2 * (((R >> 1) + 1) >> 1);
    case kUGrid:
        mx = u.cx + (t < 12 ? tab4(kGridXa, t) : tab4(kGridXb, t - 12));
        my = u.cy + (t < 12 ? tab4(kGridYa, t) : tab4(kGridYb, t - 12));
        return t < 24;
    default:   // kUBig
        mx = u.cx + u.level * (big == 2 ? tab4(kHex4X, t & 15) : tab4(kFHex4X, t & 15));
        my = u.cy + u.level * (big == 2 ? tab4(kHex4Y, t & 15) : tab4(kFHex4Y, t & 15));
        return t < (big == 2 ? 16 : 8);
    }
}

__device__ __forceinline__ int umh_stage_points(int st, int R, int big)
{
    return st == kUMvp1 ? 1 : st == kUMvp23 ? 2 : st == kUCross ? 2 * ((R + 1) >> 1) + 2 * (((R >> 1) + 1) >> 1)
         : st == kUGrid ? 24 : st == kUBig ? (big == 2 ? 16 : 8) : 4;
}

// EARLY_TERMINATION (me.c:165-170): the search's phase to go on with, or -1.
__device__ __forceinline__ int umh_early(int bcost, int pred_sad, double beta2, double beta3)
{
    if ((double)bcost < (double)pred_sad * beta3)
        return kDiaFirst;
    if ((double)bcost < (double)pred_sad * beta2)
        return kHexFirst;
    return -1;
}

// Moves a group to its next evaluating UMH stage through the decisions
// between stages, or out of the prefix (st = kUOff, phase set).
__device__ __forceinline__ void umh_next(UmhState &u, const xme_umh_ctx &c, int R, int big, int bcost, int bmx,
                                         int bmy, int &phase)
{
    const double beta2 = c.beta2 + 1, beta3 = c.beta3 + 1;
    for (;;) {
        int et = -1;
        switch (u.st) {
        case kUDia1:       // entered after kUDia0
            if ((u.pmx | u.pmy) != 0 && (bmx != u.p0x || bmy != u.p0y)) {
                u.cx = bmx;
                u.cy = bmy;
                return;
            }
            u.st = kUMvp1;
            continue;
        case kUMvp1:
            if ((c.mvp1_x | c.mvp1_y) != 0)
                return;
            u.st = kUMvp23 + 100;      // the first early termination comes first
            continue;
        case kUMvp23 + 100:
            et = umh_early(bcost, c.pred_sad_uplayer, beta2, beta3);
            if (et >= 0)
                break;
            u.st = kUMvp23;
            if (c.ref_idx > 0 || (c.mvp3_x | c.mvp3_y) != 0)
                return;
            u.st = kUDia2;
            continue;
        case kUDia2:
            if (u.pmx != bmx || u.pmy != bmy) {
                u.cx = bmx;
                u.cy = bmy;
                return;
            }
            u.st = kUCross + 100;
            continue;
        case kUCross + 100:
            et = umh_early(bcost, c.pred_sad, beta2, beta3);
            if (et >= 0)
                break;
            u.st = kUCross;
            u.cx = bmx;
            u.cy = bmy;
            return;
        case kUGrid + 100:
            et = umh_early(bcost, c.pred_sad, beta2, beta3);
            if (et >= 0)
                break;
            u.st = kUGrid;
            u.cx = bmx;
            u.cy = bmy;
            return;
        case kUBig + 100:
            et = umh_early(bcost, c.pred_sad, beta2, beta3);
            if (et >= 0)
                break;
            if (big == 0 || R / 4 < 1) {
                et = kHexFirst;
                break;
            }
            u.st = kUBig;
            u.level = 1;
            u.cx = bmx;
            u.cy = bmy;
            return;
        case kUBig + 200:  // after a level: early termination if it moved
            if (bmx != u.cx || bmy != u.cy) {
                et = umh_early(bcost, c.pred_sad, beta2, beta3);
                if (et >= 0)
                    break;
            }
            if (++u.level > R / 4) {
                et = kHexFirst;
                break;
            }
            u.st = kUBig;
            u.cx = bmx;
            u.cy = bmy;
            return;
        default:
            return;
        }
        // out of the prefix
        u.st = kUOff;
        phase = et;
        return;
    }
}

// After the last batch of an evaluating stage: the stage that follows.
__device__ __forceinline__ void umh_done_stage(UmhState &u, int bmx, int bmy)
{
    switch (u.st) {
    case kUDia0:
        u.st = kUDia1;
        break;
    case kUDia1:
        u.pmx = bmx;
        u.pmy = bmy;
        u.st = kUMvp1;
        break;
    case kUMvp1:
        u.st = kUMvp23 + 100;
        break;
    case kUMvp23:
        u.st = kUDia2;
        break;
    case kUDia2:
        u.st = kUCross + 100;
        break;
    case kUCross:
        u.st = kUGrid + 100;
        break;
    case kUGrid:
        u.st = kUBig + 100;
        break;
    default:   // kUBig
        u.st = kUBig + 200;
        break;
    }
    u.bi = 0;
}

// TZ (me.c:926-1028): the stages before umh_step_2.  Evaluating stages hold
// their points; the decisions between them are tz_next's.
enum {
    kT6 = 0,         // ME_COST_IPEL_X3 twice around the best candidate (unchecked)
    kTDia,           // DIA_ITER around the best, if it lies in the window
    kTRound,         // one round of tz_pattern_search
    kTTwo,           // the two points next to a best at distance 1
    kTRaster,        // the raster refinement
    kTOff            // left for the hexagon, or done
};

struct TzState {
    int st, bi;                 // stage, batch of points within it
    int cx, cy;                 // centre of the stage's points
    int ox, oy;                 // omx, omy of xavs2_me_search: the raster's offset
    int dist, rounds, early;    // the round's distance, rounds without a gain, EarlyExitIters or MaxIters
    int r0;                     // the best cost when the round started
    int bdist, bdir;            // mv_info's bdist and bdir
    int call_cost;              // mvinfo.bcost after the first pattern search
    int rx, ry, rnx, rny;       // raster: first cell, cells a row and rows
    bool first;                 // the two points follow the first pattern search
    bool dfree;                 // the window holds the offsets (0,-1), (-1,0), (1,0), (0,1)
    bool undef;                 // a raster read left the reference's arrays
};

// The two ME_COST_IPEL_X3 of the TZ start (me.c:937-938) in order, and the
// Two Point Search offsets (me.c:144-153), 16 entries.
constexpr uint32_t kTz6X = pack8(-2, -1, 1, 2, 1, -1, 0, 0, 2);
constexpr uint32_t kTz6Y = pack8(0, 2, 2, 0, -2, -2, 0, 0, 2);
constexpr int kTzOffXV[16] = { -1, 0, -1, 1, -1, 1, -1, -1, 1, 1, -1, 0, -1, 1, 1, 0 };
constexpr int kTzOffYV[16] = { 0, -1, -1, -1, 0, 0, 1, -1, -1, 1, 0, 1, 1, 1, 0, 1 };
constexpr uint64_t kTzOffX = pack16(kTzOffXV), kTzOffY = pack16(kTzOffYV);
static_assert(!((kTz6X | kTz6Y) >> 31) && fits16(kTzOffXV) && fits16(kTzOffYV), "TZ table entry out of range");
// The rounds at distance d <= 8 (me.c:700-719, :745-782), in evaluation
// order: (0,-d) (-h,-h) (h,-h) (-d,0) (d,0) (-h,h) (h,h) (0,d) with h = d/2,
// as multiples of h; the round at distance 1 takes entries 0, 3, 4 and 7.
// Their direction codes (1..8, 4-bit entries).
constexpr uint32_t kTz8X = pack8(0, -1, 1, -2, 2, -1, 1, 0, 2);
constexpr uint32_t kTz8Y = pack8(-2, -1, -1, 0, 0, 1, 1, 2, 2);
constexpr int kTz8DV[16] = { 2, 1, 3, 4, 5, 6, 8, 7, 0, 0, 0, 0, 0, 0, 0, 0 };
constexpr uint64_t kTz8D = pack16(kTz8DV);
static_assert(!((kTz8X | kTz8Y) >> 31) && fits16(kTz8DV), "TZ table entry out of range");

// Point t of the current TZ stage: 0 none, 1 taken only inside the MV window
// (ME_COST_IPEL, ME_COST_IPEL_DIR_DIST), 2 taken without a range check, 3 a
// raster point at (omx + i, omy + j) (the sad_x4 branch of ME_COST_IPEL_X4,
// which offsets the absolute MV (i, j) by pix_base, me.c:245), taken
// unchecked if the reference's arrays hold it.
__device__ __forceinline__ int tz_stage_points(const TzState &z)
{
    return z.st == kT6 ? 6 : z.st == kTDia ? 4 : z.st == kTTwo ? 2 : z.st == kTRaster ? 4 * z.rnx * z.rny
         : z.dist == 1 ? 4 : z.dist <= 8 ? 8 : 16;
}

__device__ __forceinline__ int tz_point(const TzState &z, const MvWindow &win, int t, int &mx, int &my)
{
    mx = my = 0;
    if (t >= tz_stage_points(z))
        return 0;
    switch (z.st) {
    case kT6:
        mx = z.cx + tab(kTz6X, t, 2);
        my = z.cy + tab(kTz6Y, t, 2);
        return 2;
    case kTDia:
        mx = (z.dfree ? z.cx : 0) + tab(kDia4X, t, 2);
        my = (z.dfree ? z.cy : 0) + tab(kDia4Y, t, 2);
        return z.dfree ? 2 : 1;
    case kTRound: {
        const int d = z.dist;
        if (d <= 8) {
            // at distance 1 the entries 0, 3, 4, 7 are (0, -+2 h) and (-+2 h, 0) with h = 1/2
            const int e = d == 1 ? (t == 0 ? 0 : t == 1 ? 3 : t == 2 ? 4 : 7) : t;
            mx = z.cx + (d == 1 ? tab(kTz8X, e, 2) / 2 : tab(kTz8X, e, 2) * (d >> 1));
            my = z.cy + (d == 1 ? tab(kTz8Y, e, 2) / 2 : tab(kTz8Y, e, 2) * (d >> 1));
            return 1;
        }
        if (t < 4) {
            mx = z.cx + (t == 1 ? -d : t == 2 ? d : 0);
            my = z.cy + (t == 0 ? -d : t == 3 ? d : 0);
        } else {
            const int idx = ((t - 4) >> 2) + 1, k = (t - 4) & 3, q = (d >> 2) * idx;
            mx = z.cx + ((k & 1) ? q : -q);
            my = z.cy + ((k & 2) ? d - q : q - d);
        }
        return 1;
    }
    case kTTwo:
        mx = z.cx + tab4(kTzOffX, (z.bdir - 1) * 2 + t);
        my = z.cy + tab4(kTzOffY, (z.bdir - 1) * 2 + t);
        return 1;
    default: {   // kTRaster: cells of 4 points (i, j) (i, j + 4) (i + 4, j) (i + 4, j + 4)
        const int c = t >> 2, k = t & 3, cy = c / z.rnx;
        const int i = z.rx + 8 * (c - cy * z.rnx), j = z.ry + 8 * cy;
        const bool all = xme_mv_legal(win, i, j) && xme_mv_legal(win, i, j + 4) && xme_mv_legal(win, i + 4, j) &&
                         xme_mv_legal(win, i + 4, j + 4);
        mx = i + ((k & 2) ? 4 : 0);
        my = j + ((k & 1) ? 4 : 0);
        if (!all)
            return 1;
        mx += z.ox;
        my += z.oy;
        return 3;
    }
    }
}

// A pattern search call (me.c:946-948, :1001): mv_info from the best so far.
__device__ __forceinline__ void tz_call(TzState &z, int bcost, int bmx, int early)
{
    z.st = kTRound;
    z.dist = 1;
    z.rounds = 0;
    z.early = early;
    z.r0 = bcost;
    z.bdist = z.bdir = 0;
    z.cx = bmx;     // tz_pattern_search's omy is mv->bmx (me.c:689-690)
    z.cy = bmx;
}

// After the first pattern search and its two-point search: the raster step
// for a best beyond RasterDistance (me.c:980-995), then the while loop's
// first test (me.c:997).
__device__ __forceinline__ void tz_after_first(TzState &z, const MvWindow &win, int bcost, int bmx, int bmy,
                                               int &phase)
{
    if (z.bdist > 16) {
        const int y0 = max(win.min_fpel[1], bmy - 14), y1 = min(win.max_fpel[1], bmy + 14);
        const int x0 = max(win.min_fpel[0], bmx - 14), x1 = min(win.max_fpel[0], bmx + 14);
        z.rny = y1 > y0 ? (y1 - y0 + 7) >> 3 : 0;
        z.rnx = x1 > x0 ? (x1 - x0 + 7) >> 3 : 0;
        if (z.rnx > 0 && z.rny > 0) {
            z.st = kTRaster;
            z.rx = x0;
            z.ry = y0;
            return;
        }
    }
    if (z.bdist > 0) {
        tz_call(z, bcost, bmx, 32);
        return;
    }
    z.st = kTOff;
    phase = kHexFirst;
}

// The decision after the last batch of a TZ stage.
__device__ __forceinline__ void tz_next(TzState &z, const DevParams &p, const MvWindow &win, int bcost, int bmx,
                                        int bmy, int &phase)
{
    z.bi = 0;
    switch (z.st) {
    case kT6:
        if (xme_mv_legal(win, bmx, bmy)) {
            z.st = kTDia;     // DIA_ITER sets omx, omy (me.c:330-331)
            z.cx = z.ox = bmx;
            z.cy = z.oy = bmy;
            return;
        }
        tz_call(z, bcost, bmx, 3);
        return;
    case kTDia:
        tz_call(z, bcost, bmx, 3);
        return;
    case kTRound: {
        bool end = false;
        if (bcost < z.r0)
            z.rounds = 0;
        else if (++z.rounds >= z.early)
            end = true;
        const int nd = z.dist == 1 ? 2 : 2 * z.dist;   // 1, 2, 4, 8, then 16 .. merange
        if (!end && (nd <= 8 || nd <= p.R)) {
            z.dist = nd;
            z.r0 = bcost;
            z.cx = bmx;
            z.cy = bmx;
            return;
        }
        if (z.early == 3) {      // the first pattern search (me.c:949-977)
            if (z.bdist == 1) {
                if (!z.bdir) {
                    z.st = kTOff;
                    phase = kDone;
                    return;
                }
                z.first = true;
                z.call_cost = bcost;
                z.st = kTTwo;
                z.cx = bmx;
                z.cy = bmy;
                return;
            }
            tz_after_first(z, win, bcost, bmx, bmy, phase);
            return;
        }
        // a pattern search of the while loop (me.c:1002-1024)
        if (z.bdist == 1 && z.bdir) {
            z.first = false;
            z.st = kTTwo;
            z.cx = bmx;
            z.cy = bmy;
            return;
        }
        if (z.bdist > 1) {
            tz_call(z, bcost, bmx, 32);
            return;
        }
        z.st = kTOff;
        phase = kHexFirst;
        return;
    }
    case kTTwo:
        if (z.first) {
            if (bcost == z.call_cost) {   // no new point: the search stops (me.c:974-976)
                z.st = kTOff;
                phase = kDone;
                return;
            }
            tz_after_first(z, win, bcost, bmx, bmy, phase);
            return;
        }
        z.st = kTOff;
        phase = kHexFirst;
        return;
    default:   // kTRaster: the while loop runs (bdist > 16)
        tz_call(z, bcost, bmx, 32);
        return;
    }
}

// The prefix stages as the lane-parallel loop below sees them: UMH's
// (umh_point, umh_done_stage, umh_next) and TZ's.
struct UmhMachine {
    UmhState u;
    xme_umh_ctx c;
    __device__ bool live() const { return u.st != kUOff; }
    __device__ int batch() const { return u.bi; }
    __device__ int points(const DevParams &p) const { return umh_stage_points(u.st, p.R, p.big_hex); }
    __device__ int point(const DevParams &p, const MvWindow &, int t, int &mx, int &my) const
    {
        return umh_point(u, c, p.R, p.big_hex, t, mx, my);
    }
    __device__ void won(int) {}
    __device__ void undefined() {}
    __device__ void next_batch(const DevParams &p, const MvWindow &, int S, int bcost, int bmx, int bmy,
                               int &phase)
    {
        u.bi++;
        if (u.bi * S >= points(p)) {
            umh_done_stage(u, bmx, bmy);
            umh_next(u, c, p.R, p.big_hex, bcost, bmx, bmy, phase);
        }
    }
};

struct TzMachine {
    TzState z;
    __device__ bool live() const { return z.st != kTOff; }
    __device__ int batch() const { return z.bi; }
    __device__ int points(const DevParams &) const { return tz_stage_points(z); }
    __device__ int point(const DevParams &, const MvWindow &win, int t, int &mx, int &my) const
    {
        return tz_point(z, win, t, mx, my);
    }
    // mv_info's bdir / bdist follow a round's winning point (COPY5_IF_LT)
    __device__ void won(int t)
    {
        if (z.st != kTRound)
            return;
        const int d = z.dist;
        if (d <= 8) {
            const int e = d == 1 ? (t == 0 ? 0 : t == 1 ? 3 : t == 2 ? 4 : 7) : t;
            z.bdir = tab4(kTz8D, e);
            z.bdist = (e == 1 || e == 2 || e == 5 || e == 6) ? d >> 1 : d;
        } else {
            z.bdir = 0;
            z.bdist = d;
        }
    }
    __device__ void undefined() { z.undef = true; }
    __device__ void next_batch(const DevParams &p, const MvWindow &win, int S, int bcost, int bmx, int bmy,
                               int &phase)
    {
        z.bi++;
        if (z.bi * S >= points(p))
            tz_next(z, p, win, bcost, bmx, bmy, phase);
    }
};

// The prefix steps (UMH's or TZ's) of every live group of the wave, up to the
// phase each one goes on with (kHexFirst, kDiaFirst, or kDone).  Each lane of
// a group takes one point of the stage per round and sums the whole PU there,
// with the current PU staged in the group's LDS share (`cur`, gw dwords a
// row, read by every lane of the group at once), so a round evaluates S
// points.  A round's points are taken as the reference takes them one by one:
// the first strict minimum, i.e. the smallest (cost, point) key, if its cost
// is below the running best (COPY3_IF_LT in order, ME_COST_IPEL, me.c:187-193).
// A point is range checked first unless it is one of the unchecked ones:
// UMH's and TZ's diamond points lie one pixel from their centre, a window
// point or the best of an earlier diamond, and TZ's six start points two
// pixels from a window point, so at most 3 pixels outside the window.  The
// window keeps 6 pixels inside the 64-pixel search margin (md_inter.c:913),
// so those reads (straight from the plane) stay inside its 80-pixel padding.
// TZ's offset raster points are tested against the padding and the mvbits
// table and skipped outside them (the reference reads outside its arrays).
//
// LP > 1 (64-lane groups): LP lanes share a point, lane j of them summing
// quads j, j + LP, ... of every row, so that a row of a point is one
// coalesced read; a round then evaluates S / LP points.
template <int S, int GWMAX, int LP, class M>
__device__ __forceinline__ void prefix_lanes(const DevParams &p, const DevPu &pu, M &m, const MvWindow &win,
                                             gbytes ref, int64_t i_ref, const uint32_t *cur, int &phase, int &bcost,
                                             int &bmx, int &bmy, uint32_t &evals)
{
    const int lane = threadIdx.x & 63;
    const int l = lane & (S - 1);
    const int gw = pu.w >> 2;
    const uint64_t gmask = (S == 64 ? ~0ull : ((1ull << S) - 1)) << (lane & ~(S - 1));
    for (;;) {
        if (__builtin_amdgcn_ballot_w64(m.live()) == 0)
            break;
        const bool live = m.live();
        const int n_st = live ? m.points(p) : 0;
        constexpr int PR = S / LP;             // points per round
        const int pi = l / LP, j = l % LP;     // point of the round, lane within it
        const int t = m.batch() * PR + pi;
        int mx = 0, my = 0;
        const int kind = live && t < n_st ? m.point(p, win, t, mx, my) : 0;
        bool inside = true;
        if (kind == 3) {
            const int ax = pu.x + mx, ay = pu.y + my;
            // a row's dword reads reach 4 bytes past the block
            inside = ax >= -XME_PAD && ax + pu.w + 4 <= p.frame_w + XME_PAD && ay >= -XME_PAD &&
                     ay + pu.h <= p.frame_h + XME_PAD && abs(4 * mx - pu.mvp_x) <= p.max_mvd &&
                     abs(4 * my - pu.mvp_y) <= p.max_mvd;
        }
        if (__builtin_amdgcn_ballot_w64(kind == 3 && !inside) & gmask)
            m.undefined();
        const bool ok = kind == 2 || (kind == 3 && inside) || (kind == 1 && xme_mv_legal(win, mx, my));
        int key = 0x7fffffff;
        if constexpr (LP > 1) {
            uint32_t sad = 0;
            if (ok && j < gw) {
                const uintptr_t a = (uintptr_t)(ref + (int64_t)my * i_ref + mx + 4 * j);
                const gdwords q0 = (gdwords)(a & ~(uintptr_t)3);
                const uint32_t sh = (uint32_t)(a & 3);
                const int64_t rstep = i_ref >> 2;
                constexpr int KK = GWMAX / LP > 1 ? GWMAX / LP : 1;   // quads of a row per lane
#pragma unroll 4
                for (int r = 0; r < pu.h; r++) {
                    const gdwords rp = q0 + r * rstep;
#pragma unroll
                    for (int kk = 0; kk < KK; kk++)
                        if (j + kk * LP < gw)
                            sad = __builtin_amdgcn_sad_u8(cur[r * gw + j + kk * LP],
                                                          __builtin_amdgcn_alignbyte(rp[kk * LP + 1], rp[kk * LP], sh), sad);
                }
            }
            const int sum = group_sum((int)sad, LP);    // in the point's last lane
            if (ok && j == LP - 1) {
                const int cost = sum + xme_weighted_cost(p.lambda, xme_mvbits(4 * mx - pu.mvp_x) + xme_mvbits(4 * my - pu.mvp_y));
                key = (cost << 6) | pi;
            }
        } else if (ok) {
            const uintptr_t a = (uintptr_t)(ref + (int64_t)my * i_ref + mx);
            const gdwords row0 = (gdwords)(a & ~(uintptr_t)3);
            const uint32_t sh = (uint32_t)(a & 3);
            const int64_t rstep = i_ref >> 2;
            uint32_t sad = 0;
            for (int r = 0; r < pu.h; r++) {
                const gdwords rp = row0 + r * rstep;
                uint32_t d[GWMAX + 1];
#pragma unroll
                for (int k = 0; k <= GWMAX; k++)
                    d[k] = k <= gw ? rp[k] : 0u;
#pragma unroll
                for (int k = 0; k < GWMAX; k++)
                    if (k < gw)
                        sad = __builtin_amdgcn_sad_u8(cur[r * gw + k], __builtin_amdgcn_alignbyte(d[k + 1], d[k], sh), sad);
            }
            const int cost = (int)sad + xme_weighted_cost(p.lambda, xme_mvbits(4 * mx - pu.mvp_x) + xme_mvbits(4 * my - pu.mvp_y));
            // costs stay below 2^21 (a 64x64 PU: 1,044,480 + lambda bits),
            // the point of the round takes the low 6 bits
            key = (cost << 6) | pi;
        }
        int kmin = group_min(key, S);
        kmin = __builtin_amdgcn_ds_bpermute((lane | (S - 1)) << 2, kmin);
        // the group's points of the round, the same count in every lane
        int nok = group_sum(ok && j == 0 ? 1 : 0, S);
        evals += (uint32_t)__builtin_amdgcn_ds_bpermute((lane | (S - 1)) << 2, nok);
        if (live && kmin != 0x7fffffff && (kmin >> 6) < bcost) {
            bcost = kmin >> 6;
            const int tw = m.batch() * PR + (kmin & 63);
            int wx = 0, wy = 0;
            m.point(p, win, tw, wx, wy);
            bmx = wx;
            bmy = wy;
            m.won(tw);
        }
        if (live)
            m.next_batch(p, win, PR, bcost, bmx, bmy, phase);
    }
}

// MODE: 0 DIA / HEX, 1 UMH (its prefix, then the hexagon or the diamond),
// 2 TZ (its prefix, then the hexagon and the diamond, or done).
template <int S, int QMAX, int NB, bool SEQ, int MODE>
__device__ __forceinline__ void pattern_wave(const DevParams &p, const FrameDesc *__restrict__ desc, const DevPu *__restrict__ pus,
                             DevInst *__restrict__ inst, uint64_t *__restrict__ best, int id,
                             uint32_t *__restrict__ points, uint32_t *__restrict__ box,
                             const xme_umh_ctx *__restrict__ umh)
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
            phase = p.method == XME_ME_DIA ? kDiaFirst : kHexFirst;
        }
    }
    const int r = id >= 0 ? id - (id / p.n_ref) * p.n_ref : 0;
    const MvWindow win = xme_mv_window(p, pu.x, pu.y, pu.w, pu.h);
    const int64_t i_ref = desc->stride[r];
    const gbytes plane = (gbytes)desc->planes[r][0];
    // the group's box in LDS: pu.h + 4 rows of DW dwords, in the group's
    // fixed share of the wave's region (the groups of a wave may hold PUs of
    // different shapes); this lane stages slots l, l + S, ... of it, starting
    // at row rr0, dword cc0 and moving on by SR rows and SC dwords per slot
    const int DW = (pu.w + 10) >> 2;
    uint32_t *gbox = box + (lane / S) * (XME_PAT_BOX_DW * S / 64);
    const int nbox = (pu.h + 4) * DW;
    const int rr0 = l / DW, cc0 = l - rr0 * DW, SR = S / DW, SC = S - SR * DW;
    const int32_t i_ref32 = (int32_t)i_ref;

    // this lane's pixel quads of the PU: current dwords and byte offsets of
    // the quads in the box
    const int gw = pu.w >> 2, quads = gw * pu.h;
    uint32_t c4[QMAX];
    int32_t qoff[QMAX];
    int nq = 0;
    {
        const gbytes cur = (gbytes)desc->cur + (int64_t)pu.y * desc->i_cur + pu.x;
#pragma unroll
        for (int q = 0; q < QMAX; q++) {
            const int u = l + q * S;
            c4[q] = 0;
            qoff[q] = 0;
            if (phase != kDone && u < quads) {
                const int yy = u / gw, xx = (u - yy * gw) * 4;
                c4[q] = *(gdwords)(cur + (int64_t)yy * desc->i_cur + xx);
                qoff[q] = yy * DW * 4 + xx;
                nq = q + 1;
            }
        }
    }

    int idx = 0, hex_i = 0, dia_i = 0;
    uint32_t evals = 0;
    bool undef = false;
    if constexpr (MODE != 0) {
        // the group's current quads, row-major, in its LDS share (free until
        // the first step box; at most 1024 dwords for 64x64 in 1224)
#pragma unroll
        for (int q = 0; q < QMAX; q++)
            if (q < nq)
                gbox[l + q * S] = c4[q];
        __builtin_amdgcn_wave_barrier();
        // 64-lane groups (PUs of 192 quads and more): 16 lanes a point.  For
        // the smaller PUs a lane a point measured faster: sharing a point
        // cuts the points of a round, and rounds carry the stage logic
        constexpr int GW = S <= 8 ? 4 : S <= 32 ? 8 : 16, LP = S == 64 ? 16 : 1;
        const gbytes ref0 = plane + (int64_t)pu.y * i_ref + pu.x;
        const bool dfree = xme_mv_legal(win, 0, -1) && xme_mv_legal(win, -1, 0) && xme_mv_legal(win, 1, 0) &&
                           xme_mv_legal(win, 0, 1);
        if constexpr (MODE == 1) {
            UmhMachine m;
            m.c = {};
            if (phase != kDone)
                m.c = umh[id];
            UmhState &u = m.u;
            u.st = phase == kDone ? kUOff : kUDia0;
            u.bi = 0;
            u.cx = u.pmx = u.p0x = in.mvc[0][0];
            u.cy = u.pmy = u.p0y = in.mvc[0][1];
            u.level = 0;
            u.dfree = dfree;
            if (u.st != kUOff)
                phase = kDone;   // set again when the prefix ends
            prefix_lanes<S, GW, LP>(p, pu, m, win, ref0, i_ref, gbox, phase, bcost, bmx, bmy, evals);
        } else {
            TzMachine m;
            TzState &z = m.z;
            z = {};
            z.st = phase == kDone ? kTOff : kT6;
            z.cx = z.ox = bmx;   // omx, omy = the best candidate (me.c:935-936)
            z.cy = z.oy = bmy;
            z.dfree = dfree;
            if (z.st != kTOff)
                phase = kDone;   // set again when the prefix ends
            prefix_lanes<S, GW, LP>(p, pu, m, win, ref0, i_ref, gbox, phase, bcost, bmx, bmy, evals);
            undef = z.undef;
            // an offset raster point may leave the best anywhere in the padded
            // plane: the hexagon's box (the block grown by 2 pixels, whole
            // dwords) must lie in it, or the reference reads outside it
            const int x0 = (pu.x + bmx - 2) & ~3;
            if (phase == kHexFirst &&
                !(x0 >= -XME_PAD && x0 + 4 * DW <= p.frame_w + XME_PAD && pu.y + bmy - 2 >= -XME_PAD &&
                  pu.y + bmy + pu.h + 2 <= p.frame_h + XME_PAD)) {
                undef = true;
                phase = kDone;
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
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
        // every point of a step lies within 2 pixels of the current best,
        // which is inside the MV window or (after UMH's unchecked diamonds)
        // at most 3 pixels outside it: the box (PU + 2 pixels a side, from
        // the dword below) stays inside the plane's padding
        const int x0 = pu.x + bmx - 2, ax0 = x0 & ~3, sh = x0 - ax0;
        {
            // in batches of at most 8 loads in flight per lane (registers)
            const gbytes rows = plane + (int64_t)(pu.y + bmy - 2) * i_ref + ax0;
            constexpr int NBB = NB < 8 ? NB : 8;
            // the slot offsets are the same every step: opaque copies of the
            // start keep the compiler from holding all NB of them in
            // registers across the loop
            int lo = l, rr = rr0, cc = cc0;
            asm volatile("" : "+v"(lo), "+v"(rr), "+v"(cc));
#pragma unroll
            for (int b0 = 0; b0 < NB; b0 += NBB) {
                uint32_t v[NBB];
#pragma unroll
                for (int b = 0; b < NBB; b++) {
                    const int t = lo + (b0 + b) * S;
                    v[b] = 0;
                    if (phase != kDone && b0 + b < NB && t < nbox)
                        v[b] = *(gdwords)(rows + (rr * i_ref32 + 4 * cc));
                    cc += SC;
                    rr += SR;
                    if (cc >= DW) {
                        cc -= DW;
                        rr++;
                    }
                }
#pragma unroll
                for (int b = 0; b < NBB; b++) {
                    const int t = lo + (b0 + b) * S;
                    if (phase != kDone && b0 + b < NB && t < nbox)
                        gbox[t] = v[b];
                }
                // keeps the next batch's loads after these stores
                __builtin_amdgcn_wave_barrier();
            }
        }
        __builtin_amdgcn_wave_barrier();
        const int n = phase == kHexFirst ? 6 : phase == kDiaFirst ? 4 : phase == kDone ? 0 : 3;
        int dir = 0, nbx = bmx, nby = bmy;
        if constexpr (SEQ) {
            // one point at a time: few registers for PUs of many quads
#pragma unroll 1
            for (int k = 0; k < 6; k++) {
                int dx, dy, dc;
                point_of(phase, idx, k, dx, dy, dc);
                uint32_t sad = 0;
                if (k < n) {
                    const int pb = (dy + 2) * DW * 4 + sh + dx + 2;
#pragma unroll
                    for (int q = 0; q < QMAX; q++) {
                        if (q < nq) {
                            const int a = qoff[q] + pb;
                            const uint32_t d0 = gbox[a >> 2], d1 = gbox[(a >> 2) + 1];
                            sad = __builtin_amdgcn_sad_u8(c4[q], __builtin_amdgcn_alignbyte(d1, d0, (uint32_t)(a & 3)), sad);
                        }
                    }
                }
                int sum = group_sum((int)sad, S);
                sum = __builtin_amdgcn_ds_bpermute((lane | (S - 1)) << 2, sum);
                if (k < n)
                    take_point(p, pu, bmx + dx, bmy + dy, dc, sum, bcost, nbx, nby, dir);
            }
        } else {
            // SADs of all the step's points, then their sums in order
            uint32_t sad[6];
            int dxs[6], dys[6], dcs[6];
#pragma unroll
            for (int k = 0; k < 6; k++) {
                point_of(phase, idx, k, dxs[k], dys[k], dcs[k]);
                const int dx = dxs[k], dy = dys[k];
                sad[k] = 0;
                if (k < n) {
                    const int pb = (dy + 2) * DW * 4 + sh + dx + 2;
#pragma unroll
                    for (int q = 0; q < QMAX; q++) {
                        if (q < nq) {
                            const int a = qoff[q] + pb;
                            const uint32_t d0 = gbox[a >> 2], d1 = gbox[(a >> 2) + 1];
                            sad[k] = __builtin_amdgcn_sad_u8(c4[q], __builtin_amdgcn_alignbyte(d1, d0, (uint32_t)(a & 3)),
                                                             sad[k]);
                        }
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < 6; k++) {
                int sum = group_sum((int)sad[k], S);
                sum = __builtin_amdgcn_ds_bpermute((lane | (S - 1)) << 2, sum);
                if (k < n)
                    take_point(p, pu, bmx + dxs[k], bmy + dys[k], dcs[k], sum, bcost, nbx, nby, dir);
            }
        }
        // the next step's box overwrites this one's only after every lane's reads
        __builtin_amdgcn_wave_barrier();
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
    if (id >= 0 && l == S - 1) {
        if (evals) {
            // the sub-pel pass reads the integer best as the centre's raster
            // candidate (0, 0): order XME_ORDER_RASTER + R * 2R + R
            inst[id].omx = (int16_t)bmx;
            inst[id].omy = (int16_t)bmy;
            if (undef)
                inst[id].flags = 1;
            best[id] = xme_key((uint32_t)bcost, (uint32_t)(XME_ORDER_RASTER + p.R * 2 * p.R + p.R));
        }
        points[id] = evals;
    }
}

// Wave ranges of one launch, largest PUs first: range k holds PUs of
// xme_pattern_class 5 - k, (64 / S) instance ids per wave from list[first[k]].
struct PatLaunch {
    int end[6];      // first wave after range k
    int first[6];    // list index of range k's first id
};

// Launch ranges by class (quads = w * h / 4 of the PU), with their largest
// box of (rows + 4) x ((w + 10) / 4) dwords per group, which fits the group's
// share XME_PAT_BOX_DW * S / 64 (76, 153, 306, 612, 1224 dwords for S = 4,
// 8, 16, 32, 64), and the staging slots per lane:
//  5: > 256 quads, S 64, <= 16 quads and <= 20 slots per lane (64x64: 1224);
//  4: <= 256, S 64, 7 slots (16x64: 408);  3: <= 128, S 32, 7 (16x32: 216);
//  2: <= 64, S 16, 9 (8x32: 144);  1: <= 32, S 8, 10 (8x16: 80);
//  0: <= 16, S 4, 15 (4x16: 60).  Every lane holds at most 4 quads of the
// PU below class 5.
template <int MODE>
__global__ void __launch_bounds__(256) pattern_kernel(DevParams p, const FrameDesc *__restrict__ desc,
                                                      const DevPu *__restrict__ pus, DevInst *__restrict__ inst,
                                                      uint64_t *__restrict__ best, const int32_t *__restrict__ list,
                                                      PatLaunch L, uint32_t *__restrict__ points,
                                                      const xme_umh_ctx *__restrict__ umh)
{
    __shared__ uint32_t boxes[4][XME_PAT_BOX_DW];
    // Ranges are whole workgroups.  Workgroup b runs on XCD b % 8; within
    // its range, the workgroups of one XCD take a contiguous run of the
    // range's list (a bijection on the range), so that one L2 sees
    // neighbouring PUs of one reference.
    int w;
    {
        const int b = blockIdx.x;
        int k = 0;
        while (k < 5 && b * 4 >= L.end[k])
            k++;
        const int gs = (k == 0 ? 0 : L.end[k - 1]) >> 2, G = (L.end[k] >> 2) - gs;
        const int g = b - gs, x = g & 7, y = g >> 3, per = G >> 3, rem = G & 7;
        const int gp = x * per + min(x, rem) + y;
        w = __builtin_amdgcn_readfirstlane((gs + gp) * 4 + (int)(threadIdx.x >> 6));
    }
    const int lane = threadIdx.x & 63;
    uint32_t *box = boxes[threadIdx.x >> 6];
    if (w < L.end[0]) {
        pattern_wave<64, 16, 20, true, MODE>(p, desc, pus, inst, best, list[L.first[0] + w], points, box, umh);
    } else if (w < L.end[1]) {
        pattern_wave<64, 4, 7, false, MODE>(p, desc, pus, inst, best, list[L.first[1] + (w - L.end[0])], points, box, umh);
    } else if (w < L.end[2]) {
        pattern_wave<32, 4, 7, false, MODE>(p, desc, pus, inst, best, list[L.first[2] + (w - L.end[1]) * 2 + lane / 32],
                                      points, box, umh);
    } else if (w < L.end[3]) {
        pattern_wave<16, 4, 9, false, MODE>(p, desc, pus, inst, best, list[L.first[3] + (w - L.end[2]) * 4 + lane / 16],
                                      points, box, umh);
    } else if (w < L.end[4]) {
        pattern_wave<8, 4, 10, false, MODE>(p, desc, pus, inst, best, list[L.first[4] + (w - L.end[3]) * 8 + lane / 8],
                                      points, box, umh);
    } else if (w < L.end[5]) {
        pattern_wave<4, 4, 15, false, MODE>(p, desc, pus, inst, best, list[L.first[5] + (w - L.end[4]) * 16 + lane / 4],
                                      points, box, umh);
    }
}

}  // namespace

int xme_pattern_class(int w, int h)
{
    const int q = w * h / 4;
    return q <= 16 ? 0 : q <= 32 ? 1 : q <= 64 ? 2 : q <= 128 ? 3 : q <= 256 ? 4 : 5;
}

int xme_launch_pattern(const DevParams &p, const FrameDesc *d_desc, const DevPu *d_pus, DevInst *d_inst,
                       uint64_t *d_best, const int32_t *d_list, const int wave_end[6], uint32_t *d_points,
                       const xme_umh_ctx *d_umh, hipStream_t s)
{
    const int waves = wave_end[5];
    if (waves <= 0)
        return XME_OK;
    // the list holds the ranges back to back, 64 / S ids per wave
    static const int per_wave[6] = { 1, 1, 2, 4, 8, 16 };
    PatLaunch L;
    int first = 0, prev = 0;
    for (int k = 0; k < 6; k++) {
        L.end[k] = wave_end[k];
        L.first[k] = first;
        first += (wave_end[k] - prev) * per_wave[k];
        prev = wave_end[k];
    }
    if (p.method == XME_ME_UMH)
        hipLaunchKernelGGL(pattern_kernel<1>, dim3((waves + 3) / 4), dim3(256), 0, s, p, d_desc, d_pus, d_inst,
                           d_best, d_list, L, d_points, d_umh);
    else if (p.method == XME_ME_TZ)
        hipLaunchKernelGGL(pattern_kernel<2>, dim3((waves + 3) / 4), dim3(256), 0, s, p, d_desc, d_pus, d_inst,
                           d_best, d_list, L, d_points, d_umh);
    else
        hipLaunchKernelGGL(pattern_kernel<0>, dim3((waves + 3) / 4), dim3(256), 0, s, p, d_desc, d_pus, d_inst,
                           d_best, d_list, L, d_points, d_umh);
    XME_HIP(hipGetLastError());
    return XME_OK;
}
