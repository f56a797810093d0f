/*
 * xme_oracle.c -- CPU restatement of xavs2's motion-estimation cost path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Used by tests/, by
 * __graft_entry__.smoke() and by the CPU-baseline leg of bench.py; the product
 * library never links or calls it.
 *
 * Reference: pkuvcl/xavs2 (XAVS2_BUILD 13), 8-bit build (BIT_DEPTH 8).  The
 * reference cannot be compiled here from its own sources without the
 * configure-generated config.h (source/common/osdep.h:64), so this file is
 * pinned by independent checks instead (tests/test_oracle_*.py):
 *   - xo_sad / sad_x3 / sad_x4 / avg : brute force in numpy.
 *   - xo_satd       : H4.D.H4^T matrix form in numpy, constant-difference closed
 *                     form (8|d| per 4x4), extreme (+-255) patterns.
 *   - xo_mvbits     : Exp-Golomb se(v) length, and the table-size formula.
 *   - xo_me_search  : numpy brute force (lexicographic (cost, order) minimum),
 *                     engineered ties (flat frame, period-2 texture), PMVR on/off.
 *   - xo_interpolate_frame : numpy separable-FIR closed form with the border
 *                     clamp model plane_k(x,y) = interp_k(clamp(x,-4,W+3),
 *                     clamp(y,-4,H+3)); flat frames; full-pel copy.
 *   - luma/chroma block filters, mc_chroma : numpy FIR closed form.
 *   - xo_dct / xo_idct : table symmetries that make the butterflies of
 *                     transform.c equal to these matrix products, orientation
 *                     (row-only variation -> first coefficient row), DC of a
 *                     flat block, round trip, int16 wrap on adversarial input.
 *   - quant/dequant : scalar formula in numpy, table relations Q*IQ ~ 2^(15+s).
 */
#include "xme_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* PU table, common/pixel.h:50-65 (enum LumaPU order).                       */
const uint8_t xo_pu_w[XO_NUM_PU] = { 4, 8, 16, 32, 64, 8, 4, 16, 8, 32, 16, 64, 32,
                                     16, 12, 16, 4, 32, 24, 32, 8, 64, 48, 64, 16 };
const uint8_t xo_pu_h[XO_NUM_PU] = { 4, 8, 16, 32, 64, 4, 8, 8, 16, 16, 32, 32, 64,
                                     12, 16, 4, 16, 24, 32, 8, 32, 48, 64, 16, 64 };

int xo_part_index(int w, int h)
{
    for (int i = 0; i < XO_NUM_PU; i++)
        if (xo_pu_w[i] == w && xo_pu_h[i] == h)
            return i;
    return -1;
}

static inline int clip3(int lo, int hi, int v) { return v < lo ? lo : (v > hi ? hi : v); }
static inline int clip_pel(int v) { return v < 0 ? 0 : (v > 255 ? 255 : v); }
static inline int iabs(int v) { return v < 0 ? -v : v; }

/* ------------------------------------------------------------------------ */
/* SAD, common/pixel.c:93-135: sum over the block of |p1 - p2|.              */
int32_t xo_sad(int w, int h, const uint8_t *p1, intptr_t s1, const uint8_t *p2, intptr_t s2)
{
    int32_t sum = 0;
    for (int y = 0; y < h; y++, p1 += s1, p2 += s2) {
        int32_t row = 0;
        for (int x = 0; x < w; x++)
            row += iabs((int)p1[x] - (int)p2[x]);
        sum += row;
    }
    return sum;
}

/* SAD_x3 / SAD_x4, pixel.c:143-243: fenc stride is fixed at FENC_STRIDE. */
void xo_sad_x3(int w, int h, const uint8_t *fenc, const uint8_t *p0, const uint8_t *p1,
               const uint8_t *p2, intptr_t i_ref, int32_t res[3])
{
    res[0] = xo_sad(w, h, fenc, XO_FENC_STRIDE, p0, i_ref);
    res[1] = xo_sad(w, h, fenc, XO_FENC_STRIDE, p1, i_ref);
    res[2] = xo_sad(w, h, fenc, XO_FENC_STRIDE, p2, i_ref);
}

void xo_sad_x4(int w, int h, const uint8_t *fenc, const uint8_t *p0, const uint8_t *p1,
               const uint8_t *p2, const uint8_t *p3, intptr_t i_ref, int32_t res[4])
{
    res[0] = xo_sad(w, h, fenc, XO_FENC_STRIDE, p0, i_ref);
    res[1] = xo_sad(w, h, fenc, XO_FENC_STRIDE, p1, i_ref);
    res[2] = xo_sad(w, h, fenc, XO_FENC_STRIDE, p2, i_ref);
    res[3] = xo_sad(w, h, fenc, XO_FENC_STRIDE, p3, i_ref);
}

/* SATD, pixel.c:253-407.  The reference packs two 16-bit lanes per 32-bit
 * word; for 8-bit input every transformed value lies in [-4080, 4080] and a
 * lane accumulates at most 16 of them (<= 65280), so the packing is lossless
 * and the result equals: for every 4x4 tile, D = A - B, T = H4 D H4^T, and
 * SATD = (sum over tiles of sum |T|) / 2.  Each tile sum is even (the last
 * butterfly stage yields pairs a+b, a-b with |a+b|+|a-b| = 2 max(|a|,|b|)),
 * so the per-4x4 (:321) and per-8x4 (:347) halvings are exact and the tiling
 * (4x4 tiles for 4/12-wide blocks, 8x4 otherwise) does not change the value. */
static int satd4x4_sum(const uint8_t *p1, intptr_t s1, const uint8_t *p2, intptr_t s2)
{
    int d[4][4], t[4][4];
    for (int y = 0; y < 4; y++)
        for (int x = 0; x < 4; x++)
            d[y][x] = (int)p1[y * s1 + x] - (int)p2[y * s2 + x];
    /* horizontal 4-point Hadamard on each row */
    for (int y = 0; y < 4; y++) {
        int a = d[y][0] + d[y][1], b = d[y][0] - d[y][1];
        int c = d[y][2] + d[y][3], e = d[y][2] - d[y][3];
        t[y][0] = a + c; t[y][1] = b + e; t[y][2] = a - c; t[y][3] = b - e;
    }
    int sum = 0;
    /* vertical 4-point Hadamard on each column */
    for (int x = 0; x < 4; x++) {
        int a = t[0][x] + t[1][x], b = t[0][x] - t[1][x];
        int c = t[2][x] + t[3][x], e = t[2][x] - t[3][x];
        sum += iabs(a + c) + iabs(b + e) + iabs(a - c) + iabs(b - e);
    }
    return sum;
}

int32_t xo_satd(int w, int h, const uint8_t *p1, intptr_t s1, const uint8_t *p2, intptr_t s2)
{
    int32_t sum = 0;
    for (int y = 0; y < h; y += 4)
        for (int x = 0; x < w; x += 4)
            sum += satd4x4_sum(p1 + y * s1 + x, s1, p2 + y * s2 + x, s2);
    return sum >> 1;
}

/* AVG, pixel.c:680-719: (a + b + 1) >> 1 (the weight argument is unused). */
void xo_avg(int w, int h, uint8_t *dst, intptr_t ds, const uint8_t *s0, intptr_t ss0,
            const uint8_t *s1, intptr_t ss1)
{
    for (int y = 0; y < h; y++, dst += ds, s0 += ss0, s1 += ss1)
        for (int x = 0; x < w; x++)
            dst[x] = (uint8_t)((s0[x] + s1[x] + 1) >> 1);
}

void xo_sad_batch(int w, int h, const uint8_t *p1, intptr_t s1, const uint8_t *p2, intptr_t s2,
                  const int64_t *off1, const int64_t *off2, int n, int32_t *out)
{
    for (int i = 0; i < n; i++)
        out[i] = xo_sad(w, h, p1 + off1[i], s1, p2 + off2[i], s2);
}

void xo_satd_batch(int w, int h, const uint8_t *p1, intptr_t s1, const uint8_t *p2, intptr_t s2,
                   const int64_t *off1, const int64_t *off2, int n, int32_t *out)
{
    for (int i = 0; i < n; i++)
        out[i] = xo_satd(w, h, p1 + off1[i], s1, p2 + off2[i], s2);
}

/* ------------------------------------------------------------------------ */
/* MV rate, encoder/me.c:592-641.  The table size follows
 * xavs2_me_get_buf_size / xavs2_me_init with me_range = max(256, R). */
static int mvbits_max_bits(int search_range)
{
    int me_range   = search_range > 256 ? search_range : 256;
    int subpel_num = 4 * (2 * me_range + 3);
    return 5 + 2 * (int)ceil(log(subpel_num + 1) / log(2) + 1e-10);
}

int xo_mvbits_max_mvd(int search_range)
{
    return (1 << (mvbits_max_bits(search_range) >> 1)) - 1;
}

int xo_mvbits(int search_range, int v)
{
    /* mvbits[0] = 1; entries with imin <= |v| < imax get 'bits', for
     * bits = 3, 5, ..., max_mv_bits and imax = 1 << (bits >> 1). */
    int a = iabs(v);
    if (a == 0)
        return 1;
    int max_bits = mvbits_max_bits(search_range);
    for (int bits = 3; bits <= max_bits; bits += 2) {
        int imax = 1 << (bits >> 1);
        if (a >= (imax >> 1) && a < imax)
            return bits;
    }
    return -1; /* outside the reference table */
}

int xo_lambda_factor(double lambda_mode)
{
    return (int)((double)(1 << 16) * sqrt(lambda_mode) + 0.5);
}

/* ------------------------------------------------------------------------ */
/* ME search.                                                                */

/* CHECK_MV_RANGE, encoder/me.h:46-48, with mv_min / mv_max packed as in
 * xavs2_me_search (me.c:896-897).  Restated bit for bit, since the packed
 * test is what the reference evaluates. */
static inline uint32_t pack_mask2(int mx, int my)
{
    return ((uint32_t)mx << 16) | ((uint32_t)my & 0x7FFFu);
}

static inline int check_mv_range(int mx, int my, uint32_t mv_min, uint32_t mv_max)
{
    uint32_t p = pack_mask2(mx, my);
    return !(((p + mv_min) | (mv_max - p)) & 0x80004000u);
}

typedef struct {
    int mv_min_fpel[2], mv_max_fpel[2];
    uint32_t mv_min, mv_max;
} mv_window_t;

/* MV limits of pred_inter_search_single, encoder/md_inter.c:910-926. */
static void pu_mv_window(const xo_me_param_t *prm, const xo_pu_t *pu, mv_window_t *win)
{
    const int m = 6;
    int lo, hi;
    lo = (-64 - pu->pix_x) * 4;
    hi = (prm->frame_w + 64 - pu->pix_x - pu->w) * 4;
    lo = clip3(prm->min_mv_range[0], prm->max_mv_range[0], lo);
    hi = clip3(prm->min_mv_range[0], prm->max_mv_range[0], hi);
    win->mv_min_fpel[0] = (lo >> 2) + m;
    win->mv_max_fpel[0] = (hi >> 2) - m;
    lo = (-64 - pu->pix_y) * 4;
    hi = (prm->frame_h + 64 - pu->pix_y - pu->h) * 4;
    lo = clip3(prm->min_mv_range[1], prm->max_mv_range[1], lo);
    hi = clip3(prm->min_mv_range[1], prm->max_mv_range[1], hi);
    win->mv_min_fpel[1] = (lo >> 2) + m;
    win->mv_max_fpel[1] = (hi >> 2) - m;
    win->mv_min = pack_mask2(-win->mv_min_fpel[0], -win->mv_min_fpel[1]);
    win->mv_max = pack_mask2(win->mv_max_fpel[0], win->mv_max_fpel[1]) | 0x8000u;
}

/* Spiral / Spiral2 patterns, encoder/me.c:129-139. */
static const int8_t SPIRAL[9][2] = {
    { 0, 0 }, { 0, -1 }, { 0, 1 }, { -1, -1 }, { 1, -1 }, { -1, 0 }, { 1, 0 }, { -1, 1 }, { 1, 1 }
};
static const int8_t SPIRAL2[9][2] = {
    { 0, 0 }, { 0, -1 }, { -1, -1 }, { -1, 0 }, { -1, 1 }, { 0, 1 }, { 1, 1 }, { 1, 0 }, { 1, -1 }
};

/* pmvr_adapt_mv, encoder/me.c:412-424 (TH_PMVR = 2, defines.h:227).
 * Returns 1 when the candidate is to be skipped. */
static int pmvr_adapt(int *mx, int *my, int ctr_x, int ctr_y, int mv_x, int mv_y, int sx, int sy)
{
    if (iabs(mv_x - ctr_x) > 2 || iabs(mv_y - ctr_y) > 2) {
        *mx = mv_x + sx * 2;
        *my = mv_y + sy * 2;
        return iabs(*mx - ctr_x) <= 2 && iabs(*my - ctr_y) <= 2;
    }
    *mx = mv_x + sx;
    *my = mv_y + sy;
    return iabs(*mx - ctr_x) > 2 || iabs(*my - ctr_y) > 2;
}

int xo_pmvr_adapt(int *mx, int *my, int ctr_x, int ctr_y, int mv_x, int mv_y, int sx, int sy)
{
    return pmvr_adapt(mx, my, ctr_x, ctr_y, mv_x, mv_y, sx, sy);
}

static inline int weighted_cost(int lambda, int bits) { return (lambda * bits) >> 16; }

/* Radius-2 hexagon and radius-1 diamond with their wrap-around entries, and
 * the (x - 1) mod 6 / mod 4 index tables, encoder/me.c:83-112. */
static const int8_t HEX2[8][2] = {
    { -1, -2 }, { -2, 0 }, { -1, 2 }, { 1, 2 }, { 2, 0 }, { 1, -2 }, { -1, -2 }, { -2, 0 }
};
static const int8_t M1MOD6[8] = { 5, 0, 1, 2, 3, 4, 5, 0 };
static const int8_t DIA1[6][2] = {
    { 0, -1 }, { -1, 0 }, { 0, 1 }, { 1, 0 }, { 0, -1 }, { -1, 0 }
};
static const int8_t M1MOD4[6] = { 3, 0, 1, 2, 3, 0 };

/* UMH: the big hexagons and the 5x5 grid, encoder/me.c:69-79, :116-126.  The
 * grid repeats (0, 2) where (0, -2) would complete it: (0, -2) is never
 * evaluated, (0, 2) twice. */
static const int8_t HEX4[16][2] = {
    { 0, -4 }, { 0, 4 }, { -2, -3 }, { 2, -3 }, { -4, -2 }, { 4, -2 }, { -4, -1 }, { 4, -1 },
    { -4, 0 }, { 4, 0 }, { -4, 1 }, { 4, 1 }, { -4, 2 }, { 4, 2 }, { -2, 3 }, { 2, 3 }
};
static const int8_t FAST_HEX4[8][2] = {
    { 0, -4 }, { 0, 4 }, { -2, -3 }, { 2, -3 }, { -4, 0 }, { 4, 0 }, { -2, 3 }, { 2, 3 }
};
static const int8_t GRID[24][2] = {
    { -1, -1 }, { 0, -1 }, { 1, -1 }, { -1, 0 }, { 1, 0 }, { -1, 1 }, { 0, 1 }, { 1, 1 },
    { -2, -2 }, { -1, -2 }, { 0, 2 }, { 1, -2 }, { 2, -2 }, { -2, -1 }, { 2, -1 }, { -2, 0 },
    { 2, 0 }, { -2, 1 }, { 2, 1 }, { -2, 2 }, { -1, 2 }, { 0, 2 }, { 1, 2 }, { 2, 2 }
};

typedef struct {
    int W, H, R, lambda, pmx, pmy;
    const uint8_t *org;
    intptr_t i_cur;
    const uint8_t *fref;
    intptr_t i_ref;
    uint32_t mv_min, mv_max;
    int bcost, bmx, bmy, dir;
    int *path, max_pts, n_pts;
    /* TZ: the MV window's full-pel limits, the PU and the picture (the raster
     * step's reads are checked against the padded plane) */
    int xmin, ymin, xmax, ymax;
    int pix_x, pix_y, fw, fh;
    int undefined;
    /* test hook: costs from a table over [-span, span]^2 instead of the
     * pictures, and the evaluated points in order */
    const int32_t *land;
    int span;
    int *trace, max_trace, n_trace;
} pat_ctx_t;

/* The cost of one integer point: SAD + MV_COST_IPEL (me.c:181-185). */
static int point_cost(pat_ctx_t *c, int mx, int my)
{
    if (c->trace) {
        if (c->n_trace < c->max_trace) {
            c->trace[2 * c->n_trace] = mx;
            c->trace[2 * c->n_trace + 1] = my;
        }
        c->n_trace++;
    }
    if (c->land) {
        if (mx < -c->span || mx > c->span || my < -c->span || my > c->span)
            return 1 << 28;
        return c->land[(my + c->span) * (2 * c->span + 1) + (mx + c->span)];
    }
    return xo_sad(c->W, c->H, c->org, c->i_cur, c->fref + (intptr_t)my * c->i_ref + mx, c->i_ref) +
           weighted_cost(c->lambda, xo_mvbits(c->R, 4 * mx - c->pmx) + xo_mvbits(c->R, 4 * my - c->pmy));
}

/* One point of ME_COST_IPEL_X3_DIR / ME_COST_IPEL_X4_DIR (me.c:223-287): the
 * SAD plus MV_COST_IPEL, evaluated without a range check, taken on a strict
 * '<' with its direction code. */
static void pat_point(pat_ctx_t *c, int mx, int my, int d)
{
    int cost = point_cost(c, mx, my);
    if (cost < c->bcost) {
        c->bcost = cost;
        c->bmx = mx;
        c->bmy = my;
        c->dir = d;
    }
}

static void pat_accept(pat_ctx_t *c)
{
    if (c->path && c->n_pts < c->max_pts) {
        c->path[2 * c->n_pts] = c->bmx;
        c->path[2 * c->n_pts + 1] = c->bmy;
    }
    c->n_pts++;
}

/* ME_COST_IPEL (me.c:187-193): a range-checked point, strict '<', no
 * direction code. */
static void ipel_point(pat_ctx_t *c, int mx, int my)
{
    if (!check_mv_range(mx, my, c->mv_min, c->mv_max))
        return;
    int cost = point_cost(c, mx, my);
    if (cost < c->bcost) {
        c->bcost = cost;
        c->bmx = mx;
        c->bmy = my;
    }
}

/* One point of the sad_x4 branch of ME_COST_IPEL_X4 (me.c:244-258): SAD plus
 * MV_COST_IPEL, no range check, COPY3_IF_LT (strict '<', no direction code). */
static void free_point(pat_ctx_t *c, int mx, int my)
{
    int cost = point_cost(c, mx, my);
    if (cost < c->bcost) {
        c->bcost = cost;
        c->bmx = mx;
        c->bmy = my;
    }
}

/* DIA_ITER (me.c:328-334): omx = mx, omy = my, then ME_COST_IPEL_X4 with the
 * offsets (0,-1), (-1,0), (1,0), (0,1) (me.c:241-267).  Its CHECK_MV_RANGE_X4
 * (me.h:51-56) tests those offsets themselves against the window, not the
 * points: when the window holds all four, the points around (mx, my) are
 * taken in order WITHOUT a range check (so a centre on the window edge
 * evaluates a point one pixel outside it); otherwise the else branch's
 * ME_COST_IPEL(m0x, m0y) ... take the offsets as absolute MVs (0,-1),
 * (-1,0), (1,0), (0,1), each range-checked, whatever the centre. */
static void dia_iter(pat_ctx_t *c, int mx, int my)
{
    static const int8_t off[4][2] = { { 0, -1 }, { -1, 0 }, { 1, 0 }, { 0, 1 } };
    int inside = 1;
    for (int k = 0; k < 4; k++)
        inside &= check_mv_range(off[k][0], off[k][1], c->mv_min, c->mv_max);
    for (int k = 0; k < 4; k++) {
        if (inside)
            free_point(c, mx + off[k][0], my + off[k][1]);
        else
            ipel_point(c, off[k][0], off[k][1]);
    }
}

static inline int ipel(int v) { return (v + 2) >> 2; }       /* IPEL, me.h:60 */

/* EARLY_TERMINATION (me.c:165-170) with the search's beta + 1 (me.c:883-884):
 * 3 jumps to the diamond (umh_step_3), 2 to the hexagon (umh_step_2). */
static int early_termination(const pat_ctx_t *c, int pred_sad, double beta2, double beta3)
{
    if (c->bcost < pred_sad * beta3)
        return 3;
    if (c->bcost < pred_sad * beta2)
        return 2;
    return 0;
}

static void accept_if_moved(pat_ctx_t *c, int *lx, int *ly)
{
    if (c->bmx != *lx || c->bmy != *ly) {
        pat_accept(c);
        *lx = c->bmx;
        *ly = c->bmy;
    }
}

/* The UMH steps before the hexagon (me.c:1030-1099), from the best MV
 * candidate; (p0x, p0y) is the first candidate (mvc[0], the MVP).  Returns
 * the stage to go on with, 2 (hexagon) or 3 (diamond), and in *where the
 * point it left: 0 after the big hexagons, 1..5 the early termination that
 * jumped (after the predictor points, the second diamond, the cross, the
 * grid, a big-hexagon level). */
static int umh_prefix(pat_ctx_t *c, const xo_umh_ctx_t *u, int p0x, int p0y, int big_hex_level, int *where)
{
    const double beta2 = u->beta2 + 1, beta3 = u->beta3 + 1;
    int lx = c->bmx, ly = c->bmy, et;
    int pmx = p0x, pmy = p0y;                     /* pmv = MAKEDWORD(mvc[0]) */
    dia_iter(c, p0x, p0y);
    if ((pmx || pmy) && (c->bmx != p0x || c->bmy != p0y)) {
        dia_iter(c, c->bmx, c->bmy);
        pmx = c->bmx;
        pmy = c->bmy;
    }
    if (u->mvp1_x || u->mvp1_y)
        ipel_point(c, ipel(u->mvp1_x), ipel(u->mvp1_y));
    accept_if_moved(c, &lx, &ly);
    *where = 1;
    if ((et = early_termination(c, u->pred_sad_uplayer, beta2, beta3)))
        return et;
    if (u->ref_idx > 0)
        ipel_point(c, ipel(u->mvp2_x), ipel(u->mvp2_y));
    if (u->mvp3_x || u->mvp3_y)
        ipel_point(c, ipel(u->mvp3_x), ipel(u->mvp3_y));
    if (pmx != c->bmx || pmy != c->bmy)
        dia_iter(c, c->bmx, c->bmy);
    accept_if_moved(c, &lx, &ly);
    *where = 2;
    if ((et = early_termination(c, u->pred_sad, beta2, beta3)))
        return et;
    /* unsymmetrical cross: odd steps to +-R horizontally, +-R/2 vertically */
    int omx = c->bmx, omy = c->bmy;
    for (int i = 1; i <= c->R; i += 2) {
        ipel_point(c, omx + i, omy);
        ipel_point(c, omx - i, omy);
    }
    for (int j = 1; j <= c->R / 2; j += 2) {
        ipel_point(c, omx, omy + j);
        ipel_point(c, omx, omy - j);
    }
    accept_if_moved(c, &lx, &ly);
    *where = 3;
    if ((et = early_termination(c, u->pred_sad, beta2, beta3)))
        return et;
    omx = c->bmx;
    omy = c->bmy;
    for (int i = 0; i < 24; i++)
        ipel_point(c, omx + GRID[i][0], omy + GRID[i][1]);
    accept_if_moved(c, &lx, &ly);
    *where = 4;
    if ((et = early_termination(c, u->pred_sad, beta2, beta3)))
        return et;
    if (big_hex_level) {
        const int n = big_hex_level == 2 ? 16 : 8;
        const int8_t (*pat)[2] = big_hex_level == 2 ? HEX4 : FAST_HEX4;
        for (int j = 1; j <= c->R / 4; j++) {
            omx = c->bmx;
            omy = c->bmy;
            for (int i = 0; i < n; i++)
                ipel_point(c, omx + pat[i][0] * j, omy + pat[i][1] * j);
            accept_if_moved(c, &lx, &ly);
            if (c->bmx != omx || c->bmy != omy) {
                *where = 5;
                if ((et = early_termination(c, u->pred_sad, beta2, beta3)))
                    return et;
            }
        }
    }
    *where = 0;
    return 2;
}

/* TZ: mv_info (me.c:51-57), the running best of tz_pattern_search. */
typedef struct {
    int bmx, bmy, bdir, bcost, bdist;
} tz_info_t;

/* ME_COST_IPEL_DIR_DIST (me.c:289-294), and one point of
 * ME_COST_IPEL_X4_DIR_DIST (me.c:298-322), which computes the four SADs but
 * takes each point only inside the window: COPY5_IF_LT into mv_info. */
static void tz_point(pat_ctx_t *c, tz_info_t *m, int mx, int my, int dir, int dist)
{
    if (!check_mv_range(mx, my, c->mv_min, c->mv_max))
        return;
    const int cost = point_cost(c, mx, my);
    if (cost < m->bcost) {
        m->bcost = cost;
        m->bmx = mx;
        m->bmy = my;
        m->bdir = dir;
        m->bdist = dist;
    }
}

/* A round's exit test (me.c:721-725 and the same after every round):
 * returns 1 when the pattern search stops. */
static int tz_round_end(const tz_info_t *m, int bcost, int *rounds, int early_exit)
{
    if (m->bcost < bcost)
        *rounds = 0;
    else if (++*rounds >= early_exit)
        return 1;
    return 0;
}

/* tz_pattern_search (me.c:669-874): a cross at distance 1, then diamonds of
 * 8 points at 2, 4, 8 (the four at dist and the four diagonal ones at
 * dist / 2), then 16-point rings at 16, 32, .. merange, each round centred on
 * the best so far, stopping after early_exit rounds without a gain.  The
 * reference sets omy = mv->bmx in every round (me.c:690, :733, :796), so the
 * rounds are centred on (bmx, bmx); restated as written. */
static void tz_pattern(pat_ctx_t *c, tz_info_t *m, int early_exit, int merange)
{
    int rounds = 0, dist = 1;
    int omx = m->bmx, omy = m->bmx;
    int bcost = m->bcost;
    int top = omy - dist, bottom = omy + dist, left = omx - dist, right = omx + dist;
    if (top >= c->ymin && left >= c->xmin && right <= c->xmax && bottom <= c->ymax) {
        tz_point(c, m, omx, top, 2, dist);
        tz_point(c, m, left, omy, 4, dist);
        tz_point(c, m, right, omy, 5, dist);
        tz_point(c, m, omx, bottom, 7, dist);
    } else {
        if (top >= c->ymin)
            tz_point(c, m, omx, top, 2, dist);
        if (left >= c->xmin)
            tz_point(c, m, left, omy, 4, dist);
        if (right <= c->xmax)
            tz_point(c, m, right, omy, 5, dist);
        if (bottom <= c->ymax)
            tz_point(c, m, omx, bottom, 7, dist);
    }
    if (tz_round_end(m, bcost, &rounds, early_exit))
        return;
    for (dist = 2; dist <= 8; dist <<= 1) {
        omx = m->bmx;
        omy = m->bmx;
        bcost = m->bcost;
        top = omy - dist;
        bottom = omy + dist;
        left = omx - dist;
        right = omx + dist;
        const int top2 = omy - (dist >> 1), bottom2 = omy + (dist >> 1);
        const int left2 = omx - (dist >> 1), right2 = omx + (dist >> 1);
        if (top >= c->ymin && left >= c->xmin && right <= c->xmax && bottom <= c->ymax) {
            tz_point(c, m, omx, top, 2, dist);
            tz_point(c, m, left2, top2, 1, dist >> 1);
            tz_point(c, m, right2, top2, 3, dist >> 1);
            tz_point(c, m, left, omy, 4, dist);
            tz_point(c, m, right, omy, 5, dist);
            tz_point(c, m, left2, bottom2, 6, dist >> 1);
            tz_point(c, m, right2, bottom2, 8, dist >> 1);
            tz_point(c, m, omx, bottom, 7, dist);
        } else {
            if (top >= c->ymin)
                tz_point(c, m, omx, top, 2, dist);
            if (top2 >= c->ymin) {
                if (left2 >= c->xmin)
                    tz_point(c, m, left2, top2, 1, dist >> 1);
                if (right2 <= c->xmax)
                    tz_point(c, m, right2, top2, 3, dist >> 1);
            }
            if (left >= c->xmin)
                tz_point(c, m, left, omy, 4, dist);
            if (right <= c->xmax)
                tz_point(c, m, right, omy, 5, dist);
            if (bottom2 <= c->ymax) {
                if (left2 >= c->xmin)
                    tz_point(c, m, left2, bottom2, 6, dist >> 1);
                if (right2 <= c->xmax)
                    tz_point(c, m, right2, bottom2, 8, dist >> 1);
            }
            if (bottom <= c->ymax)
                tz_point(c, m, omx, bottom, 7, dist);
        }
        if (tz_round_end(m, bcost, &rounds, early_exit))
            return;
    }
    for (dist = 16; dist <= merange; dist <<= 1) {
        omx = m->bmx;
        omy = m->bmx;
        bcost = m->bcost;
        top = omy - dist;
        bottom = omy + dist;
        left = omx - dist;
        right = omx + dist;
        const int inside = top >= c->ymin && left >= c->xmin && right <= c->xmax && bottom <= c->ymax;
        /* the four at dist, then per idx the four at (dist / 4) * idx in from
         * the corners of the ring, in the same order in both branches */
        if (inside || top >= c->ymin)
            tz_point(c, m, omx, top, 0, dist);
        if (inside || left >= c->xmin)
            tz_point(c, m, left, omy, 0, dist);
        if (inside || right <= c->xmax)
            tz_point(c, m, right, omy, 0, dist);
        if (inside || bottom <= c->ymax)
            tz_point(c, m, omx, bottom, 0, dist);
        for (int idx = 1; idx < 4; idx++) {
            const int yt = top + (dist >> 2) * idx, yb = bottom - (dist >> 2) * idx;
            const int xl = omx - (dist >> 2) * idx, xr = omx + (dist >> 2) * idx;
            if (inside || yt >= c->ymin) {
                if (inside || xl >= c->xmin)
                    tz_point(c, m, xl, yt, 0, dist);
                if (inside || xr <= c->xmax)
                    tz_point(c, m, xr, yt, 0, dist);
            }
            if (inside || yb <= c->ymax) {
                if (inside || xl >= c->xmin)
                    tz_point(c, m, xl, yb, 0, dist);
                if (inside || xr <= c->xmax)
                    tz_point(c, m, xr, yb, 0, dist);
            }
        }
        if (tz_round_end(m, bcost, &rounds, early_exit))
            return;
    }
}

/* Two Point Search offsets (me.c:144-153): for direction d (1..8), the two
 * outer points next to it. */
static const int8_t TZ_OFFSETS[16][2] = {
    { -1, 0 }, { 0, -1 }, { -1, -1 }, { 1, -1 }, { -1, 0 }, { 1, 0 }, { -1, 1 }, { -1, -1 },
    { 1, -1 }, { 1, 1 }, { -1, 0 }, { 0, 1 }, { -1, 1 }, { 1, 1 }, { 1, 0 }, { 0, 1 }
};

static void tz_two_points(pat_ctx_t *c, int dir)
{
    const int bx = c->bmx, by = c->bmy;
    ipel_point(c, bx + TZ_OFFSETS[(dir - 1) * 2][0], by + TZ_OFFSETS[(dir - 1) * 2][1]);
    ipel_point(c, bx + TZ_OFFSETS[(dir - 1) * 2 + 1][0], by + TZ_OFFSETS[(dir - 1) * 2 + 1][1]);
}

/* One point of the sad_x4 branch of ME_COST_IPEL_X4 as the raster step calls
 * it (me.c:990-991): the macro offsets the absolute MV (mx, my) by (omx, omy)
 * once more (pix_base, me.c:245), so the SAD, the MV cost and the stored MV
 * are those of (omx + mx, omy + my), unchecked.  Where that block leaves the
 * padded reference (XO_PAD pixels around the picture) or its MV difference
 * leaves the mvbits table, the reference reads outside its arrays
 * (undefined): the point is not taken and the search is marked undefined.
 * A block within 4 pixels of the padded reference's right edge counts as
 * leaving it (the dword reads of a row reach 4 bytes past the block). */
static void tz_raster_point(pat_ctx_t *c, int omx, int omy, int mx, int my)
{
    const int x = omx + mx, y = omy + my, lim = xo_mvbits_max_mvd(c->R);
    if (c->pix_x + x < -XO_PAD || c->pix_x + x + c->W + 4 > c->fw + XO_PAD ||
        c->pix_y + y < -XO_PAD || c->pix_y + y + c->H > c->fh + XO_PAD ||
        iabs(4 * x - c->pmx) > lim || iabs(4 * y - c->pmy) > lim) {
        c->undefined = 1;
        return;
    }
    free_point(c, x, y);
}

/* The hexagon after TZ starts unchecked at +-2 pixels around the best, which
 * an offset raster point may have put anywhere in the padded reference: it
 * runs only if the block grown by 2 pixels, widened to whole dwords
 * ([x - 2 rounded down to 4, x + w + 8)), lies in the padded reference;
 * otherwise the reference may read outside it and the search stops
 * undefined. */
static int tz_hexagon_box_ok(const pat_ctx_t *c)
{
    const int x0 = (c->pix_x + c->bmx - 2) & ~3;
    return x0 >= -XO_PAD && x0 + 4 * ((c->W + 10) >> 2) <= c->fw + XO_PAD &&
           c->pix_y + c->bmy - 2 >= -XO_PAD && c->pix_y + c->bmy + c->H + 2 <= c->fh + XO_PAD;
}

/* The TZ branch of xavs2_me_search (me.c:926-1028) from the best MV candidate:
 * six points around it (ME_COST_IPEL_X3, unchecked), DIA_ITER, the pattern
 * search with EarlyExitIters = 3, the two-point search after a best at
 * distance 1 (which may stop the search), the raster refinement after a best
 * beyond RasterDistance = 16, pattern searches with MaxIters = 32 while the
 * best moves by 2 or more, then umh_step_2 (the hexagon, then the diamond).
 * Returns 1 to go on with the hexagon, 0 when the search stops here; *route
 * gets bit 0 for a stop, bit 1 when the raster ran, bit 2 when one of its
 * cells took the offset sad_x4 branch with (omx, omy) != (0, 0), and the
 * number of pattern searches in bits 8..15. */
static int tz_prefix(pat_ctx_t *c, int *route)
{
    const int R = c->R;
    int omx = c->bmx, omy = c->bmy, lx = c->bmx, ly = c->bmy;
    int calls = 0;
    *route = 0;
    free_point(c, omx - 2, omy);
    free_point(c, omx - 1, omy + 2);
    free_point(c, omx + 1, omy + 2);
    free_point(c, omx + 2, omy);
    free_point(c, omx + 1, omy - 2);
    free_point(c, omx - 1, omy - 2);
    if (check_mv_range(c->bmx, c->bmy, c->mv_min, c->mv_max)) {
        omx = c->bmx;                         /* DIA_ITER sets omx, omy (me.c:330-331) */
        omy = c->bmy;
        dia_iter(c, c->bmx, c->bmy);
    }
    accept_if_moved(c, &lx, &ly);
    tz_info_t m = { c->bmx, c->bmy, 0, c->bcost, 0 };
    tz_pattern(c, &m, 3, R);
    calls++;
    c->bcost = m.bcost;
    c->bmx = m.bmx;
    c->bmy = m.bmy;
    int bdist = m.bdist, dir = m.bdir;
    accept_if_moved(c, &lx, &ly);
    if (bdist == 1) {
        if (!dir) {
            *route |= 1 | (calls << 8);
            return 0;
        }
        tz_two_points(c, dir);
        accept_if_moved(c, &lx, &ly);
        if (c->bcost == m.bcost) {
            *route |= 1 | (calls << 8);
            return 0;
        }
    }
    if (bdist > 16) {
        const int rd = 16 >> 1, rd2 = 16 >> 2;
        const int rymin = c->ymin > c->bmy - 16 + 2 ? c->ymin : c->bmy - 16 + 2;
        const int rymax = c->ymax < c->bmy + 16 - 2 ? c->ymax : c->bmy + 16 - 2;
        const int rxmin = c->xmin > c->bmx - 16 + 2 ? c->xmin : c->bmx - 16 + 2;
        const int rxmax = c->xmax < c->bmx + 16 - 2 ? c->xmax : c->bmx + 16 - 2;
        *route |= 2;
        for (int j = rymin; j < rymax; j += rd) {
            for (int i = rxmin; i < rxmax; i += rd) {
                const int px[4] = { i, i, i + rd2, i + rd2 }, py[4] = { j, j + rd2, j, j + rd2 };
                int all = 1;
                for (int k = 0; k < 4; k++)
                    all &= check_mv_range(px[k], py[k], c->mv_min, c->mv_max);
                for (int k = 0; k < 4; k++) {
                    if (all)
                        tz_raster_point(c, omx, omy, px[k], py[k]);
                    else
                        ipel_point(c, px[k], py[k]);
                }
                if (all && (omx || omy))
                    *route |= 4;
            }
        }
        accept_if_moved(c, &lx, &ly);
    }
    while (bdist > 0) {
        tz_info_t n = { c->bmx, c->bmy, 0, c->bcost, 0 };
        tz_pattern(c, &n, 32, R);
        calls++;
        c->bcost = n.bcost;
        c->bmx = n.bmx;
        c->bmy = n.bmy;
        bdist = n.bdist;
        dir = n.bdir;
        accept_if_moved(c, &lx, &ly);
        if (bdist == 1) {
            if (dir)
                tz_two_points(c, dir);
            accept_if_moved(c, &lx, &ly);
            break;
        }
    }
    *route |= calls << 8;
    return 1;
}

/* The integer stage of XAVS2_ME_HEX (me.c:1107-1131, falling through to the
 * diamond), XAVS2_ME_DIA (me.c:1133-1154), XAVS2_ME_UMH (me.c:1030-1099,
 * then the hexagon or the diamond) and XAVS2_ME_TZ (me.c:926-1028, then the
 * hexagon, or nothing more), from the best MV candidate. */
static int pattern_search(pat_ctx_t *c, int method, const xo_umh_ctx_t *u, int p0x, int p0y, int big_hex_level)
{
    int omx, omy, route = 0;
    pat_accept(c);
    int hexagon = method == XO_ME_HEX;
    if (method == XO_ME_TZ) {
        if (!tz_prefix(c, &route))
            return route;
        if (!tz_hexagon_box_ok(c)) {
            c->undefined = 1;
            return route;
        }
        hexagon = 1;
    }
    if (method == XO_ME_UMH) {
        static const xo_umh_ctx_t zero;
        int where = 0;
        const int to = umh_prefix(c, u ? u : &zero, p0x, p0y, big_hex_level, &where);
        hexagon = to == 2;
        route = where | (to << 4);
    }
    if (hexagon) {
        /* the full hexagon around the start, then half hexagons that skip
         * the points the previous step already evaluated */
        c->dir = 0;
        omx = c->bmx;
        omy = c->bmy;
        pat_point(c, omx - 1, omy - 2, 6);
        pat_point(c, omx + 1, omy - 2, 5);
        pat_point(c, omx - 2, omy, 1);
        pat_point(c, omx + 2, omy, 4);
        pat_point(c, omx - 1, omy + 2, 2);
        pat_point(c, omx + 1, omy + 2, 3);
        if (c->dir) {
            pat_accept(c);
            int idx = c->dir - 1;
            for (int i = 0; i < c->R && check_mv_range(c->bmx, c->bmy, c->mv_min, c->mv_max); i++) {
                c->dir = 0;
                omx = c->bmx;
                omy = c->bmy;
                for (int k = 0; k < 3; k++)
                    pat_point(c, omx + HEX2[idx + k][0], omy + HEX2[idx + k][1], k + 1);
                if (!c->dir)
                    break;
                pat_accept(c);
                idx = M1MOD6[c->dir + idx - 1];
            }
        }
    }
    /* small diamond, then half diamonds */
    c->dir = 0;
    if (check_mv_range(c->bmx, c->bmy, c->mv_min, c->mv_max)) {
        omx = c->bmx;
        omy = c->bmy;
        pat_point(c, omx, omy - 1, 4);
        pat_point(c, omx - 1, omy, 1);
        pat_point(c, omx + 1, omy, 3);
        pat_point(c, omx, omy + 1, 2);
    }
    if (c->dir) {
        pat_accept(c);
        int idx = c->dir - 1;
        for (int i = 0; i < c->R && check_mv_range(c->bmx, c->bmy, c->mv_min, c->mv_max); i++) {
            c->dir = 0;
            omx = c->bmx;
            omy = c->bmy;
            for (int k = 0; k < 3; k++)
                pat_point(c, omx + DIA1[idx + k][0], omy + DIA1[idx + k][1], k + 1);
            if (!c->dir)
                break;
            pat_accept(c);
            idx = M1MOD4[c->dir + idx - 1];
        }
    }
    return route;
}

static void pat_init(pat_ctx_t *c, const xo_me_param_t *prm, const xo_pu_t *pu, const uint8_t *cur,
                     intptr_t i_cur, const uint8_t *plane0, intptr_t i_ref, const mv_window_t *win)
{
    memset(c, 0, sizeof(*c));
    c->W = pu->w;
    c->H = pu->h;
    c->R = prm->search_range;
    c->lambda = prm->lambda_factor;
    c->pmx = pu->mvp_x;
    c->pmy = pu->mvp_y;
    c->org = cur + (intptr_t)pu->pix_y * i_cur + pu->pix_x;
    c->i_cur = i_cur;
    c->fref = plane0 + (intptr_t)pu->pix_y * i_ref + pu->pix_x;
    c->i_ref = i_ref;
    c->mv_min = win->mv_min;
    c->mv_max = win->mv_max;
    c->xmin = win->mv_min_fpel[0];
    c->ymin = win->mv_min_fpel[1];
    c->xmax = win->mv_max_fpel[0];
    c->ymax = win->mv_max_fpel[1];
    c->pix_x = pu->pix_x;
    c->pix_y = pu->pix_y;
    c->fw = prm->frame_w;
    c->fh = prm->frame_h;
}

int xo_me_pattern_path(const xo_me_param_t *prm, const xo_pu_t *pu, const xo_umh_ctx_t *ctx,
                       const uint8_t *cur, intptr_t i_cur,
                       const uint8_t *plane0, intptr_t i_ref,
                       int sx, int sy, int scost, int p0x, int p0y, int *path, int max_pts, int *route)
{
    mv_window_t win;
    pat_ctx_t c;
    pu_mv_window(prm, pu, &win);
    pat_init(&c, prm, pu, cur, i_cur, plane0, i_ref, &win);
    c.bcost = scost;
    c.bmx = sx;
    c.bmy = sy;
    c.path = path;
    c.max_pts = max_pts;
    const int r = pattern_search(&c, prm->me_method, ctx, p0x, p0y, prm->umh_big_hex_level);
    if (route)
        *route = r | (c.undefined ? 8 : 0);
    return c.n_pts;
}

static int pattern_trace(const xo_me_param_t *prm, const xo_pu_t *pu, const xo_umh_ctx_t *ctx,
                         const uint8_t *cur, intptr_t i_cur, const uint8_t *plane0, intptr_t i_ref,
                         const int32_t *costs, int span, int sx, int sy, int scost, int p0x, int p0y,
                         int *trace, int max_trace, int *bx, int *by, int *bcost, int *route)
{
    mv_window_t win;
    pat_ctx_t c;
    pu_mv_window(prm, pu, &win);
    pat_init(&c, prm, pu, cur, i_cur, plane0, i_ref, &win);
    c.land = costs;
    c.span = span;
    c.trace = trace;
    c.max_trace = max_trace;
    c.bcost = scost;
    c.bmx = sx;
    c.bmy = sy;
    const int r = pattern_search(&c, prm->me_method, ctx, p0x, p0y, prm->umh_big_hex_level);
    *bx = c.bmx;
    *by = c.bmy;
    *bcost = c.bcost;
    if (route)
        *route = r | (c.undefined ? 8 : 0);
    return c.n_trace;
}

int xo_pattern_landscape(const xo_me_param_t *prm, const xo_pu_t *pu, const xo_umh_ctx_t *ctx,
                         const int32_t *costs, int span, int sx, int sy, int scost, int p0x, int p0y,
                         int *trace, int max_trace, int *bx, int *by, int *bcost, int *route)
{
    static const uint8_t none[1] = { 0 };
    return pattern_trace(prm, pu, ctx, none, 0, none, 0, costs, span, sx, sy, scost, p0x, p0y, trace, max_trace,
                         bx, by, bcost, route);
}

int xo_me_pattern_trace(const xo_me_param_t *prm, const xo_pu_t *pu, const xo_umh_ctx_t *ctx,
                        const uint8_t *cur, intptr_t i_cur, const uint8_t *plane0, intptr_t i_ref,
                        int sx, int sy, int scost, int p0x, int p0y,
                        int *trace, int max_trace, int *bx, int *by, int *bcost, int *route)
{
    return pattern_trace(prm, pu, ctx, cur, i_cur, plane0, i_ref, NULL, 0, sx, sy, scost, p0x, p0y, trace,
                         max_trace, bx, by, bcost, route);
}

static int32_t me_search_impl(const xo_me_param_t *prm, const xo_pu_t *pu, const xo_umh_ctx_t *ctx,
                              const uint8_t *cur, intptr_t i_cur,
                              const uint8_t *const planes[16], intptr_t i_ref,
                              xo_me_result_t *res, int32_t *cost_map)
{
    const int R = prm->search_range;
    const int W = pu->w, H = pu->h;
    const int lambda = prm->lambda_factor;
    const int pmx = pu->mvp_x, pmy = pu->mvp_y;
    const uint8_t *org = cur + (intptr_t)pu->pix_y * i_cur + pu->pix_x;
    const intptr_t i_offset = (intptr_t)pu->pix_y * i_ref + pu->pix_x;   /* p_me->i_bias */
    const uint8_t *fref = planes[0] + i_offset;
    mv_window_t win;
    pu_mv_window(prm, pu, &win);

    /* MV candidate list: add_one_mv_candidate(mvp) then (0,0),
     * md_inter.c:840-865, :957-959.  IPEL(v) = (v + 2) >> 2, me.h:60. */
    int mvc[2][2], n_mvc = 0;
    const int cand_q[2][2] = { { pmx, pmy }, { 0, 0 } };
    for (int c = 0; c < 2; c++) {
        int x = (cand_q[c][0] + 2) >> 2, y = (cand_q[c][1] + 2) >> 2;
        x = clip3(win.mv_min_fpel[0], win.mv_max_fpel[0], x);
        y = clip3(win.mv_min_fpel[1], win.mv_max_fpel[1], y);
        int dup = 0;
        for (int k = 0; k < n_mvc; k++)
            if (mvc[k][0] == x && mvc[k][1] == y)
                dup = 1;
        if (!dup) {
            mvc[n_mvc][0] = x;
            mvc[n_mvc][1] = y;
            n_mvc++;
        }
    }

#define IPEL_COST(mx, my) \
    (xo_sad(W, H, org, i_cur, fref + (intptr_t)(my) * i_ref + (mx), i_ref) + \
     weighted_cost(lambda, xo_mvbits(R, 4 * (mx) - pmx) + xo_mvbits(R, 4 * (my) - pmy)))

    /* xavs2_me_search, me.c:879-921: mvc candidates in order, strict '<'. */
    int bcost = XO_MAX_DISTORTION, bmx = 0, bmy = 0;
    for (int c = 0; c < n_mvc; c++) {
        int mx = mvc[c][0], my = mvc[c][1];
        if (check_mv_range(mx, my, win.mv_min, win.mv_max)) {
            int cost = IPEL_COST(mx, my);
            if (cost < bcost) { bcost = cost; bmx = mx; bmy = my; }
        }
    }
    if (bcost == XO_MAX_DISTORTION) {         /* _me_error, me.c:919-921, :1181 */
        memset(res, 0, sizeof(*res));
        res->cost = XO_MAX_DISTORTION;
        res->status = 1;
        if (cost_map)
            for (int k = 0; k < 4 * R * R; k++)
                cost_map[k] = -1;
        return XO_MAX_DISTORTION;
    }

    int undefined = 0;
    if (prm->me_method == XO_ME_DIA || prm->me_method == XO_ME_HEX || prm->me_method == XO_ME_UMH ||
        prm->me_method == XO_ME_TZ) {
        pat_ctx_t c;
        pat_init(&c, prm, pu, cur, i_cur, planes[0], i_ref, &win);
        c.bcost = bcost;
        c.bmx = bmx;
        c.bmy = bmy;
        pattern_search(&c, prm->me_method, ctx, mvc[0][0], mvc[0][1], prm->umh_big_hex_level);
        bcost = c.bcost;
        bmx = c.bmx;
        bmy = c.bmy;
        undefined = c.undefined;
        goto store;
    }

    /* Full search, me.c:1156-1164: rows j in [-R, R), columns i in [-R, R). */
    const int omx = bmx, omy = bmy;
    for (int j = -R; j < R; j++) {
        for (int i = -R; i < R; i++) {
            int mx = omx + i, my = omy + j;
            int cost = -1;
            if (check_mv_range(mx, my, win.mv_min, win.mv_max)) {
                cost = IPEL_COST(mx, my);
                if (cost < bcost) { bcost = cost; bmx = mx; bmy = my; }
            }
            if (cost_map)
                cost_map[(j + R) * 2 * R + (i + R)] = cost;
        }
    }
#undef IPEL_COST

store:
    /* Store the full-pel result, me.c:1169-1173. */
    res->bmv_x  = (int16_t)(4 * bmx);
    res->bmv_y  = (int16_t)(4 * bmy);
    res->bmv2_x = (int16_t)bmx;
    res->bmv2_y = (int16_t)bmy;
    res->bcost2 = bcost;
    res->mvcost = weighted_cost(lambda, xo_mvbits(R, 4 * bmx - pmx) + xo_mvbits(R, 4 * bmy - pmy));
    res->status = undefined ? 2 : 0;
    res->cost   = bcost;

    if (!prm->use_fractional_me)
        return bcost;

    /* me_subpel_refine, me.c:479-579. */
#define QPEL_COST(mx, my) \
    ((prm->use_hadamard ? xo_satd : xo_sad)(W, H, org, i_cur, \
        planes[(((my) & 3) << 2) + ((mx) & 3)] + i_offset + (intptr_t)((my) >> 2) * i_ref + ((mx) >> 2), \
        i_ref) + weighted_cost(lambda, xo_mvbits(R, (mx) - pmx) + xo_mvbits(R, (my) - pmy)))

    const int step = prm->use_fast_sub_me ? 2 : 1;
    const int8_t (*pat)[2] = prm->use_fast_sub_me ? SPIRAL2 : SPIRAL;
    const int ctr_x = (pmx >> 1) << 1, ctr_y = (pmy >> 1) << 1;
    int qbx = 4 * bmx, qby = 4 * bmy;         /* search centre in qpel */
    int best_x = qbx, best_y = qby;
    int cost;
    if (prm->use_hadamard)
        bcost = QPEL_COST(qbx, qby);
    /* else: bcost stays the integer cost (me.c:511-516) */

    for (int pos = 1; pos < 9; pos += step) {  /* half-pel around the integer best */
        int mx = qbx + pat[pos][0] * 2, my = qby + pat[pos][1] * 2;
        cost = QPEL_COST(mx, my);
        if (cost < bcost) { bcost = cost; best_x = mx; best_y = my; }
    }
    qbx = best_x;
    qby = best_y;
    if (prm->use_fractional_me > 1) {          /* quarter-pel around the half-pel best */
        for (int pos = 1; pos < 9; pos += step) {
            int mx, my;
            if (prm->enable_pmvr) {
                if (pmvr_adapt(&mx, &my, ctr_x, ctr_y, qbx, qby, pat[pos][0], pat[pos][1]))
                    continue;
            } else {
                mx = qbx + pat[pos][0];
                my = qby + pat[pos][1];
            }
            cost = QPEL_COST(mx, my);
            if (cost < bcost) { bcost = cost; best_x = mx; best_y = my; }
        }
    }
#undef QPEL_COST
    res->bmv_x  = (int16_t)best_x;
    res->bmv_y  = (int16_t)best_y;
    res->cost   = bcost;
    res->mvcost = weighted_cost(lambda, xo_mvbits(R, best_x - pmx) + xo_mvbits(R, best_y - pmy));
    return bcost;
}

int32_t xo_me_search(const xo_me_param_t *prm, const xo_pu_t *pu,
                     const uint8_t *cur, intptr_t i_cur,
                     const uint8_t *const planes[16], intptr_t i_ref,
                     xo_me_result_t *res, int32_t *cost_map)
{
    return me_search_impl(prm, pu, NULL, cur, i_cur, planes, i_ref, res, cost_map);
}

int32_t xo_me_search_ctx(const xo_me_param_t *prm, const xo_pu_t *pu, const xo_umh_ctx_t *ctx,
                         const uint8_t *cur, intptr_t i_cur,
                         const uint8_t *const planes[16], intptr_t i_ref,
                         xo_me_result_t *res)
{
    return me_search_impl(prm, pu, ctx, cur, i_cur, planes, i_ref, res, NULL);
}

/* ------------------------------------------------------------------------ */
/* Symmetric bi-prediction refinement.                                       */

int xo_scale_mv_skip(int mv, int dist_dst, int dist_src)
{
    /* (int16_t)((mv * dist_dst * (MULTI / dist_src) + HALF_MULTI) >> OFFSET),
     * MULTI 16384, OFFSET 14 (defines.h:215-218) */
    return (int16_t)((mv * dist_dst * (16384 / dist_src) + 8192) >> 14);
}

/* ME_COST_QPEL_SYM (me.c:353-382): both MVs range-checked against the qpel
 * window, both planes present, fpel_cmp of the current PU against the
 * rounded average + MV_COST_FPEL of the first MV; else XO_MAX_DISTORTION. */
static int sym_cost(const xo_me_param_t *prm, const xo_sym_job_t *j, const mv_window_t *qw,
                    const uint8_t *org, intptr_t i_cur, const uint8_t *const p1[16], const uint8_t *const p2[16],
                    intptr_t i_ref, int mx, int my)
{
    int sx = xo_scale_mv_skip(mx, j->dist_bwd, j->dist_fwd);
    int sy = xo_scale_mv_skip(my, j->dist_bwd, j->dist_fwd);
    if (j->b_slice) {
        sx = -sx;
        sy = -sy;
    }
    if (!check_mv_range(mx, my, qw->mv_min, qw->mv_max) || !check_mv_range(sx, sy, qw->mv_min, qw->mv_max))
        return XO_MAX_DISTORTION;
    const uint8_t *a = p1[((my & 3) << 2) + (mx & 3)];
    const uint8_t *b = p2[((sy & 3) << 2) + (sx & 3)];
    if (!a || !b)
        return XO_MAX_DISTORTION;
    const intptr_t off = (intptr_t)j->pix_y * i_ref + j->pix_x;
    a += off + (intptr_t)(my >> 2) * i_ref + (mx >> 2);
    b += off + (intptr_t)(sy >> 2) * i_ref + (sx >> 2);
    uint8_t pred[64 * 64];
    xo_avg(j->w, j->h, pred, 64, a, i_ref, b, i_ref);
    const int d = (prm->use_hadamard ? xo_satd : xo_sad)(j->w, j->h, org, i_cur, pred, 64);
    return d + weighted_cost(prm->lambda_factor, xo_mvbits(prm->search_range, mx - j->mvp_x) +
                                                  xo_mvbits(prm->search_range, my - j->mvp_y));
}

int32_t xo_me_search_sym(const xo_me_param_t *prm, const xo_sym_job_t *job,
                         const uint8_t *cur, intptr_t i_cur,
                         const uint8_t *const planes1[16], const uint8_t *const planes2[16], intptr_t i_ref,
                         xo_sym_result_t *res)
{
    /* p_me->mv_min / mv_max: the qpel limits of md_inter.c:913-924 */
    mv_window_t qw;
    {
        int lo = clip3(prm->min_mv_range[0], prm->max_mv_range[0], (-64 - job->pix_x) * 4);
        int hi = clip3(prm->min_mv_range[0], prm->max_mv_range[0], (prm->frame_w + 64 - job->pix_x - job->w) * 4);
        int lo_y = clip3(prm->min_mv_range[1], prm->max_mv_range[1], (-64 - job->pix_y) * 4);
        int hi_y = clip3(prm->min_mv_range[1], prm->max_mv_range[1], (prm->frame_h + 64 - job->pix_y - job->h) * 4);
        qw.mv_min = pack_mask2(-lo, -lo_y);
        qw.mv_max = pack_mask2(hi, hi_y) | 0x8000u;
    }
    const uint8_t *org = cur + (intptr_t)job->pix_y * i_cur + job->pix_x;
    const int ctr_x = (job->mvp1_x >> 1) << 1, ctr_y = (job->mvp1_y >> 1) << 1;
    int bx = job->mv_x, by = job->mv_y, bcost = XO_MAX_DISTORTION;
    res->mvcost = -1;
    if (!prm->use_fractional_me) {
        bcost = sym_cost(prm, job, &qw, org, i_cur, planes1, planes2, i_ref, bx, by);
        res->cost = bcost;
        res->mv_x = bx;
        res->mv_y = by;
        return bcost;
    }
    /* half-pel: Spiral[0..4] when hadamard, else Spiral[1..4] (me.c:1210-1211) */
    const int cx = job->mv_x, cy = job->mv_y;
    for (int pos = prm->use_hadamard ? 0 : 1; pos < 5; pos++) {
        const int mx = cx + SPIRAL[pos][0] * 2, my = cy + SPIRAL[pos][1] * 2;
        const int c = sym_cost(prm, job, &qw, org, i_cur, planes1, planes2, i_ref, mx, my);
        if (c < bcost) { bcost = c; bx = mx; by = my; }
    }
    if (prm->use_fractional_me >= 2) {
        const int qx = bx, qy = by;           /* mv->v = bmv.v */
        for (int pos = 1; pos < 5; pos++) {
            int mx, my;
            if (prm->enable_pmvr) {
                if (pmvr_adapt(&mx, &my, ctr_x, ctr_y, qx, qy, SPIRAL[pos][0], SPIRAL[pos][1]))
                    continue;
            } else {
                mx = qx + SPIRAL[pos][0];
                my = qy + SPIRAL[pos][1];
            }
            const int c = sym_cost(prm, job, &qw, org, i_cur, planes1, planes2, i_ref, mx, my);
            if (c < bcost) { bcost = c; bx = mx; by = my; }
        }
    }
    res->cost = bcost;
    res->mv_x = bx;
    res->mv_y = by;
    res->mvcost = weighted_cost(prm->lambda_factor, xo_mvbits(prm->search_range, bx - job->mvp_x) +
                                                    xo_mvbits(prm->search_range, by - job->mvp_y));
    return bcost;
}

/* check_mvd (md_inter.c:68-76). */
static int gate_mvd(const xo_gate_param_t *g, int mvd_x, int mvd_y)
{
    if (g->frame_threads > 1)
        return mvd_x < 4096 && mvd_x >= -4096 && mvd_y < ((1 << g->lcu_level) << 2) &&
               mvd_y >= -((1 << g->lcu_level) << 2);
    return mvd_x < 4096 && mvd_x >= -4096 && mvd_y < 4096 && mvd_y >= -4096;
}

/* The qpel box of check_mv_range and check_mv_range_sym (predict.h:68-78,
 * md_inter.c:83-94): the PU may reach 1 << lcu_level pixels past the
 * picture, inside h->min/max_mv_range. */
static void gate_box(const xo_me_param_t *prm, const xo_gate_param_t *g, int pix_x, int pix_y, int w, int h,
                     int *min_x, int *min_y, int *max_x, int *max_y)
{
    const int bsize = 1 << g->lcu_level;
    *min_x = -((pix_x + bsize) << 2);
    *min_y = -((pix_y + bsize) << 2);
    *max_x = ((prm->frame_w - (pix_x + w)) + bsize) << 2;
    *max_y = ((prm->frame_h - (pix_y + h)) + bsize) << 2;
    if (*min_x < prm->min_mv_range[0]) *min_x = prm->min_mv_range[0];
    if (*min_y < prm->min_mv_range[1]) *min_y = prm->min_mv_range[1];
    if (*max_x > prm->max_mv_range[0]) *max_x = prm->max_mv_range[0];
    if (*max_y > prm->max_mv_range[1]) *max_y = prm->max_mv_range[1];
}

/* check_mv_range (predict.h:68-96) with every reference row reconstructed. */
static int gate_range(const xo_me_param_t *prm, const xo_gate_param_t *g, int pix_x, int pix_y, int w, int h,
                      int mx, int my)
{
    int min_x, min_y, max_x, max_y;
    gate_box(prm, g, pix_x, pix_y, w, h, &min_x, &min_y, &max_x, &max_y);
    return mx <= max_x && mx >= min_x && my <= max_y && my >= min_y;
}

/* check_mv_range_sym (md_inter.c:83-110), frame sequences: the MV and its
 * mirror scaled to the second reference (negated in B slices). */
static int gate_range_sym(const xo_me_param_t *prm, const xo_gate_param_t *g, int pix_x, int pix_y, int w, int h,
                          int mx, int my, int dist_fwd, int dist_bwd, int b_slice)
{
    int min_x, min_y, max_x, max_y;
    gate_box(prm, g, pix_x, pix_y, w, h, &min_x, &min_y, &max_x, &max_y);
    int bx = xo_scale_mv_skip(mx, dist_bwd, dist_fwd), by = xo_scale_mv_skip(my, dist_bwd, dist_fwd);
    if (b_slice) {
        bx = -bx;
        by = -by;
    }
    if (mx > max_x || mx < min_x || my > max_y || my < min_y)
        return 0;
    if (bx > max_x || bx < min_x || by > max_y || by < min_y)
        return 0;
    return 1;
}

void xo_pred_inter_search_bi(const xo_me_param_t *prm, const xo_gate_param_t *g, const xo_bi_job_t *job,
                             int dist_fwd, int dist_bwd, const uint8_t *cur, intptr_t i_cur,
                             const uint8_t *const planes_fwd[16], const uint8_t *const planes_bwd[16],
                             intptr_t i_ref, xo_bi_result_t *res)
{
    const int px = job->pix_x, py = job->pix_y, w = job->w, h = job->h;
    /* mvp = p_me->mvp1 = all_mvp[B_FWD]; mv = fwd_mv = all_best_imv[B_FWD] << 2,
     * shifted in the int16 components of mv_t (md_inter.c:1068-1069) */
    int mv_x = (int16_t)(job->fwd_imv_x * 4), mv_y = (int16_t)(job->fwd_imv_y * 4);
    int fwd_x = mv_x, fwd_y = mv_y;
    const int bwd_x = job->bwd_x, bwd_y = job->bwd_y;
    int cost = XO_MAX_DISTORTION, cost_bid = XO_MAX_DISTORTION;
    res->sym_mvcost = -1;
    res->bid_mvcost = -1;
    int ok = gate_range_sym(prm, g, px, py, w, h, job->mvp1_x, job->mvp1_y, dist_fwd, dist_bwd, 1);
    ok &= gate_range_sym(prm, g, px, py, w, h, mv_x, mv_y, dist_fwd, dist_bwd, 1);
    ok &= gate_mvd(g, job->mvp1_x, job->mvp1_y);
    if (ok) {
        xo_sym_job_t sj = { px, py, w, h, mv_x, mv_y, job->mvp_x, job->mvp_y, job->mvp1_x, job->mvp1_y,
                            dist_fwd, dist_bwd, 1 };
        xo_sym_result_t sr;
        cost = xo_me_search_sym(prm, &sj, cur, i_cur, planes_fwd, planes_bwd, i_ref, &sr);
        mv_x = sr.mv_x;
        mv_y = sr.mv_y;
        res->sym_mvcost = sr.mvcost;
    }
    ok = gate_range(prm, g, px, py, w, h, fwd_x, fwd_y) && gate_range(prm, g, px, py, w, h, bwd_x, bwd_y) &&
         gate_mvd(g, job->mvp1_x, job->mvp1_y) && gate_mvd(g, job->mvp2_x, job->mvp2_y);
    if (ok) {
        xo_bid_job_t bj = { px, py, w, h, fwd_x, fwd_y, bwd_x, bwd_y, job->mvp1_x, job->mvp1_y, job->mvp2_x,
                            job->mvp2_y };
        xo_sym_result_t br;
        cost_bid = xo_me_search_bid(prm, &bj, cur, i_cur, planes_fwd, planes_bwd, i_ref, &br);
        fwd_x = br.mv_x;
        fwd_y = br.mv_y;
        res->bid_mvcost = br.mvcost;
    }
    /* the stored MVs' own checks (md_inter.c:1107-1122) */
    if (!(gate_range(prm, g, px, py, w, h, fwd_x, fwd_y) &&
          gate_mvd(g, fwd_x - job->mvp1_x, fwd_y - job->mvp1_y)))
        cost_bid = XO_MAX_DISTORTION;
    if (!(gate_range(prm, g, px, py, w, h, bwd_x, bwd_y) &&
          gate_mvd(g, bwd_x - job->mvp2_x, bwd_y - job->mvp2_y)))
        cost_bid = XO_MAX_DISTORTION;
    if (!(gate_range_sym(prm, g, px, py, w, h, mv_x, mv_y, dist_fwd, dist_bwd, 1) &&
          gate_mvd(g, mv_x - job->mvp1_x, mv_y - job->mvp1_y)))
        cost = XO_MAX_DISTORTION;
    res->sym_cost = cost;
    res->sym_mv_x = mv_x;
    res->sym_mv_y = mv_y;
    res->bid_cost = cost_bid;
    res->fwd_x = fwd_x;
    res->fwd_y = fwd_y;
    res->bwd_x = bwd_x;
    res->bwd_y = bwd_y;
}

/* REF_BITS (predict.h:47-52) for REF_COST(ref_idx) = WEIGHTED_COST(lambda,
 * REF_BITS[ref_idx + 1]). */
static const int REF_BITS[8] = { 1, 1, 3, 3, 5, 5, 5, 5 };

void xo_pred_inter_search_dual(const xo_me_param_t *prm, const xo_gate_param_t *g, const xo_dual_job_t *job,
                               int n_ref, const int dist[4], const uint8_t *cur, intptr_t i_cur,
                               const uint8_t *const *const planes[4], intptr_t i_ref, xo_dual_result_t *res)
{
    const int px = job->pix_x, py = job->pix_y, w = job->w, h = job->h;
    res->cost = XO_MAX_DISTORTION;
    res->fst_ref = res->snd_ref = -1;
    res->mvcost = -1;
    memset(res->fst, 0, sizeof(res->fst));
    memset(res->snd, 0, sizeof(res->snd));
    for (int r = 0; r < n_ref; r++) {
        const int snd = !r;
        const int d_fwd = dist[r], d_bwd = dist[snd];
        const int m1x = job->mvp1[r][0], m1y = job->mvp1[r][1];
        int fx = (int16_t)(job->imv[r][0] * 4), fy = (int16_t)(job->imv[r][1] * 4);   /* mv_t (:1177-1178) */
        int cost = XO_MAX_DISTORTION, mvcost = -1;
        int ok = gate_range_sym(prm, g, px, py, w, h, fx, fy, d_fwd, d_bwd, 0);
        ok &= gate_mvd(g, fx - m1x, fy - m1y);
        ok &= gate_mvd(g, m1x, m1y);
        ok &= gate_mvd(g, job->mvp_x, job->mvp_y);
        if (ok) {
            xo_sym_job_t sj = { px, py, w, h, fx, fy, job->mvp_x, job->mvp_y, m1x, m1y, d_fwd, d_bwd, 0 };
            xo_sym_result_t sr;
            cost = xo_me_search_sym(prm, &sj, cur, i_cur, planes[r], planes[snd], i_ref, &sr);
            fx = sr.mv_x;
            fy = sr.mv_y;
            mvcost = sr.mvcost;
        }
        res->fst[r][0] = fx;
        res->fst[r][1] = fy;
        res->snd[r][0] = (int16_t)xo_scale_mv_skip(fx, d_bwd, d_fwd);
        res->snd[r][1] = (int16_t)xo_scale_mv_skip(fy, d_bwd, d_fwd);
        ok &= gate_range_sym(prm, g, px, py, w, h, fx, fy, d_fwd, d_bwd, 0);
        ok &= gate_mvd(g, fx - m1x, fy - m1y);
        if (ok) {
            cost += weighted_cost(prm->lambda_factor, REF_BITS[r + 1]);
            if (cost < res->cost) {
                res->cost = cost;
                res->fst_ref = r;
                res->snd_ref = snd;
                res->mvcost = mvcost;
            }
        }
    }
}

/* The qpel window of md_inter.c:913-924 (p_me->mv_min / mv_max), packed as
 * the refinement functions pack it (me.c:1207-1208, :1300-1301). */
static mv_window_t qpel_window(const xo_me_param_t *prm, int pix_x, int pix_y, int w, int h)
{
    mv_window_t qw;
    int lo = clip3(prm->min_mv_range[0], prm->max_mv_range[0], (-64 - pix_x) * 4);
    int hi = clip3(prm->min_mv_range[0], prm->max_mv_range[0], (prm->frame_w + 64 - pix_x - w) * 4);
    int lo_y = clip3(prm->min_mv_range[1], prm->max_mv_range[1], (-64 - pix_y) * 4);
    int hi_y = clip3(prm->min_mv_range[1], prm->max_mv_range[1], (prm->frame_h + 64 - pix_y - h) * 4);
    qw.mv_min = pack_mask2(-lo, -lo_y);
    qw.mv_max = pack_mask2(hi, hi_y) | 0x8000u;
    return qw;
}

/* ME_COST_QPEL_BID (me.c:388-398) of forward candidate (mx, my). */
static int bid_cost(const xo_me_param_t *prm, const xo_bid_job_t *j, const mv_window_t *qw,
                    const uint8_t *target, const uint8_t *const p1[16], intptr_t i_ref, int mx, int my, int bid_bits)
{
    if (!check_mv_range(mx, my, qw->mv_min, qw->mv_max) || !check_mv_range(j->bwd_x, j->bwd_y, qw->mv_min, qw->mv_max))
        return XO_MAX_DISTORTION;
    const uint8_t *a = p1[((my & 3) << 2) + (mx & 3)];
    if (!a)
        return XO_MAX_DISTORTION;
    a += (intptr_t)j->pix_y * i_ref + j->pix_x + (intptr_t)(my >> 2) * i_ref + (mx >> 2);
    const int d = (prm->use_hadamard ? xo_satd : xo_sad)(j->w, j->h, target, 64, a, i_ref) >> 1;
    return d + weighted_cost(prm->lambda_factor, xo_mvbits(prm->search_range, mx - j->mvp1_x) +
                                                  xo_mvbits(prm->search_range, my - j->mvp1_y)) + bid_bits;
}

int32_t xo_me_search_bid(const xo_me_param_t *prm, const xo_bid_job_t *job,
                         const uint8_t *cur, intptr_t i_cur,
                         const uint8_t *const planes1[16], const uint8_t *const planes2[16], intptr_t i_ref,
                         xo_sym_result_t *res)
{
    const mv_window_t qw = qpel_window(prm, job->pix_x, job->pix_y, job->w, job->h);
    const uint8_t *org = cur + (intptr_t)job->pix_y * i_cur + job->pix_x;
    const int bid_bits = weighted_cost(prm->lambda_factor, xo_mvbits(prm->search_range, job->bwd_x - job->mvp2_x) +
                                                           xo_mvbits(prm->search_range, job->bwd_y - job->mvp2_y));
    /* the target 2 * cur - pred_bwd, clipped (me.c:1322-1337); the backward
     * prediction comes from its plane, or from mc_luma at the clamped MV
     * (get_mv_for_mc, mc.h:70-75) when the plane is not built */
    uint8_t target[64 * 64];
    memset(target, 0, sizeof(target));
    if (check_mv_range(job->bwd_x, job->bwd_y, qw.mv_min, qw.mv_max)) {
        uint8_t pred[64 * 64];
        const uint8_t *p2 = planes2[((job->bwd_y & 3) << 2) + (job->bwd_x & 3)];
        if (p2) {
            p2 += (intptr_t)job->pix_y * i_ref + job->pix_x + (intptr_t)(job->bwd_y >> 2) * i_ref + (job->bwd_x >> 2);
            for (int y = 0; y < job->h; y++)
                memcpy(pred + y * 64, p2 + y * i_ref, (size_t)job->w);
        } else {
            const int qx = (int16_t)xo_get_mc_pos(prm->frame_w, job->w, job->pix_x, job->bwd_x);
            const int qy = (int16_t)xo_get_mc_pos(prm->frame_h, job->h, job->pix_y, job->bwd_y);
            xo_mc_luma(pred, 64, qx, qy, job->w, job->h, planes2, i_ref);
        }
        for (int y = 0; y < job->h; y++)
            for (int x = 0; x < job->w; x++) {
                const int16_t d = (int16_t)(org[y * i_cur + x] - pred[y * 64 + x]);
                target[y * 64 + x] = (uint8_t)clip_pel(org[y * i_cur + x] + d);
            }
    }
    const int ctr_x = (job->mvp1_x >> 1) << 1, ctr_y = (job->mvp1_y >> 1) << 1;
    int bx = job->fwd_x, by = job->fwd_y, bcost = XO_MAX_DISTORTION;
    res->mvcost = -1;
    if (!prm->use_fractional_me) {
        bcost = bid_cost(prm, job, &qw, target, planes1, i_ref, bx, by, bid_bits);
        res->cost = bcost;
        res->mv_x = bx;
        res->mv_y = by;
        return bcost;
    }
    /* half-pel: Spiral[0..8] when hadamard, else Spiral[1..8] (me.c:1340-1349) */
    const int cx = job->fwd_x, cy = job->fwd_y;
    for (int pos = prm->use_hadamard ? 0 : 1; pos < 9; pos++) {
        const int mx = cx + SPIRAL[pos][0] * 2, my = cy + SPIRAL[pos][1] * 2;
        const int c = bid_cost(prm, job, &qw, target, planes1, i_ref, mx, my, bid_bits);
        if (c < bcost) { bcost = c; bx = mx; by = my; }
    }
    if (prm->use_fractional_me >= 2) {
        const int qx = bx, qy = by;           /* fwd_mv->v = bmv.v */
        for (int pos = 1; pos < 9; pos++) {
            int mx, my;
            if (prm->enable_pmvr) {
                if (pmvr_adapt(&mx, &my, ctr_x, ctr_y, qx, qy, SPIRAL[pos][0], SPIRAL[pos][1]))
                    continue;
            } else {
                mx = qx + SPIRAL[pos][0];
                my = qy + SPIRAL[pos][1];
            }
            const int c = bid_cost(prm, job, &qw, target, planes1, i_ref, mx, my, bid_bits);
            if (c < bcost) { bcost = c; bx = mx; by = my; }
        }
    }
    res->cost = bcost;
    res->mv_x = bx;
    res->mv_y = by;
    res->mvcost = weighted_cost(prm->lambda_factor, xo_mvbits(prm->search_range, bx - job->mvp1_x) +
                                                    xo_mvbits(prm->search_range, by - job->mvp1_y)) + bid_bits;
    return bcost;
}

typedef struct {
    const xo_me_param_t *prm;
    const xo_pu_t *pus;
    const xo_umh_ctx_t *ctxs;
    int n;
    const uint8_t *cur;
    intptr_t i_cur;
    const uint8_t *const *planes;
    intptr_t i_ref;
    xo_me_result_t *res;
    int next;
    pthread_mutex_t lock;
} batch_ctx_t;

static void *batch_worker(void *arg)
{
    batch_ctx_t *ctx = (batch_ctx_t *)arg;
    for (;;) {
        pthread_mutex_lock(&ctx->lock);
        int i = ctx->next;
        ctx->next += 4;
        pthread_mutex_unlock(&ctx->lock);
        if (i >= ctx->n)
            break;
        int e = i + 4 < ctx->n ? i + 4 : ctx->n;
        for (; i < e; i++)
            me_search_impl(ctx->prm, &ctx->pus[i], ctx->ctxs ? &ctx->ctxs[i] : NULL, ctx->cur, ctx->i_cur,
                           ctx->planes, ctx->i_ref, &ctx->res[i], NULL);
    }
    return NULL;
}

void xo_me_search_batch(const xo_me_param_t *prm, const xo_pu_t *pus, int n,
                        const uint8_t *cur, intptr_t i_cur,
                        const uint8_t *const planes[16], intptr_t i_ref,
                        xo_me_result_t *res, int nthreads)
{
    xo_me_search_batch_ctx(prm, pus, NULL, n, cur, i_cur, planes, i_ref, res, nthreads);
}

void xo_me_search_batch_ctx(const xo_me_param_t *prm, const xo_pu_t *pus, const xo_umh_ctx_t *ctxs, int n,
                            const uint8_t *cur, intptr_t i_cur,
                            const uint8_t *const planes[16], intptr_t i_ref,
                            xo_me_result_t *res, int nthreads)
{
    batch_ctx_t ctx = { prm, pus, ctxs, n, cur, i_cur, planes, i_ref, res, 0, PTHREAD_MUTEX_INITIALIZER };
    if (nthreads < 1)
        nthreads = 1;
    if (nthreads == 1) {
        batch_worker(&ctx);
        return;
    }
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
    for (int t = 0; t < nthreads; t++)
        pthread_create(&th[t], NULL, batch_worker, &ctx);
    for (int t = 0; t < nthreads; t++)
        pthread_join(th[t], NULL);
    free(th);
}

/* ------------------------------------------------------------------------ */
/* Borders and interpolation.                                                */

/* plane_expand_border, common/frame.c:470-507: replicate the left/right
 * columns over pad_h pixels, then the (widened) first/last rows over pad_v. */
static void expand_border(uint8_t *pix, intptr_t stride, int w, int h, int padh, int padv,
                          int top, int bottom)
{
    for (int y = 0; y < h; y++) {
        uint8_t *row = pix + y * stride;
        memset(row - padh, row[0], (size_t)padh);
        memset(row + w, row[w - 1], (size_t)padh);
    }
    int wide = w + 2 * padh;
    if (top)
        for (int y = 1; y <= padv; y++)
            memcpy(pix - padh - y * stride, pix - padh, (size_t)wide);
    if (bottom)
        for (int y = 1; y <= padv; y++)
            memcpy(pix - padh + (h - 1 + y) * stride, pix - padh + (h - 1) * stride, (size_t)wide);
}

void xo_expand_border(uint8_t *plane, intptr_t stride, int w, int h, int pad)
{
    expand_border(plane, stride, w, h, pad, pad, 1, 1);
}

/* The rows xavs2_frame_expand_border_lcurow (frame.c:537-575) expands for
 * LCU row lcu_y of a plane h rows high, one slice: [lcu_y << row_log2,
 * (lcu_y + 1) << row_log2), each end UP_SHIFT = 4 rows up except at the
 * picture's first / last LCU row, the end clipped to h. */
void xo_border_row_window(int h, int row_log2, int lcu_y, int *y0, int *y1)
{
    const int rows = (h + (1 << row_log2) - 1) >> row_log2;
    const int first = lcu_y == 0, last = lcu_y == rows - 1;
    *y0 = (lcu_y << row_log2) - (first ? 0 : 4);
    *y1 = ((lcu_y + 1) << row_log2) - (last ? 0 : 4);
    if (*y1 > h)
        *y1 = h;
}

void xo_expand_border_lcu_row(uint8_t *plane, intptr_t stride, int w, int h, int pad, int row_log2, int lcu_y)
{
    const int rows = (h + (1 << row_log2) - 1) >> row_log2;
    int y0, y1;
    xo_border_row_window(h, row_log2, lcu_y, &y0, &y1);
    if (y1 > y0)
        expand_border(plane + y0 * stride, stride, w, y1 - y0, pad, pad, lcu_y == 0, lcu_y == rows - 1);
}

/* Luma 8-tap and chroma 4-tap phases, common/mc.c:73-91. */
const int8_t xo_luma_taps[4][8] = {
    { 0, 0, 0, 64, 0, 0, 0, 0 },
    { -1, 4, -10, 57, 19, -7, 3, -1 },
    { -1, 4, -11, 40, 40, -11, 4, -1 },
    { -1, 3, -7, 19, 57, -10, 4, -1 },
};
const int8_t xo_chroma_taps[8][4] = {
    { 0, 64, 0, 0 }, { -4, 62, 6, 0 }, { -6, 56, 15, -1 }, { -5, 47, 25, -3 },
    { -4, 36, 36, -4 }, { -3, 25, 47, -5 }, { -1, 15, 56, -6 }, { 0, 6, 62, -4 },
};

/* 8-tap sum at s[0] with taps on s[-3*step .. 4*step] (FLT_8TAP_HOR/VER, mc.c:128-146). */
static inline int fir8_u8(const uint8_t *s, intptr_t step, const int8_t *c)
{
    int v = 0;
    for (int k = 0; k < 8; k++)
        v += c[k] * s[(k - 3) * step];
    return v;
}
static inline int fir8_s16(const int16_t *s, intptr_t step, const int8_t *c)
{
    int v = 0;
    for (int k = 0; k < 8; k++)
        v += c[k] * s[(k - 3) * step];
    return v;
}

/* interpolate_sample_rows, mc.c:593-744, for output rows [start_y, start_y+height).
 * tmp[3] are int16 planes of stride i_tmp whose (x=0,y=0) sits at tmp_org. */
static void interp_rows(uint8_t *const planes[16], intptr_t stride, int w,
                        int16_t *const tmp_org[3], intptr_t i_tmp,
                        int start_y, int height, int b_start, int b_end, int frac)
{
    const int wide = w + 8;                   /* columns [-4, w+4), PAD_OFFSET = 4 */
    /* horizontal a,b,c with 4 extra rows above and below (mc.c:636-667) */
    for (int y = start_y - 4; y < start_y + height + 4; y++) {
        const uint8_t *src = planes[0] + y * stride;
        for (int x = -4; x < -4 + wide; x++) {
            for (int dx = 1; dx <= 3; dx++) {
                if (frac < 2 && dx != 2)
                    continue;
                int v = fir8_u8(src + x, 1, xo_luma_taps[dx]);
                tmp_org[dx - 1][y * i_tmp + x] = (int16_t)v;
                planes[dx][y * stride + x] = (uint8_t)clip_pel((v + 32) >> 6);
            }
        }
    }
    /* vertical d,h,n (mc.c:669-682) */
    for (int y = start_y; y < start_y + height; y++) {
        for (int x = -4; x < -4 + wide; x++) {
            for (int dy = 1; dy <= 3; dy++) {
                if (frac < 2 && dy != 2)
                    continue;
                int v = fir8_u8(planes[0] + y * stride + x, stride, xo_luma_taps[dy]);
                planes[dy << 2][y * stride + x] = (uint8_t)clip_pel((v + 32) >> 6);
            }
        }
    }
    /* diagonal e..r from the int16 horizontal sums, MC_SHIFT = 12 (mc.c:684-724) */
    for (int y = start_y; y < start_y + height; y++) {
        for (int x = -4; x < -4 + wide; x++) {
            for (int dx = 1; dx <= 3; dx++) {
                for (int dy = 1; dy <= 3; dy++) {
                    if (frac < 2 && (dx != 2 || dy != 2))
                        continue;
                    int v = fir8_s16(tmp_org[dx - 1] + y * i_tmp + x, i_tmp, xo_luma_taps[dy]);
                    planes[(dy << 2) + dx][y * stride + x] = (uint8_t)clip_pel((v + 2048) >> 12);
                }
            }
        }
    }
    /* border of the 15 planes over this window (mc.c:726-742) */
    for (int k = 1; k < 16; k++) {
        if (!planes[k])
            continue;
        expand_border(planes[k] + start_y * stride - 4, stride, wide, height,
                      XO_PAD - 4, XO_PAD - 4, b_start, b_end);
    }
}

/* The rows interpolate_lcu_row (mc.c:748-785) produces for LCU row lcu_y,
 * one slice: [lcu_y << lcu_level, (lcu_y + 1) << lcu_level) moved up by
 * MC_OFFSET = 8 (mc.c:95), the first row from -PAD_OFFSET = -4 and the last
 * to h + MC_OFFSET - PAD_OFFSET = h + 4. */
void xo_lcu_row_window(int h, int lcu_level, int lcu_y, int *y0, int *y1)
{
    const int lcu = 1 << lcu_level, rows = (h + lcu - 1) / lcu;
    const int b_start = lcu_y == 0, b_end = lcu_y == rows - 1;
    *y0 = lcu_y * lcu - (b_start ? 4 : 8);
    *y1 = b_end ? h + 8 - 4 : (lcu_y + 1) * lcu - 8;
}

static int interpolate_lcu_rows(uint8_t *const planes_in[16], intptr_t stride, int w, int h, int lcu_level,
                                int row_first, int row_end, int use_fractional_me)
{
    if (!use_fractional_me)
        return 0;
    uint8_t *planes[16];
    for (int k = 0; k < 16; k++)
        planes[k] = planes_in[k];
    if (use_fractional_me == 1)               /* only b, h, j exist (frame.c:343-350) */
        for (int k = 1; k < 16; k++)
            if (k != 2 && k != 8 && k != 10)
                planes[k] = NULL;
    const intptr_t i_tmp = w + 2 * XO_PAD;
    const size_t tmp_sz = (size_t)i_tmp * (size_t)(h + 2 * XO_PAD);
    int16_t *buf = (int16_t *)calloc(3 * tmp_sz, sizeof(int16_t));
    if (!buf)
        return -1;
    int16_t *tmp_org[3];
    for (int k = 0; k < 3; k++)
        tmp_org[k] = buf + k * tmp_sz + XO_PAD * i_tmp + XO_PAD;

    /* interpolate_lcu_row, mc.c:748-785 (one slice) */
    const int lcu = 1 << lcu_level;
    const int rows = (h + lcu - 1) / lcu;
    for (int r = row_first; r < row_end && r < rows; r++) {
        int y0, y1;
        xo_lcu_row_window(h, lcu_level, r, &y0, &y1);
        interp_rows(planes, stride, w, tmp_org, i_tmp, y0, y1 - y0, r == 0, r == rows - 1, use_fractional_me);
    }
    free(buf);
    return 0;
}

int xo_interpolate_frame(uint8_t *const planes[16], intptr_t stride, int w, int h, int lcu_level,
                         int use_fractional_me)
{
    return interpolate_lcu_rows(planes, stride, w, h, lcu_level, 0, 1 << 30, use_fractional_me);
}

int xo_interpolate_lcu_row(uint8_t *const planes[16], intptr_t stride, int w, int h, int lcu_level, int lcu_y,
                           int use_fractional_me)
{
    return interpolate_lcu_rows(planes, stride, w, h, lcu_level, lcu_y, lcu_y + 1, use_fractional_me);
}

/* Block filters, mc.c:275-398 (8-bit: first-pass shift 0, second >> 12). */
/* The interpolation entries of g_funcs with their own taps (primitives.h:
 * 157-171, the C versions at mc.c:275-547).  Sums are formed in int; the
 * luma filters' intermediate tmp is mct_t (int16), the chroma one int32. */

/* intpl_luma_hor_c (mc.c:402-418): tmp = int16(8-tap sum), dst = clip((sum + 32) >> 6). */
void xo_intpl_luma_hor(uint8_t *dst, intptr_t i_dst, int16_t *tmp, intptr_t i_tmp, const uint8_t *src,
                       intptr_t i_src, int w, int h, const int8_t *c)
{
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            const int v = fir8_u8(src + y * i_src + x, 1, c);
            if (tmp)
                tmp[y * i_tmp + x] = (int16_t)v;
            dst[y * i_dst + x] = (uint8_t)clip_pel((v + 32) >> 6);
        }
}

/* intpl_luma_ver_c (mc.c:422-435), also intpl_luma_block_ver (mc.c:358). */
void xo_intpl_luma_ver(uint8_t *dst, intptr_t i_dst, const uint8_t *src, intptr_t i_src, int w, int h,
                       const int8_t *c)
{
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++)
            dst[y * i_dst + x] = (uint8_t)clip_pel((fir8_u8(src + y * i_src + x, i_src, c) + 32) >> 6);
}

/* intpl_luma_ext_c (mc.c:500-515): the 8-tap column sum of the int16 tmp
 * rows, (sum + 2048) >> 12, clipped. */
void xo_intpl_luma_ext(uint8_t *dst, intptr_t i_dst, const int16_t *tmp, intptr_t i_tmp, int w, int h,
                       const int8_t *c)
{
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++)
            dst[y * i_dst + x] = (uint8_t)clip_pel((fir8_s16(tmp + y * i_tmp + x, i_tmp, c) + 2048) >> 12);
}

/* intpl_luma_block_ext_c (mc.c:364-398): rows -3 .. h+3 filtered across into
 * int16, then down. */
void xo_intpl_luma_block_ext(uint8_t *dst, intptr_t i_dst, const uint8_t *src, intptr_t i_src, int w, int h,
                             const int8_t *ch, const int8_t *cv)
{
    int16_t tmp[(64 + 7) * 64];
    for (int y = -3; y < h + 4; y++)
        for (int x = 0; x < w; x++)
            tmp[(y + 3) * 64 + x] = (int16_t)fir8_u8(src + y * i_src + x, 1, ch);
    xo_intpl_luma_ext(dst, i_dst, tmp + 3 * 64, 64, w, h, cv);
}

void xo_luma_block_hor(uint8_t *dst, intptr_t i_dst, const uint8_t *src, intptr_t i_src, int w, int h, int dx)
{
    xo_intpl_luma_hor(dst, i_dst, NULL, 0, src, i_src, w, h, xo_luma_taps[dx]);
}

void xo_luma_block_ver(uint8_t *dst, intptr_t i_dst, const uint8_t *src, intptr_t i_src, int w, int h, int dy)
{
    xo_intpl_luma_ver(dst, i_dst, src, i_src, w, h, xo_luma_taps[dy]);
}

void xo_luma_block_ext(uint8_t *dst, intptr_t i_dst, const uint8_t *src, intptr_t i_src, int w, int h, int dx, int dy)
{
    xo_intpl_luma_block_ext(dst, i_dst, src, i_src, w, h, xo_luma_taps[dx], xo_luma_taps[dy]);
}

static inline int fir4_u8(const uint8_t *s, intptr_t step, const int8_t *c)
{
    return c[0] * s[-step] + c[1] * s[0] + c[2] * s[step] + c[3] * s[2 * step];
}

/* intpl_chroma_block_hor_c / _ver_c (mc.c:275-303): 4 taps at -1 .. 2. */
void xo_intpl_chroma_block_hor(uint8_t *dst, intptr_t i_dst, const uint8_t *src, intptr_t i_src, int w, int h,
                               const int8_t *c)
{
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++)
            dst[y * i_dst + x] = (uint8_t)clip_pel((fir4_u8(src + y * i_src + x, 1, c) + 32) >> 6);
}

void xo_intpl_chroma_block_ver(uint8_t *dst, intptr_t i_dst, const uint8_t *src, intptr_t i_src, int w, int h,
                               const int8_t *c)
{
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++)
            dst[y * i_dst + x] = (uint8_t)clip_pel((fir4_u8(src + y * i_src + x, i_src, c) + 32) >> 6);
}

/* intpl_chroma_block_ext_c (mc.c:307-337): rows -1 .. h+1 across into int32,
 * then down with (sum + 2048) >> 12. */
void xo_intpl_chroma_block_ext(uint8_t *dst, intptr_t i_dst, const uint8_t *src, intptr_t i_src, int w, int h,
                               const int8_t *ch, const int8_t *cv)
{
    int32_t tmp[(32 + 3) * 32];
    for (int y = -1; y < h + 2; y++)
        for (int x = 0; x < w; x++)
            tmp[(y + 1) * 32 + x] = fir4_u8(src + y * i_src + x, 1, ch);
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            const int32_t *t = tmp + (y + 1) * 32 + x;
            int v = cv[0] * t[-32] + cv[1] * t[0] + cv[2] * t[32] + cv[3] * t[64];
            dst[y * i_dst + x] = (uint8_t)clip_pel((v + 2048) >> 12);
        }
}

void xo_chroma_block_hor(uint8_t *dst, intptr_t i_dst, const uint8_t *src, intptr_t i_src, int w, int h, int dx)
{
    xo_intpl_chroma_block_hor(dst, i_dst, src, i_src, w, h, xo_chroma_taps[dx]);
}

void xo_chroma_block_ver(uint8_t *dst, intptr_t i_dst, const uint8_t *src, intptr_t i_src, int w, int h, int dy)
{
    xo_intpl_chroma_block_ver(dst, i_dst, src, i_src, w, h, xo_chroma_taps[dy]);
}

void xo_chroma_block_ext(uint8_t *dst, intptr_t i_dst, const uint8_t *src, intptr_t i_src, int w, int h, int dx, int dy)
{
    xo_intpl_chroma_block_ext(dst, i_dst, src, i_src, w, h, xo_chroma_taps[dx], xo_chroma_taps[dy]);
}

/* mc_luma, mc.c:559-589: copy from the interpolated plane when it exists. */
void xo_mc_luma(uint8_t *dst, intptr_t i_dst, int qx, int qy, int w, int h,
                const uint8_t *const planes[16], intptr_t i_ref)
{
    int x = qx >> 2, y = qy >> 2, dx = qx & 3, dy = qy & 3;
    const uint8_t *p = planes[(dy << 2) + dx];
    if (p) {
        p += (intptr_t)y * i_ref + x;
        for (int r = 0; r < h; r++)
            memcpy(dst + r * i_dst, p + r * i_ref, (size_t)w);
        return;
    }
    const uint8_t *src = planes[0] + (intptr_t)y * i_ref + x;
    if (dy == 0)
        xo_luma_block_hor(dst, i_dst, src, i_ref, w, h, dx);
    else if (dx == 0)
        xo_luma_block_ver(dst, i_dst, src, i_ref, w, h, dy);
    else
        xo_luma_block_ext(dst, i_dst, src, i_ref, w, h, dx, dy);
}

/* cu_get_mc_pos, common/mc.h:52-66: absolute qpel position of a block,
 * clamped for MVs that point far outside the picture (keeps mv & 7). */
int xo_get_mc_pos(int img_size, int blk_size, int blk_pos, int mv)
{
    int imv = mv >> 2, fmv = mv & 7;
    if (blk_pos + imv < -blk_size - 8)
        return (-blk_size - 8) * 4 + fmv;
    if (blk_pos + imv > img_size + 4)
        return (img_size + 4) * 4 + fmv;
    return blk_pos * 4 + mv;
}

/* One n x n block of an inter CU, prediction p given: sub_ps, the residual
 * path (dct or dct_half, quant, dequant, idct), add_ps with the clip, and the
 * SSD of the reconstruction (cu_recon_chroma, rdo.c:807-838; the luma TUs of
 * cu_recon_luma the same way). */
static void recon_block(int n, int half, int qp, const uint8_t *c, intptr_t i_cur, const uint8_t *p, intptr_t i_p,
                        int16_t *level, uint8_t *d, intptr_t i_rec, int32_t *nnz, int32_t *ssd)
{
    int16_t res[64 * 64], rr[64 * 64];
    for (int y = 0; y < n; y++)                          /* sub_ps */
        for (int x = 0; x < n; x++)
            res[y * n + x] = (int16_t)(c[y * i_cur + x] - p[y * i_p + x]);
    if (half || n == 64) {
        /* dct_half (32x32), or a 64x64 TU: dct / dct_half of LUMA_64x64 (the
         * wavelet's 32x32 low band transformed with the wavelet flag,
         * transform.c:1514-1541), quantised at level 6 (i_level of
         * tu_recon_inter_luma, rdo.c:1351-1356), 32x32 levels */
        const int l = n == 64 ? 6 : 5;
        int qs, qsh, qa, ds, dsh;
        int16_t co[32 * 32];
        xo_quant_params(l, qp, 0, &qs, &qsh, &qa);
        xo_dequant_params(l, qp, &ds, &dsh);
        if (half)
            xo_dct_half_part(n, n, res, co, n);
        else
            xo_dct_part(n, n, res, co, n);
        *nnz = xo_quant(co, 1024, qs, qsh, qa);
        memcpy(level, co, sizeof(co));
        xo_dequant(co, 1024, ds, dsh);
        xo_idct_part(n, n, co, rr, n);
    } else {
        xo_residual_batch(n, qp, 0, res, level, rr, nnz, 1);
    }
    int e = 0;
    for (int y = 0; y < n; y++)                          /* add_ps, then pixf.ssd */
        for (int x = 0; x < n; x++) {
            const int v = clip_pel(p[y * i_p + x] + rr[y * n + x]);
            d[y * i_rec + x] = (uint8_t)v;
            e += (c[y * i_cur + x] - v) * (c[y * i_cur + x] - v);
        }
    *ssd = e;
}

static void inter_recon(int n_arg, int qp, const uint8_t *cur, intptr_t i_cur, const uint8_t *const planes[16],
                        const uint8_t *const planes2[16], intptr_t i_ref, int frame_w, int frame_h,
                        const int32_t *tus, const int32_t *mv2, int ntu, int16_t *level, uint8_t *rec,
                        intptr_t i_rec, int32_t *nnz, int32_t *ssd)
{
    /* tu_get_dct_coeff (rdo.c:764-771): dct_half for a 32x32 or 64x64 TU when
     * asked; a 64x64 TU keeps 32x32 levels */
    const int n = n_arg & ~XO_TU_DCT_HALF, half = (n_arg & XO_TU_DCT_HALF) && n >= 32;
    const size_t ncoef = n == 64 ? 1024 : (size_t)n * n;
    for (int i = 0; i < ntu; i++) {
        const int32_t *r = tus + 8 * i;
        /* the PU's prediction (get_mv_for_mc with the PU size), then the TU's
         * part of it; bi-prediction averages the second reference's
         * (rdo_get_pred_inter, rdo.c:1616-1631) */
        uint8_t pred[64 * 64];
        const int qx = (int16_t)xo_get_mc_pos(frame_w, r[4], r[2], r[6]);
        const int qy = (int16_t)xo_get_mc_pos(frame_h, r[5], r[3], r[7]);
        xo_mc_luma(pred, 64, qx, qy, r[4], r[5], planes, i_ref);
        if (planes2) {
            uint8_t tmp[64 * 64];
            const int qx2 = (int16_t)xo_get_mc_pos(frame_w, r[4], r[2], mv2[2 * i]);
            const int qy2 = (int16_t)xo_get_mc_pos(frame_h, r[5], r[3], mv2[2 * i + 1]);
            xo_mc_luma(tmp, 64, qx2, qy2, r[4], r[5], planes2, i_ref);
            xo_avg(r[4], r[5], pred, 64, pred, 64, tmp, 64);
        }
        const uint8_t *p = pred + (r[1] - r[3]) * 64 + (r[0] - r[2]);
        recon_block(n, half, qp, cur + (intptr_t)r[1] * i_cur + r[0], i_cur, p, 64, level + (size_t)i * ncoef,
                    rec + (intptr_t)r[1] * i_rec + r[0], i_rec, nnz + i, ssd + i);
    }
}

void xo_recon_pred_batch(int n_arg, int qp, const uint8_t *cur, intptr_t i_cur, const uint8_t *pred, intptr_t i_pred,
                         const int32_t *blocks, int nblk, int16_t *level, uint8_t *rec, intptr_t i_rec, int32_t *nnz,
                         int32_t *ssd)
{
    const int n = n_arg & ~XO_TU_DCT_HALF, half = (n_arg & XO_TU_DCT_HALF) && n >= 32;
    const size_t ncoef = n == 64 ? 1024 : (size_t)n * n;
    for (int i = 0; i < nblk; i++) {
        const int x = blocks[2 * i], y = blocks[2 * i + 1];
        recon_block(n, half, qp, cur + (intptr_t)y * i_cur + x, i_cur, pred + (intptr_t)y * i_pred + x, i_pred,
                    level + i * ncoef, rec + (intptr_t)y * i_rec + x, i_rec, nnz + i, ssd + i);
    }
}

/* cu_get_chroma_qp (cudata.h:118-125): the luma QP clipped to 0..63, then
 * tab_qp_scale_chroma (cudata.c:98-106), written here as its runs: the
 * identity to 41, then each value twice from 42 to 47, 48..50 three times
 * each, and 51. */
int xo_chroma_qp(int luma_qp)
{
    const int q = luma_qp < 0 ? 0 : (luma_qp > 63 ? 63 : luma_qp);
    if (q < 42)
        return q;
    if (q < 54)
        return 42 + (q - 42) / 2;
    if (q < 63)
        return 48 + (q - 54) / 3;
    return 51;
}

void xo_inter_recon_batch(int n, int qp, const uint8_t *cur, intptr_t i_cur, const uint8_t *const planes[16],
                          intptr_t i_ref, int frame_w, int frame_h, const int32_t *tus, int ntu, int16_t *level,
                          uint8_t *rec, intptr_t i_rec, int32_t *nnz, int32_t *ssd)
{
    inter_recon(n, qp, cur, i_cur, planes, NULL, i_ref, frame_w, frame_h, tus, NULL, ntu, level, rec, i_rec, nnz,
                ssd);
}

void xo_inter_recon_bi_batch(int n, int qp, const uint8_t *cur, intptr_t i_cur, const uint8_t *const planes1[16],
                             const uint8_t *const planes2[16], intptr_t i_ref, int frame_w, int frame_h,
                             const int32_t *tus, const int32_t *mv2, int ntu, int16_t *level, uint8_t *rec,
                             intptr_t i_rec, int32_t *nnz, int32_t *ssd)
{
    inter_recon(n, qp, cur, i_cur, planes1, planes2, i_ref, frame_w, frame_h, tus, mv2, ntu, level, rec, i_rec, nnz,
                ssd);
}

/* ------------------------------------------------------------------------ */
/* MV predictor (get_mvp_default, md_inter.c:470-524).                       */

/* get_mvp_type_default (md_inter.c:333-368): 0 median, 1 left, 2 top, 3 top-right. */
static int mvp_type(int ref_frame, int rl, int ru, int rur, int cb_x, int cb_y, int cb_w, int cb_h)
{
    if (rl != -1 && ru == -1 && rur == -1)
        return 1;
    if (rl == -1 && ru != -1 && rur == -1)
        return 2;
    if (rl == -1 && ru == -1 && rur != -1)
        return 3;
    if (cb_w < cb_h)
        return cb_x == 0 ? (rl == ref_frame ? 1 : 0) : (rur == ref_frame ? 3 : 0);
    if (cb_w > cb_h)
        return cb_y == 0 ? (ru == ref_frame ? 2 : 0) : (rl == ref_frame ? 1 : 0);
    return 0;
}

/* scale_mv_default (md_inter.c:405-410); -1 in *bad where the product leaves int. */
static int scale_default(int mv, int dist_dst, int dist_src_scale, int *bad)
{
    const int64_t prod = (int64_t)iabs(mv) * dist_dst * dist_src_scale;
    if (prod + 8192 > INT32_MAX || prod < 0)
        *bad = 1;
    const int sign = mv > 0 ? 1 : (mv < 0 ? -1 : 0);
    const int v = sign * (int)((prod + 8192) >> 14);
    return v < -32768 ? -32768 : (v > 32767 ? 32767 : v);
}

/* derive_median_mv (md_inter.c:372-401): C division truncates toward zero. */
static int median_mv(int a, int b, int c)
{
    if ((a < 0 && b > 0 && c > 0) || (a > 0 && b < 0 && c < 0))
        return (int16_t)((b + c) / 2);
    if ((b < 0 && a > 0 && c > 0) || (b > 0 && a < 0 && c < 0))
        return (int16_t)((c + a) / 2);
    if ((c < 0 && a > 0 && b > 0) || (c > 0 && a < 0 && b < 0))
        return (int16_t)((a + b) / 2);
    const int dab = iabs(a - b), dbc = iabs(b - c), dca = iabs(c - a);
    const int m = dab < dbc ? (dab < dca ? dab : dca) : (dbc < dca ? dbc : dca);
    if (m == dab)
        return (int16_t)((a + b) / 2);
    if (m == dbc)
        return (int16_t)((b + c) / 2);
    return (int16_t)((c + a) / 2);
}

int xo_get_mvp_default(const xo_mvp_job_t *j, int b_slice, const int ref_dpoc[4], const int ref_dpoc_multi[4],
                       int *pmv_x, int *pmv_y)
{
    /* the reference lists hold MAX_REFS = 4 entries (defines.h:301, :364) and
     * the reference asserts blkref < MAX_REFS (predict.h:100) */
    const int rn = j->ur_available ? j->ref[3] : j->ref[2];
    if (j->ref_idx < 0 || j->ref_idx >= 4 || j->ref[0] < -1 || j->ref[0] >= 4 || j->ref[1] < -1 || j->ref[1] >= 4 ||
        rn < -1 || rn >= 4) {
        *pmv_x = *pmv_y = 0;
        return -2;
    }
    /* neighbours: 0 left (a), 1 top (b), 3 top-right (c), or top-left (2) when
     * the top-right is unavailable */
    const int rl = j->ref[0], ru = j->ref[1], rur = j->ur_available ? j->ref[3] : j->ref[2];
    int mv[3][2] = { { j->mv[0][0], j->mv[0][1] }, { j->mv[1][0], j->mv[1][1] },
                     { j->ur_available ? j->mv[3][0] : j->mv[2][0], j->ur_available ? j->mv[3][1] : j->mv[2][1] } };
    const int refs[3] = { rl, ru, rur };
    const int type = mvp_type(j->ref_idx, rl, ru, rur, j->cb_x, j->cb_y, j->cb_w, j->cb_h);
    int bad = 0;
    for (int k = 0; k < 3; k++) {
        if (refs[k] < 0) {
            mv[k][0] = mv[k][1] = 0;
            continue;
        }
        int dist, scale;
        if (b_slice) {                    /* B_BWD = 0, B_FWD = 1 (avs2_defs.h:138-139) */
            dist = ref_dpoc[j->bwd_2nd ? 0 : 1];
            scale = ref_dpoc_multi[j->bwd_2nd ? 0 : 1];
        } else {
            dist = ref_dpoc[j->ref_idx];
            scale = ref_dpoc_multi[refs[k]];
        }
        mv[k][1] = scale_default(mv[k][1], dist, scale, &bad);
        mv[k][0] = scale_default(mv[k][0], dist, scale, &bad);
    }
    if (type == 0) {
        *pmv_x = median_mv(mv[0][0], mv[1][0], mv[2][0]);
        *pmv_y = median_mv(mv[0][1], mv[1][1], mv[2][1]);
    } else {
        *pmv_x = mv[type - 1][0];
        *pmv_y = mv[type - 1][1];
    }
    return bad ? -1 : 0;
}

/* derive_median_mv (md_inter.c:370-402) returning which neighbour's pair it
 * took: 0 (a, b), 1 (b, c), 2 (c, a), as the reference's mvp_type. */
static int median_src(int a, int b, int c)
{
    if ((a < 0 && b > 0 && c > 0) || (a > 0 && b < 0 && c < 0))
        return 1;
    if ((b < 0 && a > 0 && c > 0) || (b > 0 && a < 0 && c < 0))
        return 2;
    if ((c < 0 && a > 0 && b > 0) || (c > 0 && a < 0 && b < 0))
        return 0;
    const int dab = iabs(a - b), dbc = iabs(b - c), dca = iabs(c - a);
    const int m = dab < dbc ? (dab < dca ? dab : dca) : (dbc < dca ? dbc : dca);
    return m == dab ? 0 : (m == dbc ? 1 : 2);
}

/* tab_umh_alpha_2nd / _3rd (md_inter.c:49-54): float literals held in doubles. */
static const float UMH_ALPHA2[8] = { 0.0f, 0.01f, 0.01f, 0.01f, 0.02f, 0.03f, 0.03f, 0.04f };
static const float UMH_ALPHA3[8] = { 0.0f, 0.06f, 0.07f, 0.07f, 0.08f, 0.12f, 0.11f, 0.15f };

/* (int16_t)(mv * dpoc_curr / (float)dpoc_src): the int product converted to
 * float, divided in float, truncated (md_inter.c:673-674, 690-691). */
static int scale_float(int mv, int dpoc_curr, int dpoc_src)
{
    const float f = (float)(mv * dpoc_curr) / (float)dpoc_src;
    return (int16_t)(int)f;
}

void xo_umh_bsize(int i_qp, int slice_type, double bsize[8])
{
    static const int quant_coef[6] = { 13107, 11916, 10082, 9362, 8192, 7282 };
    const int per = i_qp / 6, rem = i_qp % 6, q_bits = 15 + per;
    const int qp_const = slice_type == 0 ? (1 << q_bits) / 3 : (1 << q_bits) / 6;
    const int threshold_4x4 = ((1 << q_bits) - qp_const) / quant_coef[rem];
    const double quantize_step = (float)threshold_4x4 / (4 * 5.61f);
    for (int m = 0; m < 8; m++)
        bsize[m] = 0;
    bsize[7] = (16 * 16) * quantize_step;        /* PRED_nRx2N */
    bsize[6] = 4 * bsize[7];                     /* PRED_nLx2N */
    bsize[5] = 4 * bsize[7];                     /* PRED_2NxnD */
    bsize[4] = 4 * bsize[5];                     /* PRED_2NxnU */
    bsize[3] = 4 * bsize[4];                     /* PRED_Nx2N */
    bsize[2] = 4 * bsize[4];                     /* PRED_2NxN */
    bsize[1] = 4 * bsize[2];                     /* PRED_2Nx2N */
}

int xo_umh_prepare(const xo_umh_job_t *j, int slice_type, const int ref_dpoc[4], const int ref_dpoc_multi[4],
                   const double umh_bsize[8], xo_umh_out_t *o)
{
    memset(o, 0, sizeof(*o));
    const int ref_idx = j->mvp.ref_idx, mode = j->mode, b_slice = slice_type == 2;
    /* get_mvp_default_sad: the MVP as get_mvp_default, and the SAD of the
     * neighbour it came from (the median's y component picks it) */
    int rc = xo_get_mvp_default(&j->mvp, b_slice, ref_dpoc, ref_dpoc_multi, &o->pmv_x, &o->pmv_y);
    if (rc == -2 || mode < 1 || mode > 7) {
        o->status = -2;
        return o->status;
    }
    o->status = rc;
    {
        const int rl = j->mvp.ref[0], ru = j->mvp.ref[1], rur = j->mvp.ur_available ? j->mvp.ref[3] : j->mvp.ref[2];
        const int type = mvp_type(ref_idx, rl, ru, rur, j->mvp.cb_x, j->mvp.cb_y, j->mvp.cb_w, j->mvp.cb_h);
        if (type == 0) {
            /* the y components after scaling, as derive_median_mv sees them */
            int my[3];
            const int refs[3] = { rl, ru, rur };
            for (int k = 0; k < 3; k++) {
                const int src = k < 2 ? k : (j->mvp.ur_available ? 3 : 2);
                int v = j->mvp.mv[src][1], bad = 0;
                if (refs[k] < 0)
                    v = 0;
                else if (b_slice)
                    v = scale_default(v, ref_dpoc[j->mvp.bwd_2nd ? 0 : 1], ref_dpoc_multi[j->mvp.bwd_2nd ? 0 : 1], &bad);
                else
                    v = scale_default(v, ref_dpoc[ref_idx], ref_dpoc_multi[refs[k]], &bad);
                my[k] = v;
            }
            o->pred_sad_space = j->nb_sad[median_src(my[0], my[1], my[2])];
        } else {
            o->pred_sad_space = j->nb_sad[type - 1];
        }
    }
    /* fast_me_prepare_info: the upper layer's MV and half its cost */
    xo_umh_ctx_t *c = &o->ctx;
    int sad_up = 0;
    if (mode == 4 || mode == 5) {                /* PRED_2NxnU, PRED_2NxnD: from PRED_2NxN */
        c->mvp1_x = j->best_up[0][0]; c->mvp1_y = j->best_up[0][1];
        sad_up = j->mincost_up[0] / 2;
    } else if (mode == 6 || mode == 7) {         /* PRED_nLx2N, PRED_nRx2N: from PRED_Nx2N */
        c->mvp1_x = j->best_up[1][0]; c->mvp1_y = j->best_up[1][1];
        sad_up = j->mincost_up[1] / 2;
    } else if (mode > 1) {                       /* PRED_2NxN, PRED_Nx2N: from PRED_2Nx2N */
        c->mvp1_x = j->best_up[2][0]; c->mvp1_y = j->best_up[2][1];
        sad_up = j->mincost_up[2] / 2;
    }
    c->pred_sad_uplayer = sad_up;
    /* the nearest reference's MV and cost (mvp2 is unused for ref_idx 0) */
    int sad_ref = 0;
    if (b_slice && ref_idx == 1) {               /* B_FWD: the negated B_BWD MV */
        c->mvp2_x = (int16_t)(-j->best_prev[0]);
        c->mvp2_y = (int16_t)(-j->best_prev[1]);
        sad_ref = j->mincost_prev;
    } else if (ref_idx > 0) {
        if (ref_dpoc[ref_idx - 1] == 0) {
            o->status = -3;
            return o->status;
        }
        c->mvp2_x = scale_float(j->best_prev[0], ref_dpoc[ref_idx], ref_dpoc[ref_idx - 1]);
        c->mvp2_y = scale_float(j->best_prev[1], ref_dpoc[ref_idx], ref_dpoc[ref_idx - 1]);
        sad_ref = j->mincost_prev;
    }
    /* the co-located MV, scaled */
    if (j->col_valid && j->col_ref >= 0) {
        if (j->col_dpoc == 0) {
            o->status = -3;
            return o->status;
        }
        c->mvp3_x = scale_float(j->col_mv[0], ref_dpoc[ref_idx], j->col_dpoc);
        c->mvp3_y = scale_float(j->col_mv[1], ref_dpoc[ref_idx], j->col_dpoc);
    }
    c->ref_idx = ref_idx;
    const int pred_sad = (!b_slice && ref_idx > 0) ? sad_ref : (mode == 1 ? o->pred_sad_space : sad_up);
    c->pred_sad = pred_sad;
    if (pred_sad != 0) {
        /* dist_t (int32) product: the reference squares in int, wrapping */
        const int32_t sq = (int32_t)((uint32_t)pred_sad * (uint32_t)pred_sad);
        const double threshold = umh_bsize[mode] / sq;
        c->beta2 = threshold - (double)UMH_ALPHA2[mode];
        c->beta3 = threshold - (double)UMH_ALPHA3[mode];
    }
    return o->status;
}

/* mc_chroma, mc.c:797-829: luma qpel MV = chroma 1/8-pel position. */
void xo_mc_chroma(uint8_t *dst_u, uint8_t *dst_v, intptr_t i_dst, int qx, int qy, int w, int h,
                  const uint8_t *plane_u, const uint8_t *plane_v, intptr_t i_src)
{
    int px = qx & 7, py = qy & 7;
    intptr_t off = (intptr_t)(qy >> 3) * i_src + (qx >> 3);
    const uint8_t *src[2] = { plane_u + off, plane_v + off };
    uint8_t *dst[2] = { dst_u, dst_v };
    for (int c = 0; c < 2; c++) {
        if (px == 0 && py == 0) {
            for (int r = 0; r < h; r++)
                memcpy(dst[c] + r * i_dst, src[c] + r * i_src, (size_t)w);
        } else if (py == 0) {
            xo_chroma_block_hor(dst[c], i_dst, src[c], i_src, w, h, px);
        } else if (px == 0) {
            xo_chroma_block_ver(dst[c], i_dst, src[c], i_src, w, h, py);
        } else {
            xo_chroma_block_ext(dst[c], i_dst, src[c], i_src, w, h, px, py);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Transform.  The AVS2 integer DCT matrices (transform.c:59-135) are
 * T_N[k][n] = 32 for k = 0 and round(32*sqrt(2)*cos((2n+1)k*pi/(2N))) otherwise,
 * and T_N[k][n] = T_32[k*32/N][n].  They are generated here from that closed
 * form; tests compare the result entry by entry with the reference tables. */
static int16_t T4m[4][4], T8m[8][8], T16m[16][16], T32m[32][32];
static pthread_once_t tables_once = PTHREAD_ONCE_INIT;

static void build_tables(void)
{
    const double pi = 3.14159265358979323846;
    for (int k = 0; k < 32; k++)
        for (int n = 0; n < 32; n++)
            T32m[k][n] = (int16_t)(k == 0 ? 32 : lround(32.0 * sqrt(2.0) * cos((2 * n + 1) * k * pi / 64.0)));
    for (int k = 0; k < 16; k++)
        for (int n = 0; n < 16; n++)
            T16m[k][n] = T32m[2 * k][n];
    for (int k = 0; k < 8; k++)
        for (int n = 0; n < 8; n++)
            T8m[k][n] = T32m[4 * k][n];
    for (int k = 0; k < 4; k++)
        for (int n = 0; n < 4; n++)
            T4m[k][n] = T32m[8 * k][n];
}

static const int16_t *tmat(int n)
{
    pthread_once(&tables_once, build_tables);
    switch (n) {
    case 4: return &T4m[0][0];
    case 8: return &T8m[0][0];
    case 16: return &T16m[0][0];
    default: return &T32m[0][0];
    }
}

/* One forward pass (partialButterflyN(src, dst, shift, line), transform.c:
 * 167-498): an n-point transform of each of `lines` input rows, written
 * transposed: out[k][j] = int16((sum_m T[k][m] * in[j][m] + rnd) >> shift),
 * out having `lines` columns.  The butterflies compute exactly this sum
 * because T[k][N-1-m] = (-1)^k T[k][m] (checked by the tests), and every
 * partial sum fits in int32. */
static void fwd_pass_n(const int16_t *in, int16_t *out, int n, int lines, int shift)
{
    const int16_t *T = tmat(n);
    const int add = (1 << shift) >> 1;
    for (int j = 0; j < lines; j++)
        for (int k = 0; k < n; k++) {
            int s = 0;
            for (int m = 0; m < n; m++)
                s += T[k * n + m] * in[j * n + m];
            out[k * lines + j] = (int16_t)((s + add) >> shift);
        }
}

static void fwd_pass(const int16_t *in, int16_t *out, int n, int shift) { fwd_pass_n(in, out, n, n, shift); }

/* One inverse pass (partialButterflyInverseN(src, dst, shift, line, clip)):
 * for each of the `lines` columns j of in (n rows),
 * out[j][m] = clip((sum_k T[k][m] * in[k][j] + rnd) >> shift). */
static void inv_pass_n(const int16_t *in, int16_t *out, int n, int lines, int shift, int clip_depth)
{
    const int16_t *T = tmat(n);
    const int add = (1 << shift) >> 1;
    const int hi = ((1 << clip_depth) >> 1) - 1, lo = -hi - 1;
    for (int j = 0; j < lines; j++)
        for (int m = 0; m < n; m++) {
            int s = 0;
            for (int k = 0; k < n; k++)
                s += T[k * n + m] * in[k * lines + j];
            out[j * n + m] = (int16_t)clip3(lo, hi, (s + add) >> shift);
        }
}

static void inv_pass(const int16_t *in, int16_t *out, int n, int shift, int clip_depth)
{
    inv_pass_n(in, out, n, n, shift, clip_depth);
}

const int16_t *xo_tmat(int n) { return tmat(n); }

static int ilog2(int n) { int l = 0; while ((1 << l) < n) l++; return l; }

/* dct_NxN_c, transform.c:1123-1270: shift1 = log2N - 2 (+1 with the 32x32
 * stride flag), shift2 = log2N + 5. */
void xo_dct(int n, const int16_t *src, int16_t *dst, int i_src)
{
    int16_t block[32 * 32], tmp[32 * 32];
    int l = ilog2(n);
    int shift1 = l + 5 + 8 + 1 - 16;
    int shift2 = l + 5;
    if (n == 32) {
        shift1 += i_src & 1;
        i_src &= 0xFE;
    }
    for (int y = 0; y < n; y++)
        memcpy(block + y * n, src + y * i_src, sizeof(int16_t) * (size_t)n);
    fwd_pass(block, tmp, n, shift1);
    fwd_pass(tmp, dst, n, shift2);
}

/* idct_NxN_c, transform.c:1145-1312: shift 5 clipped to 16 bits, then
 * shift 12 clipped to 9 bits (the 32x32 stride flag trades one bit). */
void xo_idct(int n, const int16_t *src, int16_t *dst, int i_dst)
{
    int16_t block[32 * 32], tmp[32 * 32];
    int a_flag = 0;
    if (n == 32) {
        a_flag = i_dst & 1;
        i_dst &= 0xFE;
    }
    inv_pass(src, tmp, n, 5, 16);
    inv_pass(tmp, block, n, 12 - a_flag, 9 + a_flag);
    for (int y = 0; y < n; y++)
        memcpy(dst + y * i_dst, block + y * n, sizeof(int16_t) * (size_t)n);
}

/* ------------------------------------------------------------------------ */
/* The dct_t table: dctf.dct / idct / dct_half for every size the reference
 * installs (transform.c:1599-1643).                                          */

/* A W x H block transform (dct_WxH_c, transform.c:1123-1490): the W-point
 * transform of each row with shift1 = log2(W) - 2, then the H-point
 * transform of each column with shift2 = log2(H) + 5 (for 8-bit video:
 * B*_IN_BIT + FACTO_BIT + bit depth + 1 - LIMIT_BIT and B*_IN_BIT +
 * FACTO_BIT, with the "- 2" of the asymmetric shapes).  The 32x32, 32x8 and
 * 8x32 transforms take a wavelet flag in the low bit of the stride and read
 * the stride masked with 0xFE; the flag adds 1 to shift1 of 32x32 and 8x32
 * and takes 1 from shift2 of 32x8. */
static int takes_flag(int w, int h) { return (w == 32 && h == 32) || (w == 32 && h == 8) || (w == 8 && h == 32); }

static void dct_rect(int w, int h, const int16_t *src, int16_t *dst, int i_src)
{
    int16_t block[32 * 32], tmp[32 * 32];
    int s1 = ilog2(w) - 2, s2 = ilog2(h) + 5;
    if (takes_flag(w, h)) {
        const int f = i_src & 1;
        if (w == 32 && h == 8)
            s2 -= f;
        else
            s1 += f;
        i_src &= 0xFE;
    }
    for (int y = 0; y < h; y++)
        memcpy(block + y * w, src + (intptr_t)y * i_src, sizeof(int16_t) * (size_t)w);
    fwd_pass_n(block, tmp, w, h, s1);
    fwd_pass_n(tmp, dst, h, w, s2);
}

/* idct_WxH_c (transform.c:1145-1512): the H-point inverse of each column with
 * shift 5 clipped to 16 bits, then the W-point inverse of each row with shift
 * 12 clipped to 9 bits (the wavelet flag: shift 11, clipped to 10 bits). */
static void idct_rect(int w, int h, const int16_t *src, int16_t *dst, int i_dst)
{
    int16_t block[32 * 32], tmp[32 * 32];
    int f = 0;
    if (takes_flag(w, h)) {
        f = i_dst & 1;
        i_dst &= 0xFE;
    }
    inv_pass_n(src, tmp, h, w, 5, 16);
    inv_pass_n(tmp, block, w, h, 12 - f, 9 + f);
    for (int y = 0; y < h; y++)
        memcpy(dst + (intptr_t)y * i_dst, block + y * w, sizeof(int16_t) * (size_t)w);
}

/* One forward lifting pass of the 64-wide transforms (wavelet_64x64_c and
 * kin, transform.c:586-785) over n samples x[0], x[step], ...: the signal is
 * extended by mirroring two samples a side (x[-1] = x[1], x[-2] = x[2],
 * x[n] = x[n-2], x[n+1] = x[n-3]), the odd samples -1, 1, ..., n-1 become
 * x - ((left + right) >> 1), then the even samples become
 * x + ((left + right + 2) >> 2) (rows) or 2x + ((left + right + 1) >> 1)
 * (columns); the n/2 even samples are the low band.  Every store is to a
 * coeff_t (int16). */
static void lift_fwd(const int16_t *in, int step_in, int16_t *out, int step_out, int n, int columns)
{
    int16_t e[64 + 4] = { 0 };
    int16_t *x = e + 2;
    for (int i = 0; i < n; i++)
        x[i] = in[i * step_in];
    x[-1] = x[1];
    x[-2] = x[2];
    x[n] = x[n - 2];
    x[n + 1] = x[n - 3];
    for (int i = -1; i < n; i += 2)
        x[i] = (int16_t)(x[i] - ((x[i - 1] + x[i + 1]) >> 1));
    for (int i = 0; i < n; i += 2)
        x[i] = columns ? (int16_t)((x[i] << 1) + ((x[i - 1] + x[i + 1] + 1) >> 1))
                       : (int16_t)(x[i] + ((x[i - 1] + x[i + 1] + 2) >> 2));
    for (int i = 0; i < n / 2; i++)
        out[i * step_out] = x[2 * i];
}

/* The forward wavelet of a W x H block (W, H in {16, 64}, one of them 64):
 * rows of the contiguous source (stride W), then columns of the half-width
 * result; out is the (W/2) x (H/2) low band, contiguous. */
static void wavelet_fwd(int w, int h, const int16_t *src, int16_t *out)
{
    int16_t mid[64 * 64];
    for (int y = 0; y < h; y++)
        lift_fwd(src + y * w, 1, mid + y * (w / 2), 1, w, 0);
    for (int x = 0; x < w / 2; x++)
        lift_fwd(mid + x, w / 2, out + x, w / 2, h, 1);
}

/* One inverse lifting pass (inv_wavelet_64x64_c and kin, transform.c:787-
 * 956): the n/2 low-band samples go to the even places, x[n] = x[n-2]; for
 * columns every even sample 0..n is halved first; each odd sample becomes
 * (left + right) >> 1. */
static void lift_inv(const int16_t *in, int step_in, int16_t *out, int step_out, int n, int columns)
{
    int16_t e[64 + 2] = { 0 };
    for (int i = 0; i < n / 2; i++)
        e[2 * i] = in[i * step_in];
    e[n] = e[n - 2];
    if (columns)
        for (int i = 0; i <= n; i += 2)
            e[i] = (int16_t)(e[i] >> 1);
    for (int i = 1; i < n; i += 2)
        e[i] = (int16_t)((e[i - 1] + e[i + 1]) >> 1);
    for (int i = 0; i < n; i++)
        out[i * step_out] = e[i];
}

/* The inverse wavelet: columns of the (W/2) x (H/2) low band (contiguous),
 * then rows; out is W x H, contiguous. */
static void wavelet_inv(int w, int h, const int16_t *low, int16_t *out)
{
    int16_t mid[64 * 64];
    for (int x = 0; x < w / 2; x++)
        lift_inv(low + x, w / 2, mid + x, w / 2, h, 1);
    for (int y = 0; y < h; y++)
        lift_inv(mid + y * (w / 2), 1, out + y * w, 1, w, 0);
}

void xo_wavelet_fwd(int w, int h, const int16_t *src, int16_t *low) { wavelet_fwd(w, h, src, low); }
void xo_wavelet_inv(int w, int h, const int16_t *low, int16_t *out) { wavelet_inv(w, h, low, out); }

static int dct_table_size(int w, int h)
{
    return (w == h && (w == 4 || w == 8 || w == 16 || w == 32 || w == 64)) || (w == 16 && h == 4) ||
           (w == 4 && h == 16) || (w == 32 && h == 8) || (w == 8 && h == 32) || (w == 64 && h == 16) ||
           (w == 16 && h == 64);
}

/* dctf.dct[PART_INDEX(w, h)]: the W x H transforms above; the 64-wide sizes
 * (dct_64x64_c, dct_64x16_c, dct_16x64_c, transform.c:1514-1580) read a
 * contiguous source whatever i_src, take the wavelet's low band and run the
 * half-size transform on it with the wavelet flag, leaving (W/2) x (H/2)
 * coefficients in dst.  Returns -1 for a size the table does not hold. */
int xo_dct_part(int w, int h, const int16_t *src, int16_t *dst, int i_src)
{
    if (!dct_table_size(w, h))
        return -1;
    if (w == 64 || h == 64) {
        int16_t low[32 * 32];
        wavelet_fwd(w, h, src, low);
        dct_rect(w / 2, h / 2, low, dst, (w / 2) | 1);
        return 0;
    }
    dct_rect(w, h, src, dst, i_src);
    return 0;
}

/* dctf.idct[PART_INDEX(w, h)]; the 64-wide sizes (idct_64x64_c and kin)
 * invert the half-size transform with the wavelet flag and then the wavelet,
 * writing W x H contiguous samples whatever i_dst. */
int xo_idct_part(int w, int h, const int16_t *src, int16_t *dst, int i_dst)
{
    if (!dct_table_size(w, h))
        return -1;
    if (w == 64 || h == 64) {
        int16_t low[32 * 32];
        idct_rect(w / 2, h / 2, src, low, (w / 2) | 1);
        wavelet_inv(w, h, low, dst);
        return 0;
    }
    idct_rect(w, h, src, dst, i_dst);
    return 0;
}

/* dctf.dct_half (dct_32x32_half_c, dct_64x64_half_c, transform.c:1276-1286,
 * :1525-1530): the 32x32 transform (on the wavelet's low band for 64x64)
 * with every coefficient outside the top-left 16x16 cleared. */
int xo_dct_half_part(int w, int h, const int16_t *src, int16_t *dst, int i_src)
{
    if (!((w == 32 && h == 32) || (w == 64 && h == 64)))
        return -1;
    xo_dct_part(w, h, src, dst, i_src);
    for (int y = 0; y < 32; y++)
        for (int x = 0; x < 32; x++)
            if (x >= 16 || y >= 16)
                dst[y * 32 + x] = 0;
    return 0;
}

/* Quantisation tables, common/quant.c:55-97. */
const uint16_t xo_Q_TAB[80] = {
    32768, 29775, 27554, 25268, 23170, 21247, 19369, 17770,
    16302, 15024, 13777, 12634, 11626, 10624,  9742,  8958,
     8192,  7512,  6889,  6305,  5793,  5303,  4878,  4467,
     4091,  3756,  3444,  3161,  2894,  2654,  2435,  2235,
     2048,  1878,  1722,  1579,  1449,  1329,  1218,  1117,
     1024,   939,   861,   790,   724,   664,   609,   558,
      512,   470,   430,   395,   362,   332,   304,   279,
      256,   235,   215,   197,   181,   166,   152,   140,
      128,   116,   108,    99,    91,    83,    76,    69,
       64,    59,    54,    49,    45,    41,    38,    35,
};
const uint16_t xo_IQ_TAB[80] = {
    32768, 36061, 38968, 42495, 46341, 50535, 55437, 60424,
    32932, 35734, 38968, 42495, 46177, 50535, 55109, 59933,
    65535, 35734, 38968, 42577, 46341, 50617, 55027, 60097,
    32809, 35734, 38968, 42454, 46382, 50576, 55109, 60056,
    65535, 35734, 38968, 42495, 46320, 50515, 55109, 60076,
    65535, 35744, 38968, 42495, 46341, 50535, 55099, 60087,
    65535, 35734, 38973, 42500, 46341, 50535, 55109, 60097,
    32771, 35734, 38965, 42497, 46341, 50535, 55109, 60099,
    32768, 36061, 38968, 42495, 46341, 50535, 55437, 60424,
    32932, 35734, 38968, 42495, 46177, 50535, 55109, 59933,
};
const uint8_t xo_IQ_SHIFT[80] = {
    15, 15, 15, 15, 15, 15, 15, 15,
    14, 14, 14, 14, 14, 14, 14, 14,
    14, 13, 13, 13, 13, 13, 13, 13,
    12, 12, 12, 12, 12, 12, 12, 12,
    12, 11, 11, 11, 11, 11, 11, 11,
    11, 10, 10, 10, 10, 10, 10, 10,
    10,  9,  9,  9,  9,  9,  9,  9,
     8,  8,  8,  8,  8,  8,  8,  8,
     7,  7,  7,  7,  7,  7,  7,  7,
     6,  6,  6,  6,  6,  6,  6,  6,
};

/* quant_c, quant.c:109-121: sign2(c) * ((|c| * scale + add) >> shift), with
 * sign2(0) = +1 (avs2_defs.h:327-330); returns the nonzero count. */
int xo_quant(int16_t *coef, int n, int scale, int shift, int add)
{
    int nz = 0;
    for (int i = 0; i < n; i++) {
        int c = coef[i];
        int s = c < 0 ? -1 : 1;
        coef[i] = (int16_t)(s * ((iabs(c) * scale + add) >> shift));
        nz += coef[i] != 0;
    }
    return nz;
}

/* dequant_c, quant.c:173-184. */
void xo_dequant(int16_t *coef, int n, int scale, int shift)
{
    const int add = 1 << (shift - 1);
    for (int i = 0; i < n; i++)
        if (coef[i] != 0)
            coef[i] = (int16_t)clip3(-32768, 32767, (coef[i] * scale + add) >> shift);
}

/* tu_quant_forward, rdo.c:658-669: shift = 15 + 16 - 9 - log2N and
 * add = floor(2^(shift + b_intra) * 5 / 31) (table at rdo.c:662-666). */
void xo_quant_params(int log2n, int qp, int b_intra, int *scale, int *shift, int *add)
{
    int s = 15 + 16 - (8 + 1) - log2n;
    *scale = xo_Q_TAB[qp];
    *shift = s;
    *add = (int)(((int64_t)1 << (s + b_intra)) * 5 / 31);
}

/* tu_quant_inverse, rdo.c:708-711. */
void xo_dequant_params(int log2n, int qp, int *scale, int *shift)
{
    *scale = xo_IQ_TAB[qp];
    *shift = xo_IQ_SHIFT[qp] + (8 + 1) + log2n - 16;
}

void xo_residual_batch(int n, int qp, int b_intra, const int16_t *in, int16_t *level,
                       int16_t *recon, int32_t *nnz, int nblk)
{
    int l = ilog2(n), qs, qsh, qa, ds, dsh;
    xo_quant_params(l, qp, b_intra, &qs, &qsh, &qa);
    xo_dequant_params(l, qp, &ds, &dsh);
    int16_t c[32 * 32];
    const int nn = n * n;
    for (int b = 0; b < nblk; b++) {
        xo_dct(n, in + (size_t)b * nn, c, n);
        nnz[b] = xo_quant(c, nn, qs, qsh, qa);
        memcpy(level + (size_t)b * nn, c, sizeof(int16_t) * (size_t)nn);
        xo_dequant(c, nn, ds, dsh);
        xo_idct(n, c, recon + (size_t)b * nn, n);
    }
}

/* CHECK_MV_RANGE of a PU's MV box, exported for tests (me.h:46-48, me.c:896-897). */
int xo_check_mv_range(int mx, int my, int min_x, int min_y, int max_x, int max_y)
{
    uint32_t mv_min = pack_mask2(-min_x, -min_y);
    uint32_t mv_max = pack_mask2(max_x, max_y) | 0x8000u;
    return check_mv_range(mx, my, mv_min, mv_max);
}
