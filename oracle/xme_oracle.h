/*
 * xme_oracle.h -- CPU restatement of xavs2's motion-estimation cost path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by
 * tests/, by __graft_entry__.smoke() and by the CPU-baseline leg of bench.py,
 * and by nothing else.  The product (libxavs2_amd.so) never links, calls or
 * falls back to it.
 *
 * Every function restates the arithmetic of the reference (pkuvcl/xavs2,
 * 8-bit build) and cites the file:line it follows.  Pixels are uint8_t
 * (pel_t), coefficients int16_t (coeff_t), distortions int32_t (cmp_dist_t),
 * see source/common/basic_types.h:50-62.
 *
 * Parity pins (which independent check covers which function) are listed in
 * the header comment of xme_oracle.c.
 */
#ifndef XME_ORACLE_H
#define XME_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XO_FENC_STRIDE    64          /* common/common.h:1350 */
#define XO_MAX_DISTORTION (1 << 30)   /* common/defines.h:295 */
#define XO_PAD            80          /* common/defines.h:298 (XAVS2_PAD) */
#define XO_NUM_PU         25          /* common/pixel.h:50-65 */

/* PU sizes in LumaPU order (common/pixel.h:50-65). */
extern const uint8_t xo_pu_w[XO_NUM_PU];
extern const uint8_t xo_pu_h[XO_NUM_PU];
int xo_part_index(int w, int h);      /* PART_INDEX, pixel.h:178 / pixel.c:60-78; -1 if invalid */

/* ---- pixel cost primitives (common/pixel.c) ---------------------------- */
int32_t xo_sad(int w, int h, const uint8_t *p1, intptr_t s1, const uint8_t *p2, intptr_t s2);
void    xo_sad_x3(int w, int h, const uint8_t *fenc, const uint8_t *p0, const uint8_t *p1,
                  const uint8_t *p2, intptr_t i_ref, int32_t res[3]);
void    xo_sad_x4(int w, int h, const uint8_t *fenc, const uint8_t *p0, const uint8_t *p1,
                  const uint8_t *p2, const uint8_t *p3, intptr_t i_ref, int32_t res[4]);
int32_t xo_satd(int w, int h, const uint8_t *p1, intptr_t s1, const uint8_t *p2, intptr_t s2);
void    xo_avg(int w, int h, uint8_t *dst, intptr_t ds, const uint8_t *s0, intptr_t ss0,
               const uint8_t *s1, intptr_t ss1);

/* Batched helpers for tests: call i compares p1 + off1[i] with p2 + off2[i]. */
void xo_sad_batch(int w, int h, const uint8_t *p1, intptr_t s1, const uint8_t *p2, intptr_t s2,
                  const int64_t *off1, const int64_t *off2, int n, int32_t *out);
void xo_satd_batch(int w, int h, const uint8_t *p1, intptr_t s1, const uint8_t *p2, intptr_t s2,
                   const int64_t *off1, const int64_t *off2, int n, int32_t *out);

/* ---- MV rate (encoder/me.c:614-641) ------------------------------------ */
int  xo_mvbits_max_mvd(int search_range);
/* Value of h->mvbits[v]; v must satisfy |v| <= xo_mvbits_max_mvd(). */
int  xo_mvbits(int search_range, int v);
/* LAMBDA_FACTOR(sqrt(lambda)), common/common.h:85, encoder/encoder.c:329 */
int  xo_lambda_factor(double lambda_mode);
/* CHECK_MV_RANGE (me.h:46-48) for the full-pel box [min_x, max_x] x [min_y, max_y]. */
int  xo_check_mv_range(int mx, int my, int min_x, int min_y, int max_x, int max_y);
/* pmvr_adapt_mv (encoder/me.c:412-424): the quarter-pel point for Spiral
 * step (sx, sy) from the half-pel best (mv_x, mv_y) around the PMVR centre
 * (ctr_x, ctr_y) = ((mvp >> 1) << 1); returns 1 when the point is skipped. */
int  xo_pmvr_adapt(int *mx, int *my, int ctr_x, int ctr_y, int mv_x, int mv_y, int sx, int sy);

/* ---- ME search (encoder/me.c, encoder/md_inter.c) ---------------------- */
typedef struct {
    int frame_w, frame_h;             /* h->i_width, h->i_height */
    int search_range;                 /* h->param->search_range (full search: [-R, R)) */
    int lambda_factor;                /* h->i_lambda_factor */
    int min_mv_range[2];              /* h->min_mv_range (qpel), encoder.c:849-892 */
    int max_mv_range[2];              /* h->max_mv_range (qpel) */
    int use_hadamard;                 /* h->param->enable_hadamard: fpel_cmp = satd : sad */
    int use_fractional_me;            /* 0 none, 1 half-pel, 2 half + quarter */
    int use_fast_sub_me;              /* Spiral2 with step 2 */
    int enable_pmvr;                  /* h->param->enable_pmvr */
    int me_method;                    /* h->param->me_method: XO_ME_FS, _DIA, _HEX, _UMH or _TZ */
    int umh_big_hex_level;            /* h->UMH_big_hex_level (presets.c:341): 0, 1 or 2 */
} xo_me_param_t;

/* enum me_mothod_e, common/common.h:145-151. */
#define XO_ME_FS  0
#define XO_ME_DIA 1
#define XO_ME_HEX 2
#define XO_ME_UMH 3
#define XO_ME_TZ  4

/* The per-(PU, reference) encoder state the UMH search reads from xavs2_me_t
 * (set by md_inter.c:640-718): predictor MVs in qpel, the reference index,
 * the predicted SADs and the beta thresholds (the search compares against
 * beta + 1, me.c:883-884). */
typedef struct {
    int mvp1_x, mvp1_y;               /* p_me->mvp1: upper-layer MV ((0,0): none) */
    int mvp2_x, mvp2_y;               /* p_me->mvp2: scaled MV of the previous reference */
    int mvp3_x, mvp3_y;               /* p_me->mvp3: scaled co-located MV ((0,0): none) */
    int ref_idx;                      /* p_me->i_ref_idx */
    int pred_sad;                     /* p_me->pred_sad */
    int pred_sad_uplayer;             /* p_me->pred_sad_uplayer */
    int pad;
    double beta2, beta3;              /* p_me->beta2, p_me->beta3 */
} xo_umh_ctx_t;

typedef struct {
    int pix_x, pix_y;                 /* PU position in the frame */
    int w, h;                         /* PU size (one of the 25 LumaPU sizes) */
    int mvp_x, mvp_y;                 /* MV predictor, qpel */
} xo_pu_t;

typedef struct {
    int32_t cost;                     /* return value of xavs2_me_search */
    int32_t bcost2;                   /* integer-pel best cost (p_me->bcost2) */
    int16_t bmv_x, bmv_y;             /* p_me->bmv, qpel */
    int16_t bmv2_x, bmv2_y;           /* p_me->bmv2, full-pel */
    int32_t mvcost;                   /* p_me->mvcost[PDIR_FWD] */
    int32_t status;                   /* 0 ok; 1 no legal MV candidate (_me_error); 2 (TZ) a raster read
                                       * outside the padded reference, undefined in the reference: the
                                       * point is skipped and the result is not the reference's */
} xo_me_result_t;

/* Integer search (full search, or the DIA / HEX / UMH pattern search by
 * prm->me_method) + sub-pel refinement for one PU against one reference.
 *   cur    : current luma plane origin, stride i_cur (PU at cur + pix_y*i_cur + pix_x)
 *   planes : planes[0] = padded full-pel reference origin; planes[1..15] the
 *            interpolated planes (may be NULL when use_fractional_me == 0).
 *   cost_map (optional, may be NULL): (2R)*(2R) integer-pel costs of the raster
 *            pass, row-major by (j+R, i+R); illegal candidates hold -1.  Full
 *            search only: the pattern searches leave it untouched.
 */
int32_t xo_me_search(const xo_me_param_t *prm, const xo_pu_t *pu,
                     const uint8_t *cur, intptr_t i_cur,
                     const uint8_t *const planes[16], intptr_t i_ref,
                     xo_me_result_t *res, int32_t *cost_map);

/* The same with the UMH context of the PU (ctx may be NULL for the other
 * methods; NULL under XO_ME_UMH reads as an all-zero context). */
int32_t xo_me_search_ctx(const xo_me_param_t *prm, const xo_pu_t *pu, const xo_umh_ctx_t *ctx,
                         const uint8_t *cur, intptr_t i_cur,
                         const uint8_t *const planes[16], intptr_t i_ref,
                         xo_me_result_t *res);

/* The integer stage of a pattern search (prm->me_method XO_ME_DIA, XO_ME_HEX,
 * XO_ME_UMH or XO_ME_TZ) from the full-pel start point (sx, sy) with cost scost, which
 * xo_me_search takes from the MV candidate list; (p0x, p0y) is the first
 * candidate (the clipped MVP), where UMH starts its first diamond.  Writes
 * the best point after each step in order to path[0..2*n) as (x, y) pairs
 * when it moved (at most max_pts, the start point first) and returns their
 * count n; the last one is the integer best.  *route receives, for UMH, the
 * stage the prefix ended in: 0 the big hexagon ran out, 1..4 an early
 * termination after the predictor points, the first cross centre, the cross,
 * the grid, or 5 one after a big-hexagon level; and 2 or 3 in bits 4..7 for a
 * jump to the hexagon or the diamond.  For TZ: bit 0 when the search stopped
 * before the hexagon, bit 1 when the raster step ran, bit 2 when a raster
 * cell read at the offset position (omx, omy) != (0, 0), bit 3 when such a
 * read left the padded plane (undefined), and the number of pattern searches
 * in bits 8..15. */
int xo_me_pattern_path(const xo_me_param_t *prm, const xo_pu_t *pu, const xo_umh_ctx_t *ctx,
                       const uint8_t *cur, intptr_t i_cur,
                       const uint8_t *plane0, intptr_t i_ref,
                       int sx, int sy, int scost, int p0x, int p0y, int *path, int max_pts, int *route);

/* Test hook for the pattern searches' control flow: the integer stage of
 * prm->me_method from (sx, sy) with cost scost over a table of costs instead
 * of pictures, costs[(y + span) * (2 span + 1) + (x + span)] for |x|, |y| <=
 * span (points beyond cost 2^28).  The PU (position, size) and the picture
 * size set the MV window.  Writes the evaluated points in order to trace (at
 * most max_trace (x, y) pairs) and returns their count; *bx, *by and *bcost
 * get the integer best, *route as for xo_me_pattern_path. */
int xo_pattern_landscape(const xo_me_param_t *prm, const xo_pu_t *pu, const xo_umh_ctx_t *ctx,
                         const int32_t *costs, int span, int sx, int sy, int scost, int p0x, int p0y,
                         int *trace, int max_trace, int *bx, int *by, int *bcost, int *route);
/* The same on pictures (as xo_me_pattern_path): the evaluated points in order. */
int xo_me_pattern_trace(const xo_me_param_t *prm, const xo_pu_t *pu, const xo_umh_ctx_t *ctx,
                        const uint8_t *cur, intptr_t i_cur, const uint8_t *plane0, intptr_t i_ref,
                        int sx, int sy, int scost, int p0x, int p0y,
                        int *trace, int max_trace, int *bx, int *by, int *bcost, int *route);

/* Symmetric bi-prediction refinement, xavs2_me_search_sym (me.c:1190-1262),
 * for frame (not field) sequences.  The second MV mirrors the first,
 * scaled by the distances (scale_mv_skip, predict.h:148-151); the prediction
 * is the rounded average of the two references' planes. */
typedef struct {
    int pix_x, pix_y, w, h;           /* the PU */
    int mv_x, mv_y;                   /* *mv on entry (qpel) */
    int mvp_x, mvp_y;                 /* p_me->mvp (MV cost) */
    int mvp1_x, mvp1_y;               /* p_me->mvp1 (PMVR centre) */
    int dist_fwd, dist_bwd;           /* p_me->i_distance_1st / _2nd (> 0) */
    int b_slice;                      /* h->i_type == SLICE_TYPE_B: the second MV is negated */
} xo_sym_job_t;

typedef struct {
    int32_t cost;                     /* return value (XO_MAX_DISTORTION if no point was legal) */
    int32_t mv_x, mv_y;               /* *mv on return (qpel) */
    int32_t mvcost;                   /* p_me->mvcost[PDIR_SYM]; -1 where the search does not set it */
} xo_sym_result_t;

/* scale_mv_skip (predict.h:148-151). */
int xo_scale_mv_skip(int mv, int dist_dst, int dist_src);
int32_t xo_me_search_sym(const xo_me_param_t *prm, const xo_sym_job_t *job,
                         const uint8_t *cur, intptr_t i_cur,
                         const uint8_t *const planes1[16], const uint8_t *const planes2[16], intptr_t i_ref,
                         xo_sym_result_t *res);

/* Bi-directional refinement, xavs2_me_search_bid (me.c:1278-1392).  The
 * backward MV stays fixed; the search refines the forward MV against the
 * target clip(2 * cur - pred_bwd) (sub_ps / add_ps, pixel.c:780-806), whose
 * fpel_cmp is halved once more (ME_COST_QPEL_BID, me.c:388-398).  MV costs
 * against mvp1 (forward) and mvp2 (backward). */
typedef struct {
    int pix_x, pix_y, w, h;           /* the PU */
    int fwd_x, fwd_y;                 /* *fwd_mv on entry (qpel) */
    int bwd_x, bwd_y;                 /* *bwd_mv (qpel, not searched) */
    int mvp1_x, mvp1_y;               /* p_me->mvp1: forward MV cost and PMVR centre */
    int mvp2_x, mvp2_y;               /* p_me->mvp2: backward MV cost */
} xo_bid_job_t;

/* res->mv = *fwd_mv on return; res->mvcost = p_me->mvcost[PDIR_BID] (-1 where
 * the search does not set it).  A forward candidate on a plane that is not
 * built has no cost here (the reference reads through a NULL plane). */
int32_t xo_me_search_bid(const xo_me_param_t *prm, const xo_bid_job_t *job,
                         const uint8_t *cur, intptr_t i_cur,
                         const uint8_t *const planes1[16], const uint8_t *const planes2[16], intptr_t i_ref,
                         xo_sym_result_t *res);

/* The callers of the two refinements, frame sequences, every reference row
 * reconstructed (so check_mv_range's num_lcu_coded_in_row test holds,
 * predict.h:86-95): pred_inter_search_bi (md_inter.c:1036-1128, B slices)
 * and pred_inter_search_dual (md_inter.c:1133-1219, F slices).  lcu_level is
 * h->i_lcu_level (the valid padding is 1 << lcu_level); frame_threads > 1
 * narrows check_mvd's vertical range (md_inter.c:68-76). */
typedef struct {
    int lcu_level;
    int frame_threads;
} xo_gate_param_t;

typedef struct {
    int pix_x, pix_y, w, h;           /* the PU */
    int fwd_imv_x, fwd_imv_y;         /* p_me->all_best_imv[B_FWD] (full-pel) */
    int bwd_x, bwd_y;                 /* all_single_mv[B_BWD] (qpel) */
    int mvp_x, mvp_y;                 /* p_me->mvp on entry: the symmetric search's MV rate */
    int mvp1_x, mvp1_y;               /* all_mvp[B_FWD] */
    int mvp2_x, mvp2_y;               /* all_mvp[B_BWD] */
} xo_bi_job_t;

typedef struct {
    int32_t sym_cost;                 /* *sym_mcost */
    int32_t sym_mv_x, sym_mv_y;       /* all_sym_mv[0] */
    int32_t sym_mvcost;               /* p_me->bmvcost[PDIR_SYM] (-1: the search did not run) */
    int32_t bid_cost;                 /* *bid_mcost */
    int32_t fwd_x, fwd_y, bwd_x, bwd_y;   /* all_dual_mv_1st / _2nd[0] */
    int32_t bid_mvcost;               /* p_me->bmvcost[PDIR_BID] (-1: the search did not run) */
} xo_bi_result_t;

void xo_pred_inter_search_bi(const xo_me_param_t *prm, const xo_gate_param_t *g, const xo_bi_job_t *job,
                             int dist_fwd, int dist_bwd, const uint8_t *cur, intptr_t i_cur,
                             const uint8_t *const planes_fwd[16], const uint8_t *const planes_bwd[16],
                             intptr_t i_ref, xo_bi_result_t *res);

typedef struct {
    int pix_x, pix_y, w, h;           /* the PU */
    int mvp_x, mvp_y;                 /* p_me->mvp on entry */
    int mvp1[4][2];                   /* all_mvp[ref_idx] */
    int imv[4][2];                    /* p_me->all_best_imv[ref_idx] (full-pel) */
} xo_dual_job_t;

typedef struct {
    int32_t cost;                     /* *dual_mcost (XO_MAX_DISTORTION: none) */
    int32_t fst_ref, snd_ref;         /* *dual_best_fst_ref / _snd_ref (-1: none) */
    int32_t mvcost;                   /* p_me->bmvcost[PDIR_DUAL] (-1: not set) */
    int32_t fst[4][2], snd[4][2];     /* all_dual_mv_1st / _2nd[ref_idx] */
} xo_dual_result_t;

/* planes[r] = the 16 planes of reference r (r < n_ref, n_ref 2..4); dist[r] =
 * h->fdec->ref_dpoc[r] (calculate_distance). */
void xo_pred_inter_search_dual(const xo_me_param_t *prm, const xo_gate_param_t *g, const xo_dual_job_t *job,
                               int n_ref, const int dist[4], const uint8_t *cur, intptr_t i_cur,
                               const uint8_t *const *const planes[4], intptr_t i_ref, xo_dual_result_t *res);

/* The same over n PUs on nthreads POSIX threads (CPU baseline); ctxs holds
 * one UMH context per PU or is NULL. */
void xo_me_search_batch(const xo_me_param_t *prm, const xo_pu_t *pus, int n,
                        const uint8_t *cur, intptr_t i_cur,
                        const uint8_t *const planes[16], intptr_t i_ref,
                        xo_me_result_t *res, int nthreads);
void xo_me_search_batch_ctx(const xo_me_param_t *prm, const xo_pu_t *pus, const xo_umh_ctx_t *ctxs, int n,
                            const uint8_t *cur, intptr_t i_cur,
                            const uint8_t *const planes[16], intptr_t i_ref,
                            xo_me_result_t *res, int nthreads);

/* ---- frame borders and interpolated planes (common/frame.c, common/mc.c) */
/* plane_expand_border over a whole plane (frame.c:470-507, :513-535). */
void xo_expand_border(uint8_t *plane, intptr_t stride, int w, int h, int pad);
/* interpolate_lcu_row for every LCU row of a frame (mc.c:593-785).
 * planes[0] must already be border-expanded by XO_PAD.  Planes 1..15 get the
 * same stride and padding. */
int  xo_interpolate_frame(uint8_t *const planes[16], intptr_t stride, int w, int h,
                          int lcu_level, int use_fractional_me);
/* One LCU row of it (interpolate_lcu_row, mc.c:748-785, one slice), and the
 * rows [*y0, *y1) that row produces; the horizontal planes a, b, c are also
 * written 4 rows above and below, as the reference's horizontal pass does. */
int  xo_interpolate_lcu_row(uint8_t *const planes[16], intptr_t stride, int w, int h, int lcu_level, int lcu_y,
                            int use_fractional_me);
void xo_lcu_row_window(int h, int lcu_level, int lcu_y, int *y0, int *y1);
/* xavs2_frame_expand_border_lcurow for one plane (frame.c:537-575, one
 * slice; row_log2 = lcu_level, less 1 for 4:2:0 chroma), and its rows. */
void xo_expand_border_lcu_row(uint8_t *plane, intptr_t stride, int w, int h, int pad, int row_log2, int lcu_y);
void xo_border_row_window(int h, int row_log2, int lcu_y, int *y0, int *y1);

/* Luma inter reconstruction of N x N TUs: the prediction error (sub_ps,
 * rdo.c:1751-1753), tu_recon_inter_luma (rdo.c:1333-1366: DCT, inter quant
 * with b_intra = 0, dequant, IDCT, add_ps) without RDOQ, and the SSD of
 * cu_recon_inter_luma (rdo.c:1497-1499); prediction as rdo_get_pred_inter,
 * rdo.c:1616-1633.  The transform is the full DCT for every N, as
 * tu_get_dct_coeff (rdo.c:764-771) uses it for TUs under 32x32 and in the
 * second RD pass (the first pass of a 32x32 TU under OPT_BIT_EST_PSZT uses
 * dct_half instead).
 * tus[8 * i] = {tu_x, tu_y, pu_x, pu_y, pu_w, pu_h, mv_x, mv_y}; the
 * prediction is mc_luma at the PU's clamped position, offset to the TU.
 * Writes level (n*n per TU), the reconstruction into rec (stride i_rec, at
 * the TU's position), nnz and the SSD against cur per TU. */
/* n = 64: the TU's transform is dct_64x64 (the wavelet's 32x32 low band,
 * transform.c:1514-1541) at quant level 6, 1024 levels per TU. */
#define XO_TU_DCT_HALF 0x100     /* n = 32 or 64 | XO_TU_DCT_HALF: dct_half (rdo.c:764-771) */
void xo_inter_recon_batch(int n, int qp, const uint8_t *cur, intptr_t i_cur, const uint8_t *const planes[16],
                          intptr_t i_ref, int frame_w, int frame_h, const int32_t *tus, int ntu, int16_t *level,
                          uint8_t *rec, intptr_t i_rec, int32_t *nnz, int32_t *ssd);
/* The same for bi-predicted PUs: the average (xo_avg) of the predictions of
 * planes1 at the record's MV and planes2 at mv2[2 i], mv2[2 i + 1]. */
void xo_inter_recon_bi_batch(int n, int qp, const uint8_t *cur, intptr_t i_cur, const uint8_t *const planes1[16],
                             const uint8_t *const planes2[16], intptr_t i_ref, int frame_w, int frame_h,
                             const int32_t *tus, const int32_t *mv2, int ntu, int16_t *level, uint8_t *rec,
                             intptr_t i_rec, int32_t *nnz, int32_t *ssd);
/* The reconstruction of n x n blocks whose prediction is a picture (pred,
 * stride i_pred) at the block's own position: the chroma blocks of inter CUs
 * (cu_recon_chroma, rdo.c:785-846), one call per plane.  blocks[2 i] =
 * {x, y}; level, nnz, ssd per block as above. */
void xo_recon_pred_batch(int n, int qp, const uint8_t *cur, intptr_t i_cur, const uint8_t *pred, intptr_t i_pred,
                         const int32_t *blocks, int nblk, int16_t *level, uint8_t *rec, intptr_t i_rec, int32_t *nnz,
                         int32_t *ssd);
/* The chroma QP of a luma QP, cu_get_chroma_qp (cudata.h:118-125). */
int xo_chroma_qp(int luma_qp);

/* MV predictor, get_mvp_default (encoder/md_inter.c:470-524), for frame
 * sequences: the neighbours left, top, top-left and top-right (index 0..3)
 * of list bwd_2nd, the MVP type rule of get_mvp_type_default (:333-368),
 * the neighbour scaling scale_mv_default (:405-410; B slices use the list's
 * distances, other slices the searched reference's and each neighbour's),
 * then the median rule (derive_median_mv, :372-401) or one neighbour.
 * ref_dpoc / ref_dpoc_multi hold MAX_REFS = 4 entries (common.h:1031-1032).
 * Returns 0, -1 where the scaling product leaves int (undefined in the
 * reference), or -2 for a searched or used neighbour reference index outside
 * 0..3 (the reference asserts blkref < MAX_REFS, predict.h:100). */
typedef struct {
    int ref_idx, bwd_2nd;
    int cb_x, cb_y, cb_w, cb_h;       /* p_cb: the PU inside its CU */
    int ur_available;
    int ref[4];                       /* neighbours' ref_idx[bwd_2nd], -1 = INVALID_REF */
    int mv[4][2];                     /* neighbours' mv[bwd_2nd] */
} xo_mvp_job_t;
int xo_get_mvp_default(const xo_mvp_job_t *j, int b_slice, const int ref_dpoc[4], const int ref_dpoc_multi[4],
                       int *pmv_x, int *pmv_y);

/* The UMH search's encoder state of one (PU, reference): get_mvp_default_sad
 * (md_inter.c:527-608: get_mvp_default's MVP plus pred_sad_space, the
 * neighbour min cost of the MVP's source) and fast_me_prepare_info
 * (md_inter.c:640-718: mvp1 from the upper layer, mvp2 from the nearest
 * reference, mvp3 co-located, pred_sad and the UMH betas).  The record holds
 * what those functions read from the encoder state. */
typedef struct {
    xo_mvp_job_t mvp;                 /* neighbours, PU in its CU, ref_idx, bwd_2nd */
    int nb_sad[3];                    /* all_mincost[...][mode][ref_idx] at the left, top, top-right 4x4 blocks
                                       * (0 where the reference reads none: left/top edge, top-right unavailable) */
    int mode;                         /* PRED_2Nx2N (1) .. PRED_nRx2N (7), common.h:198-207 */
    int mincost_up[3];                /* mincosts[PRED_2NxN / PRED_Nx2N / PRED_2Nx2N][ref_idx] of the CU */
    int best_up[3][2];                /* all_best_mv[PRED_2NxN / Nx2N / 2Nx2N][pu_idx][ref_idx] */
    int mincost_prev;                 /* mincosts[mode][ref_idx - 1] */
    int best_prev[2];                 /* all_best_mv[mode][pu_idx][ref_idx - 1] */
    int col_valid;                    /* fref[0] exists and is not an I frame */
    int col_ref;                      /* fref[0]->pu_ref at the PU's 16x16 block (-1: none) */
    int col_mv[2];                    /* fref[0]->pu_mv there */
    int col_dpoc;                     /* fref[0]->ref_dpoc[col_ref] */
} xo_umh_job_t;

typedef struct {
    int pmv_x, pmv_y;                 /* p_me->mvp */
    int pred_sad_space;
    int status;                       /* 0; -1 a scaling product leaves int; -2 a reference index outside 0..3;
                                       * -3 a distance of 0 divides */
    xo_umh_ctx_t ctx;
} xo_umh_out_t;

/* slice_type: SLICE_TYPE_I 0, P 1, B 2, F 3 (common.h:158-161); umh_bsize =
 * h->umh_bsize (xo_umh_bsize). */
int xo_umh_prepare(const xo_umh_job_t *j, int slice_type, const int ref_dpoc[4], const int ref_dpoc_multi[4],
                   const double umh_bsize[8], xo_umh_out_t *out);
/* xavs2_me_init_umh_threshold (me.c:647-665). */
void xo_umh_bsize(int i_qp, int slice_type, double bsize[8]);

/* cu_get_mc_pos (common/mc.h:52-66). */
int  xo_get_mc_pos(int img_size, int blk_size, int blk_pos, int mv);
/* Block MC (mc.c:275-398, :559-589) and chroma MC (mc.c:797-829).
 * pix_quad_x/y are absolute positions: 4 * luma position + MV (qpel). */
void xo_mc_luma(uint8_t *dst, intptr_t i_dst, int pix_quad_x, int pix_quad_y, int w, int h,
                const uint8_t *const planes[16], intptr_t i_ref);
void xo_mc_chroma(uint8_t *dst_u, uint8_t *dst_v, intptr_t i_dst, int pix_quad_x, int pix_quad_y,
                  int w, int h, const uint8_t *plane_u, const uint8_t *plane_v, intptr_t i_src);
void xo_luma_block_hor(uint8_t *dst, intptr_t i_dst, const uint8_t *src, intptr_t i_src, int w, int h, int dx);
void xo_luma_block_ver(uint8_t *dst, intptr_t i_dst, const uint8_t *src, intptr_t i_src, int w, int h, int dy);
void xo_luma_block_ext(uint8_t *dst, intptr_t i_dst, const uint8_t *src, intptr_t i_src, int w, int h, int dx, int dy);
/* The interpolation members of g_funcs (primitives.h:157-171; C versions
 * mc.c:275-547) with their own 8 (luma) or 4 (chroma) taps.  tmp of
 * xo_intpl_luma_hor may be NULL (intpl_luma_block_hor). */
void xo_intpl_luma_hor(uint8_t *dst, intptr_t i_dst, int16_t *tmp, intptr_t i_tmp, const uint8_t *src,
                       intptr_t i_src, int w, int h, const int8_t *coeff);
void xo_intpl_luma_ver(uint8_t *dst, intptr_t i_dst, const uint8_t *src, intptr_t i_src, int w, int h,
                       const int8_t *coeff);
void xo_intpl_luma_ext(uint8_t *dst, intptr_t i_dst, const int16_t *tmp, intptr_t i_tmp, int w, int h,
                       const int8_t *coeff);
void xo_intpl_luma_block_ext(uint8_t *dst, intptr_t i_dst, const uint8_t *src, intptr_t i_src, int w, int h,
                             const int8_t *coeff_h, const int8_t *coeff_v);
void xo_intpl_chroma_block_hor(uint8_t *dst, intptr_t i_dst, const uint8_t *src, intptr_t i_src, int w, int h,
                               const int8_t *coeff);
void xo_intpl_chroma_block_ver(uint8_t *dst, intptr_t i_dst, const uint8_t *src, intptr_t i_src, int w, int h,
                               const int8_t *coeff);
void xo_intpl_chroma_block_ext(uint8_t *dst, intptr_t i_dst, const uint8_t *src, intptr_t i_src, int w, int h,
                               const int8_t *coeff_h, const int8_t *coeff_v);
void xo_chroma_block_hor(uint8_t *dst, intptr_t i_dst, const uint8_t *src, intptr_t i_src, int w, int h, int dx);
void xo_chroma_block_ver(uint8_t *dst, intptr_t i_dst, const uint8_t *src, intptr_t i_src, int w, int h, int dy);
void xo_chroma_block_ext(uint8_t *dst, intptr_t i_dst, const uint8_t *src, intptr_t i_src, int w, int h, int dx, int dy);

/* ---- transform and quantisation (common/transform.c, common/quant.c) --- */
void xo_dct(int n, const int16_t *src, int16_t *dst, int i_src);   /* dct_NxN_c, dst contiguous */
void xo_idct(int n, const int16_t *src, int16_t *dst, int i_dst);  /* idct_NxN_c, src contiguous */
/* The dct_t table (transform.c:1599-1643): dctf.dct / idct / dct_half of a
 * W x H block for the sizes the reference installs (4x4 .. 64x64, 16x4,
 * 4x16, 32x8, 8x32, 64x16, 16x64; dct_half: 32x32 and 64x64), with the
 * reference's stride conventions (the wavelet flag in the low bit of the
 * 32x32 / 32x8 / 8x32 stride, contiguous blocks for the 64-wide sizes, whose
 * coefficients are the (W/2) x (H/2) transform of the wavelet's low band).
 * Coefficients are contiguous.  Return -1 for a size without an entry. */
int  xo_dct_part(int w, int h, const int16_t *src, int16_t *dst, int i_src);
/* The lifting wavelet of the 64-wide sizes alone (64x64, 64x16, 16x64):
 * W x H contiguous samples <-> the (W/2) x (H/2) low band. */
void xo_wavelet_fwd(int w, int h, const int16_t *src, int16_t *low);
void xo_wavelet_inv(int w, int h, const int16_t *low, int16_t *out);
int  xo_idct_part(int w, int h, const int16_t *src, int16_t *dst, int i_dst);
int  xo_dct_half_part(int w, int h, const int16_t *src, int16_t *dst, int i_src);
int  xo_quant(int16_t *coef, int n, int scale, int shift, int add);
void xo_dequant(int16_t *coef, int n, int scale, int shift);
/* Parameters of tu_quant_forward / tu_quant_inverse (encoder/rdo.c:658-718). */
void xo_quant_params(int log2n, int qp, int b_intra, int *scale, int *shift, int *add);
void xo_dequant_params(int log2n, int qp, int *scale, int *shift);
/* DCT -> quant -> dequant -> IDCT on nblk contiguous n*n blocks.
 * level = quantised coefficients, recon = reconstructed residual, nnz per block. */
void xo_residual_batch(int n, int qp, int b_intra, const int16_t *in, int16_t *level,
                       int16_t *recon, int32_t *nnz, int nblk);

/* Row-major N x N AVS2 transform matrix, N in {4, 8, 16, 32}. */
const int16_t *xo_tmat(int n);
extern const uint16_t xo_Q_TAB[80], xo_IQ_TAB[80];
extern const uint8_t xo_IQ_SHIFT[80];
extern const int8_t xo_luma_taps[4][8], xo_chroma_taps[8][4];

#ifdef __cplusplus
}
#endif
#endif
