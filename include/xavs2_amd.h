/*
 * xavs2_amd.h -- C ABI of libxavs2_amd.so, the MI355X (gfx950) motion-estimation
 * cost path for the AVS2 encoder xavs2.
 *
 * The reference encoder reaches its ME arithmetic only through the global
 * primitive table `intrinsic_func_t g_funcs` (source/common/primitives.h:131-223,
 * filled by xavs2_init_all_primitives, source/common/primitives.c:53-80) and the
 * search driver xavs2_me_search (source/encoder/me.c:879).  This header exposes
 *
 *   1. the per-call primitive entry points with exactly the g_funcs signatures
 *      (host pointers, synchronous; for installing into g_funcs as a drop-in),
 *   2. batched device-side versions of the same primitives (device pointers,
 *      asynchronous on a caller stream) -- one call replaces n table calls,
 *   3. the frame-level ME job that replaces the per-PU xavs2_me_search calls of
 *      one frame (full search + sub-pel refinement over every PU and reference)
 *      in one persistent launch, replayable from a captured HIP graph,
 *   4. the reference-plane producer (interpolate_lcu_row over a frame), chroma
 *      MC, and the residual path DCT -> quant -> dequant -> IDCT,
 *   5. a continuous-batching ME server for many concurrent encodes.
 *
 * Every entry point returns XME_OK (0) or a negative XME_ERR_* code; the text
 * of the last error on the calling thread is available from xme_last_error().
 * Pointers named d_* are device pointers (from xme_malloc or any HIP
 * allocation on the current device); `stream` is a hipStream_t or NULL.
 * Results are bit-exact with the reference's C primitives for 8-bit video.
 */
#ifndef XAVS2_AMD_H
#define XAVS2_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XME_OK           0
#define XME_ERR_ARG     -1   /* invalid argument (size, shape, range) */
#define XME_ERR_HIP     -2   /* HIP runtime error */
#define XME_ERR_NOMEM   -3   /* allocation failed */
#define XME_ERR_BUDGET  -4   /* server: job does not fit the memory budget */
#define XME_ERR_STATE   -5   /* object used in the wrong state */

#define XME_NUM_PU      25   /* enum LumaPU, source/common/pixel.h:50-65 */
#define XME_FENC_STRIDE 64   /* FENC_STRIDE, source/common/common.h:1350 */
#define XME_PAD         80   /* XAVS2_PAD, source/common/defines.h:298 */
#define XME_MAX_REFS    8
#define XME_MVP_REFS    4            /* MAX_REFS (source/common/defines.h:301, :364) */
#define XME_MAX_DISTORTION (1 << 30)  /* source/common/defines.h:295 */

typedef uint8_t pel_t;       /* source/common/basic_types.h:50 */
typedef int16_t coeff_t;     /* source/common/basic_types.h:60 */
typedef int16_t mct_t;       /* source/common/basic_types.h:59 (8-bit build) */

const char *xme_last_error(void);
/* Library version string. */
const char *xme_version(void);

/* ---- device, memory and streams --------------------------------------- */
int xme_device_count(int *count);
int xme_set_device(int device);
int xme_malloc(void **d_ptr, size_t bytes);
int xme_free(void *d_ptr);
int xme_host_alloc(void **h_ptr, size_t bytes);    /* pinned host memory */
int xme_host_free(void *h_ptr);
int xme_memcpy_h2d(void *d_dst, const void *h_src, size_t bytes, void *stream);
int xme_memcpy_d2h(void *h_dst, const void *d_src, size_t bytes, void *stream);
int xme_memcpy_d2d(void *d_dst, const void *d_src, size_t bytes, void *stream);
int xme_memset(void *d_dst, int value, size_t bytes, void *stream);
int xme_stream_create(void **stream);
int xme_stream_destroy(void *stream);
int xme_stream_sync(void *stream);
int xme_device_sync(void);

/* ---- 1. per-call primitives with the g_funcs signatures ---------------- *
 * pixel_cmp_t     (source/common/pixel.h:124): pixf.sad[part], pixf.satd[part]
 * pixel_cmp_x3_t  (pixel.h:127):               pixf.sad_x3[part] (fenc stride FENC_STRIDE)
 * pixel_cmp_x4_t  (pixel.h:128):               pixf.sad_x4[part]
 * Each call stages the blocks to the GPU, computes there, and returns the
 * reference's value.  The table getters fill arrays indexed by enum LumaPU so
 * an encoder can install them as a new g_funcs tier after
 * xavs2_init_all_primitives (source/common/primitives.c:53-80). */
typedef int32_t (*xme_pixel_cmp_t)(const pel_t *pix1, intptr_t i_pix1, const pel_t *pix2, intptr_t i_pix2);
typedef void (*xme_pixel_cmp_x3_t)(const pel_t *fenc, const pel_t *pix0, const pel_t *pix1, const pel_t *pix2,
                                   intptr_t i_stride, int32_t scores[3]);
typedef void (*xme_pixel_cmp_x4_t)(const pel_t *fenc, const pel_t *pix0, const pel_t *pix1, const pel_t *pix2,
                                   const pel_t *pix3, intptr_t i_stride, int32_t scores[4]);
int xme_get_sad_table(xme_pixel_cmp_t table[XME_NUM_PU]);      /* replaces pixf.sad    (pixel.c:1141) */
int xme_get_satd_table(xme_pixel_cmp_t table[XME_NUM_PU]);     /* replaces pixf.satd   (pixel.c:1144) */
int xme_get_sad_x3_table(xme_pixel_cmp_x3_t table[XME_NUM_PU]);/* replaces pixf.sad_x3 (pixel.c:1142) */
int xme_get_sad_x4_table(xme_pixel_cmp_x4_t table[XME_NUM_PU]);/* replaces pixf.sad_x4 (pixel.c:1143) */

/* dct_t (source/common/primitives.h:92): dctf.dct / idct / dct_half[part].
 * The getter fills the entries xavs2_dct_init installs (transform.c:1605-1643:
 * 4x4, 8x8, 16x16, 32x32, 64x64, 16x4, 4x16, 32x8, 8x32, 64x16, 16x64;
 * dct_half at 32x32 and 64x64) and NULL elsewhere, as the reference leaves
 * them.  Conventions as the reference's: the low bit of the 32x32 / 32x8 /
 * 8x32 stride is the wavelet flag (the stride read masked with 0xFE); the
 * 64-wide sizes read and write contiguous blocks, their coefficients the
 * (W/2) x (H/2) transform of the lifting wavelet's low band; src == dst is
 * allowed. */
typedef void (*xme_dct_t)(const coeff_t *src, coeff_t *dst, int i_src);
int xme_get_dct_table(xme_dct_t dct[XME_NUM_PU], xme_dct_t idct[XME_NUM_PU], xme_dct_t dct_half[XME_NUM_PU]);
/* dctf.quant / dctf.dequant (primitives.h:116-117; quant_c / dequant_c,
 * source/common/quant.c:109-184): in place on i_coef coefficients; quant
 * returns the nonzero count. */
typedef int (*xme_quant_t)(coeff_t *coef, const int i_coef, const int scale, const int shift, const int add);
typedef void (*xme_dequant_t)(coeff_t *coef, const int i_coef, const int scale, const int shift);
int xme_get_quant_funcs(xme_quant_t *quant, xme_dequant_t *dequant);
/* pixel_avg_pp_t (source/common/pixel.h:137): pixf.avg[part] (pixel.c:680-719),
 * (src0 + src1 + 1) >> 1 over the part's W x H; weight is unused. */
typedef void (*xme_pixel_avg_pp_t)(pel_t *dst, intptr_t dstride, const pel_t *src0, intptr_t sstride0,
                                   const pel_t *src1, intptr_t sstride1, int weight);
int xme_get_avg_table(xme_pixel_avg_pp_t table[XME_NUM_PU]);
/* The interpolation members of g_funcs (source/common/primitives.h:157-171,
 * types :62-71; C versions source/common/mc.c:275-547), any block size, with
 * the caller's taps: luma 8 taps at -3 .. +4, chroma 4 at -1 .. +2; the
 * intermediate of the luma filters is mct_t (int16), of the chroma one int32.
 * intpl_luma_hor also writes the raw 8-tap sums to tmp; intpl_luma_ext filters
 * such int16 rows down. */
typedef struct {
    void (*intpl_luma_hor)(pel_t *dst, int i_dst, mct_t *tmp, int i_tmp, pel_t *src, int i_src, int width, int height,
                           const int8_t *coeff);                                          /* intpl_luma_hor_t */
    void (*intpl_luma_ver)(pel_t *dst, int i_dst, pel_t *src, int i_src, int width, int height,
                           int8_t const *coeff);                                          /* intpl_luma_ver_t */
    void (*intpl_luma_ext)(pel_t *dst, int i_dst, mct_t *tmp, int i_tmp, int width, int height,
                           const int8_t *coeff);                                          /* intpl_luma_ext_t */
    void (*intpl_luma_ver_x3)(pel_t *const dst[3], int i_dst, pel_t *src, int i_src, int width, int height,
                              int8_t const **coeff);                                      /* intpl_luma_ver_x3_t */
    void (*intpl_luma_hor_x3)(pel_t *const dst[3], int i_dst, mct_t *const tmp[3], int i_tmp, pel_t *src,
                              int i_src, int width, int height, const int8_t **coeff);    /* intpl_luma_hor_x3_t */
    void (*intpl_luma_ext_x3)(pel_t *const dst[3], int i_dst, mct_t *tmp, int i_tmp, int width, int height,
                              const int8_t **coeff);                                      /* intpl_luma_ext_x3_t */
    void (*intpl_luma_block_hor)(pel_t *dst, int i_dst, pel_t *src, int i_src, int width, int height,
                                 const int8_t *coeff);                                    /* intpl_t */
    void (*intpl_luma_block_ver)(pel_t *dst, int i_dst, pel_t *src, int i_src, int width, int height,
                                 const int8_t *coeff);                                    /* intpl_t */
    void (*intpl_luma_block_ext)(pel_t *dst, int i_dst, pel_t *src, int i_src, int width, int height,
                                 const int8_t *coeff_x, const int8_t *coeff_y);           /* intpl_ext_t */
    void (*intpl_chroma_block_hor)(pel_t *dst, int i_dst, pel_t *src, int i_src, int width, int height,
                                   const int8_t *coeff);                                  /* intpl_t */
    void (*intpl_chroma_block_ver)(pel_t *dst, int i_dst, pel_t *src, int i_src, int width, int height,
                                   const int8_t *coeff);                                  /* intpl_t */
    void (*intpl_chroma_block_ext)(pel_t *dst, int i_dst, pel_t *src, int i_src, int width, int height,
                                   const int8_t *coeff_x, const int8_t *coeff_y);         /* intpl_ext_t */
} xme_intpl_funcs;
int xme_get_intpl_funcs(xme_intpl_funcs *funcs);

/* ---- 2. batched primitives on device buffers --------------------------- *
 * Call i of a batch evaluates the g_funcs member on (base + off[i]) pointers. */
/* pixf.sad[part] / pixf.satd[part] (pixel.c:93-135, :297-407). */
int xme_sad_batch(int part, const pel_t *d_fenc, intptr_t i_fenc, const pel_t *d_fref, intptr_t i_fref,
                  const int64_t *d_fenc_off, const int64_t *d_fref_off, int n, int32_t *d_out, void *stream);
int xme_satd_batch(int part, const pel_t *d_fenc, intptr_t i_fenc, const pel_t *d_fref, intptr_t i_fref,
                   const int64_t *d_fenc_off, const int64_t *d_fref_off, int n, int32_t *d_out, void *stream);
/* pixf.sad_x3/sad_x4[part] (pixel.c:143-243): fenc has stride XME_FENC_STRIDE;
 * d_fref_off holds 3 (or 4) offsets per call; d_scores 3 (or 4) per call. */
int xme_sad_x3_batch(int part, const pel_t *d_fenc, const pel_t *d_fref, intptr_t i_fref,
                     const int64_t *d_fenc_off, const int64_t *d_fref_off, int n, int32_t *d_scores, void *stream);
int xme_sad_x4_batch(int part, const pel_t *d_fenc, const pel_t *d_fref, intptr_t i_fref,
                     const int64_t *d_fenc_off, const int64_t *d_fref_off, int n, int32_t *d_scores, void *stream);

/* ---- 3. frame-level motion estimation ---------------------------------- */
typedef struct {
    int32_t frame_w, frame_h;        /* h->i_width, h->i_height */
    int32_t search_range;            /* full search over [-R, R) (me.c:1159-1160) */
    int32_t lambda_factor;           /* h->i_lambda_factor (encoder.c:329) */
    int32_t min_mv_range[2];         /* h->min_mv_range, qpel (encoder.c:849-892) */
    int32_t max_mv_range[2];         /* h->max_mv_range, qpel */
    int32_t use_hadamard;            /* enable_hadamard: fpel_cmp = satd, else sad (encoder.c:1840-1846) */
    int32_t use_fractional_me;       /* 0 integer only, 1 + half-pel, 2 + quarter-pel */
    int32_t use_fast_sub_me;         /* Spiral2 with step 2 (me.c:501-502) */
    int32_t enable_pmvr;             /* PMVR quarter-pel adaptation (me.c:549-556) */
    int32_t me_method;               /* h->param->me_method: XME_ME_FS, _DIA, _HEX, _UMH or _TZ */
    int32_t umh_big_hex_level;       /* h->UMH_big_hex_level (presets.c:341): 0, 1 or 2 */
} xme_me_params;

/* Integer search methods, enum me_mothod_e (source/common/common.h:145-151).
 * XME_ME_FS is the full search over [-R, R)^2 (me.c:1156-1164); XME_ME_DIA
 * the small-diamond descent (me.c:1133-1154); XME_ME_HEX the hexagon descent
 * followed by the diamond (me.c:1107-1131); XME_ME_UMH the UMH steps
 * (me.c:1030-1099) followed by the hexagon or the diamond; XME_ME_TZ the TZ
 * steps (me.c:926-1028, tz_pattern_search me.c:669-874) followed by the
 * hexagon and the diamond, or stopping after its two-point search.  TZ's
 * raster step evaluates its points offset once more by the diamond's centre
 * (pix_base of ME_COST_IPEL_X4, me.c:245); where such a block leaves the
 * padded reference (XME_PAD pixels; within 4 pixels of its right edge counts
 * as leaving) or its MV difference leaves the mvbits table, the reference
 * reads outside its arrays: the point is skipped and the result's status is
 * 2.  So is a search whose best after such a point is too close to the
 * padding's edge for the hexagon's reads: it stops there. */
#define XME_ME_FS  0
#define XME_ME_DIA 1
#define XME_ME_HEX 2
#define XME_ME_UMH 3
#define XME_ME_TZ  4

/* Per (PU, reference) encoder state that the UMH search reads from
 * xavs2_me_t (set by pred_inter_search_single's helpers,
 * source/encoder/md_inter.c:640-718).  MVs in qpel; the search compares the
 * best cost against pred_sad * (beta + 1) (me.c:165-170, :883-884). */
typedef struct {
    int32_t mvp1_x, mvp1_y;          /* p_me->mvp1 ((0,0): none) */
    int32_t mvp2_x, mvp2_y;          /* p_me->mvp2 (tried when ref_idx > 0) */
    int32_t mvp3_x, mvp3_y;          /* p_me->mvp3 ((0,0): none) */
    int32_t ref_idx;                 /* p_me->i_ref_idx */
    int32_t pred_sad;                /* p_me->pred_sad */
    int32_t pred_sad_uplayer;        /* p_me->pred_sad_uplayer */
    int32_t reserved;
    double beta2, beta3;             /* p_me->beta2, p_me->beta3 */
} xme_umh_ctx;

typedef struct {
    int16_t x, y;                    /* PU position (luma pixels) */
    uint8_t w, h;                    /* one of the 25 LumaPU sizes */
    uint8_t reserved[2];
    int16_t mvp_x, mvp_y;            /* MV predictor, qpel (p_me->mvp) */
} xme_pu;

typedef struct {
    int32_t cost;                    /* xavs2_me_search return value */
    int32_t bcost2;                  /* p_me->bcost2: integer-pel best cost */
    int16_t bmv_x, bmv_y;            /* p_me->bmv (qpel) */
    int16_t bmv2_x, bmv2_y;          /* p_me->bmv2 (full-pel) */
    int32_t mvcost;                  /* p_me->mvcost[PDIR_FWD] */
    int32_t status;                  /* 0 ok, 1 no legal candidate (returns XME_MAX_DISTORTION), 2 TZ read
                                      * outside the reference's arrays (undefined there; see XME_ME_TZ) */
} xme_result;

/* One reference picture resident in device memory: planes[0] is the full-pel
 * luma plane (origin pointer; XME_PAD pixels of replicated border around it),
 * planes[1..15] the interpolated planes at the same stride
 * (xavs2_frame_t::filtered, source/common/frame.c:335-367).  Unused planes may
 * be NULL when use_fractional_me < 2. */
typedef struct {
    const pel_t *planes[16];
    intptr_t stride;
} xme_ref_desc;

#define XME_PLAN_BATCHED 0   /* one persistent launch for every PU shape and reference */
#define XME_PLAN_PLAIN   1   /* one launch per PU shape and reference (comparison path) */

typedef struct xme_plan xme_plan;

/* Prepare a frame layout: n_pu PUs searched against n_ref references.
 * The plan owns the device work lists and is reused for every frame with the
 * same layout.  Fails with XME_ERR_ARG when a PU is not a LumaPU size, when
 * a motion-vector difference could leave the reference's mvbits table, or
 * when n_ref > XME_MAX_REFS. */
int xme_plan_create(const xme_me_params *params, const xme_pu *pus, int n_pu, int n_ref, int mode,
                    xme_plan **plan);
int xme_plan_destroy(xme_plan *plan);
/* Per-plan counters: number of PUs x refs, integer candidates evaluated per
 * run and their byte absdiffs (full search; 0 for the pattern searches, whose
 * work depends on the pictures: see xme_plan_search_count). */
int xme_plan_info(const xme_plan *plan, int64_t *n_inst, int64_t *n_candidates, int64_t *n_absdiff);
/* Pattern-search plans: the integer points the last run evaluated, and their
 * byte absdiffs (a blocking read after the plan's work on the null stream; 0
 * for full-search plans). */
int xme_plan_search_count(const xme_plan *plan, int64_t *n_points, int64_t *n_absdiff);
/* UMH plans: the encoder state of every (PU, reference), ctx[pu * n_ref + r],
 * for the following runs (copied to the device on `stream`; the caller's
 * array may be reused once the stream reaches that point).  A UMH plan fails
 * to run with XME_ERR_STATE until it has been set.  No reference
 * counterpart: the reference reads these fields from xavs2_me_t per call
 * (me.c:1030-1099). */
int xme_plan_set_umh_context(xme_plan *plan, const xme_umh_ctx *ctx, void *stream);
/* Caps the workgroups of the batched plan's persistent search launch (0 =
 * as many as the device holds at once), so that searches of concurrent
 * frames share the compute units instead of queueing behind each other. */
int xme_plan_set_max_workgroups(xme_plan *plan, int max_workgroups);
int xme_plan_workgroups(const xme_plan *plan, int *workgroups, int *device_fill);
/* Splits the batched plan's search into launches of at most this many
 * byte absdiffs each (0 = one launch per kernel variant), so that the kernels
 * of other frames can start between them. */
int xme_plan_set_launch_split(xme_plan *plan, int64_t max_absdiff_per_launch);

/* Run the search for one frame.  d_cur: current luma plane origin (no padding
 * needed).  refs: host array of n_ref descriptors.  d_results: n_pu * n_ref
 * entries, entry pu * n_ref + r.  d_cost_map (optional, full search only):
 * per entry (2R)^2 integer costs of the raster pass, -1 where CHECK_MV_RANGE
 * rejects.  */
int xme_me_run(xme_plan *plan, const pel_t *d_cur, intptr_t i_cur, const xme_ref_desc *refs,
               xme_result *d_results, int32_t *d_cost_map, void *stream);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture elsewhere in the process, tight
 * serving loops, harnesses that forbid hipMalloc): the frame-level ME job of
 * xme_plan_create (batched) + xme_me_run in one call, with the plan's device
 * buffers (work lists, per-instance state and keys, queue counters, frame
 * descriptor) in a static device scratch of XME_STATIC_SCRATCH_BYTES per GPU,
 * held from the library's load until the process ends.  The call makes no
 * device allocation; it returns once the results are in d_results (it waits
 * for `stream`).  Full search, DIA and HEX (UMH needs its context: XME_ERR_ARG).
 * XME_ERR_BUDGET when the job's plan does not fit the scratch (about 100
 * bytes per PU and reference: a 1080p frame of 8x8 PUs with 2 references
 * fits).  Calls of one process are serialised; calls from several processes
 * on one GPU each have their own scratch. */
#define XME_STATIC_SCRATCH_BYTES (64u << 20)
int xme_me_search_static(const xme_me_params *prm, const xme_pu *pus, int n_pu, int n_ref, const pel_t *d_cur,
                         intptr_t i_cur, const xme_ref_desc *refs, xme_result *d_results, void *stream);
/* Capture the kernels of xme_me_run into a HIP graph on the first call and
 * replay it afterwards (steady-state frames).  Pointers may change between
 * calls: they travel through a device-side frame descriptor. */
int xme_me_run_graph(xme_plan *plan, const pel_t *d_cur, intptr_t i_cur, const xme_ref_desc *refs,
                     xme_result *d_results, void *stream);

/* Measurement hook: runs the job `reps` times on `stream` with HIP events
 * around the integer-search kernel and around the whole job; returns the mean
 * durations in milliseconds.  Per-run job durations go to job_ms[reps] when
 * it is not NULL. */
int xme_me_time(xme_plan *plan, const pel_t *d_cur, intptr_t i_cur, const xme_ref_desc *refs,
                xme_result *d_results, void *stream, int reps, double *search_ms, double *job_ms_mean,
                double *job_ms);
/* A plain streaming device copy (one read and one write per byte), used to
 * calibrate memory-traffic counters. */
int xme_copy_kernel(void *d_dst, const void *d_src, size_t bytes, void *stream);

/* Autotuner: time the integer-search tile shapes for (PU shape, search range)
 * on this device and store the winners for later plans.  Returns the number
 * of shapes tuned.  xme_tune_get reports the chosen rows-per-lane (K) and
 * lanes-per-row (LC) for a shape. */
int xme_tune(const xme_me_params *params, const int *parts, int n_parts, void *stream);
int xme_tune_get(int part, int search_range, int *k_rows, int *lanes_per_row);
/* Installs a choice measured earlier (a shipped tuning table); K rows per lane
 * is 4 or 8 (8 falls back to 4 for PU heights it does not divide), or 16 for
 * PUs 8 or 16 pixels wide and at most 16 rows high (the short-PU kernel);
 * LC is 8, 16, 32 or 64 lanes per candidate row. */
int xme_tune_set(int part, int search_range, int k_rows, int lanes_per_row);

/* Symmetric bi-prediction refinement, xavs2_me_search_sym
 * (source/encoder/me.c:1190-1262), for a batch of PUs of one frame against one
 * pair of references (ref1 = p_fref_1st, ref2 = p_fref_2nd; the reference
 * addresses both with ref1's stride, so the strides must match).  Each job
 * starts from mv (qpel) and evaluates ME_COST_QPEL_SYM (me.c:353-382): the
 * rounded average of ref1 at mv and ref2 at the mirrored MV
 * scale_mv_skip(mv, dist_bwd, dist_fwd) (source/common/predict.h:148-151,
 * negated when b_slice), both range-checked against the PU's qpel window
 * (md_inter.c:913-924), costed with SATD (use_hadamard) or SAD plus the MV
 * rate against mvp.  Half-pel Spiral[0..4] (Spiral[1..4] with SAD), then
 * quarter-pel Spiral[1..4] with PMVR centred on mvp1 when enable_pmvr.
 * Frame sequences (b_field_sequence == 0, scale_mv_skip_y = scale_mv_skip).
 * Uses frame_w/h, search_range (mvbits table), lambda_factor, min/max_mv_range,
 * use_hadamard, use_fractional_me and enable_pmvr of prm.  d_cur must be
 * 4-byte aligned with i_cur % 4 == 0. */
typedef struct {
    int16_t x, y;                    /* PU position (luma pixels, x % 4 == 0) */
    uint8_t w, h;                    /* PU size, multiples of 4 up to 64 */
    uint8_t reserved[2];
    int16_t mv_x, mv_y;              /* start MV, qpel (*mv) */
    int16_t mvp_x, mvp_y;            /* p_me->mvp */
    int16_t mvp1_x, mvp1_y;          /* p_me->mvp1 (PMVR centre) */
    int16_t dist_fwd, dist_bwd;      /* p_me->i_distance_1st, i_distance_2nd */
    int32_t b_slice;                 /* h->i_type == SLICE_TYPE_B: mirror the scaled MV */
} xme_sym_job;

typedef struct {
    int32_t cost;                    /* return value (XME_MAX_DISTORTION when no point had a cost) */
    int16_t mv_x, mv_y;              /* *mv on return (qpel) */
    int32_t mvcost;                  /* p_me->mvcost[PDIR_SYM]; -1 when use_fractional_me == 0 (not set) */
    int32_t status;                  /* 0 ok; 2 job not evaluated: outside the frame, a distance of 0 or
                                        beyond +-255, an MV beyond +-8191 qpel (also once scaled) or an
                                        MV difference outside the mvbits table (the caller's
                                        check_mv_range / check_mvd gate, md_inter.c:1177-1184) */
} xme_sym_result;

int xme_me_search_sym_batch(const xme_me_params *prm, const xme_sym_job *d_jobs, int n, const pel_t *d_cur,
                            intptr_t i_cur, const xme_ref_desc *ref1, const xme_ref_desc *ref2,
                            xme_sym_result *d_results, void *stream);

/* Bi-directional refinement, xavs2_me_search_bid (source/encoder/me.c:1278-1392),
 * for a batch of PUs against one pair of references (ref1 = p_fref_1st,
 * forward; ref2 = p_fref_2nd, backward; equal strides).  The backward MV stays
 * fixed: the target is clip(2 * cur - pred2 at bwd) (sub_ps / add_ps,
 * pixel.c:780-806), and the forward MV is refined around fwd by
 * ME_COST_QPEL_BID (me.c:388-398): fpel_cmp(target, ref1 at the candidate)
 * >> 1 plus the MV rates of the candidate against mvp1 and of bwd against
 * mvp2.  Half-pel Spiral[0..8] (Spiral[1..8] with SAD), then quarter-pel
 * Spiral[1..8] with PMVR centred on mvp1.  Results in xme_sym_result: mv =
 * *fwd_mv on return, mvcost = p_me->mvcost[PDIR_BID].  Status 2 also marks a
 * backward MV inside the window whose plane is not built (the reference
 * predicts it with mc_luma, me.c:1329-1335; not provided).  A forward
 * candidate on a plane that is not built has no cost (the reference reads
 * through a NULL plane there). */
typedef struct {
    int16_t x, y;                    /* PU position (luma pixels, x % 4 == 0) */
    uint8_t w, h;                    /* PU size, multiples of 4 up to 64 */
    uint8_t reserved[2];
    int16_t fwd_x, fwd_y;            /* *fwd_mv on entry, qpel */
    int16_t bwd_x, bwd_y;            /* *bwd_mv, qpel (not searched) */
    int16_t mvp1_x, mvp1_y;          /* p_me->mvp1: forward rate and PMVR centre */
    int16_t mvp2_x, mvp2_y;          /* p_me->mvp2: backward rate */
} xme_bid_job;

int xme_me_search_bid_batch(const xme_me_params *prm, const xme_bid_job *d_jobs, int n, const pel_t *d_cur,
                            intptr_t i_cur, const xme_ref_desc *ref1, const xme_ref_desc *ref2,
                            xme_sym_result *d_results, void *stream);

/* The callers of the two refinements for a batch of PUs of one frame, frame
 * sequences, every reference row reconstructed (so check_mv_range's
 * num_lcu_coded_in_row test holds, source/common/predict.h:86-95), one wave
 * per PU: pred_inter_search_bi (source/encoder/md_inter.c:1036-1128, B
 * slices) and pred_inter_search_dual (md_inter.c:1133-1219, F slices).  Each
 * PU goes through the reference's gates: check_mvd (md_inter.c:68-76: +-4096,
 * vertically +-(4 << lcu_level) when frame_threads > 1), check_mv_range's box
 * (predict.h:68-78: the PU may reach 1 << lcu_level pixels past the picture,
 * inside min/max_mv_range) and check_mv_range_sym (md_inter.c:83-110: the MV
 * and its mirror scaled to the second reference, negated in B slices).  The
 * searches they let through are those of xme_me_search_sym_batch and
 * xme_me_search_bid_batch; then the stored MVs pass the reference's own
 * checks (md_inter.c:1107-1122, :1206-1217).  A gate that fails leaves its
 * search out: XME_MAX_DISTORTION and the MV as it came in.  lcu_level is
 * h->i_lcu_level (3..6); frame_threads is h->param->i_frame_threads.
 * Distances are calculate_distance's (nonzero, within +-255); strides of the
 * references must match; d_cur is 4-byte aligned with i_cur % 4 == 0.
 *
 * Status 2 marks a PU that is not evaluated: a PU shape the kernels do not
 * take (xme_sym_job's), a refinement that reported status 2, or a gate whose
 * scale_mv_skip product leaves int (undefined in the reference). */
typedef struct {
    int16_t x, y;                    /* PU position (luma pixels, x % 4 == 0) */
    uint8_t w, h;                    /* PU size, multiples of 4 up to 64 */
    uint8_t reserved[2];
    int16_t fwd_imv_x, fwd_imv_y;    /* p_me->all_best_imv[B_FWD] (full-pel; x 4 in int16 as mv_t) */
    int16_t bwd_x, bwd_y;            /* all_single_mv[B_BWD] (qpel) */
    int16_t mvp_x, mvp_y;            /* p_me->mvp on entry: the symmetric search's MV rate */
    int16_t mvp1_x, mvp1_y;          /* all_mvp[B_FWD] */
    int16_t mvp2_x, mvp2_y;          /* all_mvp[B_BWD] */
} xme_bi_job;

typedef struct {
    int32_t sym_cost;                /* *sym_mcost */
    int16_t sym_mv_x, sym_mv_y;      /* all_sym_mv[0] */
    int32_t sym_mvcost;              /* p_me->bmvcost[PDIR_SYM]; -1 when the search did not run or set it */
    int32_t bid_cost;                /* *bid_mcost */
    int16_t fwd_x, fwd_y;            /* all_dual_mv_1st[0] */
    int16_t bwd_x, bwd_y;            /* all_dual_mv_2nd[0] */
    int32_t bid_mvcost;              /* p_me->bmvcost[PDIR_BID]; -1 likewise */
    int32_t status;                  /* 0 ok, 2 not evaluated */
} xme_bi_result;

int xme_pred_inter_search_bi_batch(const xme_me_params *prm, int lcu_level, int frame_threads,
                                   const xme_bi_job *d_jobs, int n, int dist_fwd, int dist_bwd,
                                   const pel_t *d_cur, intptr_t i_cur, const xme_ref_desc *ref_fwd,
                                   const xme_ref_desc *ref_bwd, xme_bi_result *d_results, void *stream);

/* pred_inter_search_dual over references 0 .. n_ref - 1 (h->i_ref, 2..4):
 * reference r is searched with the second reference !r, distances dist[r]
 * and dist[!r] (h->fdec->ref_dpoc), and costs REF_COST(r) = (lambda *
 * REF_BITS[r + 1]) >> 16 on top (predict.h:47-52); the first strict minimum
 * wins.  refs holds n_ref descriptors. */
typedef struct {
    int16_t x, y;                    /* PU position (luma pixels, x % 4 == 0) */
    uint8_t w, h;                    /* PU size, multiples of 4 up to 64 */
    uint8_t reserved[2];
    int16_t mvp_x, mvp_y;            /* p_me->mvp on entry */
    int16_t mvp1[4][2];              /* all_mvp[ref_idx] */
    int16_t imv[4][2];               /* p_me->all_best_imv[ref_idx] (full-pel) */
} xme_dual_job;

typedef struct {
    int32_t cost;                    /* *dual_mcost (XME_MAX_DISTORTION: none passed) */
    int8_t fst_ref, snd_ref;         /* *dual_best_fst_ref / _snd_ref (-1: none) */
    int16_t reserved;
    int32_t mvcost;                  /* p_me->bmvcost[PDIR_DUAL] (-1: not set) */
    int16_t fst[4][2];               /* all_dual_mv_1st[ref_idx] */
    int16_t snd[4][2];               /* all_dual_mv_2nd[ref_idx] */
    int32_t status;                  /* 0 ok, 2 not evaluated */
} xme_dual_result;

int xme_pred_inter_search_dual_batch(const xme_me_params *prm, int lcu_level, int frame_threads,
                                     const xme_dual_job *d_jobs, int n, int n_ref, const int dist[4],
                                     const pel_t *d_cur, intptr_t i_cur, const xme_ref_desc *refs,
                                     xme_dual_result *d_results, void *stream);

/* MV predictors, get_mvp_default (source/encoder/md_inter.c:470-524), for a
 * batch of PUs of one slice, frame sequences.  Per PU: the searched
 * reference ref_idx and list bwd_2nd, the PU inside its CU (p_cb), and the
 * left, top, top-left and top-right neighbours' ref_idx[bwd_2nd]
 * (-1 = INVALID_REF) and mv[bwd_2nd]; the top-left stands in when the
 * top-right is unavailable.  ref_dpoc / ref_dpoc_multi are
 * h->fdec->ref_dpoc / ref_dpoc_multi (int[MAX_REFS], common.h:1031-1032,
 * set at encoder.c:1795-1810): XME_MVP_REFS entries each are read.  Status
 * 2: a searched or neighbour reference index outside 0..XME_MVP_REFS-1 (the
 * reference asserts blkref < MAX_REFS, predict.h:100); 3: a scaling product
 * that leaves int (undefined in the reference; the clipped value is
 * returned). */
typedef struct {
    int8_t ref_idx, bwd_2nd, ur_available, reserved;
    int8_t cb_x, cb_y;               /* p_cb->x, y: the PU inside its CU */
    uint8_t cb_w, cb_h;              /* p_cb->w, h */
    int8_t ref[4];                   /* left, top, top-left, top-right */
    int16_t mv[4][2];
} xme_mvp_job;

typedef struct {
    int16_t mvp_x, mvp_y;
    int32_t status;
} xme_mvp_result;

int xme_get_mvp_batch(const xme_mvp_job *d_jobs, int n, int b_slice, const int32_t ref_dpoc[XME_MVP_REFS],
                      const int32_t ref_dpoc_multi[XME_MVP_REFS], xme_mvp_result *d_out, void *stream);

/* The UMH search's encoder state of a batch of (PU, reference) jobs, replacing
 * get_mvp_default_sad (source/encoder/md_inter.c:527-608: the MVP of
 * get_mvp_default and pred_sad_space, the neighbour cost of the MVP's source)
 * and fast_me_prepare_info (md_inter.c:640-718: mvp1 from the upper
 * partition layer, mvp2 from the nearest reference, mvp3 from the co-located
 * block, pred_sad and the betas), which pred_inter_search_single calls per
 * reference before xavs2_me_search (md_inter.c:779-790).  A job holds what
 * those functions read from the encoder state. */
typedef struct {
    xme_mvp_job mvp;                 /* neighbours, PU in its CU, ref_idx, bwd_2nd */
    int32_t nb_sad[3];               /* all_mincost[..][mode][ref_idx] at the left, top and top-right 4x4 blocks
                                      * (0 where the reference reads none) */
    int32_t mode;                    /* PRED_2Nx2N (1) .. PRED_nRx2N (7), common.h:198-207 */
    int32_t mincost_up[3];           /* mincosts[PRED_2NxN / PRED_Nx2N / PRED_2Nx2N][ref_idx] of the CU */
    int16_t best_up[3][2];           /* all_best_mv[PRED_2NxN / Nx2N / 2Nx2N][pu_idx][ref_idx] */
    int32_t mincost_prev;            /* mincosts[mode][ref_idx - 1] */
    int16_t best_prev[2];            /* all_best_mv[mode][pu_idx][ref_idx - 1] */
    int16_t col_mv[2];               /* fref[0]->pu_mv at the PU's 16x16 block */
    int8_t col_valid;                /* fref[0] exists and is not an I picture */
    int8_t col_ref;                  /* fref[0]->pu_ref there (-1: none) */
    int16_t reserved;
    int32_t col_dpoc;                /* fref[0]->ref_dpoc[col_ref] */
} xme_umh_job;

/* Status: 0; 2 an index outside 0..XME_MVP_REFS-1 or a mode outside 1..7;
 * 3 a scaling product that leaves int (as xme_mvp_result); 4 a division by
 * a zero POC distance (undefined in the reference: the fields after it stay
 * 0).  ctx is what xme_plan_set_umh_context takes. */
typedef struct {
    int16_t mvp_x, mvp_y;            /* p_me->mvp */
    int32_t pred_sad_space;          /* get_mvp_default_sad's return */
    int32_t status;
    int32_t reserved;
    xme_umh_ctx ctx;
} xme_umh_state;

/* slice_type: SLICE_TYPE_I 0, P 1, B 2, F 3 (common.h:158-161); ref_dpoc and
 * ref_dpoc_multi as for xme_get_mvp_batch; umh_bsize = h->umh_bsize. */
int xme_umh_prepare_batch(const xme_umh_job *d_jobs, int n, int slice_type, const int32_t ref_dpoc[XME_MVP_REFS],
                          const int32_t ref_dpoc_multi[XME_MVP_REFS], const double umh_bsize[8],
                          xme_umh_state *d_out, void *stream);
/* h->umh_bsize for a picture's QP and slice type, xavs2_me_init_umh_threshold
 * (source/encoder/me.c:647-665). */
int xme_umh_bsize(int i_qp, int slice_type, double umh_bsize[8]);
/* A UMH plan's context from device-resident states, as
 * xme_plan_set_umh_context: the ctx of d_states[pu * n_ref + r] as
 * xme_umh_prepare_batch wrote them (a copy on the device, on `stream`).  No
 * reference counterpart, as xme_plan_set_umh_context. */
int xme_plan_set_umh_states(xme_plan *plan, const xme_umh_state *d_states, void *stream);

/* ---- 4. reference planes, MC and the residual path ---------------------- */
/* plane_expand_border over a whole plane (source/common/frame.c:470-507). */
int xme_expand_border(pel_t *d_plane, intptr_t stride, int w, int h, int pad, void *stream);
/* interpolate_lcu_row for every LCU row (source/common/mc.c:593-785): fills
 * d_planes[1..15] (host array of device origin pointers, same stride and
 * XME_PAD border as d_planes[0], which must already be border-expanded). */
int xme_interpolate_frame(pel_t *const d_planes[16], intptr_t stride, int w, int h,
                          int use_fractional_me, void *stream);
/* The same, one LCU row at a time, for an encoder that keeps its reference
 * planes resident in device memory and finishes a picture row by row
 * (source/encoder/slice.c:535-543): xavs2_frame_expand_border_lcurow for one
 * plane (source/common/frame.c:537-575: rows [i << row_log2, (i + 1) <<
 * row_log2), 4 rows up except at the top and bottom of the picture, with the
 * left / right pads, and the band above the first row / below the last;
 * row_log2 = lcu_level for luma, lcu_level - 1 for 4:2:0 chroma), and
 * interpolate_lcu_row (source/common/mc.c:748-785: planes 1..15 on rows
 * [(i << lcu_level) - 8, ((i + 1) << lcu_level) - 8), the first row from -4,
 * the last to h + 4, with their pads).  Called in row order after row i of
 * the picture is final, each leaves every plane equal to xme_expand_border /
 * xme_interpolate_frame on all rows above its last one; interpolate_lcu_row
 * reads the full-pel plane only up to 4 rows below its last one, which the
 * expansion of the same LCU row has completed.  One slice (slice_num 1). */
int xme_expand_border_lcu_row(pel_t *d_plane, intptr_t stride, int w, int h, int pad, int row_log2, int lcu_y,
                              void *stream);
int xme_interpolate_lcu_row(pel_t *const d_planes[16], intptr_t stride, int w, int h, int use_fractional_me,
                            int lcu_level, int lcu_y, void *stream);
/* mc_chroma (source/common/mc.c:797-829) for a batch of blocks given in luma
 * units: block i (luma_w x luma_h at luma_x, luma_y with its qpel MV) has the
 * chroma size luma_w / 2 x luma_h / 2 at the qpel position 4 * luma position
 * + MV, clamped by get_mv_for_mc (mc.h:69-75) when frame_w_luma > 0, read as
 * a chroma integer position (>> 3) and eighth-pel phase (& 7) as
 * rdo_get_pred_inter does (rdo.c:1641-1655).  Outputs are written at
 * d_dst_u/v + dst_off[i] (a dword at a time where the chroma width is a
 * multiple of 4 and the destinations and i_dst are 4-byte aligned).  Source
 * reads stay inside the planes' 40-pixel chroma border, except that a row's
 * last reads may pass its padded end by one byte, into the next row's
 * border. */
int xme_mc_chroma_batch(const pel_t *d_u, const pel_t *d_v, intptr_t i_src,
                        const int32_t *d_blocks /* n x {luma_x, luma_y, luma_w, luma_h, mv_x, mv_y} */,
                        int n, int frame_w_luma, int frame_h_luma,
                        pel_t *d_dst_u, pel_t *d_dst_v, intptr_t i_dst, const int64_t *d_dst_off, void *stream);
/* mc_luma (source/common/mc.c:559-589) from the interpolated planes, same block format. */
int xme_mc_luma_batch(const xme_ref_desc *ref, const int32_t *d_blocks, int n, int frame_w, int frame_h,
                      pel_t *d_dst, intptr_t i_dst, const int64_t *d_dst_off, void *stream);
/* The bi-predicted forms (rdo_get_pred_inter, source/encoder/rdo.c:1616-1666):
 * the rounded average (a + b + 1) >> 1 of the predictions from the first
 * reference at the first MV and the second at the second (pixf.avg and
 * pixf.average, source/common/pixel.c:844-856), blocks as
 * n x {x, y, w, h, mv1_x, mv1_y, mv2_x, mv2_y} in luma units, for the
 * prediction pictures of bi-predicted PUs (xme_recon_pred_batch,
 * xme_recon_chroma_batch).  Both references share one stride. */
int xme_mc_luma_bi_batch(const xme_ref_desc *ref1, const xme_ref_desc *ref2, const int32_t *d_blocks, int n,
                         int frame_w, int frame_h, pel_t *d_dst, intptr_t i_dst, const int64_t *d_dst_off,
                         void *stream);
int xme_mc_chroma_bi_batch(const pel_t *d_u1, const pel_t *d_v1, const pel_t *d_u2, const pel_t *d_v2,
                           intptr_t i_src, const int32_t *d_blocks, int n, int frame_w_luma, int frame_h_luma,
                           pel_t *d_dst_u, pel_t *d_dst_v, intptr_t i_dst, const int64_t *d_dst_off, void *stream);

/* dct_NxN / idct_NxN (source/common/transform.c:1123-1312) on nblk contiguous
 * n x n blocks (n in 4, 8, 16, 32). */
int xme_dct_batch(int n, const coeff_t *d_src, coeff_t *d_dst, int nblk, void *stream);
int xme_idct_batch(int n, const coeff_t *d_src, coeff_t *d_dst, int nblk, void *stream);
/* quant_c / dequant_c (source/common/quant.c:109-184) with the parameters of
 * tu_quant_forward / tu_quant_inverse (source/encoder/rdo.c:658-718). */
int xme_quant_batch(int n, coeff_t *d_coef, int nblk, int qp, int b_intra, int32_t *d_nnz, void *stream);
int xme_dequant_batch(int n, coeff_t *d_coef, int nblk, int qp, void *stream);
/* DCT -> quant -> dequant -> IDCT fused: level = quantised coefficients,
 * recon = reconstructed residual, nnz per block (d_nnz may be NULL). */
int xme_residual_batch(int n, int qp, int b_intra, const coeff_t *d_in, coeff_t *d_level, coeff_t *d_recon,
                       int32_t *d_nnz, int nblk, void *stream);
/* Luma inter reconstruction of ntu N x N transform blocks (n in 4, 8, 16, 32)
 * at one QP: the TU's prediction by mc_luma from ref at its PU's clamped
 * position (rdo_get_pred_inter, source/encoder/rdo.c:1616-1633), the
 * prediction error (sub_ps), DCT -> quant -> dequant -> IDCT
 * (xme_residual_batch, inter), the reconstruction clip(pred + residual)
 * (add_ps) written into d_rec and its SSD against d_cur: the prediction
 * error of rdo.c:1751-1753, then tu_recon_inter_luma
 * (source/encoder/rdo.c:1333-1366, inter quant with b_intra = 0) without
 * RDOQ, and the SSD of cu_recon_inter_luma (rdo.c:1497-1499) per TU.
 * d_tus: ntu x {tu_x, tu_y, pu_x, pu_y, pu_w, pu_h, mv_x, mv_y} (MV in qpel,
 * the TU inside the PU).  d_level holds the quantised levels (ntu * n * n),
 * d_nnz / d_ssd one value per TU.  d_resid (prediction error) and d_rres
 * (reconstructed residual), ntu * n * n coefficients each, are written when
 * given; with every interpolated plane of ref resident the path is one
 * launch and they may be NULL, otherwise (block-filter prediction) they are
 * required as the stages' intermediates.  A TU outside the picture, not
 * inside its PU, at a column that is not a multiple of 4, or whose PU is not
 * 4..64 on each side gets ssd -1, zero levels and no pixels.  d_cur / d_rec:
 * 4-byte aligned with strides % 4 == 0; d_resid / d_rres: 8-byte aligned.
 * n = 64: the TU of an unsplit 64x64 CU (its PU 64x64), transformed as
 * dct_64x64 / idct_64x64 (source/common/transform.c:1514-1541: the lifting
 * wavelet's 32x32 low band through the 32x32 transform with the wavelet
 * flag) and quantised at level 6 (tu_recon_inter_luma's i_level); d_level
 * then holds 32 * 32 levels per TU, and d_resid / d_rres are not written.
 * n = 64 needs every interpolated plane of the reference(s).
 * n = 32 or 64 | XME_TU_DCT_HALF transforms with dct_half instead, as
 * tu_get_dct_coeff (rdo.c:764-771) does for TUs of 32 and 64 in the first RD
 * pass under OPT_BIT_EST_PSZT (the default preset's, presets.c:277). */
#define XME_TU_DCT_HALF 0x100
int xme_inter_recon_batch(int n, int qp, const pel_t *d_cur, intptr_t i_cur, const xme_ref_desc *ref,
                          int frame_w, int frame_h, const int32_t *d_tus, int ntu, coeff_t *d_resid,
                          coeff_t *d_level, coeff_t *d_rres, pel_t *d_rec, intptr_t i_rec, int32_t *d_nnz,
                          int32_t *d_ssd, void *stream);
/* The same for TUs of bi-predicted PUs: the prediction is the rounded
 * average (pixf.avg) of mc_luma from ref1 at the record's MV and from ref2 at
 * d_mv2[2 i], d_mv2[2 i + 1], each clamped with the PU's size
 * (rdo_get_pred_inter, rdo.c:1616-1631).  ref1 and ref2 share one stride. */
int xme_inter_recon_bi_batch(int n, int qp, const pel_t *d_cur, intptr_t i_cur, const xme_ref_desc *ref1,
                             const xme_ref_desc *ref2, int frame_w, int frame_h, const int32_t *d_tus,
                             const int32_t *d_mv2, int ntu, coeff_t *d_resid, coeff_t *d_level, coeff_t *d_rres,
                             pel_t *d_rec, intptr_t i_rec, int32_t *d_nnz, int32_t *d_ssd, void *stream);
/* Reconstruction of nblk N x N blocks (n in 4, 8, 16, 32, 64, or 32 / 64 |
 * XME_TU_DCT_HALF) whose prediction is a picture d_pred (stride i_pred)
 * at the block's own position: the chroma of an inter CU, one call per
 * plane (cu_recon_chroma, source/encoder/rdo.c:785-846: sub_ps, the
 * transform of tu_get_dct_coeff, inter quant, dequant, IDCT, add_ps, and
 * pixf.ssd against d_cur), with the prediction of xme_mc_chroma_batch
 * written into d_pred; and the luma TUs of a CU whose prediction
 * (buf_pred_inter of cu_recon_inter_luma, rdo.c:1378-1493) is written by
 * xme_mc_luma_batch, such as an unsplit TU over two PUs, with n = 64 as in
 * xme_inter_recon_batch.  qp is the chroma QP (xme_chroma_qp) or the luma
 * one.  d_blocks: nblk x {x, y} in the plane_w x plane_h plane.  d_level
 * (nblk * n * n, 32 * 32 for n = 64; 16-byte aligned), d_nnz and d_ssd per
 * block; the reconstruction goes into d_rec at the block.  A block outside
 * the plane or at a column that is not a multiple of 4 gets ssd -1, zero
 * levels and no pixels.  d_cur / d_pred / d_rec: 4-byte aligned with strides
 * % 4 == 0. */
int xme_recon_pred_batch(int n, int qp, const pel_t *d_cur, intptr_t i_cur, const pel_t *d_pred, intptr_t i_pred,
                         int plane_w, int plane_h, const int32_t *d_blocks, int nblk, coeff_t *d_level, pel_t *d_rec,
                         intptr_t i_rec, int32_t *d_nnz, int32_t *d_ssd, void *stream);
/* cu_recon_chroma (source/encoder/rdo.c:785-846) for nblk CUs at once: both
 * planes' blocks in one launch, as xme_recon_pred_batch on U (d_cur_u,
 * d_pred_u, d_rec_u) and then on V with the same blocks, QP and strides.
 * d_level holds 2 * nblk blocks (U's, then V's), d_nnz and d_ssd 2 * nblk
 * values.  One chroma QP serves both planes: the reference adds a per-plane
 * delta only under ENABLE_WQUANT (rdo.c:818-820), which its build disables
 * (common/defines.h:134). */
int xme_recon_chroma_batch(int n, int qp, const pel_t *d_cur_u, const pel_t *d_cur_v, intptr_t i_cur,
                           const pel_t *d_pred_u, const pel_t *d_pred_v, intptr_t i_pred, int plane_w, int plane_h,
                           const int32_t *d_blocks, int nblk, coeff_t *d_level, pel_t *d_rec_u, pel_t *d_rec_v,
                           intptr_t i_rec, int32_t *d_nnz, int32_t *d_ssd, void *stream);
/* The chroma QP of a luma QP: tab_qp_scale_chroma at clip(0, 63, luma_qp),
 * cu_get_chroma_qp (source/common/cudata.h:118-125, table cudata.c:98-106).
 * Host only. */
int xme_chroma_qp(int luma_qp);

/* ---- 5. serving: continuous-batching ME for many concurrent encodes ------ *
 * The frame-level ME of many encodes (streams) share one GPU.  Each stream
 * keeps its last n_ref reconstructed references, with their 15 interpolated
 * planes, resident in device memory.  A submitted frame is queued; the
 * scheduler admits queued frames in arrival order whenever one of the
 * max_inflight device slots is free and the frame's working memory fits the
 * memory budget beside everything resident, so frames that arrive while
 * others run start as soon as a slot frees and their kernels overlap on the
 * device.  A frame is searched against the min(frame index, n_ref) previous
 * frames of its stream and then becomes the stream's newest reference. */
typedef struct xme_server xme_server;
typedef struct {
    int64_t memory_budget;           /* bytes of device memory the server may hold */
    int32_t max_inflight;            /* frame jobs on the device at once (1..16) */
    int32_t search_range;            /* full search over [-R, R) */
    int32_t lambda_factor;           /* h->i_lambda_factor */
    int32_t n_ref;                   /* references per job (1..XME_MAX_REFS) */
    int32_t use_hadamard;            /* as in xme_me_params */
    int32_t use_fractional_me;
    int32_t use_fast_sub_me;
    int32_t enable_pmvr;
    int32_t workgroup_share_pct;     /* % of a full-device search launch per frame (1..100; 0 = 100) */
    int64_t launch_split_absdiff;    /* search work per launch (0: 8e9; < 0: no split) */
    int32_t me_method;               /* XME_ME_FS, _DIA, _HEX or _TZ (UMH needs per-job state) */
    int32_t reserved;
    int64_t full_share_below_absdiff; /* streams whose search work per reference (absdiffs) is below
                                         this run on the whole device; workgroup_share_pct caps the
                                         others (0: it caps every stream) */
} xme_server_config;
typedef struct {
    int64_t submitted, completed, rejected;
    double p50_ms, p99_ms, mean_ms, max_ms;   /* submit -> results on the host */
    double throughput_fps;                    /* completed / (last completion - first submit) */
    int64_t resident_bytes;                   /* references, planes and plans of open streams */
    int64_t inflight_bytes;                   /* working memory of admitted frames */
    int64_t peak_bytes;                       /* largest resident + inflight seen */
} xme_server_stats;

int xme_server_create(const xme_server_config *cfg, xme_server **srv);
/* Waits for every admitted frame, then releases all device memory. */
int xme_server_destroy(xme_server *srv);
/* Register an encode of the given luma size whose frames search the n_pu PUs
 * (positions, sizes, MV predictors) of `pus`.  Fails with XME_ERR_BUDGET when
 * its resident references and plans do not fit the memory budget. */
int xme_server_open_stream(xme_server *srv, int width, int height, const xme_pu *pus, int n_pu, int *stream_id);
/* Waits for the stream's frames, then frees its references and plans.  Jobs
 * of the stream that were not waited for stay readable until waited. */
int xme_server_close_stream(xme_server *srv, int stream_id);
/* Submit one frame: the host luma picture (width x height at `stride`) is
 * copied, so the caller may reuse it on return.  Returns a job id, or
 * XME_ERR_BUDGET (counted as rejected) when the frame could not fit the
 * budget even with nothing else in flight. */
int xme_server_submit(xme_server *srv, int stream_id, const pel_t *luma, intptr_t stride, int64_t *job_id);
/* Result count of a job: n_pu * refs searched (0 for a stream's first frame). */
int xme_server_job_size(xme_server *srv, int64_t job_id, int *n_results);
/* Waits for a job, copies its results (entry pu * refs + r, reference r = 0
 * the previous frame) and forgets the job.  latency_ms: submit -> done. */
int xme_server_wait(xme_server *srv, int64_t job_id, xme_result *results, double *latency_ms);
int xme_server_stats_get(xme_server *srv, xme_server_stats *stats);
/* Clears the latency record and counters (e.g. after a warm-up). */
int xme_server_stats_reset(xme_server *srv);

#ifdef __cplusplus
}
#endif
#endif
