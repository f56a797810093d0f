#include "xme_internal.h"
extern "C" int xme_dct_batch(int, const coeff_t *, coeff_t *, int, void *) { xme_set_error("pending"); return XME_ERR_STATE; }
extern "C" int xme_idct_batch(int, const coeff_t *, coeff_t *, int, void *) { xme_set_error("pending"); return XME_ERR_STATE; }
extern "C" int xme_quant_batch(int, coeff_t *, int, int, int, int32_t *, void *) { xme_set_error("pending"); return XME_ERR_STATE; }
extern "C" int xme_dequant_batch(int, coeff_t *, int, int, void *) { xme_set_error("pending"); return XME_ERR_STATE; }
extern "C" int xme_residual_batch(int, int, int, const coeff_t *, coeff_t *, coeff_t *, int32_t *, int, void *) { xme_set_error("pending"); return XME_ERR_STATE; }
extern "C" int xme_mc_chroma_batch(const pel_t *, const pel_t *, intptr_t, const int32_t *, int, int, int, pel_t *, pel_t *, intptr_t, const int64_t *, void *) { xme_set_error("pending"); return XME_ERR_STATE; }
extern "C" int xme_mc_luma_batch(const xme_ref_desc *, const int32_t *, int, int, int, pel_t *, intptr_t, const int64_t *, void *) { xme_set_error("pending"); return XME_ERR_STATE; }
