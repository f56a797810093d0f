#include "xme_internal.h"
extern "C" int xme_get_sad_table(xme_pixel_cmp_t *) { return XME_ERR_STATE; }
extern "C" int xme_get_satd_table(xme_pixel_cmp_t *) { return XME_ERR_STATE; }
extern "C" int xme_get_sad_x3_table(xme_pixel_cmp_x3_t *) { return XME_ERR_STATE; }
extern "C" int xme_get_sad_x4_table(xme_pixel_cmp_x4_t *) { return XME_ERR_STATE; }
