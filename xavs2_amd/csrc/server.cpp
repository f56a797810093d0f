#include "xme_internal.h"
extern "C" int xme_server_create(const xme_server_config *, xme_server **) { return XME_ERR_STATE; }
extern "C" int xme_server_destroy(xme_server *) { return XME_ERR_STATE; }
extern "C" int xme_server_open_stream(xme_server *, int, int, int *) { return XME_ERR_STATE; }
extern "C" int xme_server_close_stream(xme_server *, int) { return XME_ERR_STATE; }
extern "C" int xme_server_submit(xme_server *, int, const pel_t *, intptr_t, int64_t *) { return XME_ERR_STATE; }
extern "C" int xme_server_job_size(xme_server *, int64_t, int *) { return XME_ERR_STATE; }
extern "C" int xme_server_wait(xme_server *, int64_t, xme_result *, double *) { return XME_ERR_STATE; }
extern "C" int xme_server_stats_get(xme_server *, xme_server_stats *) { return XME_ERR_STATE; }
