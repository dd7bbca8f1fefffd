// cullst stages on the GPU (placeholder).
#include "st_ctx.h"

extern "C" int st_cull_st_data(st_ctx *c, st_cull_params *cp) { (void)c; (void)cp; return ST_ERR_UNSUPPORTED; }
extern "C" int st_cull_stats(st_ctx *c, st_cull_params *cp) { (void)c; (void)cp; return ST_ERR_UNSUPPORTED; }
extern "C" int st_cull_smooth(st_ctx *c, st_cull_params *cp) { (void)c; (void)cp; return ST_ERR_UNSUPPORTED; }
extern "C" int st_cull_islands(st_ctx *c, st_cull_params *cp) { (void)c; (void)cp; return ST_ERR_UNSUPPORTED; }
extern "C" int st_cull_run(st_ctx *c, st_cull_params *cp) { (void)c; (void)cp; return ST_ERR_UNSUPPORTED; }
