// strackw on the GPU (placeholder until the NCC tracker lands).
#include "st_ctx.h"

void st_corr_track_fast_impl(st_ctx *ctx, StPair &pr, const st_track_params &par, st_track_out *out,
                             st_track_diag *diag)
{
    (void)ctx;
    (void)pr;
    (void)par;
    (void)out;
    (void)diag;
    throw StError{ST_ERR_UNSUPPORTED, "strackw is not implemented yet"};
}
