// Learned lambda* warm start for the ECP contact solve (placeholder until the
// network kernels land: no network is loaded, the learned solver is refused).
#include "sx_internal.h"

const float *sx_ecp_mlp_predict(sx_ctx *c, const double *pp, long ld, const int *work, const int *nwork, long cap)
{
    (void)c; (void)pp; (void)ld; (void)work; (void)nwork; (void)cap;
    return nullptr;
}

void sx_ecp_mlp_free(sx_ctx *c) { c->mlp = nullptr; }
