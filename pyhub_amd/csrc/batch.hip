// batch.hip -- placeholder translation unit for the batched cluster solver.
#include "phub_internal.h"

extern "C" int phub_batch_hubbard(phub_ctx *c, const phub_cluster *, int, const phub_batch_opts *,
                                  double *, double *, double *, double *) {
    return phub_fail(c, PHUB_ERR_UNSUPPORTED, "batched solver not built yet");
}
