// Learned start factors (sketch -> MLP -> QR).  Placeholder until the MLP
// kernels are in: reports DT_EUNSUPPORTED so callers fall back explicitly.
#include "dt_internal.hpp"

int dt_learned_init(dt_ctx *ctx, const void *model, const void *A, int dtA, int N, const int64_t *I,
                    const int64_t *R, uint64_t seed, double *const *U64) {
    (void)model; (void)A; (void)dtA; (void)N; (void)I; (void)R; (void)seed; (void)U64;
    return dt_fail(ctx, DT_EUNSUPPORTED, "learned init: no model support built");
}
