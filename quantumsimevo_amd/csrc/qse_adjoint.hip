// qse_adjoint.hip — placeholder entry points (filled in with the adjoint engine).
#include "qse_internal.hpp"

extern "C" {
int qse_adjoint_expect(qse_ctx*, int, int, const qse_op*, int, const double*, int, int, const uint64_t*,
                       const double*, int, int, double*, double*) {
    return qse::set_error(QSE_ERR_UNSUPPORTED, "adjoint not built yet");
}
int qse_batch_expect(qse_ctx*, int, int, const qse_op*, int, const double*, int, int, const uint64_t*, const double*,
                     int, double*) {
    return qse::set_error(QSE_ERR_UNSUPPORTED, "batch expect not built yet");
}
int qse_batch_run_small(qse_ctx*, int, int, const int*, const int64_t*, const qse_op*, const int64_t*, const double*,
                        const uint64_t*, const double*, int, double*, double*) {
    return qse::set_error(QSE_ERR_UNSUPPORTED, "batch run not built yet");
}
}
