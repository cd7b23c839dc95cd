// Batched persistent L-BFGS (one workgroup per line).
#include "qew_internal.hpp"

using namespace qew_host;

extern "C" qew_status qew_batch_lbfgs(qew_batch b, int dtype, const qew_lbfgs_params *lp,
                                      double *h, qew_lbfgs_result *results) {
    return guard([&] {
        (void)b; (void)dtype; (void)lp; (void)h; (void)results;
        throw Error(QEW_ERR_UNSUPPORTED, "qew_batch_lbfgs is not available in this build");
    });
}
