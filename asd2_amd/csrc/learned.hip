// learned.hip -- placeholder for the learned-preconditioner flexible PCG.
#include "solver.h"

namespace asd2 {

int learned_solve(asd2_ctx* ctx, const asd2_solve_opts* opts, double* d_ps, asd2_solve_stats* stats) {
    (void)ctx; (void)opts; (void)d_ps; (void)stats;
    set_error("learned preconditioner not available in this build");
    return ASD2_ERR_UNSUPPORTED;
}

}  // namespace asd2
