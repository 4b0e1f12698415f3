// qsv_fusion.cpp -- placeholder until the block scheduler lands.
#include "qsv_fusion.h"

namespace qsv {

int run_fused(qsv_state *, const qsv_gate *, int64_t, const double *, const qsv_run_opts &, int32_t *,
              int *, int64_t *, double *)
{
    return fail(QSV_ERR_UNSUPPORTED, "fused path not built");
}

}  // namespace qsv
