// qsv_costmodel.hip -- learned scheduler (placeholder: greedy at the requested k).
#include "qsv_sched.h"

namespace qsv {

int plan_learned(qsv_state *, const qsv_gate *gates, int64_t ng, const double *mats, const qsv_run_opts &o,
                 std::vector<PlanItem> &plan, SchedChoice *choice)
{
    choice->fuse_k = o.fuse_k;
    return plan_greedy(gates, ng, mats, o.fuse_k, o.exact, plan);
}

}  // namespace qsv
