// qsv_fusion.h -- gate-block scheduling for the fused path.
#pragma once

#include "qsv_internal.h"

namespace qsv {

int validate_gates(const qsv_state *s, const qsv_gate *g, int64_t ng);
int upload_mats(qsv_state *s, const double *mats, size_t count_complex);
int record_class(const qsv_state *s, const qsv_gate &r, double *bytes, double *flops);

// Per-launch event timing on the state's stream: begin() opens the first
// interval, each end() closes one launch and opens the next.
class Profiler {
public:
    Profiler(qsv_state *s, bool on);
    ~Profiler();
    int begin();
    // tag >= 0 records this launch's time under that index (collect -> tag_ms)
    int end(int cls, double bytes, double flops, int64_t tag = -1);
    int collect(qsv_run_stats *st);
    bool on() const { return on_; }
    const std::vector<float> &tag_ms() const { return tag_ms_; }

private:
    qsv_state *s_;
    bool on_;
    std::vector<hipEvent_t> events_;
    std::vector<int> cls_;
    std::vector<double> bytes_, flops_;
    std::vector<int64_t> tags_;
    std::vector<float> tag_ms_;
};

// Fuse `gates` into blocks of at most opts.fuse_k qubits (the learned
// scheduler picks the block size itself) and apply them; *k_used = the block
// size of the plan.
int run_fused(qsv_state *s, const qsv_gate *gates, int64_t ngates, const double *mats,
              const qsv_run_opts &opts, int32_t *rng, int *outcomes, int64_t *blocks,
              double *pred_ms, Profiler *prof, hipEvent_t start, double *plan_ms, int *k_used, int *planner);

}  // namespace qsv
