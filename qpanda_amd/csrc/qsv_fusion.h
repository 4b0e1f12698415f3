// qsv_fusion.h -- gate-block scheduling for the fused path.
#pragma once

#include "qsv_internal.h"

namespace qsv {

int validate_gates(const qsv_state *s, const qsv_gate *g, int64_t ng);
int upload_mats(qsv_state *s, const double *mats, size_t count_complex);

// Fuse `gates` into blocks of at most opts.fuse_k qubits and apply them.
int run_fused(qsv_state *s, const qsv_gate *gates, int64_t ngates, const double *mats,
              const qsv_run_opts &opts, int32_t *rng, int *outcomes, int64_t *blocks,
              double *pred_ms);

}  // namespace qsv
