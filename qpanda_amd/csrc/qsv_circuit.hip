// qsv_circuit.hip -- whole-circuit execution.
//
// The per-gate path issues one sweep per record, stream-ordered with no host
// synchronisation between gates (QPanda's CUDA back end synchronises the
// device after every gate, GPUGates.cu:1126-1127).  Random-number bookkeeping
// follows the reference executor: one Park-Miller draw per gate call
// (CPUQuantumGates.cpp:97 and every other gate), one per measurement
// (:1086), and Reset = measure plus a NOT (one more draw) on outcome 1
// (:126-136).  Draws whose values are not needed are skipped with the
// generator's jump-ahead.
//
// The fused path (opts->fuse_k >= 2) hands the gate list to the block
// scheduler (qsv_fusion.cpp) and applies its blocks.
#include "qsv_internal.h"
#include "qsv_fusion.h"

#include <cmath>
#include <cstring>

namespace qsv {

int upload_mats(qsv_state *s, const double *mats, size_t count_complex)
{
    const size_t bytes = count_complex * 16;
    if (bytes == 0) return QSV_OK;
    if (s->mats_bytes < bytes) {
        if (s->mats) {
            QSV_HIP(hipStreamSynchronize(s->stream));
            QSV_HIP(hipFree(s->mats));
            s->mats = nullptr;
            s->mats_bytes = 0;
        }
        if (hipMalloc((void **)&s->mats, bytes) != hipSuccess) {
            s->mats = nullptr;
            return fail(QSV_ERR_ALLOC, "cannot allocate matrix pool");
        }
        s->mats_bytes = bytes;
    }
    QSV_HIP(hipMemcpyAsync(s->mats, mats, bytes, hipMemcpyHostToDevice, s->stream));
    QSV_HIP(hipStreamSynchronize(s->stream));
    return QSV_OK;
}

static int max_mat_index(const qsv_gate *g, int64_t ng, size_t *need)
{
    size_t m = 0;
    for (int64_t i = 0; i < ng; i++) {
        if (g[i].kind == QSV_GATE_U1 || g[i].kind == QSV_GATE_U2) {
            if (g[i].mat < 0) return fail(QSV_ERR_ARG, "dense gate without matrix");
            const size_t len = g[i].kind == QSV_GATE_U1 ? 4 : 16;
            m = std::max(m, (size_t)g[i].mat + len);
        }
    }
    *need = m;
    return QSV_OK;
}

// Validate every record before the first launch.
int validate_gates(const qsv_state *s, const qsv_gate *g, int64_t ng)
{
    for (int64_t i = 0; i < ng; i++) {
        const qsv_gate &r = g[i];
        int nq = 0;
        switch (r.kind) {
        case QSV_GATE_H: case QSV_GATE_X: case QSV_GATE_RX: case QSV_GATE_RY: case QSV_GATE_RZ:
        case QSV_GATE_U1: case QSV_GATE_RESET: nq = 1; break;
        case QSV_GATE_MEASURE:
            nq = 1;
            if (r.q1 < 0 || r.q1 >= 64) return fail(QSV_ERR_ARG, "classical register out of range");
            break;
        case QSV_GATE_CNOT: case QSV_GATE_CZ: case QSV_GATE_CR: case QSV_GATE_ISWAP:
        case QSV_GATE_SQISWAP: case QSV_GATE_U2: nq = 2; break;
        case QSV_GATE_TOFFOLI: case QSV_GATE_CSWAP: nq = 3; break;
        default: return fail(QSV_ERR_UNSUPPORTED, "unknown gate kind");
        }
        const int qs[3] = {r.q0, r.q1, r.q2};
        for (int j = 0; j < nq; j++)
            if (qs[j] < 0 || qs[j] >= s->n) return fail(QSV_ERR_QUBIT, "gate qubit out of range");
        for (int a = 0; a < nq; a++)
            for (int b = a + 1; b < nq; b++)
                if (qs[a] == qs[b] && r.kind != QSV_GATE_CR && r.kind != QSV_GATE_ISWAP &&
                    r.kind != QSV_GATE_SQISWAP && r.kind != QSV_GATE_CNOT && r.kind != QSV_GATE_CZ)
                    return fail(QSV_ERR_QUBIT, "repeated qubit in gate");
    }
    return QSV_OK;
}

// One record on the per-gate path.  `draw` is the value of this record's
// first random number (used by MEASURE / RESET).
static int apply_record(qsv_state *s, const qsv_gate &r, const double *mats_host, int exact,
                        int32_t *rng, int *outcomes)
{
    auto next = [&]() -> double {
        if (!rng) return 1.0;
        *rng = qsv_pm_next(*rng);
        return (double)*rng / 2147483647;
    };
    const uint64_t b0 = uint64_t(1) << r.q0, b1 = uint64_t(1) << r.q1;
    switch (r.kind) {
    case QSV_GATE_H: next(); return qsv_h(s, r.q0, 0, exact);
    case QSV_GATE_X: next(); return qsv_x(s, r.q0, 0);
    case QSV_GATE_RX: next(); return qsv_rx(s, r.q0, r.theta, r.dagger, 0);
    case QSV_GATE_RY: next(); return qsv_ry(s, r.q0, r.theta, r.dagger, 0);
    case QSV_GATE_RZ: next(); return qsv_rz(s, r.q0, r.theta, r.dagger, 0);
    case QSV_GATE_CNOT:
        next();
        if (r.q0 == r.q1) return QSV_OK;        // no-op in the reference loop
        return qsv_x(s, r.q1, b0);
    case QSV_GATE_CZ: next(); return qsv_cphase(s, r.q0, r.q1, -1.0, 0.0, 0);
    case QSV_GATE_CR: next(); return qsv_cr(s, r.q0, r.q1, r.theta, r.dagger);
    case QSV_GATE_ISWAP: next(); return qsv_iswap(s, r.q0, r.q1, r.dagger);
    case QSV_GATE_SQISWAP: next(); return qsv_sqiswap(s, r.q0, r.q1, r.dagger, exact);
    case QSV_GATE_TOFFOLI: next(); return qsv_x(s, r.q2, b0 | b1);
    case QSV_GATE_CSWAP: next(); return qsv_cswap(s, r.q0, r.q1, r.q2);
    case QSV_GATE_U1: {
        next();
        const double *u = mats_host + 2 * (size_t)r.mat;
        return qsv_apply_1q(s, r.q0, u, 0);
    }
    case QSV_GATE_U2: {
        next();
        const int qs[2] = {r.q1, r.q0};
        return launch_block_dev(s, qs, 2, s->mats + 2 * (size_t)r.mat, 0);
    }
    case QSV_GATE_MEASURE: {
        int o = 0;
        int rc = qsv_measure(s, r.q0, next(), &o);
        if (rc) return rc;
        if (outcomes) outcomes[r.q1] = o;
        return QSV_OK;
    }
    case QSV_GATE_RESET: {
        int o = 0;
        int rc = qsv_measure(s, r.q0, next(), &o);
        if (rc) return rc;
        if (o) { next(); return qsv_x(s, r.q0, 0); }
        return QSV_OK;
    }
    default: return fail(QSV_ERR_UNSUPPORTED, "unknown gate kind");
    }
}

// Kernel class and algorithmic cost of one record on the per-gate path:
// bytes = 2 * S * (amplitudes the gate may change), flops = arithmetic on them.
int record_class(const qsv_state *s, const qsv_gate &r, double *bytes, double *flops)
{
    const double S = (double)amp_bytes(s), N = (double)s->dim;
    double touched = N, fl = 0;
    int cls = QSV_K_DENSE1;              // every kind is listed below (kinds are validated first)
    switch (r.kind) {
    case QSV_GATE_H: case QSV_GATE_RX: case QSV_GATE_RY: case QSV_GATE_U1:
        cls = QSV_K_DENSE1; fl = 14 * N; break;
    case QSV_GATE_RZ: cls = QSV_K_DIAG; fl = 6 * N; break;
    case QSV_GATE_X: cls = QSV_K_PERM; break;
    case QSV_GATE_CNOT: cls = QSV_K_PERM; touched = r.q0 == r.q1 ? 0 : N / 2; break;
    case QSV_GATE_TOFFOLI: case QSV_GATE_CSWAP: cls = QSV_K_PERM; touched = N / 4; break;
    case QSV_GATE_ISWAP: cls = QSV_K_PERM; touched = N / 2; fl = 0; break;
    case QSV_GATE_SQISWAP: cls = QSV_K_DENSE1; touched = N / 2; fl = 14 * N / 2; break;
    case QSV_GATE_CZ: case QSV_GATE_CR: cls = QSV_K_PHASE; touched = N / 4; fl = 6 * N / 4; break;
    case QSV_GATE_U2: cls = QSV_K_BLOCK64; fl = 32 * N; break;
    case QSV_GATE_MEASURE: case QSV_GATE_RESET: cls = QSV_K_READOUT; touched = N; break;
    default: break;
    }
    *bytes = 2 * S * touched;
    *flops = fl;
    return cls;
}

Profiler::Profiler(qsv_state *s, bool on) : s_(s), on_(on) {}

Profiler::~Profiler()
{
    for (hipEvent_t e : events_) (void)hipEventDestroy(e);
}

int Profiler::begin()
{
    if (!on_) return QSV_OK;
    if (events_.size() == cls_.size()) {       // need an opening event
        hipEvent_t e;
        QSV_HIP(hipEventCreate(&e));
        QSV_HIP(hipEventRecord(e, s_->stream));
        events_.push_back(e);
        cls_.push_back(-1);
        bytes_.push_back(0);
        flops_.push_back(0);
        tags_.push_back(-1);
    }
    return QSV_OK;
}

int Profiler::end(int cls, double bytes, double flops, int64_t tag)
{
    if (!on_) return QSV_OK;
    hipEvent_t e;
    QSV_HIP(hipEventCreate(&e));
    QSV_HIP(hipEventRecord(e, s_->stream));
    events_.push_back(e);
    cls_.push_back(cls);
    bytes_.push_back(bytes);
    flops_.push_back(flops);
    tags_.push_back(tag);
    return QSV_OK;
}

int Profiler::collect(qsv_run_stats *st)
{
    if (!on_ || events_.empty()) return QSV_OK;
    QSV_HIP(hipEventSynchronize(events_.back()));
    for (size_t i = 1; i < events_.size(); i++) {
        if (cls_[i] < 0) continue;
        float ms = 0;
        QSV_HIP(hipEventElapsedTime(&ms, events_[i - 1], events_[i]));
        st->class_launches[cls_[i]] += 1;
        st->class_ms[cls_[i]] += ms;
        st->class_bytes[cls_[i]] += bytes_[i];
        st->class_flops[cls_[i]] += flops_[i];
        if (tags_[i] >= 0) {
            if ((size_t)tags_[i] >= tag_ms_.size()) tag_ms_.resize(tags_[i] + 1, 0.0f);
            tag_ms_[tags_[i]] = ms;
        }
    }
    return QSV_OK;
}

}  // namespace qsv

using namespace qsv;

extern "C" int qsv_run_circuit(qsv_state *s, const qsv_gate *gates, int64_t ngates, const double *mats,
                               const qsv_run_opts *opts, int32_t *rng, int *outcomes,
                               qsv_run_stats *stats)
{
    if (!s || !s->amps) return fail(QSV_ERR_ARG, "null state");
    if (ngates < 0 || (ngates > 0 && !gates)) return fail(QSV_ERR_ARG, "bad gate list");
    qsv_run_opts o{};
    if (opts) o = *opts;
    int rc = validate_gates(s, gates, ngates);
    if (rc) return rc;
    size_t need = 0;
    if ((rc = max_mat_index(gates, ngates, &need))) return rc;
    if (need && !mats) return fail(QSV_ERR_ARG, "matrix pool missing");
    if ((rc = upload_mats(s, mats, need))) return rc;
    if ((rc = tiles_flush(s))) return rc;
    // The fused path launches its own sweeps; a per-record profile times each
    // record's own sweep.  Otherwise gate tiling (when on) queues the gates.
    TileSuspend suspend(s, o.fuse_k >= 2 || o.profile != 0);

    hipEvent_t e0, e1;
    QSV_HIP(hipEventCreate(&e0));
    QSV_HIP(hipEventCreate(&e1));
    const int64_t launches0 = s->launches, ties0 = s->tie_resolutions;
    Profiler prof(s, o.profile != 0);
    QSV_HIP(hipEventRecord(e0, s->stream));
    int64_t blocks = 0;
    double pred_ms = 0, plan_ms = 0;
    int k_used = 0, planner = 0;
    if (o.fuse_k >= 2) {
        rc = run_fused(s, gates, ngates, mats, o, rng, outcomes, &blocks, &pred_ms, &prof, e0, &plan_ms, &k_used,
                       &planner);
    } else {
        for (int64_t i = 0; i < ngates && !rc; i++) {
            rc = prof.begin();
            if (!rc) rc = apply_record(s, gates[i], mats, o.exact, rng, outcomes);
            if (!rc) {
                double bytes, flops;
                const int cls = record_class(s, gates[i], &bytes, &flops);
                rc = prof.end(cls, bytes, flops);
            }
        }
    }
    if (!rc) rc = tiles_flush(s);
    hipError_t er = hipEventRecord(e1, s->stream);
    if (!rc && er != hipSuccess) rc = hip_fail(er, "hipEventRecord");
    if (!rc) {
        er = hipEventSynchronize(e1);
        if (er != hipSuccess) rc = hip_fail(er, "hipEventSynchronize");
    }
    float ms = 0;
    if (!rc) (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->gates = ngates;
        stats->launches = s->launches - launches0;
        stats->blocks = blocks;
        stats->tie_resolutions = s->tie_resolutions - ties0;
        stats->device_ms = ms;
        stats->model_ms_pred = pred_ms;
        stats->plan_ms = plan_ms;
        stats->fuse_k = k_used;
        stats->planner = planner;
        if (!rc) rc = prof.collect(stats);
    }
    if (!rc && prof.on() && o.fuse_k >= 2) {
        s->tlog_ms = prof.tag_ms();
        s->tlog_ms.resize(s->tlog_feats.size() / QSV_CM_FEATURES, 0.0f);
    }
    return rc;
}
