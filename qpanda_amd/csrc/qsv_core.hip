// qsv_core.hip -- the C ABI of include/qsv.h: state lifecycle, gate entry
// points, readouts, the Park-Miller generator and the per-gate circuit runner.
//
// Every entry point validates its qubits on the host before any launch, so no
// kernel ever receives an index outside the state.
#include "qsv_internal.h"
#include "qsv_gatemath.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

namespace qsv {

static thread_local std::string g_last_error;

void set_error(const std::string &m) { g_last_error = m; }

int fail(int code, const std::string &m)
{
    g_last_error = m;
    return code;
}

int hip_fail(hipError_t e, const char *what)
{
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
    return QSV_ERR_DEVICE;
}

size_t amp_bytes(const qsv_state *s) { return s->prec == QSV_C128 ? 16 : 8; }

int check_qubit(const qsv_state *s, int q)
{
    if (q < 0 || q >= s->n) return fail(QSV_ERR_QUBIT, "qubit index out of range");
    return QSV_OK;
}

int ensure_aux(qsv_state *s, size_t bytes)
{
    if (s->aux_bytes >= bytes) return QSV_OK;
    if (s->aux) {
        QSV_HIP(hipStreamSynchronize(s->stream));
        QSV_HIP(hipFree(s->aux));
        s->aux = nullptr;
        s->aux_bytes = 0;
    }
    if (hipMalloc(&s->aux, bytes) != hipSuccess) {
        s->aux = nullptr;
        return fail(QSV_ERR_ALLOC, "cannot allocate readout scratch");
    }
    s->aux_bytes = bytes;
    return QSV_OK;
}

// Controls must be inside the state and must not include the target bits.
static int check_mask(const qsv_state *s, uint64_t ctrl, uint64_t targets)
{
    if (s->n < 64 && (ctrl >> s->n)) return fail(QSV_ERR_QUBIT, "control qubit out of range");
    if (ctrl & targets) return fail(QSV_ERR_QUBIT, "control overlaps target");
    return QSV_OK;
}

}  // namespace qsv

using namespace qsv;

#define CHECK_STATE(s) QSV_CHECK_STATE(s)
#define CHECK_GATE_STATE(s) QSV_CHECK_GATE_STATE(s)

extern "C" {

const char *qsv_version(void) { return "qsv 0.1 gfx950"; }
const char *qsv_last_error(void) { return g_last_error.c_str(); }

int qsv_device_count(int *count)
{
    if (!count) return fail(QSV_ERR_ARG, "null");
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *count = c;
    return QSV_OK;
}

int qsv_create(int n, int precision, qsv_state **out)
{
    if (!out) return fail(QSV_ERR_ARG, "null out");
    *out = nullptr;
    if (n < 1 || n > QSV_MAX_QUBITS) return fail(QSV_ERR_QUBIT, "qubit count out of range");
    if (precision != QSV_C128 && precision != QSV_C64) return fail(QSV_ERR_PRECISION, "bad precision");
    qsv_state *s = new qsv_state();
    s->n = n;
    s->prec = precision;
    s->dim = uint64_t(1) << n;
    int rc = QSV_OK;
    if (hipGetDevice(&s->device) != hipSuccess) rc = fail(QSV_ERR_DEVICE, "no HIP device");
    if (!rc && hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
        rc = fail(QSV_ERR_DEVICE, "stream creation failed");
    if (!rc && hipMalloc(&s->amps, s->dim * amp_bytes(s)) != hipSuccess) {
        s->amps = nullptr;
        rc = fail(QSV_ERR_ALLOC, "cannot allocate the state in HBM");
    }
    if (!rc && hipMalloc((void **)&s->red, sizeof(double) * kRedSlots) != hipSuccess) {
        s->red = nullptr;
        rc = fail(QSV_ERR_ALLOC, "cannot allocate reduction scratch");
    }
    if (!rc && hipHostMalloc((void **)&s->host_red, sizeof(double) * 64, hipHostMallocDefault) != hipSuccess) {
        s->host_red = nullptr;
        rc = fail(QSV_ERR_ALLOC, "cannot allocate pinned host buffer");
    }
    if (!rc) rc = launch_fill_zero(s);
    if (!rc && hipStreamSynchronize(s->stream) != hipSuccess) rc = fail(QSV_ERR_DEVICE, "init failed");
    if (rc) {
        if (s->amps) (void)hipFree(s->amps);
        if (s->red) (void)hipFree(s->red);
        if (s->host_red) (void)hipHostFree(s->host_red);
        if (s->stream) (void)hipStreamDestroy(s->stream);
        delete s;
        return rc;
    }
    *out = s;
    return QSV_OK;
}

int qsv_destroy(qsv_state *s)
{
    if (!s) return QSV_OK;
    tiles_free(s);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    free_operand_ring(s);
    if (s->amps) (void)hipFree(s->amps);
    if (s->red) (void)hipFree(s->red);
    if (s->aux) (void)hipFree(s->aux);
    if (s->host_red) (void)hipHostFree(s->host_red);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return QSV_OK;
}

int qsv_reset_zero(qsv_state *s)
{
    CHECK_GATE_STATE(s);
    tiles_discard(s);                         // queued gates would act on the old state
    return launch_fill_zero(s);
}

int qsv_num_qubits(const qsv_state *s) { return s ? s->n : -1; }
int qsv_precision(const qsv_state *s) { return s ? s->prec : -1; }

// qsv_set_state / qsv_get_state: qsv_io.cpp

int qsv_gather(qsv_state *s, const uint64_t *idx, int64_t count, double *amps)
{
    CHECK_STATE(s);
    if (count < 0 || (count && (!idx || !amps))) return fail(QSV_ERR_ARG, "bad gather arguments");
    for (int64_t i = 0; i < count; i++)
        if (idx[i] >= s->dim) return fail(QSV_ERR_ARG, "gather index out of range");
    QSV_HIP(hipStreamSynchronize(s->stream));
    return gather(s, idx, count, amps);
}

int qsv_copy_state(qsv_state *dst, qsv_state *src)
{
    CHECK_STATE(dst);
    CHECK_STATE(src);
    if (dst->n != src->n) return fail(QSV_ERR_ARG, "qubit counts differ");
    if (dst == src) return QSV_OK;
    return copy_state(dst, src);
}

int qsv_inner(qsv_state *a, qsv_state *b, double *re, double *im)
{
    CHECK_STATE(a);
    CHECK_STATE(b);
    if (!re || !im) return fail(QSV_ERR_ARG, "null");
    if (a->n != b->n) return fail(QSV_ERR_ARG, "qubit counts differ");
    return inner(a, b, re, im);
}

int qsv_device_ptr(qsv_state *s, void **dptr)
{
    CHECK_STATE(s);
    if (!dptr) return fail(QSV_ERR_ARG, "null");
    *dptr = s->amps;
    return QSV_OK;
}

int qsv_stream(qsv_state *s, void **stream)
{
    CHECK_STATE(s);
    if (!stream) return fail(QSV_ERR_ARG, "null");
    *stream = (void *)s->stream;
    return QSV_OK;
}

int qsv_sync(qsv_state *s)
{
    CHECK_STATE(s);
    QSV_HIP(hipStreamSynchronize(s->stream));
    return QSV_OK;
}

// ---------------------------------------------------------------- gates ----
static const double kSq2Ref = 0.707106781186548;     // CPUQuantumGates.cpp:36
static const double kSq2Exact = 0.70710678118654752440;

static int one_qubit(qsv_state *s, int op, int t, uint64_t ctrl, const double *u, double h)
{
    CHECK_GATE_STATE(s);
    int rc = check_qubit(s, t);
    if (rc) return rc;
    const uint64_t tb = uint64_t(1) << t;
    if ((rc = check_mask(s, ctrl, tb))) return rc;
    InsertSet ins = insert_set_from_mask(ctrl | tb);
    return launch_pair(s, op, ins, ctrl, ctrl | tb, u, h);
}

int qsv_h(qsv_state *s, int target, uint64_t ctrl_mask, int exact)
{
    return one_qubit(s, OP_H, target, ctrl_mask, nullptr, exact ? kSq2Exact : kSq2Ref);
}

int qsv_x(qsv_state *s, int target, uint64_t ctrl_mask)
{
    return one_qubit(s, OP_X, target, ctrl_mask, nullptr, 0);
}

int qsv_apply_1q(qsv_state *s, int target, const double *u, uint64_t ctrl_mask)
{
    if (!u) return fail(QSV_ERR_ARG, "null matrix");
    return one_qubit(s, OP_U, target, ctrl_mask, u, 0);
}

int qsv_diag(qsv_state *s, int target, const double *d, uint64_t ctrl_mask)
{
    if (!d) return fail(QSV_ERR_ARG, "null diagonal");
    return one_qubit(s, OP_DIAG, target, ctrl_mask, d, 0);
}

// RX: [[c, -i s], [-i s, c]], dagger +i (CPUQuantumGates.cpp:271-309).
int qsv_rx(qsv_state *s, int target, double theta, int dagger, uint64_t ctrl_mask)
{
    const double c = std::cos(theta / 2), sn = std::sin(theta / 2);
    const double k = dagger ? 1.0 : -1.0;
    const double u[8] = {c, 0, 0, k * sn, 0, k * sn, c, 0};
    return qsv_apply_1q(s, target, u, ctrl_mask);
}

// RY: [[c, -s], [s, c]], dagger transposed (CPUQuantumGates.cpp:473-511).
int qsv_ry(qsv_state *s, int target, double theta, int dagger, uint64_t ctrl_mask)
{
    const double c = std::cos(theta / 2), sn = std::sin(theta / 2);
    const double g = dagger ? -1.0 : 1.0;
    const double u[8] = {c, 0, -g * sn, 0, g * sn, 0, c, 0};
    return qsv_apply_1q(s, target, u, ctrl_mask);
}

// RZ: diag(e^{-i t/2}, e^{+i t/2}), dagger conjugated (CPUQuantumGates.cpp:524-570).
int qsv_rz(qsv_state *s, int target, double theta, int dagger, uint64_t ctrl_mask)
{
    const double c = std::cos(theta / 2), sn = std::sin(theta / 2);
    const double s1 = dagger ? -sn : sn;
    const double d[4] = {c, -s1, c, s1};
    return qsv_diag(s, target, d, ctrl_mask);
}

int qsv_cphase(qsv_state *s, int q1, int q2, double re, double im, uint64_t ctrl_mask)
{
    CHECK_GATE_STATE(s);
    int rc;
    if ((rc = check_qubit(s, q1)) || (rc = check_qubit(s, q2))) return rc;
    if (q1 == q2) return QSV_OK;            // the reference's loop is empty (CPUQuantumGates.cpp:813)
    const uint64_t m = (uint64_t(1) << q1) | (uint64_t(1) << q2);
    if ((rc = check_mask(s, ctrl_mask, m))) return rc;
    InsertSet ins = insert_set_from_mask(m | ctrl_mask);
    return launch_phase(s, ins, m | ctrl_mask, re, im);
}

int qsv_cr(qsv_state *s, int q1, int q2, double theta, int dagger)
{
    const double c = std::cos(theta), sn = std::sin(theta);
    return qsv_cphase(s, q1, q2, c, dagger ? -sn : sn, 0);
}

static int iswap_like(qsv_state *s, int q1, int q2, const double *u)
{
    CHECK_GATE_STATE(s);
    int rc;
    if ((rc = check_qubit(s, q1)) || (rc = check_qubit(s, q2))) return rc;
    if (q1 == q2) return QSV_OK;            // empty loop in the reference
    const int hi = std::max(q1, q2), lo = std::min(q1, q2);
    InsertSet ins = insert_set_from_mask((uint64_t(1) << hi) | (uint64_t(1) << lo));
    // member 0 = |hi=0, lo=1>, member 1 = |hi=1, lo=0> (CPUQuantumGates.cpp:877-895)
    return launch_pair(s, OP_U, ins, uint64_t(1) << lo, uint64_t(1) << hi, u, 0);
}

int qsv_iswap(qsv_state *s, int q1, int q2, int dagger)
{
    const double k = dagger ? -1.0 : 1.0;
    const double u[8] = {0, 0, 0, k, 0, k, 0, 0};
    return iswap_like(s, q1, q2, u);
}

int qsv_sqiswap(qsv_state *s, int q1, int q2, int dagger, int exact)
{
    const double h = exact ? kSq2Exact : kSq2Ref, k = dagger ? -h : h;
    const double u[8] = {h, 0, 0, k, 0, k, h, 0};
    return iswap_like(s, q1, q2, u);
}

int qsv_cswap(qsv_state *s, int c, int a, int b)
{
    CHECK_GATE_STATE(s);
    int rc;
    if ((rc = check_qubit(s, c)) || (rc = check_qubit(s, a)) || (rc = check_qubit(s, b))) return rc;
    if (c == a || c == b || a == b) return fail(QSV_ERR_QUBIT, "controlSwap qubits must differ");
    const uint64_t bc = uint64_t(1) << c, ba = uint64_t(1) << a, bb = uint64_t(1) << b;
    InsertSet ins = insert_set_from_mask(bc | ba | bb);
    return launch_pair(s, OP_X, ins, bc | ba, bc | bb, nullptr, 0);
}

int qsv_apply_2q(qsv_state *s, int qa, int qb, const double *u)
{
    CHECK_GATE_STATE(s);
    if (!u) return fail(QSV_ERR_ARG, "null matrix");
    int rc;
    if ((rc = check_qubit(s, qa)) || (rc = check_qubit(s, qb))) return rc;
    if (qa == qb) return fail(QSV_ERR_QUBIT, "two-qubit gate on one qubit");
    const int qs[2] = {qb, qa};              // block bit 0 = qb, bit 1 = qa
    return launch_block(s, qs, 2, u, 0);
}

int qsv_apply_block(qsv_state *s, const int *qubits, int k, const double *u, int split)
{
    CHECK_STATE(s);
    if (!qubits || !u) return fail(QSV_ERR_ARG, "null argument");
    if (k < 1 || k > 5 || k > s->n) return fail(QSV_ERR_ARG, "block size must be 1..5");
    uint64_t seen = 0;
    for (int j = 0; j < k; j++) {
        int rc = check_qubit(s, qubits[j]);
        if (rc) return rc;
        if (seen >> qubits[j] & 1) return fail(QSV_ERR_QUBIT, "repeated qubit in block");
        seen |= uint64_t(1) << qubits[j];
    }
    if (s->prec == QSV_C64 && k >= 2) return launch_block_mfma(s, qubits, k, u, split);
    return launch_block(s, qubits, k, u, 0);
}

// -------------------------------------------------------------- readout ----
int qsv_prob0(qsv_state *s, int q, double *p0)
{
    CHECK_STATE(s);
    if (!p0) return fail(QSV_ERR_ARG, "null");
    int rc = check_qubit(s, q);
    if (rc) return rc;
    return reduce_prob0(s, q, p0);
}

int qsv_norm2(qsv_state *s, double *v)
{
    CHECK_STATE(s);
    if (!v) return fail(QSV_ERR_ARG, "null");
    return reduce_norm2(s, v);
}

int qsv_expect_zz_chain(qsv_state *s, double *v)
{
    CHECK_STATE(s);
    if (!v) return fail(QSV_ERR_ARG, "null");
    return reduce_zz_chain(s, v);
}

int qsv_collapse(qsv_state *s, int q, int outcome, double scale)
{
    CHECK_STATE(s);
    int rc = check_qubit(s, q);
    if (rc) return rc;
    return launch_collapse(s, q, outcome ? 1 : 0, scale);
}

// Largest state for which a tie is re-decided by the serial-order sum: the
// reference runs programs of at most 30 qubits (QPandaAPI.cpp:87), so above
// that no reference order exists to reproduce and the FP64 tree decides.  At
// 30 qubits the host sum takes a few seconds (a tie needs |(float)r - p0| <=
// 2^31 eps, about one measurement in 10^6).
static const int kSerialTieMaxQubits = 30;

int qsv_measure(qsv_state *s, int q, double r, int *outcome)
{
    CHECK_STATE(s);
    if (!outcome) return fail(QSV_ERR_ARG, "null");
    int rc = check_qubit(s, q);
    if (rc) return rc;
    double p0;
    if ((rc = reduce_prob0(s, q, &p0))) return rc;
    const float fr = (float)r;                       // CPUQuantumGates.cpp:1086
    // The CPU simulator's serial sum can differ from this one by at most
    // about 2^n rounding steps; inside that band decide with its own order.
    const double band = std::ldexp(DBL_EPSILON, s->n + 1) * std::max(1.0, (double)fr);
    if (std::fabs((double)fr - p0) <= band && s->n <= kSerialTieMaxQubits) {
        if ((rc = serial_prob0(s, q, &p0))) return rc;
        s->tie_resolutions++;
    }
    const int out = ((double)fr > p0) ? 1 : 0;       // :1088
    const double scale = out ? 1 / std::sqrt(1 - p0) : 1 / std::sqrt(p0);   // :1096, :1111
    if ((rc = launch_collapse(s, q, out, scale))) return rc;
    *outcome = out;
    return QSV_OK;
}

static int check_sample_args(qsv_state *s, const int *qubits, int m, int64_t gate_draws, int32_t *rng,
                             int64_t shots, uint64_t *out)
{
    if (!qubits || !rng || (!out && shots) || shots < 0 || gate_draws < 0) return fail(QSV_ERR_ARG, "bad argument");
    if (m < 1 || m > 64 || m > s->n) return fail(QSV_ERR_ARG, "1..min(64, n) measured qubits");
    uint64_t seen = 0;
    for (int j = 0; j < m; j++) {
        int rc = check_qubit(s, qubits[j]);
        if (rc) return rc;
        if (seen >> qubits[j] & 1) return fail(QSV_ERR_QUBIT, "repeated qubit");
        seen |= uint64_t(1) << qubits[j];
    }
    return QSV_OK;
}

int qsv_sample_shots(qsv_state *s, const int *qubits, int m, int64_t gate_draws, int32_t *rng,
                     int64_t shots, uint64_t *out, int64_t *ties)
{
    CHECK_STATE(s);
    int rc = check_sample_args(s, qubits, m, gate_draws, rng, shots, out);
    if (rc) return rc;
    QSV_HIP(hipStreamSynchronize(s->stream));
    return sample_tree(s, qubits, m, gate_draws, rng, shots, out, ties);
}

int qsv_sample_shots_copy(qsv_state *s, const int *qubits, int m, int64_t gate_draws, int32_t *rng,
                          int64_t shots, uint64_t *out, int64_t *ties)
{
    CHECK_STATE(s);
    int rc = check_sample_args(s, qubits, m, gate_draws, rng, shots, out);
    if (rc) return rc;
    QSV_HIP(hipStreamSynchronize(s->stream));
    return sample_copy(s, qubits, m, gate_draws, rng, shots, out, ties);
}

static int check_marginal_args(qsv_state *s, const int *qubits, int m)
{
    if (!qubits) return fail(QSV_ERR_ARG, "null");
    if (m < 1 || m > s->n) return fail(QSV_ERR_ARG, "marginal over 1..n qubits");
    uint64_t seen = 0;
    for (int j = 0; j < m; j++) {
        int rc = check_qubit(s, qubits[j]);
        if (rc) return rc;
        if (seen >> qubits[j] & 1) return fail(QSV_ERR_QUBIT, "repeated qubit");
        seen |= uint64_t(1) << qubits[j];
    }
    return QSV_OK;
}

int qsv_marginals(qsv_state *s, const int *qubits, int m, double *out)
{
    CHECK_STATE(s);
    if (!out) return fail(QSV_ERR_ARG, "null");
    int rc = check_marginal_args(s, qubits, m);
    if (rc) return rc;
    return marginals(s, qubits, m, out);
}

int qsv_pmeasure(qsv_state *s, const int *qubits, int m, uint64_t *idx, double *prob, int *count)
{
    CHECK_STATE(s);
    if (!idx || !prob || !count) return fail(QSV_ERR_ARG, "null");
    int rc = check_marginal_args(s, qubits, m);
    if (rc) return rc;
    return pmeasure_top(s, qubits, m, idx, prob, count);
}

// ------------------------------------------------------------------ RNG ----
int32_t qsv_pm_next(int32_t st)
{
    const int32_t ia = 16807, im = 2147483647, iq = 127773, ir = 2836;
    int32_t t = ia * (st % iq) - ir * (st / iq);
    if (t < 0) t += im;
    return t;
}

int32_t qsv_pm_jump(int32_t st, uint64_t k)
{
    uint64_t base = 16807, acc = 1;
    const uint64_t m = 2147483647ull;
    while (k) {
        if (k & 1) acc = acc * base % m;
        base = base * base % m;
        k >>= 1;
    }
    return (int32_t)((uint64_t)st * acc % m);
}

int32_t qsv_pm_seed_from_tm(int y, int mo, int d, int h, int mi, int se)
{
    return y + 70 * (mo + 1 + 12 * (d + 31 * (h + 23 * (mi + 59 * se))));
}

}  // extern "C"
