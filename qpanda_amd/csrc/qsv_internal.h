// qsv_internal.h -- engine internals shared by the HIP translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/qsv.h"

namespace qsv {

constexpr int kMaxIns = 40;          // zero-bit insert positions per launch
constexpr int kBlock = 256;          // threads per workgroup for sweeps

// Interleaved complex amplitude types (16 B and 8 B).
using c128 = double2;
using c64 = float2;

// Zero-bit insertion: given a compact index i over the free bits, place a 0
// at each sorted position.  Positions are in final-index coordinates and
// ascending, so each step shifts the high part left by one.
struct InsertSet {
    int n;                 // number of positions
    int pos[kMaxIns];      // ascending
};

__host__ __device__ inline uint64_t expand_index(uint64_t i, const InsertSet &s)
{
    for (int j = 0; j < s.n; j++) {
        const int p = s.pos[j];
        const uint64_t lo = i & ((uint64_t(1) << p) - 1);
        i = ((i >> p) << (p + 1)) | lo;
    }
    return i;
}

// Scalar 2x2 complex matrix in double precision (row-major).
struct Mat2 {
    double m[8];
};

struct qsv_error_state {
    std::string msg;
};

void set_error(const std::string &m);
int fail(int code, const std::string &m);
int hip_fail(hipError_t e, const char *what);

#define QSV_HIP(call)                                              \
    do {                                                           \
        hipError_t e_ = (call);                                    \
        if (e_ != hipSuccess) return ::qsv::hip_fail(e_, #call);   \
    } while (0)

}  // namespace qsv

struct qsv_state {
    int n = 0;
    int prec = QSV_C128;
    uint64_t dim = 0;
    void *amps = nullptr;            // device, interleaved complex
    hipStream_t stream = nullptr;
    double *red = nullptr;           // device reduction scratch (kRedSlots doubles)
    double *host_red = nullptr;      // pinned host mirror
    void *aux = nullptr;             // device scratch for readouts (marginal partials)
    void *op_ring = nullptr;         // operand rings of the streamed window schedules (qsv_fusion.cpp)
    size_t aux_bytes = 0;
    double *mats = nullptr;          // device copy of a circuit's matrix pool
    size_t mats_bytes = 0;
    int device = 0;
    int64_t launches = 0;
    int64_t tie_resolutions = 0;
    float f16_scale = 32768.0f;      // FP16x3 products: state scale (f16_state_scale)
    struct qsv_costmodel *cm = nullptr;   // attached cost model (not owned)
    struct TileQueue *tq = nullptr;       // gate tiling (qsv_set_gate_tiling), owned
    std::vector<float> tlog_feats;        // per-block features of the last profiled fused run
    std::vector<float> tlog_ms;           // and their measured device ms
};

namespace qsv {

constexpr int kRedSlots = 8192;      // partial sums per reduction (device scratch)

size_t amp_bytes(const qsv_state *s);
int check_qubit(const qsv_state *s, int q);
int ensure_aux(qsv_state *s, size_t bytes);

// Gate launchers (qsv_gates.hip); pair kernel operations in qsv_gatemath.h.
int launch_pair(qsv_state *s, int op, const InsertSet &ins, uint64_t or0, uint64_t or1,
                const double *u, double hscale);
int launch_phase(qsv_state *s, const InsertSet &ins, uint64_t orm, double re, double im);
int launch_block(qsv_state *s, const int *qubits, int k, const double *u, uint64_t ctrl_mask);
int launch_block_dev(qsv_state *s, const int *qubits, int k, const double *u_dev, uint64_t ctrl_mask);
int launch_fill_zero(qsv_state *s);

// Gate tiling (qsv_tiles.hip): while a state's queue is active, the pair /
// phase / small-block launchers hand their launch to tiles_enqueue_*, which
// returns 0 when it queued it and 1 when the caller must launch it itself.
// Every entry point that is not a gate flushes the queue first.
int tiles_enqueue_pair(qsv_state *s, int op, const InsertSet &ins, uint64_t or0, uint64_t or1, const double *u,
                       double hscale);
int tiles_enqueue_phase(qsv_state *s, const InsertSet &ins, uint64_t orm, double re, double im);
int tiles_enqueue_block(qsv_state *s, const int *qubits, int k, const double *u_dev, uint64_t ctrl_mask);
int tiles_flush(qsv_state *s);
void tiles_discard(qsv_state *s);
void tiles_free(qsv_state *s);
// Entry-point state check: a null state is an argument error, and gates that
// tiling has queued are applied before anything else reads or writes the
// state.  The gate entry points check with QSV_CHECK_GATE_STATE and queue.
#define QSV_CHECK_STATE(s)                                                   \
    do {                                                                     \
        if (!(s) || !(s)->amps) return fail(QSV_ERR_ARG, "null state");    \
        if ((s)->tq) {                                                       \
            const int rc_flush_ = tiles_flush(s);                            \
            if (rc_flush_) return rc_flush_;                                 \
        }                                                                    \
    } while (0)
#define QSV_CHECK_GATE_STATE(s) \
    do { if (!(s) || !(s)->amps) return fail(QSV_ERR_ARG, "null state"); } while (0)
// Suspends tiling while it lives when `on` (the fused path launches directly).
struct TileSuspend {
    TileSuspend(qsv_state *s, bool on);
    ~TileSuspend();
    qsv_state *s_;
    bool was_;
};

// Reductions (qsv_readout.hip).
int reduce_prob0(qsv_state *s, int q, double *p0);
int reduce_norm2(qsv_state *s, double *v);
int reduce_zz_chain(qsv_state *s, double *v);
int serial_prob0(qsv_state *s, int q, double *p0);
int serial_prob0_host(qsv_state *s, int q, double *p0);   // qsv_io.cpp
int inner(qsv_state *a, qsv_state *b, double *re, double *im);
int copy_state(qsv_state *d, qsv_state *s);
int sample_tree(qsv_state *s, const int *qubits, int m, int64_t gate_draws, int32_t *rng, int64_t shots,
                uint64_t *out, int64_t *ties);
int sample_copy(qsv_state *s, const int *qubits, int m, int64_t gate_draws, int32_t *rng, int64_t shots,
                uint64_t *out, int64_t *ties);
int gather(qsv_state *s, const uint64_t *idx, int64_t count, double *out);
int launch_collapse(qsv_state *s, int q, int outcome, double scale);
int marginals(qsv_state *s, const int *qubits, int m, double *out);
int pmeasure_top(qsv_state *s, const int *qubits, int m, uint64_t *idx, double *prob, int *count);

// Releases the operand rings of the streamed window schedules (qsv_fusion.cpp).
void free_operand_ring(qsv_state *s);

// Fused FP32 MFMA block (qsv_fused.hip).
int launch_block_mfma(qsv_state *s, const int *qubits, int k, const double *u, int split);
// Sets s->f16_scale from the state's norm (one FP64 reduction).
int f16_state_scale(qsv_state *s);

// Build an insert set from a bit mask (ascending positions).
inline InsertSet insert_set_from_mask(uint64_t mask)
{
    InsertSet s{};
    s.n = 0;
    for (int b = 0; b < 64; b++)
        if (mask >> b & 1) s.pos[s.n++] = b;
    return s;
}

}  // namespace qsv
