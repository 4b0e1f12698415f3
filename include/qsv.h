/*
 * qsv.h -- C ABI of the MI355X (gfx950) state-vector engine.
 *
 * This is the drop-in boundary for QPanda's `QuantumGates` gate API
 * (QPandaSDK/QuantumInstructionHandle/QuantumGates.h:31-474).  QPanda's GPU
 * back end reaches its CUDA library through dlsym() on `extern "C"` functions
 * whose parameters are C++ types (QPandaSDK/GPUEmulator/GPUGatesDecl.h:38-75);
 * this header replaces that table with a true C ABI: an opaque handle, plain
 * pointers and sizes, `int` status codes.  Each entry point names the
 * reference function it replaces.
 *
 * Conventions (these follow QPanda's CPU simulator, the parity target):
 *  - amplitude index bit q is qubit q (qubit 0 = least significant bit);
 *  - complex numbers are interleaved (re, im) pairs of doubles on the host;
 *  - matrices are row-major complex, U[r][c] = (m[2*(r*d+c)], m[2*(r*d+c)+1]);
 *  - for a two-qubit matrix on (qa, qb) the row index is 2*bit(qa) + bit(qb);
 *  - for a k-qubit block on qubits[0..k-1] the row index bit j is bit(qubits[j]);
 *  - H uses the constant 0.707106781186548 unless `exact` asks for 1/sqrt(2)
 *    (CPUQuantumGates.cpp:36);
 *  - RZ is diag(e^{-i theta/2}, e^{+i theta/2}) (CPUQuantumGates.cpp:524-570);
 *  - iSWAP / sqiSWAP use +i (CPUQuantumGates.cpp:838-899, 978-1039).
 *
 * Every call is synchronous with respect to the host unless its comment says
 * otherwise: readouts wait for the device, gate calls are stream-ordered on the
 * handle's stream (a readout or qsv_sync() orders them), so callers observe the
 * reference's synchronous semantics.
 */
#ifndef QSV_H
#define QSV_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (0 = OK, negative = error) ------------------------------ */
#define QSV_OK              0
#define QSV_ERR_ARG        -1   /* bad argument (null pointer, bad size)        */
#define QSV_ERR_QUBIT      -2   /* qubit index out of range or repeated         */
#define QSV_ERR_ALLOC      -3   /* device or host allocation failed             */
#define QSV_ERR_DEVICE     -4   /* HIP runtime error                            */
#define QSV_ERR_PRECISION  -5   /* operation not available at this precision   */
#define QSV_ERR_UNSUPPORTED -6  /* gate kind / option not supported             */
#define QSV_ERR_NUMERIC    -7   /* probability of the requested branch is zero  */

/* ---- precisions ------------------------------------------------------------ */
#define QSV_C128 0   /* complex<double> amplitudes, 16 B each */
#define QSV_C64  1   /* complex<float> amplitudes, 8 B each; reductions in FP64 */

#define QSV_MAX_QUBITS 34

/* ---- gate records (circuit interchange format, 32 bytes each) -------------
 * Shared by the engine, the CPU oracle and the reference driver.
 *   kind        one of QSV_GATE_*
 *   dagger      nonzero = inverse (the reference's iDagger flag)
 *   q0,q1,q2    qubits; roles per kind below
 *   mat         for U1/U2: offset, in complex entries, into the matrix pool
 *   theta       rotation angle (RX/RY/RZ half-angle convention; CR full angle)
 * Roles: H/X/RX/RY/RZ/U1: q0 = target.  CNOT: q0 = control, q1 = target.
 * CZ/CR/ISWAP/SQISWAP: q0, q1.  U2: matrix row = 2*bit(q0)+bit(q1).
 * TOFFOLI: q0, q1 = controls, q2 = target.  CSWAP: q0 = control, swaps q1/q2.
 * MEASURE: q0 = qubit, q1 = classical register.  RESET: q0 = qubit.
 */
enum {
    QSV_GATE_H = 1,
    QSV_GATE_X = 2,
    QSV_GATE_RX = 3,
    QSV_GATE_RY = 4,
    QSV_GATE_RZ = 5,
    QSV_GATE_CNOT = 6,
    QSV_GATE_CZ = 7,
    QSV_GATE_CR = 8,
    QSV_GATE_ISWAP = 9,
    QSV_GATE_SQISWAP = 10,
    QSV_GATE_TOFFOLI = 11,
    QSV_GATE_CSWAP = 12,
    QSV_GATE_U1 = 13,
    QSV_GATE_U2 = 14,
    QSV_GATE_MEASURE = 15,
    QSV_GATE_RESET = 16
};

typedef struct qsv_gate {
    int32_t kind;
    int32_t dagger;
    int32_t q0, q1, q2;
    int32_t mat;
    double  theta;
} qsv_gate;

typedef struct qsv_state qsv_state;   /* opaque device state */

/* ---- lifecycle ------------------------------------------------------------- */

/* Allocate 2^n amplitudes in HBM and set |0...0>.
 * Replaces CPUQuantumGates::initState (CPUQuantumGates.cpp:1136-1157) and
 * GATEGPU::initstate (GPUGates.cu:1208-1299).  precision = QSV_C128 | QSV_C64. */
int qsv_create(int n, int precision, qsv_state **out);

/* Free the state.  Replaces destroyState (GPUQuantumGates.cpp:857-871,
 * GPUGates.cu:1129-1161). */
int qsv_destroy(qsv_state *s);

/* |0...0> again without reallocating.  Replaces endGate
 * (CPUQuantumGates.cpp:1445-1453; GPUGates.cu:1167-1206 clearState). */
int qsv_reset_zero(qsv_state *s);

int qsv_num_qubits(const qsv_state *s);
int qsv_precision(const qsv_state *s);

/* Host <-> device copies of the whole state as interleaved (re, im) doubles
 * (converted to/from FP32 for QSV_C64).  qsv_get_state replaces getState
 * (GPUGates.cu:2160-2169) / the CPU's mvQuantumStat read-out
 * (CPUQuantumGates.cpp:64-82).  qsv_set_state has no reference counterpart
 * (the reference always restarts from |0...0>, QuantumInstructionHandleAPI.cpp:96-100). */
int qsv_set_state(qsv_state *s, const double *amps);
int qsv_get_state(qsv_state *s, double *amps);

/* State files, streamed through pinned staging buffers (no whole-state host
 * copy).  QSV_FILE_BINARY: 2^n records {f64 re, f64 im} in index order, the
 * record layout of the reference CLI's .fullstate.binary
 * (QCommand.cpp:437-487); QSV_FILE_TEXT: one line per amplitude,
 * "state[i].real = X state[i].imag = Y", the getQState text
 * (CPUQuantumGates.cpp:64-82) that the CLI saves as .fullstate
 * (QCommand.cpp:402-428).  digits: significant digits of the text numbers
 * (0 = 6, as the reference prints them; 17 reads back exactly); unused for
 * binary.  A complex64 state is written as FP64 and read back rounded to FP32.
 * Loading checks the size (binary) or every index (text); a failed load
 * leaves the state partly overwritten.  The reference has no load path. */
#define QSV_FILE_BINARY 0
#define QSV_FILE_TEXT 1
int qsv_save_state(qsv_state *s, const char *path, int format, int digits);
int qsv_load_state(qsv_state *s, const char *path, int format);

/* amps[2i], amps[2i+1] = amplitude idx[i] (FP64; indices < 2^n, checked).
 * No reference counterpart: reads single amplitudes of a state too large to
 * copy out, e.g. the probabilities of sampled outcomes. */
int qsv_gather(qsv_state *s, const uint64_t *idx, int64_t count, double *amps);

/* Device-side copy of src into dst (same qubit count; precisions may differ,
 * converting c128 <-> c64).  A streaming kernel, used for snapshots and for
 * comparing paths.  No reference counterpart. */
int qsv_copy_state(qsv_state *dst, qsv_state *src);

/* Moves every qubit q to position perm[q] (perm a permutation of 0..n-1): the
 * amplitude of basis state x ends at the index whose bit perm[q] is bit q of
 * x, for every q.  Pure data movement (bit-exact), in sweeps of 13-qubit tiles
 * that hold qubits 0..3 (one sweep per 12 qubits a cycle of the permutation
 * moves outside 0..3, at most ceil((n - 1) / 12) per cycle).  complex64 states
 * of >= 13 qubits.  No reference counterpart: the window schedules relabel
 * qubits this way between sweeps and restore the order at the end. */
int qsv_permute_qubits(qsv_state *s, const int *perm);

/* <a|b> in FP64 (either precision; same qubit count).  Fidelity of two states is
 * |<a|b>|^2 / (<a|a><b|b>).  No reference counterpart. */
int qsv_inner(qsv_state *a, qsv_state *b, double *re, double *im);

/* Raw device pointer of the amplitude array (interleaved complex of the state's
 * precision) and the handle's HIP stream, for callers that keep data in HBM. */
int qsv_device_ptr(qsv_state *s, void **dptr);
int qsv_stream(qsv_state *s, void **stream);
int qsv_sync(qsv_state *s);

/* Gate tiling (on != 0): the gate entry points below (and the per-gate path
 * of qsv_run_circuit) queue their sweeps instead of launching them, as long as
 * the queued gates' qubits fit one 64 KiB tile together with the qubits of a
 * 128-byte run (12 qubits for complex128 with qubits 0..2, 13 for complex64
 * with qubits 0..3); the next call that is not a gate
 * (any readout, measurement, copy, qsv_sync, qsv_device_ptr, qsv_set_gate_tiling
 * with on = 0, ...) applies the queue in one sweep over the state's tiles,
 * with the same arithmetic, so the state is bit-identical to the
 * one-sweep-per-gate result.  Gates wider than a tile, dense blocks of 3+
 * qubits and states smaller than a tile sweep gate by gate.  Off by default;
 * the first call with on != 0 allocates 208 KiB of device staging.  A caller
 * that writes the amplitude array through qsv_device_ptr must call qsv_sync
 * (or any readout) before each such write while tiling is on.  The reference
 * sweeps the state once per gate (CPUQuantumGates.cpp:92-1435,
 * GPUGates.cu:129-1020); this call replaces nothing there. */
int qsv_set_gate_tiling(qsv_state *s, int on);

/* ---- single-qubit and controlled single-qubit gates -------------------------
 * ctrl_mask: bit c set = qubit c is a control that must be 1 (0 = uncontrolled).
 * The target bit must not be in ctrl_mask.  The reference passes controls as a
 * list with the target appended last (Instruction.cpp:62,155,243,305,415); the
 * C++ adapter turns that list into this mask. */

/* Hadamard.  Replaces Hadamard(qn,err) (CPUQuantumGates.cpp:92-117) and the
 * controlled form (CPUQuantumGates.cpp:202-258).  exact != 0 uses 1/sqrt(2). */
int qsv_h(qsv_state *s, int target, uint64_t ctrl_mask, int exact);

/* Pauli-X / NOT / CNOT / Toffoli / multi-controlled X.  Replaces NOT
 * (CPUQuantumGates.cpp:1277-1365), CNOT (:668-757), toffoli (:1378-1435). */
int qsv_x(qsv_state *s, int target, uint64_t ctrl_mask);

/* General 2x2 complex matrix u[8] (row-major re,im).  Serves RX
 * (CPUQuantumGates.cpp:271-383), RY (:396-511) and a correct
 * unitarySingleQubitGate (the reference's targets qn-1, :1188). */
int qsv_apply_1q(qsv_state *s, int target, const double *u, uint64_t ctrl_mask);

/* Diagonal diag(d0, d1), d[4] = {re0, im0, re1, im1}.  Serves RZ
 * (CPUQuantumGates.cpp:524-655). */
int qsv_diag(qsv_state *s, int target, const double *d, uint64_t ctrl_mask);

/* Convenience wrappers with the reference's angle conventions. */
int qsv_rx(qsv_state *s, int target, double theta, int dagger, uint64_t ctrl_mask);
int qsv_ry(qsv_state *s, int target, double theta, int dagger, uint64_t ctrl_mask);
int qsv_rz(qsv_state *s, int target, double theta, int dagger, uint64_t ctrl_mask);

/* ---- two- and three-qubit gates ------------------------------------------- */

/* |11> *= (re + i im) on qubits (q1, q2), with extra controls.  CR(theta) is
 * (cos theta, +-sin theta) (CPUQuantumGates.cpp:773-825); CZ is (-1, 0). */
int qsv_cphase(qsv_state *s, int q1, int q2, double re, double im, uint64_t ctrl_mask);
int qsv_cr(qsv_state *s, int q1, int q2, double theta, int dagger);

/* iSWAP / sqiSWAP with the CPU simulator's +i phase (dagger: -i).
 * Replaces iSWAP (CPUQuantumGates.cpp:838-899), sqiSWAP (:978-1039). */
int qsv_iswap(qsv_state *s, int q1, int q2, int dagger);
int qsv_sqiswap(qsv_state *s, int q1, int q2, int dagger, int exact);

/* Fredkin: swap qubits a and b where control c is 1.  Replaces controlSwap
 * (CPUQuantumGates.cpp:934-965). */
int qsv_cswap(qsv_state *s, int c, int a, int b);

/* Dense two-qubit unitary u[32] on (qa, qb), row = 2*bit(qa) + bit(qb).
 * No reference counterpart (the reference has matReverse only,
 * CPUQuantumGates.cpp:1221-1244). */
int qsv_apply_2q(qsv_state *s, int qa, int qb, const double *u);

/* Dense k-qubit block (k = 1..5), u = 2^k x 2^k row-major complex doubles.
 * C128 states apply it in FP64; C64 states use the fused MFMA kernel
 * (FP32 complex GEMM with FP32 accumulation; split = 1 adds the lo part of
 * the matrix, U = U_hi + U_lo; split = 2 runs the product as BF16x6 on the
 * BF16 matrix cores: matrix and state split into three bf16 pieces, the six
 * leading piece products accumulated in FP32; split = 3 runs it as FP16x3 on
 * the FP16 matrix cores: matrix and state (scaled by a power of two derived
 * from the state's norm) split into two fp16 pieces, the three leading piece
 * products accumulated in FP32).  No reference counterpart. */
int qsv_apply_block(qsv_state *s, const int *qubits, int k, const double *u, int split);

/* ---- readout --------------------------------------------------------------- */

/* Sum of |a|^2 over amplitudes whose bit q is 0 (FP64 on device,
 * deterministic order).  The quantity qubitMeasure computes
 * (CPUQuantumGates.cpp:1074-1083; GPUGates.cu:597-682). */
int qsv_prob0(qsv_state *s, int q, double *p0);

/* Squared norm in FP64. */
int qsv_norm2(qsv_state *s, double *norm2);

/* Keep branch `outcome` of qubit q, zero the other, scale by `scale`.
 * (CPUQuantumGates.cpp:1092-1124; GPUGates.cu:684-718) */
int qsv_collapse(qsv_state *s, int q, int outcome, double scale);

/* Measure qubit q with draw r in [0,1] (the value of the reference's
 * randGenerator() call): outcome = ((float)r > p0), then collapse and
 * renormalise by 1/sqrt(p).  Replaces qubitMeasure (CPUQuantumGates.cpp:1071-1126).
 * When (float)r lies within the rounding band of p0 (2^(n+1) eps), p0 is
 * recomputed in the CPU simulator's serial summation order (on the host,
 * with the C library's hypot) before deciding, for states of up to 30 qubits
 * (the reference's limit, QPandaAPI.cpp:87); larger states decide on the
 * FP64 device sum. */
int qsv_measure(qsv_state *s, int q, double r, int *outcome);

/* Full marginal distribution over qubits[0..m-1] (1 <= m <= n, distinct)
 * into out[2^m]; outcome bit j = bit(qubits[j]) (caller order,
 * CPUQuantumGates.cpp:171-174).  FP64, deterministic order. */
int qsv_marginals(qsv_state *s, const int *qubits, int m, double *out);

/* pMeasure compatibility view: the distribution sorted by probability
 * descending, truncated to 10 entries (CPUQuantumGates.cpp:159-190), for any
 * 1 <= m <= n.  The top 10 are selected on the device; equal probabilities
 * are ordered by index (the reference's std::sort leaves their order
 * unspecified).  Writes *count = min(10, 2^m) entries of (index, probability). */
int qsv_pmeasure(qsv_state *s, const int *qubits, int m, uint64_t *idx, double *prob, int *count);

/* sum_{i=0}^{n-2} <Z_i Z_{i+1}> in FP64. */
int qsv_expect_zz_chain(qsv_state *s, double *value);

/* Allocation-free form of qsv_norm2 / qsv_prob0 / qsv_expect_zz_chain.  No
 * counterpart in the reference: it is for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  amps: device pointer to 2^n interleaved complex amplitudes of
 * `precision`; out: device pointer to one double that receives the FP64 result
 * (the same value the handle-based call returns: same kernels, same
 * summation order).  Both must be device memory of the current device.  The
 * kernels are queued on `stream` (a hipStream_t; NULL = the null stream) and
 * the call returns without waiting.  It calls no hipMalloc / hipFree: the
 * partial sums (at most 4096 doubles, 32 KiB) live in a library-owned device
 * array, one per GPU, located with hipGetSymbolAddress on the first call on
 * that device.  Calls that overlap on one GPU (other streams, other host
 * threads) race on that array; the array keeps the last call's partial sums.
 * what: QSV_READ_NORM2, QSV_READ_PROB0 (q = the qubit) or QSV_READ_ZZ_CHAIN.
 * n = 1..34. */
#define QSV_READ_NORM2    0
#define QSV_READ_PROB0    1
#define QSV_READ_ZZ_CHAIN 2
int qsv_readout_noalloc(const void *amps, int n, int precision, int what, int q, double *out, void *stream);

/* Shot loop of QPandaSDK::run (QPandaAPI.cpp:131-211, QuantumInstructionHandleAPI.cpp:96-114)
 * without re-simulating the program per shot.  The reference restarts every
 * shot from |0...0>, runs the program (one draw per gate call) and then measures;
 * because the gates do not depend on the draws, the pre-measurement state is the
 * same every shot.  With that state in `s` (left unchanged), each shot skips
 * `gate_draws` draws by jump-ahead and measures qubits[0..m-1] (distinct) in
 * order with the reference's decision rule (CPUQuantumGates.cpp:1071-1126);
 * bit j of out[shot] is the outcome of qubits[j].  *rng is the generator state
 * before the first shot's gates and is advanced past the last shot.
 * No state copy per shot: the decisions are evaluated from FP64 prefix sums of
 * |a|^2 (a marginal tree plus per-shot LDS reductions); a shot with a decision
 * inside the measurement tie band is re-run by the copy path below, so the
 * outcomes equal qsv_sample_shots_copy's (for a complex64 state: on its FP64
 * conversion).  *ties (may be NULL) counts shots re-run that way plus serial
 * re-decisions inside them. */
int qsv_sample_shots(qsv_state *s, const int *qubits, int m, int64_t gate_draws, int32_t *rng,
                     int64_t shots, uint64_t *out, int64_t *ties);

/* The same shot loop shaped as the reference's: per shot, copy the state and
 * measure it qubit by qubit (collapse and renormalise each time).  Cost: one
 * state copy plus 2m sweeps per shot; kept as the cross-check of
 * qsv_sample_shots.  *ties counts serial re-decisions. */
int qsv_sample_shots_copy(qsv_state *s, const int *qubits, int m, int64_t gate_draws, int32_t *rng,
                          int64_t shots, uint64_t *out, int64_t *ties);

/* ---- whole circuits ---------------------------------------------------------
 * Apply gates[0..ngates-1] in order.  mats = matrix pool (complex doubles) for
 * U1/U2 records.  For C128 each record is one sweep (the per-gate path).
 * rng: Park-Miller state; each gate consumes one draw exactly as the reference
 * executor does (QuantumGates.cpp:46-85, e.g. CPUQuantumGates.cpp:97) and
 * MEASURE consumes one more; NULL skips RNG bookkeeping.  outcomes[creg] receives
 * measurement results (may be NULL when the circuit has none). */
typedef struct qsv_run_opts {
    int exact;          /* 1: H / sqiSWAP use exact 1/sqrt(2)                   */
    int fuse_k;         /* 0: per-gate path; 2..5: fused blocks of <= k qubits  */
    int split;          /* fused blocks: 0 = FP32 product, 1 = FP32 + BF16 lo   */
                        /* correction, 2 = BF16x6 product, 3 = FP16x3 product   */
    int scheduler;      /* 0: greedy fusion, 1: learned cost model              */
    int profile;        /* 1: time every launch with events (per-class stats)   */
    int tile;           /* fused blocks: tile column bits 6, 7 or 8 (0 = 7);   */
                        /* the learned scheduler picks per block instead        */
    int group;          /* scheduler 2: at most this many consecutive blocks    */
                        /* share one HBM sweep (0 or 1 = one block per sweep)   */
    int pass_tb;        /* scheduler 2: tile bits of such a pass (0 = minimal)  */
    int pass_kernel;    /* multi-block sweeps: 0 or 1 = one workgroup per tile  */
                        /* (other values: QSV_ERR_UNSUPPORTED)                  */
    int fixed_low;      /* window schedules: 0 = relabel qubits between sweeps  */
                        /* from 28 qubits, 1 = never (qubits 0..3 stay in      */
                        /* place), -1 = from 14 qubits (see below)              */
} qsv_run_opts;
/* opts.scheduler: 0 = greedy blocks at fuse_k, one sweep each;
 *                 1 = learned (planner, block size, grouping into sweeps,
 *                     tile shapes);
 *                 2 = greedy blocks grouped greedily into multi-block sweeps;
 *                 3 = window planner: per sweep a window of pass_tb (0 = 13)
 *                     tile bits holding qubits 0..3, every runnable gate
 *                     inside it fused into blocks of <= fuse_k (4, 5)
 *                     qubits, one sweep per window (states of >= 11 qubits;
 *                     smaller ones fall back to scheduler 0).
 * Relabelling (schedulers 1 and 3, 13-bit tiles; >= 28 qubits, or >= 14 with
 * fixed_low = -1): a sweep's tile always holds the amplitudes' bits 0..3 (128-byte
 * runs), but which qubits sit there changes: each window's sweep stores its
 * tile with four of the window's qubits moved to bits 0..3 (those the next
 * window starts from), so consecutive windows share four qubits instead of
 * every window holding qubits 0..3.  Before a MEASURE / RESET record and at
 * the end of the run the qubits return to their own bits (qsv_permute_qubits'
 * sweeps), so the state leaves in the usual order. */

/* Kernel classes for per-launch profiling. */
enum {
    QSV_K_DENSE1 = 0,   /* 2x2 pair sweep (H, RX, RY, U1, controlled forms)      */
    QSV_K_DIAG = 1,     /* diagonal pair sweep (RZ)                              */
    QSV_K_PERM = 2,     /* permutations (X, CNOT, Toffoli, Fredkin) and iSWAPs    */
    QSV_K_PHASE = 3,    /* one-subspace phase (CR, CZ)                           */
    QSV_K_BLOCK64 = 4,  /* dense block in FP64 (two-qubit unitaries, c128 blocks) */
    QSV_K_FUSED = 5,    /* fused FP32 MFMA block (c64)                           */
    QSV_K_READOUT = 6,  /* reductions, collapse                                  */
    QSV_K_PASS = 7,     /* several fused blocks in one sweep (c64)               */
    QSV_K_CLASSES = 8
};

typedef struct qsv_run_stats {
    int64_t gates;          /* logical gates applied                     */
    int64_t launches;       /* kernel launches issued                    */
    int64_t blocks;         /* fused blocks (0 on the per-gate path)     */
    int64_t tie_resolutions;/* measurements decided in serial order      */
    double  device_ms;      /* device time between events around the run */
    double  model_ms_pred;  /* scheduler's predicted time (fused path)    */
    /* filled when opts.profile != 0: per kernel class, launches, device time
     * (event to event on the state's stream), algorithmic HBM bytes (every
     * amplitude the gate may change read once and written once) and flops */
    int64_t class_launches[QSV_K_CLASSES];
    double  class_ms[QSV_K_CLASSES];
    double  class_bytes[QSV_K_CLASSES];
    double  class_flops[QSV_K_CLASSES];
    double  plan_ms;        /* host time spent planning (fused path); the   */
                            /* device time above starts after planning      */
    int64_t fuse_k;         /* block size the fused plan used (the learned  */
                            /* scheduler picks it), 0 on the per-gate path  */
    int64_t planner;        /* fused plan: 0 block-first (plan_greedy),     */
                            /* 1 window planner                             */
} qsv_run_stats;

/* Greedy fusion plan of a gate list (host only, no device needed): blocks of at
 * most max_k qubits in emission order.  For block b: k_out[b] qubits (0 for a
 * MEASURE/RESET record that runs alone), qubits_out[5b..5b+4] (row bit j =
 * bit(qubits_out[5b+j]), -1 padded), mats_out[2048b..] the 2^k x 2^k complex
 * matrix (row-major re, im), ngates_out[b] the logical gates it holds (the
 * record index for k = 0).  Any output pointer may be NULL.  No reference
 * counterpart (the reference applies one gate per sweep). */
int qsv_fuse_plan(const qsv_gate *gates, int64_t ngates, const double *mats, int max_k, int exact,
                  int64_t max_blocks, int *k_out, int *qubits_out, double *mats_out,
                  int64_t *ngates_out, int64_t *nblocks);

/* Window plan of a gate list for an n-qubit state (host only, no device
 * needed): the blocks (<= max_k = 4 or 5 qubits) of one sweep each lie in a
 * window of tb qubits that holds qubits 0..min(3, n-1) (tb = 7..13, or
 * 4..13 when tb >= n: a window needs room for a 3-qubit record beside qubits
 * 0..3); outputs as
 * qsv_fuse_plan plus window_out[b], the window of block b (a MEASURE/RESET
 * record has a window of its own).  No reference counterpart. */
int qsv_window_plan(const qsv_gate *gates, int64_t ngates, const double *mats, int max_k, int exact,
                    int tb, int n, int64_t max_blocks, int *k_out, int *qubits_out, double *mats_out,
                    int64_t *ngates_out, int *window_out, int64_t *nblocks);

/* The window plan found by a beam search over window sequences (host only):
 * width states kept per step, each expanded into its `branch` best windows
 * among the lookahead growth's own and the completions from its `first` best
 * first growth candidates; a window is fixed once the best state is `delay`
 * windows past it.  Each window's records are fused in the better of the
 * closure's order and program order, and of `fuse_tries` random dependency
 * orders.  relabel != 0: each window grows from the four qubits the previous
 * one leaves at bits 0..3 (see qsv_run_opts.fixed_low) instead of from qubits
 * 0..3; the blocks keep their logical qubits.  width = 1, fuse_tries = 0 and
 * relabel = 0 give qsv_window_plan's plan.  Outputs as qsv_window_plan.  The
 * learned scheduler (opts.scheduler = 1) plans this way while the first
 * windows run, relabelled from 14 qubits: width 8, branch 4, first 8, delay 6
 * for >= 30 qubits, width 4, branch 3, first 6, delay 4 at 29, width 1 below;
 * 24 fusion tries from 29 qubits, 8 at 28, none below.  No reference
 * counterpart. */
int qsv_window_plan_beam(const qsv_gate *gates, int64_t ngates, const double *mats, int max_k, int exact,
                         int tb, int n, int width, int branch, int first, int delay, int fuse_tries,
                         int relabel, int64_t max_blocks, int *k_out, int *qubits_out, double *mats_out,
                         int64_t *ngates_out, int *window_out, int64_t *nblocks);

int qsv_run_circuit(qsv_state *s, const qsv_gate *gates, int64_t ngates,
                    const double *mats, const qsv_run_opts *opts,
                    int32_t *rng, int *outcomes, qsv_run_stats *stats);

/* ---- learned cost model for the fused path ----------------------------------
 * A small MLP (QSV_CM_FEATURES -> 32 -> 32 -> 1, ReLU) predicting
 * log(device ms of one fused block, scaled to 2^30 amplitudes) from the block's
 * features.  Forward, backward and Adam run as HIP kernels on the device;
 * training data are the per-block timings the engine records itself
 * (qsv_timing_log).  The scheduler (opts.scheduler = 1) plans the circuit at
 * every block size k = 2..5, predicts every block at every tile shape in one
 * batched forward pass, and keeps the plan with the least predicted time; it
 * falls back to the greedy plan at opts.fuse_k when no model is attached.
 * Plans differ only in grouping and tiling, never in the gates applied.
 * No reference counterpart. */
#define QSV_CM_FEATURES 12
typedef struct qsv_costmodel qsv_costmodel;
int qsv_cm_create(uint64_t seed, qsv_costmodel **out);
int qsv_cm_destroy(qsv_costmodel *m);
int qsv_cm_nparams(void);
int qsv_cm_get_params(qsv_costmodel *m, float *params);
int qsv_cm_set_params(qsv_costmodel *m, const float *params);
/* features of one fused block (host computation) */
int qsv_cm_block_features(int n, int kq, int split, int tile, const int *qubits, float *features);
/* batched inference: y[i] = model(x[i]) for i < batch (x: batch x QSV_CM_FEATURES) */
int qsv_cm_forward(qsv_costmodel *m, const float *x, int64_t batch, float *y);
/* gradient of 0.5 * mean((model(x) - t)^2) with respect to the parameters */
int qsv_cm_gradients(qsv_costmodel *m, const float *x, const float *t, int64_t batch, float *grad,
                     double *loss);
/* full-batch Adam for `epochs` steps; *loss = loss after the last step */
int qsv_cm_train(qsv_costmodel *m, const float *x, const float *t, int64_t batch, int epochs, float lr,
                 double *loss);
/* attach (or detach with NULL) the model used by opts.scheduler = 1 */
int qsv_set_costmodel(qsv_state *s, qsv_costmodel *m);
/* per-block records of the last profiled fused run: features (count x
 * QSV_CM_FEATURES) and measured device ms (count); *count = records available */
int qsv_timing_log(qsv_state *s, float *features, float *ms, int64_t cap, int64_t *count);

/* ---- batched variational circuits (one persistent launch) -------------------
 * count instances; instance i has nq[i] qubits (2..22) and layers[i] layers of
 * { RY(theta) then RZ(phi) on every qubit q; CNOT(q, q+1) for q = 0..n-2 } on a
 * complex64 state from |0...0>.  angles: for each instance, layer, qubit in order,
 * (theta, phi).  Output zz[i] = sum_{q=0}^{n-2} <Z_q Z_{q+1}> in FP64.
 * One persistent workgroup per CU takes instances from an atomic work counter
 * (largest first); instances of <= 14 qubits run entirely in LDS.  No reference
 * counterpart (QPanda runs one program per process). */
typedef struct qsv_batch qsv_batch;
int qsv_batch_create(const int *nq, const int *layers, int64_t count, const double *angles, qsv_batch **out);
int qsv_batch_destroy(qsv_batch *b);
/* time each workgroup size (256, 512, 1024 threads) on the batch and keep the fastest */
int qsv_batch_autotune(qsv_batch *b, int *best_threads, double *best_ms);
/* one eager launch (threads 0 = tuned); latency_us[i] = device time of instance i */
int qsv_batch_run(qsv_batch *b, int threads, double *zz, double *latency_us, double *total_ms);
/* the same launch captured once in a hipGraph and replayed `replays` times;
 * *total_ms = mean device time per replay, zz = results of the last replay */
int qsv_batch_graph_run(qsv_batch *b, int replays, double *zz, double *total_ms);

/* ---- Park-Miller "minimal standard" generator (QuantumGates.cpp:46-85) ----- */
/* Next state from `state` (1..2^31-2); the draw is state / 2147483647.0. */
int32_t qsv_pm_next(int32_t state);
/* state after k steps, O(log k) (multiplier 16807^k mod 2^31-1). */
int32_t qsv_pm_jump(int32_t state, uint64_t k);
/* The seed the reference derives from a broken-down local time
 * (QuantumGates.cpp:68-72). */
int32_t qsv_pm_seed_from_tm(int tm_year, int tm_mon, int tm_mday, int tm_hour, int tm_min, int tm_sec);

/* ---- library info --------------------------------------------------------- */
const char *qsv_version(void);
const char *qsv_last_error(void);   /* thread-local message of the last failure */
int qsv_device_count(int *count);

#ifdef __cplusplus
}
#endif
#endif /* QSV_H */
