// qsv_tiles.hip -- the per-gate FP64 path with runs of gates applied in one
// HBM round trip.
//
// With gate tiling on (qsv_set_gate_tiling), the gate entry points do not
// launch their sweep: the pair, phase and two-qubit-block launches are queued
// as long as every queued gate's qubits fit one tile together with the
// lowest qubits that make a 128-byte run (0..2 for complex128, 0..3 for
// complex64; tiles of 2^12 or 2^13 amplitudes: 64 KiB of LDS).
// The next entry point that is not a gate (a readout, a measurement, a state
// copy, the end of a circuit) flushes the queue as one sweep over the state's
// tiles: each workgroup loads a tile (128-byte runs of amplitudes),
// applies the queued gates to it in program order with the per-gate kernels'
// own arithmetic (qsv_gatemath.h, so the state is bit-identical to the
// one-sweep-per-gate path), and stores it.  A gate whose qubits cannot join
// the queued ones flushes first; a gate wider than a tile, a larger dense
// block, or a queue of one gate launches its own sweep.  QPanda's simulators
// sweep the state once per gate (CPUQuantumGates.cpp:92-1435,
// GPUGates.cu:129-1020).
#include "qsv_gatemath.h"
#include "qsv_stream.h"

#include <algorithm>
#include <vector>

namespace qsv {

constexpr int kMaxTileBits = 13;
constexpr int kMaxTileOps = 512;      // queued gates per sweep
constexpr int kTileThreads = 512;
constexpr int kTileMatDoubles = 32;   // one 4 x 4 complex matrix

struct TileOp {                        // one queued gate in tile coordinates
    int kind;                          // PairOp, or kPhase / kBlock
    int nins;                          // tile positions the gate's index skips
    int d;                             // block: rows (2 or 4)
    int moff;                          // block: matrix offset in the staging buffer (doubles)
    int ins[kMaxTileBits];
    uint32_t or0, or1;                 // pair members; phase mask (or0); block controls (or0)
    uint32_t off[4];                   // block: member offsets
    double c[8];                       // pair coefficients (FP32 values for complex64); phase re, im
};
constexpr int kPhase = 4, kBlock2 = 5;

struct HostOp {
    int kind;
    uint64_t insmask, or0, or1;
    double c[8];
    int k, q[2];                       // block: qubits (row bit j = bit(q[j]))
    int moff;
};

struct TileArgs {
    InsertSet tile_ins;                // the tile's qubits, ascending
    int lpos[kMaxTileBits];            // tile position -> qubit
    uint64_t hioff[16];                // pdep(NT i, tile): address part of element NT i
};

}  // namespace qsv

struct TileQueue {
    bool active = false;
    std::vector<qsv::HostOp> ops;
    uint64_t qmask = 0;                // qubits of the queued gates
    int nmats = 0;                     // matrices in the staging buffer
    qsv::TileOp *dops = nullptr;       // device copy of the ops of a sweep
    double *dmats = nullptr;           // device staging of the queued blocks' matrices
};

namespace qsv {

namespace {

// 64 KiB tiles: 2^12 complex128 or 2^13 complex64 amplitudes
int tile_bits(const qsv_state *s) { return s->prec == QSV_C128 ? 12 : 13; }

// the qubits of one 128-byte run of amplitudes: every tile holds them
uint64_t run_mask(const qsv_state *s) { return s->prec == QSV_C128 ? 0x7 : 0xF; }

// the run's qubits and the gates' qubits must fit one tile together
bool fits(const qsv_state *s, uint64_t mask) { return __builtin_popcountll(mask | run_mask(s)) <= tile_bits(s); }

__device__ __forceinline__ uint32_t tile_expand(uint32_t p, const TileOp &op)
{
    for (int j = 0; j < op.nins; j++) {
        const int pos = op.ins[j];
        const uint32_t lo = p & ((1u << pos) - 1);
        p = ((p >> pos) << (pos + 1)) | lo;
    }
    return p;
}

template <typename T, int TB>
__global__ __launch_bounds__(kTileThreads) void k_gate_tile(T *__restrict__ a, TileArgs ta,
                                                             const TileOp *__restrict__ ops, int nops,
                                                             const double *__restrict__ dmats)
{
    using R = typename Real<T>::type;
    constexpr int TS = 1 << TB, PER = TS / kTileThreads;
    static_assert(PER >= 1 && PER <= 16, "tile size and thread count");
    __shared__ T t[TS];
    const int tid = threadIdx.x;
    const uint64_t base = expand_index(xcd_block(blockIdx.x, gridDim.x), ta.tile_ins);
    uint64_t toff = 0;
    for (int j = 0; j < TB && (1 << j) < kTileThreads; j++)
        if ((tid >> j) & 1) toff |= uint64_t(1) << ta.lpos[j];
    T v[PER];
#pragma unroll
    for (int i = 0; i < PER; i++) v[i] = ld_nt(a + base + toff + ta.hioff[i]);
#pragma unroll
    for (int i = 0; i < PER; i++) t[tid + kTileThreads * i] = v[i];
    __syncthreads();
    for (int g = 0; g < nops; g++) {
        const TileOp &op = ops[g];
        const uint32_t count = (uint32_t)TS >> op.nins;
        if (op.kind == kPhase) {
            const R re = (R)op.c[0], im = (R)op.c[1];
            for (uint32_t p = tid; p < count; p += kTileThreads) {
                const uint32_t i = tile_expand(p, op) | op.or0;
                t[i] = phase_math(t[i], re, im);
            }
        } else if (op.kind == kBlock2) {
            const double *u = dmats + op.moff;
            for (uint32_t p = tid; p < count; p += kTileThreads) {
                const uint32_t b = tile_expand(p, op) | op.or0;
                double xr[4], xi[4];
#pragma unroll
                for (int c = 0; c < 4; c++)
                    if (c < op.d) {
                        const T x = t[b + op.off[c]];
                        xr[c] = x.x;
                        xi[c] = x.y;
                    }
                T y[4];
#pragma unroll
                for (int r = 0; r < 4; r++)
                    if (r < op.d) y[r] = block_row_math<T, 4>(u, op.d, r, xr, xi);
#pragma unroll
                for (int r = 0; r < 4; r++)
                    if (r < op.d) t[b + op.off[r]] = y[r];
            }
        } else {
            Coef<R> c;
#pragma unroll
            for (int k = 0; k < 8; k++) c.u[k] = (R)op.c[k];
            for (uint32_t p = tid; p < count; p += kTileThreads) {
                const uint32_t i0 = tile_expand(p, op);
                T y0, y1;
                const T x0 = t[i0 | op.or0], x1 = t[i0 | op.or1];
                switch (op.kind) {
                case OP_H: pair_math<T, OP_H>(x0, x1, c, y0, y1); break;
                case OP_X: pair_math<T, OP_X>(x0, x1, c, y0, y1); break;
                case OP_DIAG: pair_math<T, OP_DIAG>(x0, x1, c, y0, y1); break;
                default: pair_math<T, OP_U>(x0, x1, c, y0, y1); break;
                }
                t[i0 | op.or0] = y0;
                t[i0 | op.or1] = y1;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < PER; i++) st_nt(a + base + toff + ta.hioff[i], t[tid + kTileThreads * i]);
}

uint32_t pext(uint64_t x, const int *lpos, int tb)
{
    uint32_t o = 0;
    for (int j = 0; j < tb; j++)
        if (x >> lpos[j] & 1) o |= 1u << j;
    return o;
}

int launch_direct(qsv_state *s, const HostOp &h)
{
    if (h.kind == kPhase) {
        const InsertSet ins = insert_set_from_mask(h.insmask);
        return launch_phase(s, ins, h.or0, h.c[0], h.c[1]);
    }
    if (h.kind == kBlock2) return launch_block_dev(s, h.q, h.k, s->tq->dmats + h.moff, h.or0);
    const InsertSet ins = insert_set_from_mask(h.insmask);
    return launch_pair(s, h.kind, ins, h.or0, h.or1, h.c, h.c[0]);
}

int flush_now(qsv_state *s)
{
    TileQueue *q = s->tq;
    if (q->ops.empty()) return QSV_OK;
    const bool was = q->active;
    q->active = false;                          // the launchers below launch directly
    int rc = QSV_OK;
    if (q->ops.size() == 1) {
        rc = launch_direct(s, q->ops[0]);
    } else {
        const int tb = tile_bits(s);
        uint64_t lmask = q->qmask | run_mask(s);
        for (int b = 0; __builtin_popcountll(lmask) < tb; b++) lmask |= uint64_t(1) << b;
        TileArgs ta{};
        ta.tile_ins = insert_set_from_mask(lmask);
        for (int j = 0; j < tb; j++) ta.lpos[j] = ta.tile_ins.pos[j];
        auto pdep = [&](uint64_t e) {
            uint64_t o = 0;
            for (int j = 0; j < tb; j++)
                if (e >> j & 1) o |= uint64_t(1) << ta.lpos[j];
            return o;
        };
        const int per = (1 << tb) / kTileThreads;
        for (int i = 0; i < per; i++) ta.hioff[i] = pdep((uint64_t)kTileThreads * i);
        std::vector<TileOp> ops(q->ops.size());
        for (size_t g = 0; g < q->ops.size(); g++) {
            const HostOp &h = q->ops[g];
            TileOp &o = ops[g];
            o = TileOp{};
            o.kind = h.kind;
            o.nins = 0;
            for (int j = 0; j < tb; j++)
                if (h.insmask >> ta.lpos[j] & 1) o.ins[o.nins++] = j;
            o.or0 = pext(h.or0, ta.lpos, tb);
            o.or1 = pext(h.or1, ta.lpos, tb);
            for (int k = 0; k < 8; k++) o.c[k] = h.c[k];
            if (h.kind == kBlock2) {
                o.d = 1 << h.k;
                o.moff = h.moff;
                for (int r = 0; r < o.d; r++) {
                    uint64_t off = 0;
                    for (int j = 0; j < h.k; j++)
                        if (r >> j & 1) off |= uint64_t(1) << h.q[j];
                    o.off[r] = pext(off, ta.lpos, tb);
                }
            }
        }
        const hipError_t e = hipMemcpyAsync(q->dops, ops.data(), ops.size() * sizeof(TileOp), hipMemcpyHostToDevice,
                                            s->stream);
        if (e != hipSuccess) rc = hip_fail(e, "tile ops upload");
        const dim3 grid((unsigned)(s->dim >> tb)), threads(kTileThreads);
        if (!rc) {
            if (s->prec == QSV_C128)
                hipLaunchKernelGGL((k_gate_tile<c128, 12>), grid, threads, 0, s->stream, (c128 *)s->amps, ta,
                                   q->dops, (int)ops.size(), q->dmats);
            else
                hipLaunchKernelGGL((k_gate_tile<c64, 13>), grid, threads, 0, s->stream, (c64 *)s->amps, ta, q->dops,
                                   (int)ops.size(), q->dmats);
            s->launches++;
            const hipError_t le = hipGetLastError();
            if (le != hipSuccess) rc = hip_fail(le, "k_gate_tile");
        }
    }
    q->ops.clear();
    q->qmask = 0;
    q->nmats = 0;
    q->active = was;
    return rc;
}

// Queue h (its qubits: opmask) or tell the caller to launch it.
int enqueue(qsv_state *s, const HostOp &h, uint64_t opmask)
{
    TileQueue *q = s->tq;
    if (!fits(s, opmask)) {                     // wider than a tile: its own sweep
        int rc = flush_now(s);
        return rc ? rc : 1;
    }
    if (!fits(s, q->qmask | opmask) || (int)q->ops.size() >= kMaxTileOps) {
        int rc = flush_now(s);
        if (rc) return rc;
    }
    q->ops.push_back(h);
    q->qmask |= opmask;
    return 0;
}

// states smaller than a tile sweep gate by gate
bool tiling(const qsv_state *s) { return s->tq && s->tq->active && s->n >= tile_bits(s); }

}  // namespace

int tiles_enqueue_pair(qsv_state *s, int op, const InsertSet &ins, uint64_t or0, uint64_t or1, const double *u,
                       double hscale)
{
    if (!tiling(s) || (s->prec != QSV_C128 && s->prec != QSV_C64)) return 1;
    HostOp h{};
    h.kind = op;
    for (int j = 0; j < ins.n; j++) h.insmask |= uint64_t(1) << ins.pos[j];
    h.or0 = or0;
    h.or1 = or1;
    if (op == OP_H) h.c[0] = hscale;
    else if (op == OP_DIAG) for (int i = 0; i < 4; i++) h.c[i] = u[i];
    else if (op == OP_U) for (int i = 0; i < 8; i++) h.c[i] = u[i];
    if (s->prec == QSV_C64)                    // the per-gate kernel's FP32 coefficients
        for (double &c : h.c) c = (double)(float)c;
    return enqueue(s, h, h.insmask | or0 | or1);
}

int tiles_enqueue_phase(qsv_state *s, const InsertSet &ins, uint64_t orm, double re, double im)
{
    if (!tiling(s)) return 1;
    HostOp h{};
    h.kind = kPhase;
    for (int j = 0; j < ins.n; j++) h.insmask |= uint64_t(1) << ins.pos[j];
    h.or0 = orm;
    h.c[0] = s->prec == QSV_C64 ? (double)(float)re : re;
    h.c[1] = s->prec == QSV_C64 ? (double)(float)im : im;
    return enqueue(s, h, h.insmask | orm);
}

int tiles_enqueue_block(qsv_state *s, const int *qubits, int k, const double *u_dev, uint64_t ctrl_mask)
{
    if (!tiling(s)) return 1;
    if (k > 2) {                                 // larger blocks: their own sweep
        int rc = tiles_flush(s);
        return rc ? rc : 1;
    }
    HostOp h{};
    h.kind = kBlock2;
    h.k = k;
    uint64_t m = ctrl_mask;
    for (int j = 0; j < k; j++) {
        h.q[j] = qubits[j];
        m |= uint64_t(1) << qubits[j];
    }
    h.insmask = m;
    h.or0 = ctrl_mask;
    TileQueue *q = s->tq;
    if (!fits(s, m)) {
        int rc = flush_now(s);
        return rc ? rc : 1;
    }
    if (!fits(s, q->qmask | m) || (int)q->ops.size() >= kMaxTileOps) {
        int rc = flush_now(s);
        if (rc) return rc;
    }
    // the matrix travels with the queue: copied (in stream order) into staging
    h.moff = q->nmats * kTileMatDoubles;
    const size_t bytes = sizeof(double) * 2 * ((size_t)1 << (2 * k));
    QSV_HIP(hipMemcpyAsync(q->dmats + h.moff, u_dev, bytes, hipMemcpyDeviceToDevice, s->stream));
    q->nmats++;
    q->ops.push_back(h);
    q->qmask |= m;
    return 0;
}

int tiles_flush(qsv_state *s)
{
    if (!s->tq || s->tq->ops.empty()) return QSV_OK;
    return flush_now(s);
}

void tiles_discard(qsv_state *s)
{
    if (!s->tq) return;
    s->tq->ops.clear();
    s->tq->qmask = 0;
    s->tq->nmats = 0;
}

void tiles_free(qsv_state *s)
{
    if (!s->tq) return;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    if (s->tq->dops) (void)hipFree(s->tq->dops);
    if (s->tq->dmats) (void)hipFree(s->tq->dmats);
    delete s->tq;
    s->tq = nullptr;
}

TileSuspend::TileSuspend(qsv_state *s, bool on) : s_(s), was_(on && s->tq && s->tq->active)
{
    if (was_) s->tq->active = false;
}

TileSuspend::~TileSuspend()
{
    if (was_) s_->tq->active = true;
}

}  // namespace qsv

using namespace qsv;

extern "C" int qsv_set_gate_tiling(qsv_state *s, int on)
{
    if (!s || !s->amps) return fail(QSV_ERR_ARG, "null state");
    if (!on) {
        int rc = tiles_flush(s);
        if (s->tq) s->tq->active = false;
        return rc;
    }
    if (!s->tq) {
        TileQueue *q = new TileQueue();
        if (hipMalloc((void **)&q->dops, sizeof(TileOp) * kMaxTileOps) != hipSuccess ||
            hipMalloc((void **)&q->dmats, sizeof(double) * kTileMatDoubles * kMaxTileOps) != hipSuccess) {
            if (q->dops) (void)hipFree(q->dops);
            delete q;
            return fail(QSV_ERR_ALLOC, "cannot allocate the gate-tiling buffers");
        }
        s->tq = q;
    }
    s->tq->active = true;
    return QSV_OK;
}
