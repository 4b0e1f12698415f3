// qsv_gates.hip -- per-gate amplitude sweeps for gfx950.
//
// One kernel family covers every gate that mixes amplitudes in pairs: the two
// members of a pair are expand(p) | or0 and expand(p) | or1, where expand()
// inserts zero bits at the target/control positions.  That single shape serves
//   1q gates           ins = {t} u C,  or0 = C,          or1 = C | 2^t
//   CNOT / Toffoli     the same with OP_X
//   iSWAP / sqiSWAP    ins = {a, b},   or0 = 2^lo,       or1 = 2^hi
//   Fredkin            ins = {c,a,b},  or0 = 2^c | 2^a,  or1 = 2^c | 2^b
// (QPanda's CPU loops for these are CPUQuantumGates.cpp:92-1435; its CUDA
// kernels launch one thread per pair with 32-bit shifts, GPUGates.cu:129-1020.)
//
// Layout: interleaved complex (double2 = 16 B, float2 = 8 B).  Each thread
// handles kPairsPerThread pairs spaced one workgroup apart, so every load
// instruction of a wave covers 64 consecutive compact indices (contiguous
// bytes whenever the lowest inserted position is >= 6) and several loads are
// in flight per thread.  The sweeps are HBM-bound: 2 x 16 B read + written per
// pair for c128.
#include "qsv_gatemath.h"
#include "qsv_stream.h"

namespace qsv {

namespace {

constexpr int kPairsPerThread = 4;


template <typename T, int OP>
__global__ __launch_bounds__(kBlock) void k_pair(T *__restrict__ a, uint64_t npairs, InsertSet ins,
                                                  uint64_t or0, uint64_t or1,
                                                  Coef<typename Real<T>::type> c)
{
    const uint64_t base = xcd_block(blockIdx.x, gridDim.x) * (kBlock * kPairsPerThread) + threadIdx.x;
    T x0[kPairsPerThread], x1[kPairsPerThread];
    const bool dense = ins.n == 1 && or0 == 0;     // every amplitude in some pair
    uint64_t i0[kPairsPerThread];
    bool ok[kPairsPerThread];
#pragma unroll
    for (int k = 0; k < kPairsPerThread; k++) {
        const uint64_t p = base + (uint64_t)k * kBlock;
        ok[k] = p < npairs;
        i0[k] = expand_index(p, ins);
        if (ok[k]) {
            if (dense) {
                x0[k] = ld_nt(a + (i0[k] | or0));
                x1[k] = ld_nt(a + (i0[k] | or1));
            } else {
                x0[k] = a[i0[k] | or0];
                x1[k] = a[i0[k] | or1];
            }
        }
    }
#pragma unroll
    for (int k = 0; k < kPairsPerThread; k++) {
        if (!ok[k]) continue;
        T y0, y1;
        pair_math<T, OP>(x0[k], x1[k], c, y0, y1);
        if (dense) {
            st_nt(a + (i0[k] | or0), y0);
            st_nt(a + (i0[k] | or1), y1);
        } else {
            a[i0[k] | or0] = y0;
            a[i0[k] | or1] = y1;
        }
    }
}

// Multiply the amplitudes expand(p) | orm by one complex factor (CR, CZ,
// controlled phases: CPUQuantumGates.cpp:773-825 touches 2^(n-2) amplitudes).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_phase(T *__restrict__ a, uint64_t count, InsertSet ins,
                                                   uint64_t orm, typename Real<T>::type re,
                                                   typename Real<T>::type im)
{
    const uint64_t base = xcd_block(blockIdx.x, gridDim.x) * (kBlock * kPairsPerThread) + threadIdx.x;
#pragma unroll
    for (int k = 0; k < kPairsPerThread; k++) {
        const uint64_t p = base + (uint64_t)k * kBlock;
        if (p >= count) continue;
        const uint64_t i = expand_index(p, ins) | orm;
        a[i] = phase_math(a[i], re, im);
    }
}

// Dense k-qubit block in FP64 arithmetic, one amplitude group per thread.
// Matrix staged in LDS (at most 32 x 32 complex doubles = 16 KiB).
struct BlockArgs {
    uint64_t off[32];     // offset of group member r (row/column r)
    uint64_t ctrl;        // OR-ed into every group base
    int d;                // 2^k
};

template <typename T>
__global__ __launch_bounds__(kBlock) void k_block_f64(T *__restrict__ a, uint64_t groups, InsertSet ins,
                                                       BlockArgs ba, const double *__restrict__ u)
{
    __shared__ double su[32 * 32 * 2];
    const int d = ba.d;
    for (int i = threadIdx.x; i < d * d * 2; i += kBlock) su[i] = u[i];
    __syncthreads();
    // grid-stride: a launch stays below 2^32 threads at any state size
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t g = (uint64_t)blockIdx.x * kBlock + threadIdx.x; g < groups; g += stride) {
        const uint64_t b = expand_index(g, ins) | ba.ctrl;
        double xr[32], xi[32];
#pragma unroll
        for (int c = 0; c < 32; c++) {
            if (c < d) {
                T v = a[b + ba.off[c]];
                xr[c] = v.x; xi[c] = v.y;
            }
        }
        for (int r = 0; r < d; r++) a[b + ba.off[r]] = block_row_math<T, 32>(su, d, r, xr, xi);
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_fill_zero(T *__restrict__ a, uint64_t dim)
{
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < dim; i += stride) {
        T v;
        v.x = i == 0 ? 1 : 0;
        v.y = 0;
        a[i] = v;
    }
}

// Workgroups of a grid-stride launch: 2^20 x 256 threads, below HIP's 2^32
// threads per launch.
constexpr unsigned kMaxGrid = 1u << 20;

inline unsigned grid_for(uint64_t items, uint64_t per_block)
{
    uint64_t g = (items + per_block - 1) / per_block;
    return (unsigned)(g ? g : 1);
}

}  // namespace

int launch_pair(qsv_state *s, int op, const InsertSet &ins, uint64_t or0, uint64_t or1,
                const double *u, double hscale)
{
    if (s->tq) {
        const int q = tiles_enqueue_pair(s, op, ins, or0, or1, u, hscale);
        if (q <= 0) return q;                  // queued, or an error
    }
    const uint64_t npairs = s->dim >> ins.n;
    const unsigned grid = grid_for(npairs, (uint64_t)kBlock * kPairsPerThread);
    if (s->prec == QSV_C128) {
        Coef<double> c{};
        if (op == OP_H) c.u[0] = hscale;
        else if (op == OP_DIAG) for (int i = 0; i < 4; i++) c.u[i] = u[i];
        else if (op == OP_U) for (int i = 0; i < 8; i++) c.u[i] = u[i];
        c128 *a = (c128 *)s->amps;
        switch (op) {
        case OP_H: hipLaunchKernelGGL((k_pair<c128, OP_H>), dim3(grid), dim3(kBlock), 0, s->stream, a, npairs, ins, or0, or1, c); break;
        case OP_X: hipLaunchKernelGGL((k_pair<c128, OP_X>), dim3(grid), dim3(kBlock), 0, s->stream, a, npairs, ins, or0, or1, c); break;
        case OP_DIAG: hipLaunchKernelGGL((k_pair<c128, OP_DIAG>), dim3(grid), dim3(kBlock), 0, s->stream, a, npairs, ins, or0, or1, c); break;
        default: hipLaunchKernelGGL((k_pair<c128, OP_U>), dim3(grid), dim3(kBlock), 0, s->stream, a, npairs, ins, or0, or1, c); break;
        }
    } else {
        Coef<float> c{};
        if (op == OP_H) c.u[0] = (float)hscale;
        else if (op == OP_DIAG) for (int i = 0; i < 4; i++) c.u[i] = (float)u[i];
        else if (op == OP_U) for (int i = 0; i < 8; i++) c.u[i] = (float)u[i];
        c64 *a = (c64 *)s->amps;
        switch (op) {
        case OP_H: hipLaunchKernelGGL((k_pair<c64, OP_H>), dim3(grid), dim3(kBlock), 0, s->stream, a, npairs, ins, or0, or1, c); break;
        case OP_X: hipLaunchKernelGGL((k_pair<c64, OP_X>), dim3(grid), dim3(kBlock), 0, s->stream, a, npairs, ins, or0, or1, c); break;
        case OP_DIAG: hipLaunchKernelGGL((k_pair<c64, OP_DIAG>), dim3(grid), dim3(kBlock), 0, s->stream, a, npairs, ins, or0, or1, c); break;
        default: hipLaunchKernelGGL((k_pair<c64, OP_U>), dim3(grid), dim3(kBlock), 0, s->stream, a, npairs, ins, or0, or1, c); break;
        }
    }
    s->launches++;
    QSV_HIP(hipGetLastError());
    return QSV_OK;
}

int launch_phase(qsv_state *s, const InsertSet &ins, uint64_t orm, double re, double im)
{
    if (s->tq) {
        const int q = tiles_enqueue_phase(s, ins, orm, re, im);
        if (q <= 0) return q;
    }
    const uint64_t count = s->dim >> ins.n;
    const unsigned grid = grid_for(count, (uint64_t)kBlock * kPairsPerThread);
    if (s->prec == QSV_C128)
        hipLaunchKernelGGL(k_phase<c128>, dim3(grid), dim3(kBlock), 0, s->stream, (c128 *)s->amps, count, ins, orm, re, im);
    else
        hipLaunchKernelGGL(k_phase<c64>, dim3(grid), dim3(kBlock), 0, s->stream, (c64 *)s->amps, count, ins, orm, (float)re, (float)im);
    s->launches++;
    QSV_HIP(hipGetLastError());
    return QSV_OK;
}

int launch_block(qsv_state *s, const int *qubits, int k, const double *u, uint64_t ctrl_mask)
{
    // matrix to device through the reduction scratch (2 * 32 * 32 doubles fit)
    const size_t mbytes = sizeof(double) * 2 * ((size_t)1 << (2 * k));
    if (k < 1 || k > 5) return fail(QSV_ERR_ARG, "block size must be 1..5");
    QSV_HIP(hipMemcpyAsync(s->red, u, mbytes, hipMemcpyHostToDevice, s->stream));
    return launch_block_dev(s, qubits, k, s->red, ctrl_mask);
}

int launch_block_dev(qsv_state *s, const int *qubits, int k, const double *u_dev, uint64_t ctrl_mask)
{
    if (k < 1 || k > 5) return fail(QSV_ERR_ARG, "block size must be 1..5");
    if (s->tq) {
        const int q = tiles_enqueue_block(s, qubits, k, u_dev, ctrl_mask);
        if (q <= 0) return q;
    }
    BlockArgs ba{};
    ba.d = 1 << k;
    ba.ctrl = ctrl_mask;
    uint64_t mask = ctrl_mask;
    for (int j = 0; j < k; j++) mask |= uint64_t(1) << qubits[j];
    for (int r = 0; r < ba.d; r++) {
        uint64_t o = 0;
        for (int j = 0; j < k; j++)
            if (r >> j & 1) o |= uint64_t(1) << qubits[j];
        ba.off[r] = o;
    }
    InsertSet ins = insert_set_from_mask(mask);
    const uint64_t groups = s->dim >> ins.n;
    const unsigned grid = std::min(grid_for(groups, kBlock), kMaxGrid);
    if (s->prec == QSV_C128)
        hipLaunchKernelGGL(k_block_f64<c128>, dim3(grid), dim3(kBlock), 0, s->stream, (c128 *)s->amps, groups, ins, ba, u_dev);
    else
        hipLaunchKernelGGL(k_block_f64<c64>, dim3(grid), dim3(kBlock), 0, s->stream, (c64 *)s->amps, groups, ins, ba, u_dev);
    s->launches++;
    QSV_HIP(hipGetLastError());
    return QSV_OK;
}

int launch_fill_zero(qsv_state *s)
{
    uint64_t blocks = (s->dim + kBlock * 8 - 1) / (kBlock * 8);
    if (blocks > 65536) blocks = 65536;
    if (blocks == 0) blocks = 1;
    if (s->prec == QSV_C128)
        hipLaunchKernelGGL(k_fill_zero<c128>, dim3((unsigned)blocks), dim3(kBlock), 0, s->stream, (c128 *)s->amps, s->dim);
    else
        hipLaunchKernelGGL(k_fill_zero<c64>, dim3((unsigned)blocks), dim3(kBlock), 0, s->stream, (c64 *)s->amps, s->dim);
    s->launches++;
    QSV_HIP(hipGetLastError());
    return QSV_OK;
}

}  // namespace qsv
