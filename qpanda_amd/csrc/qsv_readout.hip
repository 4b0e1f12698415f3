// qsv_readout.hip -- FP64 reductions, collapse and marginal distributions.
//
// All sums are deterministic: the partition of the index range into
// workgroups, each thread's stride, the wave butterfly (xor shuffles) and the
// final pass are fixed functions of the state size, so a readout is
// bit-identical from run to run.  Reductions accumulate in FP64 for both
// precisions.  QPanda's CUDA counterpart is a shared-memory tree per block
// followed by one thread summing the partials serially
// (GPUGates.cu:597-682); its CPU simulator sums serially in index order
// (CPUQuantumGates.cpp:1077-1083), which serial_prob0 reproduces on the host
// (serial_prob0_host, qsv_io.cpp) for the measurement tie band.
#include "qsv_internal.h"
#include "qsv_stream.h"

#include <algorithm>
#include <cstring>
#include <mutex>
#include <cmath>

namespace qsv {

namespace {

enum RedMode { RED_NORM = 0, RED_PROB0 = 1, RED_ZZ = 2 };

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Block sum with a fixed order: wave butterflies, then wave 0 adds the four
// wave totals in index order.
__device__ __forceinline__ double block_sum(double v, double *lds)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    double t = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < kBlock / 64; w++) t += lds[w];
    return t;
}

template <typename T>
__device__ __forceinline__ double norm_of(T v)
{
    const double x = v.x, y = v.y;
    return x * x + y * y;
}

template <typename T, int MODE>
__global__ __launch_bounds__(kBlock) void k_reduce(const T *__restrict__ a, uint64_t count, uint64_t chunk,
                                                    InsertSet ins, int nq, double *__restrict__ partial)
{
    __shared__ double lds[kBlock / 64];
    // chunk c of the partition, wherever its workgroup runs (XCD mapping;
    // the sums and their order do not change)
    const uint64_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint64_t start = blk * chunk;
    uint64_t end = start + chunk;
    if (end > count) end = count;
    const uint64_t zmask = nq > 1 ? ((uint64_t(1) << (nq - 1)) - 1) : 0;
    double acc = 0;
    for (uint64_t i = start + threadIdx.x; i < end; i += kBlock) {
        if (MODE == RED_PROB0) {
            acc += norm_of(a[expand_index(i, ins)]);
        } else if (MODE == RED_ZZ) {
            const int flips = __popcll((i ^ (i >> 1)) & zmask);
            acc += norm_of(ld_nt(a + i)) * (double)(nq - 1 - 2 * flips);
        } else {
            acc += norm_of(ld_nt(a + i));
        }
    }
    const double t = block_sum(acc, lds);
    if (threadIdx.x == 0) partial[blk] = t;
}

__global__ __launch_bounds__(kBlock) void k_reduce_final(const double *__restrict__ partial, int n,
                                                          double *__restrict__ out)
{
    __shared__ double lds[kBlock / 64];
    double acc = 0;
    for (int i = threadIdx.x; i < n; i += kBlock) acc += partial[i];
    const double t = block_sum(acc, lds);
    if (threadIdx.x == 0) *out = t;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_collapse(T *__restrict__ a, uint64_t npairs, InsertSet ins,
                                                      uint64_t bit, int outcome, double scale)
{
    // grid-stride: a launch stays below 2^32 threads at any state size
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t p = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x; p < npairs; p += stride) {
        const uint64_t i0 = expand_index(p, ins), i1 = i0 | bit;
        const uint64_t keep = outcome ? i1 : i0, drop = outcome ? i0 : i1;
        using R = decltype(T{}.x);
        T v = ld_nt(a + keep), z;
        z.x = 0; z.y = 0;
        v.x = (R)((double)v.x * scale);
        v.y = (R)((double)v.y * scale);
        st_nt(a + keep, v);
        st_nt(a + drop, z);
    }
}

// <a|b> over two states of possibly different precisions: per-block partial
// (re, im) pairs in a fixed order, then the same fixed final pass.
template <typename TA, typename TB>
__global__ __launch_bounds__(kBlock) void k_inner(const TA *__restrict__ a, const TB *__restrict__ b,
                                                   uint64_t count, uint64_t chunk, double *__restrict__ partial)
{
    __shared__ double lds[kBlock / 64];
    const uint64_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint64_t start = blk * chunk;
    uint64_t end = start + chunk;
    if (end > count) end = count;
    double re = 0, im = 0;
    for (uint64_t i = start + threadIdx.x; i < end; i += kBlock) {
        const TA va = ld_nt(a + i);
        const TB vb = ld_nt(b + i);
        const double ar = va.x, ai = va.y, br = vb.x, bi = vb.y;
        re += ar * br + ai * bi;          // conj(a) * b
        im += ar * bi - ai * br;
    }
    const double tr = block_sum(re, lds);
    __syncthreads();
    const double ti = block_sum(im, lds);
    if (threadIdx.x == 0) {
        partial[2 * blk] = tr;
        partial[2 * blk + 1] = ti;
    }
}

__global__ __launch_bounds__(kBlock) void k_inner_final(const double *__restrict__ partial, int n,
                                                         double *__restrict__ out)
{
    __shared__ double lds[kBlock / 64];
    double re = 0, im = 0;
    for (int i = threadIdx.x; i < n; i += kBlock) {
        re += partial[2 * i];
        im += partial[2 * i + 1];
    }
    const double tr = block_sum(re, lds);
    __syncthreads();
    const double ti = block_sum(im, lds);
    if (threadIdx.x == 0) {
        out[0] = tr;
        out[1] = ti;
    }
}

template <typename TD, typename TS>
__global__ __launch_bounds__(kBlock) void k_copy(TD *__restrict__ d, const TS *__restrict__ s, uint64_t count)
{
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    using R = decltype(TD{}.x);
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < count; i += stride) {
        const TS v = s[i];
        TD w;
        w.x = (R)v.x;
        w.y = (R)v.y;
        d[i] = w;
    }
}

// ---- marginals --------------------------------------------------------------
constexpr int kTileBits = 12;
// Marginals over more qubits use k_marg_direct (the tile path's final pass
// launches 2^m workgroups of kBlock threads, within the grid limits only for
// small m).
constexpr int kMargTileMaxQubits = 16;

// Stage 1: one tile of 2^tb consecutive amplitudes per workgroup; probabilities
// into LDS, then pairwise sums over every unmeasured tile bit in ascending bit
// order.  Writes 2^ml sums (ml = measured bits inside the tile).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_marg_tile(const T *__restrict__ a, int tb, uint64_t meas_low_mask,
                                                       int ml, InsertSet low_pos, double *__restrict__ partial)
{
    __shared__ double p[1 << kTileBits];
    const int tsize = 1 << tb;
    const uint64_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint64_t base = blk << tb;
    for (int i = threadIdx.x; i < tsize; i += kBlock) p[i] = norm_of(ld_nt(a + base + i));
    __syncthreads();
    int done = 0;
    for (int u = 0; u < tb; u++) {
        if (meas_low_mask >> u & 1) continue;
        const int ub = 1 << u;
        for (int i = threadIdx.x; i < tsize; i += kBlock)
            if ((i & (done | ub)) == 0) p[i] += p[i | ub];
        done |= ub;
        __syncthreads();
    }
    const int nb = 1 << ml;
    for (int lb = threadIdx.x; lb < nb; lb += kBlock) {
        int idx = 0;
        for (int r = 0; r < ml; r++)
            if (lb >> r & 1) idx |= 1 << low_pos.pos[r];
        partial[blk * nb + lb] = p[idx];
    }
}

struct MargMap {
    int m;
    int ml;                  // measured bits inside a tile
    int low_rank[24];        // caller bit j -> rank among low positions, or -1
    int high_shift[24];      // caller bit j -> tile-index bit, for high positions
};

// Stage 2: output o sums the tile partials of every tile whose measured high
// bits match o, in ascending tile order (strided per thread, then the fixed
// block sum).
__global__ __launch_bounds__(kBlock) void k_marg_final(const double *__restrict__ partial, MargMap mm,
                                                        InsertSet high_ins, uint64_t ntiles_each,
                                                        double *__restrict__ out)
{
    __shared__ double lds[kBlock / 64];
    const uint64_t o = blockIdx.x;
    int lb = 0;
    uint64_t hp = 0;
    for (int j = 0; j < mm.m; j++) {
        const uint64_t v = (o >> j) & 1;
        if (mm.low_rank[j] >= 0) lb |= (int)v << mm.low_rank[j];
        else hp |= v << mm.high_shift[j];
    }
    const int nb = 1 << mm.ml;
    double acc = 0;
    for (uint64_t t = threadIdx.x; t < ntiles_each; t += kBlock) {
        const uint64_t tile = expand_index(t, high_ins) | hp;
        acc += partial[tile * nb + lb];
    }
    const double s = block_sum(acc, lds);
    if (threadIdx.x == 0) out[o] = s;
}

// Marginals over many qubits (m > 16, up to m = n): one thread per outcome
// t, whose measured bits (in ascending qubit order) are t's bits; it sums
// |a|^2 over the 2^(n-m) settings u of the unmeasured qubits in ascending
// order (u steps through the subsets of the unmeasured mask) and writes the
// sum at the caller-order index (bit j = bit of qubits[j]).  QPanda's CUDA
// back end has its own branch for large marginals (GPUGates.cu:2033-2102).
struct CallerMap {
    int m;
    int pos[QSV_MAX_QUBITS];     // k-th smallest measured qubit -> caller bit j
};

template <typename T>
__global__ __launch_bounds__(kBlock) void k_marg_direct(const T *__restrict__ a, uint64_t nout, InsertSet unmeas_ins,
                                                         uint64_t unmeas_mask, CallerMap cm, double *__restrict__ out)
{
    const uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= nout) return;
    const uint64_t base = expand_index(t, unmeas_ins);
    double acc = 0;
    uint64_t u = 0;
    do {
        acc += norm_of(a[base | u]);
        u = ((u | ~unmeas_mask) + 1) & unmeas_mask;
    } while (u != 0);
    uint64_t o = 0;
    for (int k = 0; k < cm.m; k++) o |= ((t >> k) & 1) << cm.pos[k];
    out[o] = acc;
}

// ---- top 10 of a distribution on the device (pMeasure's view) ---------------
// Order: probability descending, ties by index ascending.  (The reference
// sorts with std::sort, whose order among equal probabilities is unspecified,
// CPUQuantumGates.cpp:178-187.)  Each thread keeps a sorted list of its 10
// best in registers (insertion with constant indices); a workgroup then draws
// its 10 best by ten rounds of argmax over the threads' list heads.
constexpr int kTopK = 10;

__device__ __forceinline__ bool topk_better(double a, uint64_t ia, double b, uint64_t ib)
{
    return a > b || (a == b && ia < ib);
}

__device__ __forceinline__ void topk_init(double (&v)[kTopK], uint64_t (&ix)[kTopK])
{
#pragma unroll
    for (int k = 0; k < kTopK; k++) {
        v[k] = -1.0;                 // below every probability
        ix[k] = ~uint64_t(0);
    }
}

__device__ __forceinline__ void topk_insert(double (&v)[kTopK], uint64_t (&ix)[kTopK], double x, uint64_t xi)
{
    if (!topk_better(x, xi, v[kTopK - 1], ix[kTopK - 1])) return;
    v[kTopK - 1] = x;
    ix[kTopK - 1] = xi;
#pragma unroll
    for (int k = kTopK - 1; k > 0; k--) {
        if (topk_better(v[k], ix[k], v[k - 1], ix[k - 1])) {
            const double tv = v[k];
            v[k] = v[k - 1];
            v[k - 1] = tv;
            const uint64_t ti = ix[k];
            ix[k] = ix[k - 1];
            ix[k - 1] = ti;
        }
    }
}

// The workgroup's 10 best of all threads' lists into ov / oi (thread 0 writes).
__device__ __forceinline__ void topk_block(double (&v)[kTopK], uint64_t (&ix)[kTopK], double *ov, uint64_t *oi)
{
    __shared__ double sv[kBlock / 64];
    __shared__ uint64_t si[kBlock / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int r = 0; r < kTopK; r++) {
        double hv = v[0];
        uint64_t hi = ix[0];
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            const double ov2 = __shfl_xor(hv, off, 64);
            const uint64_t oi2 = __shfl_xor(hi, off, 64);
            if (topk_better(ov2, oi2, hv, hi)) {
                hv = ov2;
                hi = oi2;
            }
        }
        if (lane == 0) {
            sv[wave] = hv;
            si[wave] = hi;
        }
        __syncthreads();
        double bv = sv[0];
        uint64_t bi = si[0];
        for (int w = 1; w < kBlock / 64; w++)
            if (topk_better(sv[w], si[w], bv, bi)) {
                bv = sv[w];
                bi = si[w];
            }
        __syncthreads();
        if (threadIdx.x == 0) {
            ov[r] = bv;
            oi[r] = bi;
        }
        if (ix[0] == bi && v[0] == bv) {             // this thread's head won: pop it
#pragma unroll
            for (int k = 0; k + 1 < kTopK; k++) {
                v[k] = v[k + 1];
                ix[k] = ix[k + 1];
            }
            v[kTopK - 1] = -1.0;
            ix[kTopK - 1] = ~uint64_t(0);
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_topk_partial(const double *__restrict__ vals, uint64_t count,
                                                          double *__restrict__ pv, uint64_t *__restrict__ pi)
{
    double v[kTopK];
    uint64_t ix[kTopK];
    topk_init(v, ix);
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < count; i += stride) topk_insert(v, ix, vals[i], i);
    topk_block(v, ix, pv + (uint64_t)blockIdx.x * kTopK, pi + (uint64_t)blockIdx.x * kTopK);
}

__global__ __launch_bounds__(kBlock) void k_topk_final(const double *__restrict__ pv, const uint64_t *__restrict__ pi,
                                                        int ncand, double *__restrict__ ov, uint64_t *__restrict__ oi)
{
    double v[kTopK];
    uint64_t ix[kTopK];
    topk_init(v, ix);
    for (int i = threadIdx.x; i < ncand; i += kBlock) topk_insert(v, ix, pv[i], pi[i]);
    topk_block(v, ix, ov, oi);
}

// Partition of a reduction: nblk <= kRedParts workgroups of `chunk`
// consecutive (compact) indices each, a fixed function of the size.
constexpr uint64_t kRedParts = 4096;

struct RedPlan {
    InsertSet ins;
    uint64_t count, chunk, nblk;
};

RedPlan reduce_plan(int n, int mode, int q)
{
    RedPlan r{};
    r.count = uint64_t(1) << n;
    if (mode == RED_PROB0) {
        r.ins.n = 1;
        r.ins.pos[0] = q;
        r.count >>= 1;
    }
    uint64_t nblk = (r.count + (uint64_t)kBlock * 16 - 1) / ((uint64_t)kBlock * 16);
    if (nblk > kRedParts) nblk = kRedParts;
    if (nblk == 0) nblk = 1;
    uint64_t chunk = (r.count + nblk - 1) / nblk;
    chunk = (chunk + kBlock - 1) / kBlock * kBlock;
    r.chunk = chunk;
    r.nblk = (r.count + chunk - 1) / chunk;
    return r;
}

// Partial sums into partial[0 .. nblk), their sum into *out (device), on st.
template <typename T>
void reduce_launch_t(const T *a, int n, int mode, const RedPlan &r, double *partial, double *out, hipStream_t st)
{
    const dim3 g((unsigned)r.nblk), t(kBlock);
    if (mode == RED_PROB0) hipLaunchKernelGGL((k_reduce<T, RED_PROB0>), g, t, 0, st, a, r.count, r.chunk, r.ins, n, partial);
    else if (mode == RED_ZZ) hipLaunchKernelGGL((k_reduce<T, RED_ZZ>), g, t, 0, st, a, r.count, r.chunk, r.ins, n, partial);
    else hipLaunchKernelGGL((k_reduce<T, RED_NORM>), g, t, 0, st, a, r.count, r.chunk, r.ins, n, partial);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kBlock), 0, st, partial, (int)r.nblk, out);
}

void reduce_launch(const void *amps, int prec, int n, int mode, int q, double *partial, double *out, hipStream_t st)
{
    const RedPlan r = reduce_plan(n, mode, q);
    if (prec == QSV_C128) reduce_launch_t((const c128 *)amps, n, mode, r, partial, out, st);
    else reduce_launch_t((const c64 *)amps, n, mode, r, partial, out, st);
}

int reduce_common(qsv_state *s, int mode, int q, double *result)
{
    double *out = s->red + (kRedSlots - 1);
    reduce_launch(s->amps, s->prec, s->n, mode, q, s->red, out, s->stream);
    s->launches += 2;
    QSV_HIP(hipGetLastError());
    QSV_HIP(hipMemcpyAsync(s->host_red, out, sizeof(double), hipMemcpyDeviceToHost, s->stream));
    QSV_HIP(hipStreamSynchronize(s->stream));
    *result = s->host_red[0];
    return QSV_OK;
}

static_assert(kRedParts < (uint64_t)kRedSlots, "state scratch holds the partials and the result");

// ---- allocation-free readout -------------------------------------------------
// Partial sums of qsv_readout_noalloc: one library-owned array per GPU (each
// device holds its own copy until the process ends).  Host code never uses the
// array's name as an address: its device address comes from
// hipGetSymbolAddress, once per device, with that device current.
static __device__ __align__(256) double scratch[kRedParts];

std::mutex g_scratch_mu;
double *g_scratch_ptr[64] = {};

int scratch_on_current_device(int *dev_out, double **ptr)
{
    int dev = 0;
    QSV_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail(QSV_ERR_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[dev] = (double *)p;
    }
    *dev_out = dev;
    *ptr = g_scratch_ptr[dev];
    return QSV_OK;
}

}  // namespace

int reduce_prob0(qsv_state *s, int q, double *p0) { return reduce_common(s, RED_PROB0, q, p0); }
int reduce_norm2(qsv_state *s, double *v) { return reduce_common(s, RED_NORM, 0, v); }
int reduce_zz_chain(qsv_state *s, double *v) { return reduce_common(s, RED_ZZ, 0, v); }

}  // namespace qsv

extern "C" int qsv_readout_noalloc(const void *amps, int n, int precision, int what, int q, double *out,
                                   void *stream)
{
    using namespace qsv;
    if (!amps || !out) return fail(QSV_ERR_ARG, "null amplitude or output pointer");
    if (n < 1 || n > QSV_MAX_QUBITS) return fail(QSV_ERR_ARG, "qubit count must be 1..34");
    if (precision != QSV_C128 && precision != QSV_C64) return fail(QSV_ERR_ARG, "precision must be QSV_C128 or QSV_C64");
    int mode;
    switch (what) {
    case QSV_READ_NORM2: mode = RED_NORM; break;
    case QSV_READ_PROB0: mode = RED_PROB0; break;
    case QSV_READ_ZZ_CHAIN: mode = RED_ZZ; break;
    default: return fail(QSV_ERR_ARG, "what must be QSV_READ_NORM2, QSV_READ_PROB0 or QSV_READ_ZZ_CHAIN");
    }
    if (mode == RED_PROB0 && (q < 0 || q >= n)) return fail(QSV_ERR_QUBIT, "qubit out of range");
    const RedPlan r = reduce_plan(n, mode, q);
    if (r.nblk > kRedParts) return fail(QSV_ERR_ARG, "reduction needs more partial sums than the scratch holds");
    int dev = 0;
    double *partial = nullptr;
    int rc = scratch_on_current_device(&dev, &partial);
    if (rc) return rc;
    // the amplitudes and the result must live on the current device
    hipPointerAttribute_t pa{}, po{};
    if (hipPointerGetAttributes(&pa, amps) != hipSuccess || hipPointerGetAttributes(&po, out) != hipSuccess) {
        (void)hipGetLastError();
        return fail(QSV_ERR_ARG, "amplitudes and result must be device memory");
    }
    if (pa.type != hipMemoryTypeDevice || po.type != hipMemoryTypeDevice || pa.device != dev || po.device != dev)
        return fail(QSV_ERR_ARG, "amplitudes and result must be device memory of the current device");
    reduce_launch(amps, precision, n, mode, q, partial, out, (hipStream_t)stream);
    QSV_HIP(hipGetLastError());
    return QSV_OK;
}

namespace qsv {

int inner(qsv_state *a, qsv_state *b, double *re, double *im)
{
    const uint64_t count = a->dim;
    uint64_t nblk = (count + (uint64_t)kBlock * 16 - 1) / ((uint64_t)kBlock * 16);
    if (nblk > 2048) nblk = 2048;
    if (nblk == 0) nblk = 1;
    uint64_t chunk = (count + nblk - 1) / nblk;
    chunk = (chunk + kBlock - 1) / kBlock * kBlock;
    nblk = (count + chunk - 1) / chunk;
    QSV_HIP(hipStreamSynchronize(b->stream));           // b's pending gates first
    double *partial = a->red;
    double *out = a->red + (kRedSlots - 4);
    const dim3 g((unsigned)nblk), t(kBlock);
    if (a->prec == QSV_C128 && b->prec == QSV_C128)
        hipLaunchKernelGGL((k_inner<c128, c128>), g, t, 0, a->stream, (const c128 *)a->amps, (const c128 *)b->amps, count, chunk, partial);
    else if (a->prec == QSV_C128)
        hipLaunchKernelGGL((k_inner<c128, c64>), g, t, 0, a->stream, (const c128 *)a->amps, (const c64 *)b->amps, count, chunk, partial);
    else if (b->prec == QSV_C128)
        hipLaunchKernelGGL((k_inner<c64, c128>), g, t, 0, a->stream, (const c64 *)a->amps, (const c128 *)b->amps, count, chunk, partial);
    else
        hipLaunchKernelGGL((k_inner<c64, c64>), g, t, 0, a->stream, (const c64 *)a->amps, (const c64 *)b->amps, count, chunk, partial);
    hipLaunchKernelGGL(k_inner_final, dim3(1), t, 0, a->stream, partial, (int)nblk, out);
    a->launches += 2;
    QSV_HIP(hipGetLastError());
    QSV_HIP(hipMemcpyAsync(a->host_red, out, 2 * sizeof(double), hipMemcpyDeviceToHost, a->stream));
    QSV_HIP(hipStreamSynchronize(a->stream));
    *re = a->host_red[0];
    *im = a->host_red[1];
    return QSV_OK;
}

int copy_state(qsv_state *d, qsv_state *s)
{
    QSV_HIP(hipStreamSynchronize(s->stream));
    uint64_t blocks = (d->dim + kBlock * 8 - 1) / (kBlock * 8);
    if (blocks > 65536) blocks = 65536;
    if (blocks == 0) blocks = 1;
    const dim3 g((unsigned)blocks), t(kBlock);
    if (d->prec == QSV_C128 && s->prec == QSV_C128)
        hipLaunchKernelGGL((k_copy<c128, c128>), g, t, 0, d->stream, (c128 *)d->amps, (const c128 *)s->amps, d->dim);
    else if (d->prec == QSV_C128)
        hipLaunchKernelGGL((k_copy<c128, c64>), g, t, 0, d->stream, (c128 *)d->amps, (const c64 *)s->amps, d->dim);
    else if (s->prec == QSV_C128)
        hipLaunchKernelGGL((k_copy<c64, c128>), g, t, 0, d->stream, (c64 *)d->amps, (const c128 *)s->amps, d->dim);
    else
        hipLaunchKernelGGL((k_copy<c64, c64>), g, t, 0, d->stream, (c64 *)d->amps, (const c64 *)s->amps, d->dim);
    d->launches++;
    QSV_HIP(hipGetLastError());
    QSV_HIP(hipStreamSynchronize(d->stream));
    return QSV_OK;
}

int serial_prob0(qsv_state *s, int q, double *p0)
{
    QSV_HIP(hipStreamSynchronize(s->stream));        // the gates before the measurement
    return serial_prob0_host(s, q, p0);
}

int launch_collapse(qsv_state *s, int q, int outcome, double scale)
{
    InsertSet ins{};
    ins.n = 1;
    ins.pos[0] = q;
    const uint64_t npairs = s->dim >> 1;
    const unsigned grid = (unsigned)std::min<uint64_t>((npairs + kBlock - 1) / kBlock, uint64_t(1) << 20);
    if (s->prec == QSV_C128)
        hipLaunchKernelGGL(k_collapse<c128>, dim3(grid), dim3(kBlock), 0, s->stream, (c128 *)s->amps, npairs, ins, uint64_t(1) << q, outcome, scale);
    else
        hipLaunchKernelGGL(k_collapse<c64>, dim3(grid), dim3(kBlock), 0, s->stream, (c64 *)s->amps, npairs, ins, uint64_t(1) << q, outcome, scale);
    s->launches++;
    QSV_HIP(hipGetLastError());
    return QSV_OK;
}

// Full marginal distribution over qubits[0..m-1] into device memory: the
// 2^m doubles at *dout, inside s->aux, followed by `extra` free bytes.
static int marginals_dev(qsv_state *s, const int *qubits, int m, size_t extra, double **dout)
{
    const size_t out_bytes = sizeof(double) * ((size_t)1 << m);
    if (m > kMargTileMaxQubits) {
        uint64_t meas = 0;
        for (int j = 0; j < m; j++) meas |= uint64_t(1) << qubits[j];
        const uint64_t unmeas = (s->dim - 1) & ~meas;
        CallerMap cm{};
        cm.m = m;
        int k = 0;
        for (int q = 0; q < s->n; q++)
            if (meas >> q & 1) {
                for (int j = 0; j < m; j++)
                    if (qubits[j] == q) cm.pos[k] = j;
                k++;
            }
        int rc = ensure_aux(s, out_bytes + extra);
        if (rc) return rc;
        double *out = (double *)s->aux;
        const uint64_t nout = uint64_t(1) << m;
        const InsertSet ins = insert_set_from_mask(unmeas);
        const dim3 g((unsigned)((nout + kBlock - 1) / kBlock)), t(kBlock);
        if (s->prec == QSV_C128)
            hipLaunchKernelGGL(k_marg_direct<c128>, g, t, 0, s->stream, (const c128 *)s->amps, nout, ins, unmeas, cm, out);
        else
            hipLaunchKernelGGL(k_marg_direct<c64>, g, t, 0, s->stream, (const c64 *)s->amps, nout, ins, unmeas, cm, out);
        s->launches++;
        QSV_HIP(hipGetLastError());
        *dout = out;
        return QSV_OK;
    }
    const int tb = std::min(s->n, kTileBits);
    uint64_t low_mask = 0, high_tile_mask = 0;
    MargMap mm{};
    mm.m = m;
    for (int j = 0; j < m; j++) {
        if (qubits[j] < tb) low_mask |= uint64_t(1) << qubits[j];
        else high_tile_mask |= uint64_t(1) << (qubits[j] - tb);
    }
    InsertSet low_pos = insert_set_from_mask(low_mask);
    mm.ml = low_pos.n;
    for (int j = 0; j < m; j++) {
        mm.low_rank[j] = -1;
        mm.high_shift[j] = 0;
        if (qubits[j] < tb) {
            for (int r = 0; r < low_pos.n; r++)
                if (low_pos.pos[r] == qubits[j]) mm.low_rank[j] = r;
        } else {
            mm.high_shift[j] = qubits[j] - tb;
        }
    }
    const uint64_t ntiles = s->dim >> tb;
    const size_t partial_bytes = sizeof(double) * (size_t)ntiles * ((size_t)1 << mm.ml);
    int rc = ensure_aux(s, partial_bytes + out_bytes + extra);
    if (rc) return rc;
    double *partial = (double *)((char *)s->aux + out_bytes + extra);
    double *out = (double *)s->aux;
    if (s->prec == QSV_C128)
        hipLaunchKernelGGL(k_marg_tile<c128>, dim3((unsigned)ntiles), dim3(kBlock), 0, s->stream, (const c128 *)s->amps, tb, low_mask, mm.ml, low_pos, partial);
    else
        hipLaunchKernelGGL(k_marg_tile<c64>, dim3((unsigned)ntiles), dim3(kBlock), 0, s->stream, (const c64 *)s->amps, tb, low_mask, mm.ml, low_pos, partial);
    InsertSet high_ins = insert_set_from_mask(high_tile_mask);
    const uint64_t each = ntiles >> high_ins.n;
    hipLaunchKernelGGL(k_marg_final, dim3((unsigned)(uint64_t(1) << m)), dim3(kBlock), 0, s->stream, partial, mm, high_ins, each, out);
    s->launches += 2;
    QSV_HIP(hipGetLastError());
    *dout = out;
    return QSV_OK;
}

int marginals(qsv_state *s, const int *qubits, int m, double *out)
{
    double *dout = nullptr;
    int rc = marginals_dev(s, qubits, m, 0, &dout);
    if (rc) return rc;
    QSV_HIP(hipMemcpyAsync(out, dout, sizeof(double) * ((size_t)1 << m), hipMemcpyDeviceToHost, s->stream));
    QSV_HIP(hipStreamSynchronize(s->stream));
    return QSV_OK;
}

// pMeasure's view (CPUQuantumGates.cpp:159-190): the 10 most probable outcomes
// of the marginal distribution, found on the device (no host copy or sort of
// the 2^m probabilities).
int pmeasure_top(qsv_state *s, const int *qubits, int m, uint64_t *idx, double *prob, int *count)
{
    const uint64_t nv = uint64_t(1) << m;
    const unsigned nblk = (unsigned)std::min<uint64_t>(1024, (nv + kBlock - 1) / kBlock);
    const size_t cand = (size_t)nblk * kTopK;
    // after the distribution: partial candidates (values, indices), then the result
    const size_t extra = cand * 16 + kTopK * 16 + 256;
    double *dv = nullptr;
    int rc = marginals_dev(s, qubits, m, extra, &dv);
    if (rc) return rc;
    char *x = (char *)dv + sizeof(double) * nv;
    x = (char *)(((uintptr_t)x + 255) & ~(uintptr_t)255);
    double *pv = (double *)x;
    uint64_t *pi = (uint64_t *)(pv + cand);
    double *ov = (double *)(pi + cand);
    uint64_t *oi = (uint64_t *)(ov + kTopK);
    hipLaunchKernelGGL(k_topk_partial, dim3(nblk), dim3(kBlock), 0, s->stream, dv, nv, pv, pi);
    hipLaunchKernelGGL(k_topk_final, dim3(1), dim3(kBlock), 0, s->stream, pv, pi, (int)cand, ov, oi);
    s->launches += 2;
    QSV_HIP(hipGetLastError());
    static_assert(2 * kTopK <= 64, "pinned mirror holds the result");
    QSV_HIP(hipMemcpyAsync(s->host_red, ov, kTopK * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    QSV_HIP(hipMemcpyAsync(s->host_red + kTopK, oi, kTopK * sizeof(uint64_t), hipMemcpyDeviceToHost, s->stream));
    QSV_HIP(hipStreamSynchronize(s->stream));
    const int keep = (int)std::min<uint64_t>(kTopK, nv);
    for (int i = 0; i < keep; i++) {
        prob[i] = s->host_red[i];
        std::memcpy(&idx[i], &s->host_red[kTopK + i], sizeof(uint64_t));
    }
    *count = keep;
    return QSV_OK;
}

}  // namespace qsv
