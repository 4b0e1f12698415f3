// qsv_batch.hip -- many small circuits in one persistent launch.
//
// Workload: hardware-efficient ansatz instances, each n qubits (n <= 22) and L
// layers of { U_q = RZ(phi) RY(theta) on every qubit q; CNOT(q, q+1) for
// q = 0..n-2 }, complex64 state from |0...0>, output
// sum_{i=0}^{n-2} <Z_i Z_{i+1}> accumulated in FP64.
//
// One persistent workgroup per CU pulls instances from an atomic counter; the
// host orders instances by decreasing work, so the largest start first.
//   n <= 14: the whole state lives in LDS (2^14 x 8 B = 128 KiB); a layer is
//            ceil(n / 4) LDS sweeps (rotations in groups of four qubits, the
//            CNOT ladder folded into the last sweep's write-back); HBM is never
//            touched.
//   n >  14: the state lives in a per-workgroup HBM slice; each layer is two
//            LDS-tiled passes over it: qubits 0..13 (contiguous 2^14 tiles: the
//            U's on qubits < 14 and CNOT(q, q+1) for q <= 12), then qubits
//            13..n-1 plus the 27-n lowest (the U's on qubits >= 14 and the rest
//            of the ladder, in order).  This is the circuit's own order up to
//            commuting disjoint gates.  The first pass of layer 0 starts from
//            |0...0> without reading it and the last pass of the last layer
//            reduces instead of storing: (4 L - 2) x 8 x 2^n bytes of HBM
//            traffic per instance.
// Instances of 20..22 qubits are split over 2, 4 or 8 workgroups ("parts",
// consecutive work items): every part takes an equal share of the tiles of
// each pass, the parts meet at a per-instance barrier (an arrival counter in
// HBM) after every pass, and each adds its share of the ZZ sum to a per-part
// slot that part 0 sums in part order.  The state lives in part 0's slice.
// Parts are handed out in item order, so at most one instance is ever partly
// handed out: every other part that waits at a barrier waits for parts that
// are already running, and the waits end (bounded spins that raise an error
// flag as a safety net).
// The counters are reset by a small kernel (no memset node), so the pair of
// launches can be captured in a hipGraph and replayed.
#include "qsv_internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <numeric>
#include <vector>

namespace qsv {

namespace {

constexpr int kLdsBits = 14;                 // 2^14 complex64 = 128 KiB
constexpr int kMaxBatchQubits = 22;

constexpr int kSplitMinQubits = 20;          // instances from here on run as several parts
constexpr int kMaxParts = 8;
constexpr unsigned kSpinLimit = 1u << 22;    // barrier polls before the error flag (~1-10 s)

struct Inst {
    int n, layers;
    int64_t mat_off;                         // float offset of this instance's U's (layers x n x 8)
    int64_t index;                           // caller's instance index
    int parts;                               // workgroups sharing the instance
};

struct Item {                                // one work item: a part of an instance
    int inst;                                // index into the (sorted) instance array
    int part;
};

struct BatchSync {
    const Item *items;
    int64_t nitems;
    unsigned int *counter;                   // next work item
    unsigned int *arrive;                    // per instance: barrier arrivals
    unsigned int *owner;                     // per instance: 1 + the workgroup whose slice holds it
    double *partial;                         // per instance: kMaxParts partial ZZ sums
    unsigned int *err;                       // set when a bounded wait gave up
};


__device__ __forceinline__ uint32_t ladder_perm(uint32_t x, int l0, uint32_t lmask)
{
    // CNOT(b, b + 1) for b = l0 .. l1 - 1 in order maps x to y with bits
    // l0..l1 of y the prefix XOR of those of x
    uint32_t f = (x >> l0) & lmask;
    f ^= f << 1;
    f ^= f << 2;
    f ^= f << 4;
    f ^= f << 8;
    f &= lmask;
    return (x & ~(lmask << l0)) | (f << l0);
}

// One LDS sweep applying the rotations of the G contiguous tile bits gb ..
// gb + G - 1 (matrices u[i]) to every group of 2^G amplitudes.  LAST: all
// reads happen before all writes (a barrier between), and amplitude x is
// written to ladder_perm(x) -- the CNOT ladder folded into the write-back;
// a thread then holds 2^tb / T <= 2^kLdsBits / T amplitudes in registers.
template <int T, int G, bool LAST>
__device__ __forceinline__ void group_sweep(float2 *t, int tb, int gb, const float (&u)[4][8], int l0, uint32_t lmask)
{
    constexpr int D = 1 << G;
    constexpr int ITER = LAST ? (1 << kLdsBits) / T / D : 1;   // groups per thread, compile-time when LAST
    const uint32_t ngroups = 1u << (tb - G);
    float2 keep[LAST ? ITER * D : 1];
    auto apply = [&](uint32_t p, float2 (&v)[D]) {
        const uint32_t base = (p & ((1u << gb) - 1)) | ((p >> gb) << (gb + G));
#pragma unroll
        for (int j = 0; j < D; j++) v[j] = t[base | ((uint32_t)j << gb)];
#pragma unroll
        for (int i = 0; i < G; i++)
#pragma unroll
            for (int j = 0; j < D; j++) {
                if (j >> i & 1) continue;
                const float2 a = v[j], b = v[j | (1 << i)];
                const float *m = u[i];
                v[j].x = m[0] * a.x - m[1] * a.y + m[2] * b.x - m[3] * b.y;
                v[j].y = m[0] * a.y + m[1] * a.x + m[2] * b.y + m[3] * b.x;
                v[j | (1 << i)].x = m[4] * a.x - m[5] * a.y + m[6] * b.x - m[7] * b.y;
                v[j | (1 << i)].y = m[4] * a.y + m[5] * a.x + m[6] * b.y + m[7] * b.x;
            }
        return base;
    };
    if (!LAST) {
        for (uint32_t p = threadIdx.x; p < ngroups; p += T) {
            float2 v[D];
            const uint32_t base = apply(p, v);
#pragma unroll
            for (int j = 0; j < D; j++) t[base | ((uint32_t)j << gb)] = v[j];
        }
        __syncthreads();
        return;
    }
#pragma unroll
    for (int k = 0; k < ITER; k++) {
        const uint32_t p = threadIdx.x + (uint32_t)k * T;
        if (p < ngroups) {
            float2 v[D];
            apply(p, v);
#pragma unroll
            for (int j = 0; j < D; j++) keep[LAST ? k * D + j : 0] = v[j];
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < ITER; k++) {
        const uint32_t p = threadIdx.x + (uint32_t)k * T;
        if (p < ngroups) {
            const uint32_t base = (p & ((1u << gb) - 1)) | ((p >> gb) << (gb + G));
#pragma unroll
            for (int j = 0; j < D; j++) t[ladder_perm(base | ((uint32_t)j << gb), l0, lmask)] = keep[LAST ? k * D + j : 0];
        }
    }
    __syncthreads();
}

template <int T, bool LAST>
__device__ __forceinline__ void group_sweep_g(float2 *t, int tb, int gb, int g, const float (&u)[4][8], int l0,
                                              uint32_t lmask)
{
    switch (g) {
    case 1: group_sweep<T, 1, LAST>(t, tb, gb, u, l0, lmask); break;
    case 2: group_sweep<T, 2, LAST>(t, tb, gb, u, l0, lmask); break;
    case 3: group_sweep<T, 3, LAST>(t, tb, gb, u, l0, lmask); break;
    default: group_sweep<T, 4, LAST>(t, tb, gb, u, l0, lmask); break;
    }
}

// One layer's share on an LDS tile of 2^tb amplitudes: the rotations on the
// contiguous tile bits [ub0, ub1) -- U for tile bit b at ubase + 8 (b + qoff)
// -- then the CNOT ladder CNOT(b, b + 1), b = l0 .. l1 - 1, in order.  The
// rotations commute (disjoint qubits), so they go in groups of up to four
// bits, one LDS sweep per group instead of one per gate (the first group takes
// the remainder, so the last is as wide as possible); the ladder is folded
// into the last group's write-back.
template <int T>
__device__ __forceinline__ void tile_layer(float2 *t, int tb, int ub0, int ub1, const float *ubase, int qoff, int l0,
                                           int l1)
{
    const uint32_t lmask = l1 > l0 ? ((1u << (l1 - l0 + 1)) - 1) : 0;
    const int nu = ub1 - ub0;
    int gb = ub0;
    int g = nu % 4 ? nu % 4 : 4;
    while (gb < ub1) {
        float u[4][8];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int k = 0; k < 8; k++) u[i][k] = i < g ? ubase[8 * (gb + i + qoff) + k] : 0.f;
        if (gb + g >= ub1) group_sweep_g<T, true>(t, tb, gb, g, u, l0, lmask);
        else group_sweep_g<T, false>(t, tb, gb, g, u, l0, lmask);
        gb += g;
        g = 4;
    }
}

__device__ __forceinline__ double block_sum_d(double v, double *red)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double s = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x >> 6); i++) s += red[i];
    __syncthreads();
    return s;
}

__device__ __forceinline__ double zz_weight(uint64_t x, int n)
{
    // __popcll returns unsigned: convert before subtracting so that negative
    // weights stay negative
    const uint64_t m = (n > 1) ? ((uint64_t(1) << (n - 1)) - 1) : 0;
    const int flips = (int)__popcll((x ^ (x >> 1)) & m);
    return (double)(n - 1 - 2 * flips);
}

__global__ void k_batch_reset(BatchSync bs, int64_t count)
{
    for (int64_t i = threadIdx.x; i < count; i += blockDim.x) {
        bs.arrive[i] = 0;
        bs.owner[i] = 0;
    }
    if (threadIdx.x == 0) {
        *bs.counter = 0;
        *bs.err = 0;
    }
}

// Thread 0 polls *word until it reaches target (acquire, device scope); the
// workgroup waits for it.  Gives up after kSpinLimit polls and raises err.
__device__ __forceinline__ unsigned int wait_at_least(unsigned int *word, unsigned int target, unsigned int *err)
{
    unsigned int v = __hip_atomic_load(word, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    for (unsigned int spins = 0; v < target; spins++) {
        if (spins >= kSpinLimit) {
            atomicOr(err, 1u);
            break;
        }
        __builtin_amdgcn_s_sleep(8);
        v = __hip_atomic_load(word, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    }
    return v;
}

// Barrier of the parts of one instance: this workgroup's global writes
// complete and are released at device scope before it arrives; it proceeds
// once all parts have arrived `round` times.
__device__ __forceinline__ void parts_barrier(unsigned int *arrive, unsigned int parts, unsigned int round,
                                              unsigned int *err)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        atomicAdd(arrive, 1u);
        wait_at_least(arrive, parts * round, err);
        __threadfence();
    }
    __syncthreads();
}

template <int T>
__global__ __launch_bounds__(T) void k_batch_hea(const Inst *__restrict__ inst, BatchSync bs,
                                                 const float *__restrict__ mats, float2 *__restrict__ scratch,
                                                 double *__restrict__ zz, uint64_t *__restrict__ t_start,
                                                 uint64_t *__restrict__ t_end)
{
    // exactly 128 KiB of LDS: the work-counter slot and the reduction scratch
    // reuse the tile while it holds no state (between instances)
    __shared__ __align__(16) float2 tile[1 << kLdsBits];
    unsigned int *const next_slot = reinterpret_cast<unsigned int *>(tile);
    double *const red = reinterpret_cast<double *>(tile);
    for (;;) {
        if (threadIdx.x == 0) *next_slot = atomicAdd(bs.counter, 1u);
        __syncthreads();
        const unsigned int k = *next_slot;
        __syncthreads();
        if (k >= bs.nitems) break;            // every wave of the workgroup sees the same k
        const Item it = bs.items[k];
        const Inst I = inst[it.inst];
        const int n = I.n;
        const unsigned int parts = (unsigned int)I.parts;
        const float *U = mats + I.mat_off;
        if (threadIdx.x == 0 && it.part == 0) t_start[I.index] = wall_clock64();
        double acc = 0;
        unsigned int round = 0;               // barriers passed (split instances)
        if (n <= kLdsBits) {
            const uint32_t dim = 1u << n;
            for (uint32_t i = threadIdx.x; i < dim; i += T) tile[i] = make_float2(i == 0 ? 1.f : 0.f, 0.f);
            __syncthreads();
            for (int l = 0; l < I.layers; l++) tile_layer<T>(tile, n, 0, n, U + (size_t)l * n * 8, 0, 0, n - 1);
            for (uint32_t i = threadIdx.x; i < dim; i += T) {
                const double x = tile[i].x, y = tile[i].y;
                acc += (x * x + y * y) * zz_weight(i, n);
            }
        } else {
            // HBM slice; each tile moves as 16-byte accesses, all of a thread's
            // loads issued before its LDS writes (and LDS reads before stores).
            // Layer 0's first pass starts from |0...0> without reading it, and
            // the last pass accumulates the ZZ sum instead of storing.
            const uint64_t dim = uint64_t(1) << n;
            constexpr uint32_t TS = 1u << kLdsBits;
            constexpr int NL = (int)TS / (2 * T);          // 16-byte accesses per thread per tile
            const int c = 27 - n;                           // low bits in the high-qubit tile
            // the slice: this workgroup's own, or part 0's for a split instance
            unsigned int owner = blockIdx.x;
            if (parts > 1) {
                if (threadIdx.x == 0) {
                    if (it.part == 0)
                        __hip_atomic_store(&bs.owner[it.inst], blockIdx.x + 1, __ATOMIC_RELEASE,
                                           __HIP_MEMORY_SCOPE_AGENT);
                    else
                        *next_slot = wait_at_least(&bs.owner[it.inst], 1u, bs.err) - 1;
                }
                __syncthreads();
                if (it.part != 0) owner = *next_slot;
                __syncthreads();
                if (owner >= gridDim.x) owner = blockIdx.x;     // only after a failed wait
            }
            float2 *slice = scratch + (size_t)owner * ((size_t)1 << kMaxBatchQubits);
            // this part's share of each pass's tiles
            const uint64_t share = (dim >> kLdsBits) / parts;
            const uint64_t t_lo = share * (uint64_t)it.part, t_hi = t_lo + share;
            if (I.layers == 0 && threadIdx.x == 0) acc = zz_weight(0, n);
            for (int l = 0; l < I.layers; l++) {
                const float *Ul = U + (size_t)l * n * 8;
                const bool last = l + 1 == I.layers;
                // pass A: qubits 0..13, contiguous tiles
                for (uint64_t base = t_lo * TS; base < t_hi * TS; base += TS) {
                    if (l == 0) {
                        for (uint32_t e = threadIdx.x; e < TS; e += T)
                            tile[e] = make_float2(base == 0 && e == 0 ? 1.f : 0.f, 0.f);
                    } else {
                        float4 buf[NL];
#pragma unroll
                        for (int i = 0; i < NL; i++)
                            buf[i] = *reinterpret_cast<const float4 *>(slice + base + 2 * (threadIdx.x + T * i));
#pragma unroll
                        for (int i = 0; i < NL; i++)
                            *reinterpret_cast<float4 *>(tile + 2 * (threadIdx.x + T * i)) = buf[i];
                    }
                    __syncthreads();
                    tile_layer<T>(tile, kLdsBits, 0, kLdsBits, Ul, 0, 0, kLdsBits - 1);
                    float4 out[NL];
#pragma unroll
                    for (int i = 0; i < NL; i++) out[i] = *reinterpret_cast<const float4 *>(tile + 2 * (threadIdx.x + T * i));
#pragma unroll
                    for (int i = 0; i < NL; i++)
                        *reinterpret_cast<float4 *>(slice + base + 2 * (threadIdx.x + T * i)) = out[i];
                    __syncthreads();
                }
                if (parts > 1) parts_barrier(&bs.arrive[it.inst], parts, ++round, bs.err);
                // pass B: tile bits 0..c-1 = qubits 0..c-1, bits c.. = qubits 13..n-1;
                // c >= 5, so tile elements e, e + 1 (e even) are adjacent in the slice
                for (uint64_t t = t_lo; t < t_hi; t++) {
                    // t enumerates qubits c..12 (13 - c = n - 14 bits)
                    const uint64_t fixed = t << c;
                    auto gaddr = [&](uint32_t e) -> uint64_t {
                        return (uint64_t)(e & ((1u << c) - 1)) | fixed | ((uint64_t)(e >> c) << 13);
                    };
                    {
                        float4 buf[NL];
#pragma unroll
                        for (int i = 0; i < NL; i++)
                            buf[i] = *reinterpret_cast<const float4 *>(slice + gaddr(2 * (threadIdx.x + T * i)));
#pragma unroll
                        for (int i = 0; i < NL; i++)
                            *reinterpret_cast<float4 *>(tile + 2 * (threadIdx.x + T * i)) = buf[i];
                    }
                    __syncthreads();
                    // qubit q at tile bit c + q - 13: rotations on qubits 14..n-1, ladder 13..n-1
                    tile_layer<T>(tile, kLdsBits, c + 1, c + n - 13, Ul, 13 - c, c, c + n - 14);
                    float4 out[NL];
#pragma unroll
                    for (int i = 0; i < NL; i++) out[i] = *reinterpret_cast<const float4 *>(tile + 2 * (threadIdx.x + T * i));
                    if (last) {
#pragma unroll
                        for (int i = 0; i < NL; i++) {
                            const uint64_t g = gaddr(2 * (threadIdx.x + T * i));
                            const double x0 = out[i].x, y0 = out[i].y, x1 = out[i].z, y1 = out[i].w;
                            acc += (x0 * x0 + y0 * y0) * zz_weight(g, n) + (x1 * x1 + y1 * y1) * zz_weight(g + 1, n);
                        }
                    } else {
#pragma unroll
                        for (int i = 0; i < NL; i++)
                            *reinterpret_cast<float4 *>(slice + gaddr(2 * (threadIdx.x + T * i))) = out[i];
                    }
                    __syncthreads();
                }
                if (parts > 1 && !last) parts_barrier(&bs.arrive[it.inst], parts, ++round, bs.err);
            }
        }
        __syncthreads();                      // the tile is read; reuse it as scratch
        const double s = block_sum_d(acc, red);
        if (parts == 1) {
            if (threadIdx.x == 0) {
                zz[I.index] = s;
                t_end[I.index] = wall_clock64();
            }
        } else {
            // every part's share, then part 0 sums them in part order (the
            // last barrier also keeps part 0's slice until all parts are done)
            if (threadIdx.x == 0) bs.partial[(size_t)it.inst * kMaxParts + it.part] = s;
            parts_barrier(&bs.arrive[it.inst], parts, ++round, bs.err);
            if (threadIdx.x == 0 && it.part == 0) {
                double z = 0;
                for (unsigned int j = 0; j < parts; j++) z += bs.partial[(size_t)it.inst * kMaxParts + j];
                zz[I.index] = z;
                t_end[I.index] = wall_clock64();
            }
        }
        __syncthreads();
    }
}

}  // namespace
}  // namespace qsv

struct qsv_batch {
    int64_t count = 0;
    qsv::Inst *d_inst = nullptr;
    qsv::Item *d_items = nullptr;
    int64_t nitems = 0;
    float *d_mats = nullptr;
    float2 *d_scratch = nullptr;
    unsigned int *d_counter = nullptr;       // [0] next item, [1] error flag, then arrivals and owners
    double *d_partial = nullptr;
    double *d_zz = nullptr;
    uint64_t *d_t0 = nullptr, *d_t1 = nullptr;
    int grid = 0;
    int threads = 0;                 // 0 = not tuned
    hipStream_t stream = nullptr;
    hipGraphExec_t graph = nullptr;
    int graph_threads = 0;
    double clock_mhz = 100.0;
};

namespace qsv {

static void batch_free(qsv_batch *b)
{
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    for (void *p : {(void *)b->d_inst, (void *)b->d_items, (void *)b->d_mats, (void *)b->d_scratch,
                    (void *)b->d_counter, (void *)b->d_partial, (void *)b->d_zz, (void *)b->d_t0, (void *)b->d_t1})
        if (p) (void)hipFree(p);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

static BatchSync batch_sync(qsv_batch *b)
{
    BatchSync bs;
    bs.items = b->d_items;
    bs.nitems = b->nitems;
    bs.counter = b->d_counter;
    bs.err = b->d_counter + 1;
    bs.arrive = b->d_counter + 2;
    bs.owner = b->d_counter + 2 + b->count;
    bs.partial = b->d_partial;
    return bs;
}

static int batch_launch(qsv_batch *b, int threads)
{
    const BatchSync bs = batch_sync(b);
    hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(256), 0, b->stream, bs, b->count);
    const dim3 g(b->grid);
    switch (threads) {
    case 256: hipLaunchKernelGGL(k_batch_hea<256>, g, dim3(256), 0, b->stream, b->d_inst, bs, b->d_mats, b->d_scratch, b->d_zz, b->d_t0, b->d_t1); break;
    case 512: hipLaunchKernelGGL(k_batch_hea<512>, g, dim3(512), 0, b->stream, b->d_inst, bs, b->d_mats, b->d_scratch, b->d_zz, b->d_t0, b->d_t1); break;
    default: hipLaunchKernelGGL(k_batch_hea<1024>, g, dim3(1024), 0, b->stream, b->d_inst, bs, b->d_mats, b->d_scratch, b->d_zz, b->d_t0, b->d_t1); break;
    }
    QSV_HIP(hipGetLastError());
    return QSV_OK;
}

static int batch_timed(qsv_batch *b, int threads, double *ms)
{
    hipEvent_t e0, e1;
    QSV_HIP(hipEventCreate(&e0));
    QSV_HIP(hipEventCreate(&e1));
    QSV_HIP(hipEventRecord(e0, b->stream));
    int rc = batch_launch(b, threads);
    if (!rc) QSV_HIP(hipEventRecord(e1, b->stream));
    if (!rc) QSV_HIP(hipEventSynchronize(e1));
    float f = 0;
    if (!rc) QSV_HIP(hipEventElapsedTime(&f, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = f;
    return rc;
}

// A bounded wait that gave up leaves the results undefined: report it.
static int batch_check(qsv_batch *b)
{
    unsigned int err = 0;
    QSV_HIP(hipMemcpyAsync(&err, b->d_counter + 1, sizeof(err), hipMemcpyDeviceToHost, b->stream));
    QSV_HIP(hipStreamSynchronize(b->stream));
    if (err) return fail(QSV_ERR_DEVICE, "batch: a part of a split instance waited too long at its barrier");
    return QSV_OK;
}

static int batch_results(qsv_batch *b, double *zz, double *lat_us)
{
    int rc = batch_check(b);
    if (rc) return rc;
    if (zz) QSV_HIP(hipMemcpyAsync(zz, b->d_zz, sizeof(double) * b->count, hipMemcpyDeviceToHost, b->stream));
    std::vector<uint64_t> t0, t1;
    if (lat_us) {
        t0.resize(b->count);
        t1.resize(b->count);
        QSV_HIP(hipMemcpyAsync(t0.data(), b->d_t0, sizeof(uint64_t) * b->count, hipMemcpyDeviceToHost, b->stream));
        QSV_HIP(hipMemcpyAsync(t1.data(), b->d_t1, sizeof(uint64_t) * b->count, hipMemcpyDeviceToHost, b->stream));
    }
    QSV_HIP(hipStreamSynchronize(b->stream));
    if (lat_us)
        for (int64_t i = 0; i < b->count; i++) lat_us[i] = (double)(t1[i] - t0[i]) / b->clock_mhz;
    return QSV_OK;
}

}  // namespace qsv

using namespace qsv;

extern "C" {

int qsv_batch_create(const int *nq, const int *layers, int64_t count, const double *angles, qsv_batch **out)
{
    if (!nq || !layers || !angles || !out || count < 1) return fail(QSV_ERR_ARG, "bad batch arguments");
    *out = nullptr;
    // host: combined RZ(phi) RY(theta) per (instance, layer, qubit) in the CPU
    // simulator's conventions, rounded to FP32
    std::vector<Inst> inst(count);
    std::vector<float> mats;
    int64_t aoff = 0;
    for (int64_t i = 0; i < count; i++) {
        const int n = nq[i], L = layers[i];
        if (n < 2 || n > kMaxBatchQubits || L < 0) return fail(QSV_ERR_ARG, "instance qubits must be 2..22");
        inst[i] = {n, L, (int64_t)mats.size(), i, 1};
        for (int l = 0; l < L; l++)
            for (int q = 0; q < n; q++) {
                const double th = angles[aoff++], ph = angles[aoff++];
                const double c = std::cos(th / 2), s = std::sin(th / 2);
                const double zr = std::cos(ph / 2), zi = std::sin(ph / 2);
                // RY = [[c, -s], [s, c]]; RZ = diag(e^{-i ph/2}, e^{i ph/2}); U = RZ * RY
                const double u[8] = {zr * c, -zi * c, -zr * s, zi * s, zr * s, zi * s, zr * c, zi * c};
                for (double v : u) mats.push_back((float)v);
            }
    }
    // largest work first: 2^n * layers
    std::vector<Inst> order = inst;
    std::stable_sort(order.begin(), order.end(), [](const Inst &a, const Inst &b) {
        return ((double)(1ull << a.n) * (a.layers + 1)) > ((double)(1ull << b.n) * (b.layers + 1));
    });
    qsv_batch *b = new qsv_batch();
    b->count = count;
    int dev = 0, cus = 0, clk_khz = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
        batch_free(b);
        return fail(QSV_ERR_DEVICE, "no HIP device");
    }
    if (hipDeviceGetAttribute(&clk_khz, hipDeviceAttributeWallClockRate, dev) == hipSuccess && clk_khz > 0)
        b->clock_mhz = clk_khz / 1000.0;
    b->grid = cus;                            // 128 KiB of LDS per workgroup: one per CU
    // split the largest instances: 2^(n - 19) parts from 20 qubits on, each
    // with 32 tiles of every pass
    std::vector<Item> items;
    for (int64_t i = 0; i < count; i++) {
        Inst &I = order[i];
        if (I.n >= kSplitMinQubits && I.layers > 0)
            I.parts = std::min({1 << (I.n - kSplitMinQubits + 1), kMaxParts, std::max(1, cus)});
        for (int p = 0; p < I.parts; p++) items.push_back({(int)i, p});
    }
    b->nitems = (int64_t)items.size();
    bool ok = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) == hipSuccess &&
              hipMalloc((void **)&b->d_inst, sizeof(Inst) * count) == hipSuccess &&
              hipMalloc((void **)&b->d_mats, sizeof(float) * std::max<size_t>(1, mats.size())) == hipSuccess &&
              hipMalloc((void **)&b->d_scratch, sizeof(float2) * ((size_t)b->grid << kMaxBatchQubits)) == hipSuccess &&
              hipMalloc((void **)&b->d_items, sizeof(Item) * items.size()) == hipSuccess &&
              hipMalloc((void **)&b->d_counter, sizeof(unsigned int) * (2 + 2 * (size_t)count)) == hipSuccess &&
              hipMalloc((void **)&b->d_partial, sizeof(double) * kMaxParts * (size_t)count) == hipSuccess &&
              hipMalloc((void **)&b->d_zz, sizeof(double) * count) == hipSuccess &&
              hipMalloc((void **)&b->d_t0, sizeof(uint64_t) * count) == hipSuccess &&
              hipMalloc((void **)&b->d_t1, sizeof(uint64_t) * count) == hipSuccess;
    if (!ok) {
        batch_free(b);
        return fail(QSV_ERR_ALLOC, "batch allocation");
    }
    if (hipMemcpy(b->d_inst, order.data(), sizeof(Inst) * count, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->d_items, items.data(), sizeof(Item) * items.size(), hipMemcpyHostToDevice) != hipSuccess ||
        (!mats.empty() && hipMemcpy(b->d_mats, mats.data(), sizeof(float) * mats.size(), hipMemcpyHostToDevice) != hipSuccess)) {
        batch_free(b);
        return fail(QSV_ERR_DEVICE, "batch upload");
    }
    *out = b;
    return QSV_OK;
}

int qsv_batch_destroy(qsv_batch *b)
{
    batch_free(b);
    return QSV_OK;
}

int qsv_batch_autotune(qsv_batch *b, int *best_threads, double *best_ms)
{
    if (!b) return fail(QSV_ERR_ARG, "null batch");
    const int cand[3] = {256, 512, 1024};
    double best = 1e300;
    int bt = 512;
    for (int t : cand) {
        double ms = 0, m2 = 0;
        int rc = batch_timed(b, t, &ms);       // warm
        if (!rc) rc = batch_timed(b, t, &m2);
        if (!rc) rc = batch_check(b);
        if (rc) return rc;
        if (m2 < best) {
            best = m2;
            bt = t;
        }
    }
    b->threads = bt;
    if (best_threads) *best_threads = bt;
    if (best_ms) *best_ms = best;
    return QSV_OK;
}

int qsv_batch_run(qsv_batch *b, int threads, double *zz, double *latency_us, double *total_ms)
{
    if (!b) return fail(QSV_ERR_ARG, "null batch");
    if (threads == 0) threads = b->threads ? b->threads : 512;
    if (threads != 256 && threads != 512 && threads != 1024) return fail(QSV_ERR_ARG, "threads must be 256/512/1024");
    double ms = 0;
    int rc = batch_timed(b, threads, &ms);
    if (rc) return rc;
    if (total_ms) *total_ms = ms;
    return batch_results(b, zz, latency_us);
}

int qsv_batch_graph_run(qsv_batch *b, int replays, double *zz, double *total_ms)
{
    if (!b || replays < 1) return fail(QSV_ERR_ARG, "bad arguments");
    const int threads = b->threads ? b->threads : 512;
    if (!b->graph || b->graph_threads != threads) {
        if (b->graph) {
            (void)hipGraphExecDestroy(b->graph);
            b->graph = nullptr;
        }
        hipGraph_t g;
        QSV_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        int rc = batch_launch(b, threads);
        hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (rc) return rc;
        if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
        e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
        b->graph_threads = threads;
    }
    hipEvent_t e0, e1;
    QSV_HIP(hipEventCreate(&e0));
    QSV_HIP(hipEventCreate(&e1));
    QSV_HIP(hipEventRecord(e0, b->stream));
    for (int r = 0; r < replays; r++) QSV_HIP(hipGraphLaunch(b->graph, b->stream));
    QSV_HIP(hipEventRecord(e1, b->stream));
    QSV_HIP(hipEventSynchronize(e1));
    float f = 0;
    QSV_HIP(hipEventElapsedTime(&f, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (total_ms) *total_ms = f / replays;
    return batch_results(b, zz, nullptr);
}

}  // extern "C"
