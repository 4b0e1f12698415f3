// qsv_batch.hip -- many small circuits in one persistent launch.
//
// Workload: hardware-efficient ansatz instances, each n qubits (n <= 22) and L
// layers of { U_q = RZ(phi) RY(theta) on every qubit q; CNOT(q, q+1) for
// q = 0..n-2 }, complex64 state from |0...0>, output
// sum_{i=0}^{n-2} <Z_i Z_{i+1}> accumulated in FP64.
//
// One persistent workgroup per CU pulls instances from an atomic counter; the
// host orders instances by decreasing work, so the largest start first.
//   n <= 14: the whole state lives in LDS (2^14 x 8 B = 128 KiB); every gate is
//            an LDS sweep between barriers; HBM is never touched.
//   n >  14: the state lives in a per-workgroup HBM slice; each layer is two
//            LDS-tiled passes over it: qubits 0..13 (contiguous 2^14 tiles: the
//            U's on qubits < 14 and CNOT(q, q+1) for q <= 12), then qubits
//            13..n-1 plus the 27-n lowest (the U's on qubits >= 14 and the rest
//            of the ladder, in order).  This is the circuit's own order up to
//            commuting disjoint gates.
// The counter is reset by a one-thread kernel (no memset node), so the pair of
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

struct Inst {
    int n, layers;
    int64_t mat_off;                         // float offset of this instance's U's (layers x n x 8)
    int64_t index;                           // caller's instance index
};

__device__ __forceinline__ uint32_t ins0(uint32_t i, int r)
{
    return ((i >> r) << (r + 1)) | (i & ((1u << r) - 1));
}

// Apply U (8 floats, row-major complex) to tile bit r of an LDS tile of 2^tb.
template <int T>
__device__ __forceinline__ void lds_u(float2 *t, int tb, int r, const float *u)
{
    const uint32_t np = 1u << (tb - 1);
    for (uint32_t p = threadIdx.x; p < np; p += T) {
        const uint32_t i0 = ins0(p, r), i1 = i0 | (1u << r);
        const float2 a = t[i0], b = t[i1];
        float2 y0, y1;
        y0.x = u[0] * a.x - u[1] * a.y + u[2] * b.x - u[3] * b.y;
        y0.y = u[0] * a.y + u[1] * a.x + u[2] * b.y + u[3] * b.x;
        y1.x = u[4] * a.x - u[5] * a.y + u[6] * b.x - u[7] * b.y;
        y1.y = u[4] * a.y + u[5] * a.x + u[6] * b.y + u[7] * b.x;
        t[i0] = y0;
        t[i1] = y1;
    }
    __syncthreads();
}

// CNOT with control tile bit rc and target tile bit rt.
template <int T>
__device__ __forceinline__ void lds_cnot(float2 *t, int tb, int rc, int rt)
{
    const int lo = rc < rt ? rc : rt, hi = rc < rt ? rt : rc;
    const uint32_t nq = 1u << (tb - 2);
    for (uint32_t p = threadIdx.x; p < nq; p += T) {
        const uint32_t i = ins0(ins0(p, lo), hi) | (1u << rc);
        const float2 a = t[i], b = t[i | (1u << rt)];
        t[i] = b;
        t[i | (1u << rt)] = a;
    }
    __syncthreads();
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

__global__ void k_batch_reset(unsigned int *counter) { *counter = 0; }

template <int T>
__global__ __launch_bounds__(T) void k_batch_hea(const Inst *__restrict__ inst, int64_t count,
                                                 const float *__restrict__ mats, float2 *__restrict__ scratch,
                                                 unsigned int *__restrict__ counter, double *__restrict__ zz,
                                                 uint64_t *__restrict__ t_start, uint64_t *__restrict__ t_end)
{
    // exactly 128 KiB of LDS: the work-counter slot and the reduction scratch
    // reuse the tile while it holds no state (between instances)
    __shared__ float2 tile[1 << kLdsBits];
    unsigned int *const next_slot = reinterpret_cast<unsigned int *>(tile);
    double *const red = reinterpret_cast<double *>(tile);
    float2 *slice = scratch + (size_t)blockIdx.x * ((size_t)1 << kMaxBatchQubits);
    for (;;) {
        if (threadIdx.x == 0) *next_slot = atomicAdd(counter, 1u);
        __syncthreads();
        const unsigned int k = *next_slot;
        __syncthreads();
        if (k >= count) break;                // every wave of the workgroup sees the same k
        const Inst I = inst[k];
        const int n = I.n;
        const float *U = mats + I.mat_off;
        if (threadIdx.x == 0) t_start[I.index] = wall_clock64();
        double acc = 0;
        if (n <= kLdsBits) {
            const uint32_t dim = 1u << n;
            for (uint32_t i = threadIdx.x; i < dim; i += T) tile[i] = make_float2(i == 0 ? 1.f : 0.f, 0.f);
            __syncthreads();
            for (int l = 0; l < I.layers; l++) {
                for (int q = 0; q < n; q++) lds_u<T>(tile, n, q, U + ((size_t)l * n + q) * 8);
                for (int q = 0; q + 1 < n; q++) lds_cnot<T>(tile, n, q, q + 1);
            }
            for (uint32_t i = threadIdx.x; i < dim; i += T) {
                const double x = tile[i].x, y = tile[i].y;
                acc += (x * x + y * y) * zz_weight(i, n);
            }
        } else {
            const uint64_t dim = uint64_t(1) << n;
            for (uint64_t i = threadIdx.x; i < dim; i += T) slice[i] = make_float2(i == 0 ? 1.f : 0.f, 0.f);
            __syncthreads();
            const int c = 27 - n;                               // low bits in the high-qubit tile
            const uint32_t ts = 1u << kLdsBits;
            for (int l = 0; l < I.layers; l++) {
                const float *Ul = U + (size_t)l * n * 8;
                // pass A: qubits 0..13, contiguous tiles
                for (uint64_t base = 0; base < dim; base += ts) {
                    for (uint32_t e = threadIdx.x; e < ts; e += T) tile[e] = slice[base + e];
                    __syncthreads();
                    for (int q = 0; q < kLdsBits; q++) lds_u<T>(tile, kLdsBits, q, Ul + q * 8);
                    for (int q = 0; q + 1 < kLdsBits; q++) lds_cnot<T>(tile, kLdsBits, q, q + 1);
                    for (uint32_t e = threadIdx.x; e < ts; e += T) slice[base + e] = tile[e];
                    __syncthreads();
                }
                // pass B: tile bits 0..c-1 = qubits 0..c-1, bits c.. = qubits 13..n-1
                const uint64_t ntile = dim >> kLdsBits;
                for (uint64_t t = 0; t < ntile; t++) {
                    // t enumerates qubits c..12 (13 - c = n - 14 bits)
                    const uint64_t fixed = t << c;
                    auto gaddr = [&](uint32_t e) -> uint64_t {
                        return (uint64_t)(e & ((1u << c) - 1)) | fixed | ((uint64_t)(e >> c) << 13);
                    };
                    for (uint32_t e = threadIdx.x; e < ts; e += T) tile[e] = slice[gaddr(e)];
                    __syncthreads();
                    for (int q = kLdsBits; q < n; q++) lds_u<T>(tile, kLdsBits, c + (q - 13), Ul + q * 8);
                    for (int q = 13; q + 1 < n; q++) lds_cnot<T>(tile, kLdsBits, c + (q - 13), c + (q + 1 - 13));
                    for (uint32_t e = threadIdx.x; e < ts; e += T) slice[gaddr(e)] = tile[e];
                    __syncthreads();
                }
            }
            for (uint64_t i = threadIdx.x; i < dim; i += T) {
                const double x = slice[i].x, y = slice[i].y;
                acc += (x * x + y * y) * zz_weight(i, n);
            }
        }
        __syncthreads();                      // the tile is read; reuse it as scratch
        const double s = block_sum_d(acc, red);
        if (threadIdx.x == 0) {
            zz[I.index] = s;
            t_end[I.index] = wall_clock64();
        }
        __syncthreads();
    }
}

}  // namespace
}  // namespace qsv

struct qsv_batch {
    int64_t count = 0;
    qsv::Inst *d_inst = nullptr;
    float *d_mats = nullptr;
    float2 *d_scratch = nullptr;
    unsigned int *d_counter = nullptr;
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
    for (void *p : {(void *)b->d_inst, (void *)b->d_mats, (void *)b->d_scratch, (void *)b->d_counter,
                    (void *)b->d_zz, (void *)b->d_t0, (void *)b->d_t1})
        if (p) (void)hipFree(p);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

static int batch_launch(qsv_batch *b, int threads)
{
    hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(1), 0, b->stream, b->d_counter);
    const dim3 g(b->grid);
    switch (threads) {
    case 256: hipLaunchKernelGGL(k_batch_hea<256>, g, dim3(256), 0, b->stream, b->d_inst, b->count, b->d_mats, b->d_scratch, b->d_counter, b->d_zz, b->d_t0, b->d_t1); break;
    case 512: hipLaunchKernelGGL(k_batch_hea<512>, g, dim3(512), 0, b->stream, b->d_inst, b->count, b->d_mats, b->d_scratch, b->d_counter, b->d_zz, b->d_t0, b->d_t1); break;
    default: hipLaunchKernelGGL(k_batch_hea<1024>, g, dim3(1024), 0, b->stream, b->d_inst, b->count, b->d_mats, b->d_scratch, b->d_counter, b->d_zz, b->d_t0, b->d_t1); break;
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

static int batch_results(qsv_batch *b, double *zz, double *lat_us)
{
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
        inst[i] = {n, L, (int64_t)mats.size(), i};
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
    bool ok = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) == hipSuccess &&
              hipMalloc((void **)&b->d_inst, sizeof(Inst) * count) == hipSuccess &&
              hipMalloc((void **)&b->d_mats, sizeof(float) * std::max<size_t>(1, mats.size())) == hipSuccess &&
              hipMalloc((void **)&b->d_scratch, sizeof(float2) * ((size_t)b->grid << kMaxBatchQubits)) == hipSuccess &&
              hipMalloc((void **)&b->d_counter, sizeof(unsigned int)) == hipSuccess &&
              hipMalloc((void **)&b->d_zz, sizeof(double) * count) == hipSuccess &&
              hipMalloc((void **)&b->d_t0, sizeof(uint64_t) * count) == hipSuccess &&
              hipMalloc((void **)&b->d_t1, sizeof(uint64_t) * count) == hipSuccess;
    if (!ok) {
        batch_free(b);
        return fail(QSV_ERR_ALLOC, "batch allocation");
    }
    if (hipMemcpy(b->d_inst, order.data(), sizeof(Inst) * count, hipMemcpyHostToDevice) != hipSuccess ||
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
