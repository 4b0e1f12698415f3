// qsv_sample.hip -- measurement shots from the final state without a state copy
// per shot.
//
// The shot loop of QPandaSDK::run (QPandaAPI.cpp:131-211) re-runs the program
// for every shot and measures qubits q_0 .. q_{m-1} in order; each measurement
// (CPUQuantumGates.cpp:1071-1126) computes p0 = sum over the current state of
// |a|^2 where q_j = 0, decides outcome = (float)r > p0 and rescales the kept
// half by 1/sqrt(p0) or 1/sqrt(1 - p0).  In terms of the pre-measurement state
// a (the same every shot) that is the recurrence
//     p0_j = S(prefix, q_j = 0) / C_j,   C_0 = 1,
//     C_{j+1} = C_j * (outcome_j ? 1 - p0_j : p0_j),
// where S(prefix, ...) sums |a_x|^2 over the amplitudes consistent with the
// outcomes so far.  The sampler evaluates exactly that recurrence:
//   top    levels j < T: S comes from a prefix-sum tree over the FP64 marginal
//          of q_0 .. q_{T-1} (one pass over the state, then pairwise sums);
//          one thread per shot walks it;
//   rest   levels j >= T: one workgroup per shot gathers |a_x|^2 of the 2^(n-T)
//          amplitudes under its T-qubit prefix into LDS and reduces level by
//          level (deterministic order).
// S differs from the per-shot simulation's sums only by rounding (~1e-16
// relative), so a decision is reproduced exactly unless (float)r lies within
// the measurement tie band of p0; such shots are flagged and re-run by the
// per-shot copy path, which makes the outcomes identical to it.
#include "qsv_internal.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

namespace qsv {

namespace {

constexpr int kRunBits = 6;          // contiguous amplitudes per run in the marginal pass
constexpr int kSubBits = 13;         // at most 2^13 amplitudes per shot below the tree (64 KiB)
constexpr int kTopFull = 16;         // m <= 16: the whole walk is in the tree

template <typename T>
__device__ __forceinline__ double prob_of(const T &v)
{
    const double x = v.x, y = v.y;
    return x * x + y * y;
}

struct TopArgs {
    int n;
    int r;                     // run bits (qubits 0..r-1)
    uint64_t qc_high;          // unmeasured qubits >= r
    uint64_t q_high_mask;      // measured (top) qubits >= r
    int ql_rank[kRunBits];     // run bit i -> caller index j if measured in the top levels, else -1
    int qh_pos[24];            // measured top qubit positions >= r, ascending
    int qh_rank[24];           // and their caller indices
    int nqh;
};

// One workgroup per assignment of the measured top qubits above the run:
// every amplitude with those bits fixed, in runs of 2^r, summed over the
// unmeasured qubits into 2^|measured run bits| bins; out is indexed by the
// caller's bit order (bit j = outcome of q_j).
template <typename T>
__global__ __launch_bounds__(256) void k_tree_top(const T *__restrict__ a, TopArgs ta, double *__restrict__ out)
{
    __shared__ double part[256];
    const int tid = threadIdx.x;
    const int runsz = 1 << ta.r;
    const int per = 256 / runsz;                   // threads per run position
    const int j = tid & (runsz - 1), k0 = tid / runsz;
    // base: the assignment blockIdx.x of the measured high qubits
    uint64_t base = 0;
    uint64_t obase = 0;                            // caller-order index bits of that assignment
    for (int i = 0; i < ta.nqh; i++)
        if (blockIdx.x >> i & 1) {
            base |= uint64_t(1) << ta.qh_pos[i];
            obase |= uint64_t(1) << ta.qh_rank[i];
        }
    const uint64_t mask = ta.qc_high;
    const uint64_t total = mask ? (uint64_t(1) << __popcll(mask)) : 1;
    // enumerate k = k0, k0 + per, ... as pdep(k, mask) with the carry trick
    uint64_t x = 0, step = 0;
    {
        uint64_t kk = (uint64_t)k0, st = (uint64_t)per, mm = mask;
        for (int b = 0; mm; b++) {
            const uint64_t low = mm & (~mm + 1);
            if (kk >> b & 1) x |= low;
            if (st >> b & 1) step |= low;
            mm ^= low;
        }
    }
    double acc = 0;
    for (uint64_t k = k0; k < total; k += per) {
        acc += prob_of(a[base | x | (uint64_t)j]);
        x = ((x | ~mask) + step) & mask;
    }
    part[tid] = acc;
    __syncthreads();
    // bins over the measured run bits, fixed summation order
    if (tid < runsz) {
        double s = 0;
        for (int t = 0; t < per; t++) s += part[tid + t * runsz];
        part[tid] = s;
    }
    __syncthreads();
    // bin b (an assignment of the measured run bits) sums its run positions
    // in ascending order
    int measured_bits = 0, run_meas_mask = 0;
    for (int i = 0; i < ta.r; i++)
        if (ta.ql_rank[i] >= 0) {
            measured_bits++;
            run_meas_mask |= 1 << i;
        }
    if (tid < (1 << measured_bits)) {
        const int b = tid;
        int pos_val = 0, bi = 0;
        uint64_t oidx = obase;
        for (int i = 0; i < ta.r; i++)
            if (ta.ql_rank[i] >= 0) {
                if (b >> bi & 1) {
                    pos_val |= 1 << i;
                    oidx |= uint64_t(1) << ta.ql_rank[i];
                }
                bi++;
            }
        double s = 0;
        for (int pos = 0; pos < runsz; pos++)
            if ((pos & run_meas_mask) == pos_val) s += part[pos];
        out[oidx] = s;
    }
}

// parent level j (2^j entries) from level j + 1: children differ in bit j
__global__ void k_tree_level(const double *__restrict__ child, double *__restrict__ parent, int j)
{
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < (uint64_t(1) << j)) parent[p] = child[p] + child[p | (uint64_t(1) << j)];
}

__device__ __forceinline__ double tie_band(int n, float fr)
{
    // the measurement's own band (qsv_measure) for n <= 30; above that the
    // per-shot path has no serial re-decision (no reference order exists
    // beyond the reference's 30 qubits), so the band only covers the
    // rounding difference between the two evaluations
    const double f = fmax(1.0, (double)fr);
    return n <= 30 ? ldexp(DBL_EPSILON, n + 1) * f : 1e-12 * f;
}

// one thread per shot over the top T levels; tree level j lives at offset 2^j - 1
__global__ void k_tree_walk(const double *__restrict__ tree, int T, int n, const float *__restrict__ fr, int m,
                            int64_t shots, uint64_t *__restrict__ bits, double *__restrict__ cnorm,
                            int *__restrict__ flag)
{
    const int64_t sh = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (sh >= shots) return;
    double c = 1.0;
    uint64_t p = 0;
    int fl = 0;
    for (int j = 0; j < T; j++) {
        const double s0 = tree[((uint64_t(1) << (j + 1)) - 1) + p];
        const double p0 = s0 / c;
        const float r = fr[sh * m + j];
        if (fabs((double)r - p0) <= tie_band(n, r)) fl = 1;
        const int o = ((double)r > p0) ? 1 : 0;
        c *= o ? 1 - p0 : p0;
        p |= (uint64_t)o << j;
    }
    bits[sh] = p;
    cnorm[sh] = c;
    flag[sh] = fl;
}

struct SubArgs {
    int n, T, m, F;
    int top_pos[24];           // q_0 .. q_{T-1}
    int free_pos[kSubBits];    // qubits not in the top levels, ascending
    int rank[64];              // caller index j >= T -> index among free_pos
};

template <typename T>
__global__ __launch_bounds__(256) void k_tree_sub(const T *__restrict__ a, SubArgs sa, const float *__restrict__ fr,
                                                  int64_t shots, uint64_t *__restrict__ bits,
                                                  const double *__restrict__ cnorm, int *__restrict__ flag)
{
    __shared__ double pr[1 << kSubBits];
    __shared__ double red[256];
    const int64_t sh = blockIdx.x;
    if (sh >= shots || flag[sh]) return;            // uniform per workgroup
    const int tid = threadIdx.x;
    uint64_t p = bits[sh];
    uint64_t base = 0;
    for (int j = 0; j < sa.T; j++)
        if (p >> j & 1) base |= uint64_t(1) << sa.top_pos[j];
    const int nsub = 1 << sa.F;
    for (int e = tid; e < nsub; e += 256) {
        uint64_t x = base;
        for (int i = 0; i < sa.F; i++)
            if (e >> i & 1) x |= uint64_t(1) << sa.free_pos[i];
        pr[e] = prob_of(a[x]);
    }
    __syncthreads();
    double c = cnorm[sh];
    int fixmask = 0, fixval = 0, fl = 0;
    for (int j = sa.T; j < sa.m; j++) {
        const int rj = sa.rank[j];
        double part = 0;
        for (int e = tid; e < nsub; e += 256)
            if ((e & fixmask) == fixval && !((e >> rj) & 1)) part += pr[e];
        red[tid] = part;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (tid < w) red[tid] += red[tid + w];
            __syncthreads();
        }
        const double s0 = red[0];
        __syncthreads();
        const double p0 = s0 / c;
        const float r = fr[sh * sa.m + j];
        if (fabs((double)r - p0) <= tie_band(sa.n, r)) fl = 1;
        const int o = ((double)r > p0) ? 1 : 0;
        c *= o ? 1 - p0 : p0;
        fixmask |= 1 << rj;
        fixval |= o << rj;
        p |= (uint64_t)o << j;
    }
    if (tid == 0) {
        bits[sh] = p;
        flag[sh] = fl;
    }
}

template <typename T>
__global__ void k_gather(const T *__restrict__ a, const uint64_t *__restrict__ idx, int64_t count,
                         double *__restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const T v = a[idx[i]];
    out[2 * i] = v.x;
    out[2 * i + 1] = v.y;
}

}  // namespace

int gather(qsv_state *s, const uint64_t *idx, int64_t count, double *out)
{
    if (count == 0) return QSV_OK;
    uint64_t *didx = nullptr;
    double *dout = nullptr;
    if (hipMalloc((void **)&didx, sizeof(uint64_t) * count) != hipSuccess ||
        hipMalloc((void **)&dout, sizeof(double) * 2 * count) != hipSuccess) {
        if (didx) (void)hipFree(didx);
        return fail(QSV_ERR_ALLOC, "gather buffers");
    }
    int rc = QSV_OK;
    hipError_t e = hipMemcpyAsync(didx, idx, sizeof(uint64_t) * count, hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) {
        const dim3 g((unsigned)((count + 255) / 256)), t(256);
        if (s->prec == QSV_C128)
            hipLaunchKernelGGL(k_gather<c128>, g, t, 0, s->stream, (const c128 *)s->amps, didx, count, dout);
        else
            hipLaunchKernelGGL(k_gather<c64>, g, t, 0, s->stream, (const c64 *)s->amps, didx, count, dout);
        s->launches++;
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, dout, sizeof(double) * 2 * count, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) rc = hip_fail(e, "gather");
    (void)hipFree(didx);
    (void)hipFree(dout);
    return rc;
}

// The per-shot copy path (the reference's shot loop shape): copy the state
// into w (same or higher precision), measure qubits[0..m-1] in order with the
// draws that follow `rng_shot`.
static int sample_one_copy(qsv_state *s, qsv_state *w, const int *qubits, int m, int64_t gate_draws,
                           int32_t rng_shot, uint64_t *bits)
{
    int rc = copy_state(w, s);
    int32_t rng = qsv_pm_jump(rng_shot, (uint64_t)gate_draws);
    uint64_t b = 0;
    for (int j = 0; j < m && !rc; j++) {
        rng = qsv_pm_next(rng);
        int o = 0;
        rc = qsv_measure(w, qubits[j], (double)rng / 2147483647, &o);
        b |= (uint64_t)o << j;
    }
    *bits = b;
    return rc;
}

// A working state sharing s's stream and scratch, with its own amplitudes.
static int make_work(qsv_state *s, int prec, qsv_state *w)
{
    *w = *s;
    w->prec = prec;
    w->amps = nullptr;
    w->tq = nullptr;                     // the work copy applies nothing deferred
    const size_t bytes = s->dim * (prec == QSV_C128 ? sizeof(c128) : sizeof(c64));
    if (hipMalloc(&w->amps, bytes) != hipSuccess) {
        w->amps = nullptr;
        return fail(QSV_ERR_ALLOC, "cannot allocate shot scratch");
    }
    return QSV_OK;
}

static void drop_work(qsv_state *s, qsv_state *w)
{
    s->launches = w->launches;
    s->tie_resolutions = w->tie_resolutions;
    (void)hipStreamSynchronize(s->stream);
    if (w->amps) (void)hipFree(w->amps);
    w->amps = nullptr;
}

int sample_copy(qsv_state *s, const int *qubits, int m, int64_t gate_draws, int32_t *rng, int64_t shots,
                uint64_t *out, int64_t *ties)
{
    qsv_state w;
    int rc = make_work(s, s->prec, &w);
    if (rc) return rc;
    const int64_t ties0 = s->tie_resolutions;
    for (int64_t sh = 0; sh < shots && !rc; sh++) {
        rc = sample_one_copy(s, &w, qubits, m, gate_draws, *rng, &out[sh]);
        *rng = qsv_pm_jump(*rng, (uint64_t)gate_draws + (uint64_t)m);
    }
    drop_work(s, &w);
    if (ties) *ties = s->tie_resolutions - ties0;
    return rc;
}

int sample_tree(qsv_state *s, const int *qubits, int m, int64_t gate_draws, int32_t *rng, int64_t shots,
                uint64_t *out, int64_t *ties)
{
    const int n = s->n;
    if (shots == 0) return QSV_OK;
    const int T = m <= kTopFull ? m : std::min(m, std::max(kTopFull, n - kSubBits));
    const int F = n - T;
    if (T < m && F > kSubBits) return fail(QSV_ERR_ARG, "internal: subtree too large");
    // draws: per shot, skip the gate draws, then one per measured qubit
    std::vector<float> fr((size_t)shots * m);
    int32_t st = *rng;
    for (int64_t sh = 0; sh < shots; sh++) {
        const int32_t start = st;
        st = qsv_pm_jump(start, (uint64_t)gate_draws);
        for (int j = 0; j < m; j++) {
            st = qsv_pm_next(st);
            fr[(size_t)sh * m + j] = (float)((double)st / 2147483647);
        }
    }
    const int32_t rng_end = st;
    // device buffers: tree (2^(T+1) doubles), draws, bits, C, flags
    const size_t tree_n = (size_t(1) << (T + 1));
    double *tree = nullptr, *cn = nullptr;
    float *dfr = nullptr;
    uint64_t *dbits = nullptr;
    int *dflag = nullptr;
    auto release = [&]() {
        for (void *p : {(void *)tree, (void *)cn, (void *)dfr, (void *)dbits, (void *)dflag})
            if (p) (void)hipFree(p);
    };
    if (hipMalloc((void **)&tree, sizeof(double) * tree_n) != hipSuccess ||
        hipMalloc((void **)&cn, sizeof(double) * shots) != hipSuccess ||
        hipMalloc((void **)&dfr, sizeof(float) * fr.size()) != hipSuccess ||
        hipMalloc((void **)&dbits, sizeof(uint64_t) * shots) != hipSuccess ||
        hipMalloc((void **)&dflag, sizeof(int) * shots) != hipSuccess) {
        release();
        return fail(QSV_ERR_ALLOC, "sampler buffers");
    }
    int rc = QSV_OK;
    do {
        hipError_t e = hipMemcpyAsync(dfr, fr.data(), sizeof(float) * fr.size(), hipMemcpyHostToDevice, s->stream);
        if (e != hipSuccess) { rc = hip_fail(e, "hipMemcpyAsync"); break; }
        // marginal of q_0..q_{T-1} into tree level T
        TopArgs ta{};
        ta.n = n;
        ta.r = std::min(kRunBits, n);
        uint64_t top_mask = 0;
        for (int j = 0; j < T; j++) top_mask |= uint64_t(1) << qubits[j];
        const uint64_t run_mask = (uint64_t(1) << ta.r) - 1;
        for (int i = 0; i < ta.r; i++) {
            ta.ql_rank[i] = -1;
            for (int j = 0; j < T; j++)
                if (qubits[j] == i) ta.ql_rank[i] = j;
        }
        for (int b = ta.r; b < n; b++)
            if (top_mask >> b & 1) {
                ta.qh_pos[ta.nqh] = b;
                for (int j = 0; j < T; j++)
                    if (qubits[j] == b) ta.qh_rank[ta.nqh] = j;
                ta.nqh++;
            }
        const uint64_t all = (n == 64) ? ~uint64_t(0) : ((uint64_t(1) << n) - 1);
        ta.q_high_mask = top_mask & ~run_mask;
        ta.qc_high = all & ~top_mask & ~run_mask;
        double *levelT = tree + ((size_t(1) << T) - 1);
        const dim3 gtop((unsigned)(uint64_t(1) << ta.nqh));
        if (s->prec == QSV_C128)
            hipLaunchKernelGGL(k_tree_top<c128>, gtop, dim3(256), 0, s->stream, (const c128 *)s->amps, ta, levelT);
        else
            hipLaunchKernelGGL(k_tree_top<c64>, gtop, dim3(256), 0, s->stream, (const c64 *)s->amps, ta, levelT);
        for (int j = T - 1; j >= 0; j--) {
            const uint64_t cnt = uint64_t(1) << j;
            hipLaunchKernelGGL(k_tree_level, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s->stream,
                               tree + ((size_t(1) << (j + 1)) - 1), tree + ((size_t(1) << j) - 1), j);
        }
        hipLaunchKernelGGL(k_tree_walk, dim3((unsigned)((shots + 255) / 256)), dim3(256), 0, s->stream, tree, T, n,
                           dfr, m, shots, dbits, cn, dflag);
        s->launches += 2 + T;
        if (T < m) {
            SubArgs sa{};
            sa.n = n;
            sa.T = T;
            sa.m = m;
            sa.F = F;
            for (int j = 0; j < T; j++) sa.top_pos[j] = qubits[j];
            int f = 0;
            for (int b = 0; b < n; b++)
                if (!(top_mask >> b & 1)) sa.free_pos[f++] = b;
            for (int j = T; j < m; j++)
                for (int i = 0; i < F; i++)
                    if (sa.free_pos[i] == qubits[j]) sa.rank[j] = i;
            if (s->prec == QSV_C128)
                hipLaunchKernelGGL(k_tree_sub<c128>, dim3((unsigned)shots), dim3(256), 0, s->stream, (const c128 *)s->amps,
                                   sa, dfr, shots, dbits, cn, dflag);
            else
                hipLaunchKernelGGL(k_tree_sub<c64>, dim3((unsigned)shots), dim3(256), 0, s->stream, (const c64 *)s->amps,
                                   sa, dfr, shots, dbits, cn, dflag);
            s->launches++;
        }
        if ((e = hipGetLastError()) != hipSuccess) { rc = hip_fail(e, "sampler launch"); break; }
        std::vector<int> flags(shots);
        if ((e = hipMemcpyAsync(out, dbits, sizeof(uint64_t) * shots, hipMemcpyDeviceToHost, s->stream)) != hipSuccess ||
            (e = hipMemcpyAsync(flags.data(), dflag, sizeof(int) * shots, hipMemcpyDeviceToHost, s->stream)) != hipSuccess ||
            (e = hipStreamSynchronize(s->stream)) != hipSuccess) {
            rc = hip_fail(e, "sampler readback");
            break;
        }
        // shots with a decision inside the tie band: re-run by the copy path
        int64_t nflag = 0;
        for (int64_t sh = 0; sh < shots; sh++) nflag += flags[sh] != 0;
        const int64_t ties0 = s->tie_resolutions;
        if (nflag) {
            // FP64 working copy: for a complex64 state this keeps the collapse
            // arithmetic in FP64, as the recurrence above
            qsv_state w;
            if ((rc = make_work(s, QSV_C128, &w))) break;
            for (int64_t sh = 0; sh < shots && !rc; sh++)
                if (flags[sh])
                    rc = sample_one_copy(s, &w, qubits, m, gate_draws,
                                         qsv_pm_jump(*rng, (uint64_t)sh * ((uint64_t)gate_draws + (uint64_t)m)), &out[sh]);
            drop_work(s, &w);
        }
        if (ties) *ties = nflag + (s->tie_resolutions - ties0);
    } while (0);
    (void)hipStreamSynchronize(s->stream);
    release();
    if (!rc) *rng = rng_end;
    return rc;
}

}  // namespace qsv
