// qsv_costmodel.hip -- learned cost model for fused blocks and the scheduler
// that uses it.
//
// Model: MLP F=12 -> 32 -> 32 -> 1 with ReLU, predicting
//   y = log(ms of one fused-block launch, scaled to 2^30 amplitudes).
// Parameters (1505 floats): W1[32][12], b1[32], W2[32][32], b2[32], W3[32], b3.
// Kernels (all deterministic):
//   k_mlp_forward   one thread per sample, activations kept for backward;
//   k_mlp_backward  one thread per sample, per-sample gradient of
//                   0.5 * mean((y - t)^2) into a [batch][P] buffer;
//   k_grad_reduce   one thread per parameter, sums the batch in index order;
//   k_adam          Adam update.
// The scheduler plans the circuit at k = 2..5, evaluates every block at every
// tile shape in one batched forward pass, and keeps the cheapest plan; the
// greedy planner at opts.fuse_k is the fallback when no model is attached.
#include "qsv_fusion.h"
#include "qsv_sched.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <thread>

struct qsv_costmodel {
    float *params = nullptr;     // device, kParams
    float *m = nullptr, *v = nullptr, *grad = nullptr;
    hipStream_t stream = nullptr;
    int64_t step = 0;
    std::mutex mu;               // one call at a time (the scheduler scores candidates on threads)
    // inference buffers, grown as needed and kept: the scheduler predicts
    // while the state's stream runs sweeps, and freeing device memory would
    // wait for them
    float *ix = nullptr, *iy = nullptr, *iact = nullptr;   // device
    float *hx = nullptr, *hy = nullptr;                     // pinned host staging
    int64_t icap = 0;
};

namespace qsv {

namespace {

constexpr int F = QSV_CM_FEATURES, H = 32;
constexpr int OW1 = 0, OB1 = OW1 + H * F, OW2 = OB1 + H, OB2 = OW2 + H * H, OW3 = OB2 + H, OB3 = OW3 + H;
constexpr int kParams = OB3 + 1;

// Layer 2's sums are accumulated as layer 1's units come out (the same terms
// in the same order as a row-by-row product, so the same values), which keeps
// the kernel within 64 VGPRs without spills: 8 waves per SIMD, so that an
// inference's workgroups fit beside two resident 64 KiB pass workgroups (4
// waves per SIMD at <= 112 VGPRs) and the scheduler's predictions do not wait
// for the sweeps they plan.
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(8)))
void k_mlp_forward(const float *__restrict__ p, const float *__restrict__ x, int64_t batch,
                   float *__restrict__ y, float *__restrict__ act)
{
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= batch) return;
    const float *xb = x + b * F;
    float xv[F], s2[H];
    for (int i = 0; i < F; i++) xv[i] = xb[i];
    for (int j = 0; j < H; j++) s2[j] = p[OB2 + j];
    for (int i = 0; i < H; i++) {
        float s = p[OB1 + i];
        for (int f = 0; f < F; f++) s += p[OW1 + i * F + f] * xv[f];
        const float h1 = s > 0 ? s : 0;
        if (act) act[b * 2 * H + i] = h1;
        for (int j = 0; j < H; j++) s2[j] += p[OW2 + j * H + i] * h1;
    }
    float out = p[OB3];
    for (int j = 0; j < H; j++) {
        const float h2 = s2[j] > 0 ? s2[j] : 0;
        out += p[OW3 + j] * h2;
        if (act) act[b * 2 * H + H + j] = h2;
    }
    y[b] = out;
}

__global__ void k_mlp_backward(const float *__restrict__ p, const float *__restrict__ x,
                               const float *__restrict__ act, const float *__restrict__ y,
                               const float *__restrict__ t, int64_t batch, float *__restrict__ g)
{
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= batch) return;
    const float *xb = x + b * F, *h1 = act + b * 2 * H, *h2 = h1 + H;
    float *gb = g + b * kParams;
    const float dy = (y[b] - t[b]) / (float)batch;
    gb[OB3] = dy;
    float dh2[H];
    for (int j = 0; j < H; j++) {
        gb[OW3 + j] = dy * h2[j];
        dh2[j] = h2[j] > 0 ? dy * p[OW3 + j] : 0.0f;
        gb[OB2 + j] = dh2[j];
    }
    for (int j = 0; j < H; j++)
        for (int i = 0; i < H; i++) gb[OW2 + j * H + i] = dh2[j] * h1[i];
    for (int i = 0; i < H; i++) {
        float s = 0;
        for (int j = 0; j < H; j++) s += p[OW2 + j * H + i] * dh2[j];
        const float dh1 = h1[i] > 0 ? s : 0.0f;
        gb[OB1 + i] = dh1;
        for (int k = 0; k < F; k++) gb[OW1 + i * F + k] = dh1 * xb[k];
    }
}

__global__ void k_grad_reduce(const float *__restrict__ g, int64_t batch, float *__restrict__ out)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kParams) return;
    float s = 0;
    for (int64_t b = 0; b < batch; b++) s += g[b * kParams + p];
    out[p] = s;
}

__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, float lr, float c1, float c2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kParams) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    m[i] = b1 * m[i] + (1 - b1) * g[i];
    v[i] = b2 * v[i] + (1 - b2) * g[i] * g[i];
    p[i] -= lr * (m[i] / c1) / (sqrtf(v[i] / c2) + eps);
}

uint64_t splitmix(uint64_t &x)
{
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Loss of the current parameters on (x, t), plus the device buffers it used.
struct Batch {
    float *x = nullptr, *t = nullptr, *y = nullptr, *act = nullptr, *g = nullptr;
    int64_t n = 0;
    ~Batch()
    {
        for (float *p : {x, t, y, act, g})
            if (p) (void)hipFree(p);
    }
};

void free_inference(qsv_costmodel *m)
{
    if (m->ix) (void)hipFree(m->ix);
    if (m->iy) (void)hipFree(m->iy);
    if (m->iact) (void)hipFree(m->iact);
    if (m->hx) (void)hipHostFree(m->hx);
    if (m->hy) (void)hipHostFree(m->hy);
    m->ix = m->iy = m->iact = m->hx = m->hy = nullptr;
    m->icap = 0;
}

int upload_batch(qsv_costmodel *m, const float *x, const float *t, int64_t batch, bool need_grad, Batch &B)
{
    B.n = batch;
    if (hipMalloc((void **)&B.x, sizeof(float) * F * batch) != hipSuccess ||
        hipMalloc((void **)&B.y, sizeof(float) * batch) != hipSuccess ||
        hipMalloc((void **)&B.act, sizeof(float) * 2 * H * batch) != hipSuccess)
        return fail(QSV_ERR_ALLOC, "cost-model batch");
    QSV_HIP(hipMemcpyAsync(B.x, x, sizeof(float) * F * batch, hipMemcpyHostToDevice, m->stream));
    if (t) {
        if (hipMalloc((void **)&B.t, sizeof(float) * batch) != hipSuccess) return fail(QSV_ERR_ALLOC, "targets");
        QSV_HIP(hipMemcpyAsync(B.t, t, sizeof(float) * batch, hipMemcpyHostToDevice, m->stream));
    }
    if (need_grad && hipMalloc((void **)&B.g, sizeof(float) * kParams * batch) != hipSuccess)
        return fail(QSV_ERR_ALLOC, "per-sample gradients");
    return QSV_OK;
}

int forward(qsv_costmodel *m, Batch &B)
{
    const unsigned grid = (unsigned)((B.n + 127) / 128);
    hipLaunchKernelGGL(k_mlp_forward, dim3(grid), dim3(128), 0, m->stream, m->params, B.x, B.n, B.y, B.act);
    QSV_HIP(hipGetLastError());
    return QSV_OK;
}

int backward(qsv_costmodel *m, Batch &B)
{
    const unsigned grid = (unsigned)((B.n + 127) / 128);
    hipLaunchKernelGGL(k_mlp_backward, dim3(grid), dim3(128), 0, m->stream, m->params, B.x, B.act, B.y, B.t, B.n, B.g);
    hipLaunchKernelGGL(k_grad_reduce, dim3((kParams + 255) / 256), dim3(256), 0, m->stream, B.g, B.n, m->grad);
    QSV_HIP(hipGetLastError());
    return QSV_OK;
}

int batch_loss(qsv_costmodel *m, Batch &B, const float *t_host, double *loss)
{
    std::vector<float> y(B.n);
    QSV_HIP(hipMemcpyAsync(y.data(), B.y, sizeof(float) * B.n, hipMemcpyDeviceToHost, m->stream));
    QSV_HIP(hipStreamSynchronize(m->stream));
    double s = 0;
    for (int64_t i = 0; i < B.n; i++) s += 0.5 * ((double)y[i] - t_host[i]) * ((double)y[i] - t_host[i]);
    *loss = s / (double)B.n;
    return QSV_OK;
}

}  // namespace

// Widened qubit list of a block (the lowest unused qubits fill it to kq).
static std::vector<int> widened_qubits(const std::vector<int> &qs, int kq)
{
    std::vector<int> out = qs;
    uint64_t used = 0;
    for (int q : qs) used |= uint64_t(1) << q;
    for (int b = 0; (int)out.size() < kq; b++)
        if (!(used >> b & 1)) out.push_back(b);
    return out;
}

void launch_features(int n, int split, const std::vector<const std::vector<int> *> &qsets, bool single, int tb,
                     float *f)
{
    uint64_t u = 0;
    int work = 0, five = 0;
    for (auto *q : qsets) {
        for (int x : *q) u |= uint64_t(1) << x;
        work += 1 << q->size();
        five += q->size() == 5;
    }
    uint64_t lmask = u;
    for (int b = 0; __builtin_popcountll(lmask) < tb && b < 64; b++) lmask |= uint64_t(1) << b;
    int run = 0;
    while (run < 64 && (lmask >> run & 1)) run++;
    const int nu = __builtin_popcountll(u);
    int mn = 64, mx = -1, top = 0;
    for (int b = 0; b < 64; b++)
        if (u >> b & 1) {
            mn = std::min(mn, b);
            mx = std::max(mx, b);
            top += b >= n - 3;
        }
    const int G = (int)qsets.size();
    f[0] = n / 34.0f;
    f[1] = single ? 1.0f : 0.0f;
    f[2] = G / 8.0f;
    f[3] = work / 128.0f;
    f[4] = (float)split;   // product mode 0, 1, 2
    f[5] = (tb - 10) / 5.0f;
    f[6] = (float)std::min(run, tb) / tb;
    f[7] = nu / 14.0f;
    f[8] = (float)five / G;
    f[9] = (float)mn / n;
    f[10] = (float)mx / n;
    f[11] = nu ? (float)top / nu : 0.0f;
}

// Single-block tile column bits that fit this state (6..8), default 7.
static int default_tc(int n, int kq)
{
    return std::max(6, std::min(kDefaultTileCols, n - kq));
}

int group_launches(qsv_state *s, const std::vector<PlanItem> &plan, const qsv_run_opts &o,
                   std::vector<Launch> &launches, double *pred_total, const std::vector<int> *window)
{
    launches.clear();
    *pred_total = 0;
    const int n = s->n;
    const bool learned = o.scheduler == 1 && s->cm;
    const int gmax = learned ? kMaxPassBlocks : std::max(1, std::min(o.group, kMaxPassBlocks));
    auto mfma_ok = [&](size_t b) { return plan[b].qubits.size() >= 4 && (int)plan[b].qubits.size() + 6 <= n; };
    // segments: runs of consecutive blocks up to the next barrier (and, for a
    // window plan, the end of the window: its blocks are planned as one sweep,
    // and the DP only chooses how to launch them); barriers are segments of
    // their own with first == last
    struct Seg { size_t i, j; size_t c0 = 0, c1 = 0; };
    std::vector<Seg> segs;
    for (size_t i = 0; i < plan.size();) {
        size_t j = i + 1;
        if (plan[i].barrier < 0)
            while (j < plan.size() && plan[j].barrier < 0 && (!window || (*window)[j] == (*window)[i])) j++;
        segs.push_back({i, j});
        i = j;
    }
    if (!learned) {
        for (const Seg &sg : segs) {
            if (plan[sg.i].barrier >= 0) {
                Launch L;
                L.first = sg.i;
                launches.push_back(L);
                continue;
            }
            size_t b = sg.i;
            while (b < sg.j) {
                Launch L;
                L.first = b;
                L.count = 1;
                L.tc = mfma_ok(b) ? (o.tile ? o.tile : default_tc(n, (int)plan[b].qubits.size())) : 0;
                if (mfma_ok(b) && (int)plan[b].qubits.size() + L.tc > n) L.tc = default_tc(n, (int)plan[b].qubits.size());
                std::vector<const std::vector<int> *> q = {&plan[b].qubits};
                while (mfma_ok(b) && gmax > 1 && b + L.count < sg.j && L.count < gmax && mfma_ok(b + L.count)) {
                    q.push_back(&plan[b + L.count].qubits);
                    int tb = pass_tile_bits(q, n);
                    if (tb < 0) {
                        q.pop_back();
                        break;
                    }
                    L.count++;
                }
                if (L.count > 1) {
                    L.tb = pass_tile_bits(q, n);
                    if (o.pass_tb > L.tb && o.pass_tb <= std::min(kMaxPassTileBits, n)) L.tb = o.pass_tb;
                }
                launches.push_back(L);
                b += L.count;
            }
        }
        return QSV_OK;
    }
    // learned: every candidate launch of every segment, predicted in one
    // batched inference, then a dynamic program per segment
    struct Cand { size_t start; int count, tc, tb; };
    std::vector<Cand> cands;
    std::vector<float> feats;
    for (Seg &sg : segs) {
        sg.c0 = cands.size();
        if (plan[sg.i].barrier < 0)
            for (size_t last = sg.i; last < sg.j; last++) {     // candidate launches ending at block `last`
                if (!mfma_ok(last)) {
                    cands.push_back({last, 1, 0, 0});
                    float f[F] = {0};
                    feats.insert(feats.end(), f, f + F);
                    continue;
                }
                const int kq = (int)plan[last].qubits.size();
                std::vector<const std::vector<int> *> q1 = {&plan[last].qubits};
                for (int tc = 6; tc <= 8; tc++) {
                    if (kq + tc > n) continue;
                    float f[F];
                    launch_features(n, o.split, q1, true, kq + tc, f);
                    cands.push_back({last, 1, tc, 0});
                    feats.insert(feats.end(), f, f + F);
                }
                std::vector<const std::vector<int> *> q = q1;
                for (int g = 2; g <= kMaxPassBlocks && (size_t)g <= last - sg.i + 1; g++) {
                    const size_t st = last + 1 - g;
                    if (!mfma_ok(st)) break;
                    q.insert(q.begin(), &plan[st].qubits);
                    const int tbmin = pass_tile_bits(q, n);
                    if (tbmin < 0) break;
                    for (int tb = tbmin; tb <= std::min(kMaxPassTileBits, n); tb++) {
                        float f[F];
                        launch_features(n, o.split, q, false, tb, f);
                        cands.push_back({st, g, 0, tb});
                        feats.insert(feats.end(), f, f + F);
                    }
                }
            }
        sg.c1 = cands.size();
    }
    std::vector<float> y(cands.size());
    if (!cands.empty()) {
        int rc = qsv_cm_forward(s->cm, feats.data(), (int64_t)cands.size(), y.data());
        if (rc) return rc;
    }
    const double scale = std::ldexp(1.0, n - 30);
    auto cost = [&](size_t c) {
        const Cand &cd = cands[c];
        return (cd.tc == 0 && cd.tb == 0) ? 0.0 : std::exp((double)y[c]) * scale;
    };
    // a window keeps its one sweep (at its minimal tile bits, or the default
    // tile shape for a lone block) unless the DP predicts a gain above the
    // model's resolution: deviations on smaller predicted gains measured
    // slower more often than faster
    constexpr double kWindowMargin = 0.05;     // ~ the per-launch prediction error
    for (const Seg &sg : segs) {
        const size_t i = sg.i, m = sg.j - sg.i;
        if (plan[i].barrier >= 0) {
            Launch L;
            L.first = i;
            launches.push_back(L);
            continue;
        }
        std::vector<double> dp(m + 1, std::numeric_limits<double>::infinity());
        std::vector<size_t> from(m + 1, 0);
        dp[0] = 0;
        // candidates come grouped by their end block, in increasing order
        for (size_t c = sg.c0; c < sg.c1; c++) {
            const Cand &cd = cands[c];
            const size_t end = cd.start + cd.count - i;      // dp index after this launch
            const size_t beg = cd.start - i;
            const double ms = cost(c);
            if (dp[beg] + ms < dp[end]) {
                dp[end] = dp[beg] + ms;
                from[end] = c;
            }
        }
        long whole = -1;
        if (window)
            for (size_t c = sg.c0; c < sg.c1; c++) {
                const Cand &cd = cands[c];
                if (cd.start != i || (size_t)cd.count != m) continue;
                const bool lone = m == 1 && cd.tc == default_tc(n, (int)plan[i].qubits.size());
                const bool pass = m > 1 && (whole < 0 || cd.tb < cands[whole].tb);
                if (lone || pass || (m == 1 && cd.tc == 0 && cd.tb == 0)) whole = (long)c;
            }
        std::vector<Launch> seg;
        if (whole >= 0 && dp[m] > cost(whole) * (1 - kWindowMargin)) {
            const Cand &cd = cands[whole];
            Launch L;
            L.first = cd.start;
            L.count = cd.count;
            L.tc = cd.tc;
            L.tb = cd.tb;
            L.pred_ms = cost(whole);
            seg.push_back(L);
            *pred_total += L.pred_ms;
        } else {
            for (size_t e = m; e > 0;) {
                const Cand &cd = cands[from[e]];
                Launch L;
                L.first = cd.start;
                L.count = cd.count;
                L.tc = cd.tc;
                L.tb = cd.tb;
                L.pred_ms = dp[e] - dp[cd.start - i];
                seg.push_back(L);
                e = cd.start - i;
            }
            std::reverse(seg.begin(), seg.end());
            *pred_total += dp[m];
        }
        launches.insert(launches.end(), seg.begin(), seg.end());
    }
    return QSV_OK;
}

int plan_learned(qsv_state *s, const qsv_gate *gates, int64_t ng, const double *mats, const qsv_run_opts &o,
                 std::vector<PlanItem> &plan, SchedChoice *choice, std::vector<int> *window)
{
    window->clear();
    if (!s->cm || s->n < 4 + 6) {
        choice->fuse_k = o.fuse_k;
        return plan_greedy(gates, ng, mats, o.fuse_k, o.exact, plan);
    }
    // candidate plans: block-first at k = 2..5 and window plans at k = 4, 5
    // (states of >= 11 qubits).  Planning is host work with no shared state,
    // so the candidates are built on threads of their own; the cost-model DP
    // (device inference on the model's stream) then scores them in turn.  The
    // candidates are planned as shapes (qubit lists, no block matrices: the
    // planners' decisions depend on the qubits alone) and only the winner is
    // planned again with its matrices (a window plan replays its windows'
    // qubit masks instead of growing them again).
    struct Cand {
        int k = 0, window = 0, rc = QSV_OK;
        std::string err;                 // the planner's message (errors are per thread)
        double total = 0;                // predicted device ms
        std::vector<PlanItem> plan;
        std::vector<int> win;
        std::vector<uint64_t> masks;     // window plans: each window's qubits
    };
    std::vector<Cand> cands;
    for (int k = 2; k <= 5; k++) {
        Cand c;
        c.k = k;
        cands.push_back(std::move(c));
    }
    if (s->n >= 11)
        for (int k = 4; k <= 5; k++) {
            Cand c;
            c.k = k;
            c.window = 1;
            cands.push_back(std::move(c));
        }
    auto score = [&](Cand &c) {
        // the DP sees the qubit lists as the run will widen them (the
        // matrices are widened later); it needs nothing else of a block
        std::vector<PlanItem> shape(c.plan.size());
        for (size_t i = 0; i < c.plan.size(); i++) {
            shape[i].barrier = c.plan[i].barrier;
            shape[i].qubits = c.plan[i].qubits;
            if (c.plan[i].barrier < 0 && s->n >= std::max(4, (int)c.plan[i].qubits.size()) + 6) {
                if (c.window) {
                    const std::vector<int> add = window_widening(c.plan, c.win, i, 4);
                    shape[i].qubits.insert(shape[i].qubits.end(), add.begin(), add.end());
                } else {
                    shape[i].qubits = widened_qubits(c.plan[i].qubits, std::max(4, (int)c.plan[i].qubits.size()));
                }
            }
        }
        std::vector<Launch> launches;
        return group_launches(s, shape, o, launches, &c.total, c.window ? &c.win : nullptr);
    };
    // each thread plans and scores its candidate (feature building is host
    // work; the model's inference calls take turns)
    std::vector<std::thread> pool;
    for (Cand &c : cands)
        pool.emplace_back([&c, gates, ng, mats, &o, s, &score]() {
            c.rc = c.window ? plan_window(gates, ng, mats, c.k, o.exact, kMaxPassTileBits, s->n, c.plan, c.win, true,
                                          nullptr, &c.masks)
                            : plan_greedy(gates, ng, mats, c.k, o.exact, c.plan, true);
            // a new thread's current device is device 0: score on the state's
            if (!c.rc && hipSetDevice(s->device) != hipSuccess) c.rc = fail(QSV_ERR_DEVICE, "cannot select the device");
            if (!c.rc) c.rc = score(c);
            if (c.rc) c.err = qsv_last_error();
        });
    for (std::thread &t : pool) t.join();
    double best_total = std::numeric_limits<double>::infinity();
    Cand *best = nullptr;
    for (Cand &c : cands) {
        if (c.rc) return fail(c.rc, c.err);
        if (c.total < best_total) {
            best_total = c.total;
            best = &c;
        }
    }
    if (!best) return fail(QSV_ERR_NUMERIC, "cost model predicted no finite schedule");
    int rc = best->window ? plan_window(gates, ng, mats, best->k, o.exact, kMaxPassTileBits, s->n, plan, *window,
                                        false, &best->masks)
                          : plan_greedy(gates, ng, mats, best->k, o.exact, plan);
    if (rc) return rc;
    bool same = plan.size() == best->plan.size() && window->size() == best->win.size();
    for (size_t i = 0; same && i < plan.size(); i++)
        same = plan[i].qubits == best->plan[i].qubits && plan[i].barrier == best->plan[i].barrier;
    if (!same) return fail(QSV_ERR_UNSUPPORTED, "planner shapes differ with and without block matrices");
    choice->fuse_k = best->k;
    choice->window = best->window;
    choice->pred_ms = best_total;
    return QSV_OK;
}

}  // namespace qsv

using namespace qsv;

extern "C" {

int qsv_cm_nparams(void) { return kParams; }

int qsv_cm_block_features(int n, int kq, int split, int tile, const int *qubits, float *f)
{
    if (!qubits || !f || n < 1 || kq < 1 || kq > 5 || tile < 6 || tile > 8) return fail(QSV_ERR_ARG, "bad feature arguments");
    std::vector<int> q(qubits, qubits + kq);
    std::vector<const std::vector<int> *> qs = {&q};
    launch_features(n, split, qs, true, kq + tile, f);
    return QSV_OK;
}

int qsv_cm_create(uint64_t seed, qsv_costmodel **out)
{
    if (!out) return fail(QSV_ERR_ARG, "null");
    *out = nullptr;
    qsv_costmodel *m = new qsv_costmodel();
    bool ok = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) == hipSuccess &&
              hipMalloc((void **)&m->params, sizeof(float) * kParams) == hipSuccess &&
              hipMalloc((void **)&m->m, sizeof(float) * kParams) == hipSuccess &&
              hipMalloc((void **)&m->v, sizeof(float) * kParams) == hipSuccess &&
              hipMalloc((void **)&m->grad, sizeof(float) * kParams) == hipSuccess;
    if (!ok) {
        qsv_cm_destroy(m);
        return fail(QSV_ERR_ALLOC, "cost model allocation");
    }
    // He-uniform weights from a seeded SplitMix64 stream, zero biases
    std::vector<float> p(kParams, 0.0f);
    uint64_t st = seed;
    auto uni = [&]() { return (double)(splitmix(st) >> 11) * (1.0 / 9007199254740992.0); };
    auto fill = [&](int off, int count, int fan_in) {
        const double a = std::sqrt(6.0 / fan_in);
        for (int i = 0; i < count; i++) p[off + i] = (float)((2 * uni() - 1) * a);
    };
    fill(OW1, H * F, F);
    fill(OW2, H * H, H);
    fill(OW3, H, H);
    int rc = qsv_cm_set_params(m, p.data());
    if (rc) {
        qsv_cm_destroy(m);
        return rc;
    }
    *out = m;
    return QSV_OK;
}

int qsv_cm_destroy(qsv_costmodel *m)
{
    if (!m) return QSV_OK;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    for (float *p : {m->params, m->m, m->v, m->grad})
        if (p) (void)hipFree(p);
    free_inference(m);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return QSV_OK;
}

int qsv_cm_get_params(qsv_costmodel *m, float *params)
{
    if (!m || !params) return fail(QSV_ERR_ARG, "null");
    QSV_HIP(hipMemcpyAsync(params, m->params, sizeof(float) * kParams, hipMemcpyDeviceToHost, m->stream));
    QSV_HIP(hipStreamSynchronize(m->stream));
    return QSV_OK;
}

int qsv_cm_set_params(qsv_costmodel *m, const float *params)
{
    if (!m || !params) return fail(QSV_ERR_ARG, "null");
    QSV_HIP(hipMemcpyAsync(m->params, params, sizeof(float) * kParams, hipMemcpyHostToDevice, m->stream));
    QSV_HIP(hipMemsetAsync(m->m, 0, sizeof(float) * kParams, m->stream));
    QSV_HIP(hipMemsetAsync(m->v, 0, sizeof(float) * kParams, m->stream));
    QSV_HIP(hipStreamSynchronize(m->stream));
    m->step = 0;
    return QSV_OK;
}

int qsv_cm_forward(qsv_costmodel *m, const float *x, int64_t batch, float *y)
{
    if (!m || !x || !y || batch < 0) return fail(QSV_ERR_ARG, "bad argument");
    if (batch == 0) return QSV_OK;
    std::lock_guard<std::mutex> lock(m->mu);
    if (batch > m->icap) {
        free_inference(m);
        const int64_t cap = std::max<int64_t>(batch, 65536);
        if (hipMalloc((void **)&m->ix, sizeof(float) * F * cap) != hipSuccess ||
            hipMalloc((void **)&m->iy, sizeof(float) * cap) != hipSuccess ||
            hipMalloc((void **)&m->iact, sizeof(float) * 2 * H * cap) != hipSuccess ||
            hipHostMalloc((void **)&m->hx, sizeof(float) * F * cap, hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc((void **)&m->hy, sizeof(float) * cap, hipHostMallocDefault) != hipSuccess) {
            free_inference(m);
            return fail(QSV_ERR_ALLOC, "cost-model inference buffers");
        }
        m->icap = cap;
    }
    std::memcpy(m->hx, x, sizeof(float) * F * batch);
    QSV_HIP(hipMemcpyAsync(m->ix, m->hx, sizeof(float) * F * batch, hipMemcpyHostToDevice, m->stream));
    const unsigned grid = (unsigned)((batch + 127) / 128);
    hipLaunchKernelGGL(k_mlp_forward, dim3(grid), dim3(128), 0, m->stream, m->params, m->ix, batch, m->iy, m->iact);
    QSV_HIP(hipGetLastError());
    QSV_HIP(hipMemcpyAsync(m->hy, m->iy, sizeof(float) * batch, hipMemcpyDeviceToHost, m->stream));
    QSV_HIP(hipStreamSynchronize(m->stream));
    std::memcpy(y, m->hy, sizeof(float) * batch);
    return QSV_OK;
}

int qsv_cm_gradients(qsv_costmodel *m, const float *x, const float *t, int64_t batch, float *grad, double *loss)
{
    if (!m || !x || !t || !grad || batch < 1) return fail(QSV_ERR_ARG, "bad argument");
    std::lock_guard<std::mutex> lock(m->mu);
    Batch B;
    int rc = upload_batch(m, x, t, batch, true, B);
    if (!rc) rc = forward(m, B);
    if (!rc) rc = backward(m, B);
    if (rc) return rc;
    QSV_HIP(hipMemcpyAsync(grad, m->grad, sizeof(float) * kParams, hipMemcpyDeviceToHost, m->stream));
    QSV_HIP(hipStreamSynchronize(m->stream));
    if (loss) return batch_loss(m, B, t, loss);
    return QSV_OK;
}

int qsv_cm_train(qsv_costmodel *m, const float *x, const float *t, int64_t batch, int epochs, float lr,
                 double *loss)
{
    if (!m || !x || !t || batch < 1 || epochs < 0) return fail(QSV_ERR_ARG, "bad argument");
    std::lock_guard<std::mutex> lock(m->mu);
    Batch B;
    int rc = upload_batch(m, x, t, batch, true, B);
    if (rc) return rc;
    for (int e = 0; e < epochs && !rc; e++) {
        rc = forward(m, B);
        if (!rc) rc = backward(m, B);
        if (rc) break;
        m->step++;
        const float c1 = 1.0f - std::pow(0.9f, (float)m->step), c2 = 1.0f - std::pow(0.999f, (float)m->step);
        hipLaunchKernelGGL(k_adam, dim3((kParams + 255) / 256), dim3(256), 0, m->stream, m->params, m->grad, m->m,
                           m->v, lr, c1, c2);
        hipError_t er = hipGetLastError();
        if (er != hipSuccess) rc = hip_fail(er, "k_adam");
    }
    if (!rc) rc = forward(m, B);
    if (!rc && loss) rc = batch_loss(m, B, t, loss);
    return rc;
}

int qsv_set_costmodel(qsv_state *s, qsv_costmodel *m)
{
    if (!s) return fail(QSV_ERR_ARG, "null state");
    s->cm = m;
    return QSV_OK;
}

int qsv_timing_log(qsv_state *s, float *features, float *ms, int64_t cap, int64_t *count)
{
    if (!s || !count) return fail(QSV_ERR_ARG, "null");
    const int64_t n = (int64_t)s->tlog_ms.size();
    *count = n;
    const int64_t k = std::min(n, cap);
    if (features && k > 0) std::memcpy(features, s->tlog_feats.data(), sizeof(float) * F * k);
    if (ms && k > 0) std::memcpy(ms, s->tlog_ms.data(), sizeof(float) * k);
    return QSV_OK;
}

}  // extern "C"
