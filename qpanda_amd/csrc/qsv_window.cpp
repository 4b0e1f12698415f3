// qsv_window.cpp -- window planning: the sweep first, then the blocks in it.
//
// plan_greedy fuses gates into k-qubit blocks, and group_launches then packs
// consecutive blocks into shared HBM sweeps; the sweeps it can form are
// limited by the order in which the blocks happened to close.  The window
// planner starts from the sweep instead.  A multi-block sweep's tile holds TB
// qubits: qubits 0..3 (the contiguous 128-byte run) and TB - 4 others, and any
// gate whose qubits all lie in that set can run in the sweep once every
// earlier gate on its qubits has run.  Per segment between MEASURE / RESET
// records, repeatedly:
//   1. grow the window from {0, 1, 2, 3}: each step adds a candidate -- the
//      qubits a gate at the front of some qubit's queue still misses -- until
//      the window holds TB qubits or nothing gains.  Greedy growth takes the
//      candidate that makes the most further gates runnable per added qubit;
//      the planner looks one decision further: for each of the kLookahead
//      best candidates by that measure it completes the growth greedily and
//      keeps the candidate whose completed window runs the most gates;
//   2. take the runnable gates in dependency order (front-of-queue closure)
//      and fuse them into blocks of <= max_k qubits with plan_greedy, so every
//      block lies inside the window (plan_greedy ends by merging narrow blocks
//      into commuting neighbours, absorb_blocks); of the closure's order and
//      program order, the one that gives fewer blocks is kept;
//   3. emit the blocks as one window.
// On 30-qubit random circuits of depth 200 (TB = 13) this applies ~38 gates
// per sweep (~31 with greedy growth) against ~21 for block-first grouping,
// with more (but cheaper per gate) blocks per sweep.  The gates a window takes commute with every gate
// left behind that precedes them in the list (they act on other qubits), so
// the state is the circuit's own up to that reordering.
#include "qsv_sched.h"

#include <algorithm>
#include <array>
#include <condition_variable>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <unordered_set>

namespace qsv {

namespace {

// Growth candidates completed per step (1 = greedy growth).  Windows on the
// 30-qubit depth-200 benchmark circuit: 1 -> 288, 3 -> 246, 5 -> 240, 6 -> 234,
// 8 -> 236; over it and three other seeded random circuits (30/200, 28/100,
// 33/60) 6 gives the fewest windows in all (661 against 669 for 5).
constexpr int kLookahead = 6;

// Qubits a record acts on (the no-op CNOT(q, q) is kept on q so that it is
// counted like every other record).
int record_qubits(const qsv_gate &r, int *qs)
{
    switch (r.kind) {
    case QSV_GATE_H: case QSV_GATE_X: case QSV_GATE_RX: case QSV_GATE_RY: case QSV_GATE_RZ: case QSV_GATE_U1:
        qs[0] = r.q0;
        return 1;
    case QSV_GATE_CNOT: case QSV_GATE_CZ: case QSV_GATE_CR: case QSV_GATE_ISWAP: case QSV_GATE_SQISWAP:
    case QSV_GATE_U2:
        qs[0] = r.q0;
        if (r.q1 == r.q0) return 1;
        qs[1] = r.q1;
        return 2;
    case QSV_GATE_TOFFOLI: case QSV_GATE_CSWAP:
        qs[0] = r.q0; qs[1] = r.q1; qs[2] = r.q2;
        return 3;
    default:
        return -1;
    }
}

struct Segment {
    const qsv_gate *gates;
    int64_t count;
    int n;
    std::vector<int> nq;                  // qubits per record
    std::vector<int> qs;                  // 3 per record
    std::vector<std::vector<int>> queue;  // per qubit: records in order
    std::vector<size_t> head;             // per qubit: first record not yet taken

    // Records runnable inside window mask w starting from queue positions h
    // (advanced past them), in a dependency order.  Runnable sets only grow
    // with w, so a candidate's gain is the closure of w | add continued from
    // the positions after w's own closure.
    // A worklist of qubits whose queue front may have become runnable: a
    // record blocked on another qubit's front is retried when that front
    // advances.
    int64_t closure(uint64_t w, size_t *h, std::vector<int64_t> *out) const
    {
        int64_t cnt = 0;
        int work[64], top = 0;
        for (uint64_t m = w; m; m &= m - 1) work[top++] = __builtin_ctzll(m);
        uint64_t queued = w;
        while (top > 0) {
            const int q = work[--top];
            queued &= ~(uint64_t(1) << q);
            while (h[q] < queue[q].size()) {
                const int64_t g = queue[q][h[q]];
                const int *gq = &qs[3 * g];
                bool ok = true;
                for (int j = 0; j < nq[g] && ok; j++) {
                    const int x = gq[j];
                    ok = (w >> x & 1) && h[x] < queue[x].size() && queue[x][h[x]] == g;
                }
                if (!ok) break;
                for (int j = 0; j < nq[g]; j++) {
                    const int x = gq[j];
                    h[x]++;
                    if (x != q && !(queued >> x & 1)) {
                        queued |= uint64_t(1) << x;
                        work[top++] = x;
                    }
                }
                if (out) out->push_back(g);
                cnt++;
            }
        }
        return cnt;
    }

    // Growth candidates of window w whose closure left the queue positions
    // h: the qubits the front record of each queue outside w still misses,
    // at most `room` of them; gain = further runnable records per qubit.
    void candidates(uint64_t w, const size_t *h, int room, std::vector<std::pair<double, uint64_t>> &out) const
    {
        out.clear();
        size_t hc[64];
        for (int q = 0; q < n; q++) {
            if ((w >> q & 1) || h[q] >= queue[q].size()) continue;
            const int64_t g = queue[q][h[q]];
            uint64_t add = 0;
            for (int x = 0; x < nq[g]; x++) add |= uint64_t(1) << qs[3 * g + x];
            add &= ~w;
            const int k = __builtin_popcountll(add);
            if (k == 0 || k > room) continue;
            bool dup = false;
            for (const auto &c : out) dup |= c.second == add;
            if (dup) continue;
            std::copy(h, h + n, hc);
            out.push_back({(double)closure(w | add, hc, nullptr) / k, add});
        }
    }

    // Greedy growth of w (h: positions after w's closure, advanced with it)
    // to tb qubits or until no candidate gains.  The result depends on w
    // alone (the positions follow from it), so memo maps every window passed
    // on the way to the grown window, and a later growth that reaches one of
    // them stops there.
    uint64_t grow_greedy(uint64_t w, size_t *h, int tb, std::unordered_map<uint64_t, uint64_t> &memo) const
    {
        std::vector<std::pair<double, uint64_t>> cs;
        std::vector<uint64_t> path;
        uint64_t out = 0;
        while (true) {
            auto it = memo.find(w);
            if (it != memo.end()) {
                out = it->second;
                break;
            }
            path.push_back(w);
            if (__builtin_popcountll(w) >= tb) {
                out = w;
                break;
            }
            candidates(w, h, tb - __builtin_popcountll(w), cs);
            double best = 0;
            uint64_t add = 0;
            for (const auto &c : cs)
                if (c.first > best) {
                    best = c.first;
                    add = c.second;
                }
            if (add == 0) {
                out = w;
                break;
            }
            w |= add;
            closure(w, h, nullptr);
        }
        for (uint64_t p : path) memo[p] = out;
        return out;
    }

    // Growth with lookahead from window w with queue positions `from`: at each
    // step, complete the growth greedily from each of the `look` best
    // candidates and take the one whose completed window runs the most records
    // (ties: the better first step).
    uint64_t grow_from(uint64_t w, const size_t *from, int tb, int look) const
    {
        size_t h[64], h2[64];
        std::copy(from, from + n, h);
        closure(w, h, nullptr);
        std::vector<std::pair<double, uint64_t>> cs;
        std::unordered_map<uint64_t, uint64_t> memo;     // greedy completions
        std::unordered_map<uint64_t, int64_t> runs;      // records a window runs
        while (__builtin_popcountll(w) < tb) {
            candidates(w, h, tb - __builtin_popcountll(w), cs);
            std::stable_sort(cs.begin(), cs.end(), [](const auto &a, const auto &b) { return a.first > b.first; });
            if (cs.empty() || cs[0].first <= 0) break;
            uint64_t add = cs[0].second;
            if (look > 1) {
                int64_t best = -1;
                for (int c = 0; c < (int)cs.size() && c < look && cs[c].first > 0; c++) {
                    const uint64_t wc = w | cs[c].second;
                    uint64_t full;
                    auto m = memo.find(wc);
                    if (m != memo.end()) {
                        full = m->second;
                    } else {
                        std::copy(h, h + n, h2);
                        closure(wc, h2, nullptr);
                        full = grow_greedy(wc, h2, tb, memo);
                    }
                    auto r = runs.find(full);
                    int64_t got;
                    if (r != runs.end()) {
                        got = r->second;
                    } else {
                        std::copy(from, from + n, h2);
                        got = runs[full] = closure(full, h2, nullptr);
                    }
                    if (got > best) {
                        best = got;
                        add = cs[c].second;
                    }
                }
            }
            w |= add;
            closure(w, h, nullptr);
        }
        return w;
    }

    uint64_t base_window() const
    {
        uint64_t w = 0;
        for (int q = 0; q < std::min(4, n); q++) w |= uint64_t(1) << q;
        return w;
    }

    uint64_t grow(int tb, int look) const { return grow_from(base_window(), head.data(), tb, look); }

    // Relabelling: the four qubits window w leaves at bits 0..3 for the next
    // window, which must grow from them.  The next window is estimated by a
    // growth from nothing (positions `from`, after w); the four are taken
    // from its qubits that w holds -- first those already at bits 0..3 (no
    // data moves for them), then the others -- and then from w's other
    // qubits.  With nothing left to run, the qubits stay.  On the 30-qubit
    // depth-200 benchmark circuit the lookahead planner needs 187 windows this
    // way against 234 with qubits 0..3 in every window.
    uint64_t next_low(const size_t *from, uint64_t w, uint64_t low, int tb, int look) const
    {
        bool any = false;
        for (int q = 0; q < n && !any; q++) any = from[q] < queue[q].size();
        if (!any) return low;
        const uint64_t f = grow_from(0, from, tb, look);
        const uint64_t order[4] = {low & f, f & w & ~low, low & ~f, w & ~f & ~low};
        uint64_t out = 0;
        int c = 0;
        for (uint64_t m : order)
            for (; m && c < 4; m &= m - 1, c++) out |= m & (~m + 1);
        return out;
    }

    // The starting windows of a beam state's expansion from positions
    // `from`: the base window w0 (whose lookahead growth is the heuristic's
    // own window) and the base plus each of its `first` best first candidates.
    void starts(const size_t *from, uint64_t w0, int tb, int first, std::vector<uint64_t> &out) const
    {
        out.assign(1, w0);
        size_t h[64];
        std::copy(from, from + n, h);
        closure(w0, h, nullptr);
        std::vector<std::pair<double, uint64_t>> cs;
        candidates(w0, h, tb - __builtin_popcountll(w0), cs);
        std::stable_sort(cs.begin(), cs.end(), [](const auto &a, const auto &b) { return a.first > b.first; });
        for (int c = 0; c < (int)cs.size() && c < first && cs[c].first > 0; c++) out.push_back(w0 | cs[c].second);
    }
};

// The records [i, j) (no MEASURE / RESET among them) as a segment: per qubit,
// the records that act on it, in order.
int build_segment(const qsv_gate *gates, int64_t i, int64_t j, int n, Segment &sg)
{
    sg.gates = gates + i;
    sg.count = j - i;
    sg.n = n;
    sg.nq.resize(sg.count);
    sg.qs.assign(3 * sg.count, 0);
    sg.queue.assign(n, {});
    sg.head.assign(n, 0);
    for (int64_t g = 0; g < sg.count; g++) {
        const int k = record_qubits(sg.gates[g], &sg.qs[3 * g]);
        if (k < 0) return fail(QSV_ERR_UNSUPPORTED, "gate kind cannot be fused");
        sg.nq[g] = k;
        for (int x = 0; x < k; x++) {
            const int q = sg.qs[3 * g + x];
            if (q < 0 || q >= n) return fail(QSV_ERR_ARG, "qubit out of range");
            sg.queue[q].push_back(g);
        }
    }
    return QSV_OK;
}

bool is_barrier(const qsv_gate &r) { return r.kind == QSV_GATE_MEASURE || r.kind == QSV_GATE_RESET; }

// A few worker threads for the beam search: run(n, f) calls f(0) .. f(n - 1)
// on the workers and the calling thread and returns when all have returned.
class JobPool {
public:
    explicit JobPool(int workers)
    {
        for (int i = 0; i < workers; i++) th_.emplace_back([this] { loop(); });
    }
    ~JobPool()
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            quit_ = true;
        }
        cv_.notify_all();
        for (std::thread &t : th_) t.join();
    }
    void run(size_t n, const std::function<void(size_t)> &f)
    {
        if (th_.empty() || n <= 1) {
            for (size_t i = 0; i < n; i++) f(i);
            return;
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            f_ = &f;
            n_ = n;
            next_ = 0;
            done_ = 0;
            gen_++;
        }
        cv_.notify_all();
        work();
        std::unique_lock<std::mutex> lk(mu_);
        done_cv_.wait(lk, [&] { return done_ == n_; });
        f_ = nullptr;
    }

private:
    void work()
    {
        for (;;) {
            size_t i;
            const std::function<void(size_t)> *f;
            {
                std::lock_guard<std::mutex> lk(mu_);
                if (!f_ || next_ >= n_) return;
                i = next_++;
                f = f_;
            }
            (*f)(i);
            std::lock_guard<std::mutex> lk(mu_);
            if (++done_ == n_) done_cv_.notify_all();
        }
    }
    void loop()
    {
        int seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return quit_ || gen_ != seen; });
                if (quit_) return;
                seen = gen_;
            }
            work();
        }
    }
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    const std::function<void(size_t)> *f_ = nullptr;
    size_t n_ = 0, next_ = 0, done_ = 0;
    int gen_ = 0;
    bool quit_ = false;
};

}  // namespace

int fuse_order(const qsv_gate *gates, const std::vector<int64_t> &order, const double *mats, int max_k, int exact,
               bool shape_only, std::vector<PlanItem> &blocks)
{
    std::vector<qsv_gate> sub;
    sub.reserve(order.size());
    for (int64_t g : order) sub.push_back(gates[g]);
    int rc = plan_greedy(sub.data(), (int64_t)sub.size(), mats, max_k, exact, blocks, shape_only);
    if (rc) return rc;
    // every record is counted by some block (no-op records as the identity on
    // their qubit), one draw each
    int64_t counted = 0;
    for (const PlanItem &b : blocks) counted += b.ngates;
    if (counted != (int64_t)sub.size()) return fail(QSV_ERR_ARG, "internal: window records not all counted");
    return QSV_OK;
}

int best_fusion_order(const qsv_gate *gates, const std::vector<int64_t> &taken, const double *mats, int max_k,
                      int exact, int tries, uint64_t seed, std::vector<int64_t> &order, std::vector<PlanItem> &shapes)
{
    // greedy fusion depends on the order it sees the gates in: fuse the
    // window's gates in the closure's order, in program order and in `tries`
    // random dependency orders (as shapes) and keep the first order that gives
    // the fewest blocks
    order = taken;
    int rc = fuse_order(gates, order, mats, max_k, exact, true, shapes);
    if (rc) return rc;
    std::vector<int64_t> cand(taken);
    std::sort(cand.begin(), cand.end());
    std::vector<PlanItem> other;
    auto consider = [&]() {
        int r = fuse_order(gates, cand, mats, max_k, exact, true, other);
        if (r) return r;
        if (other.size() < shapes.size()) {
            shapes.swap(other);
            order = cand;
        }
        return QSV_OK;
    };
    if ((rc = consider())) return rc;
    if (tries <= 0) return QSV_OK;
    // random orders: each step takes a uniformly chosen record among those at
    // the front of every queue of their qubits (per-qubit queues in program
    // order), so every order respects the records' dependencies
    const std::vector<int64_t> prog(cand);
    const size_t m = prog.size();
    std::vector<std::vector<int>> queue(64);
    std::vector<std::array<int, 3>> rq(m);
    std::vector<int> nq(m);
    for (size_t i = 0; i < m; i++) {
        nq[i] = record_qubits(gates[prog[i]], rq[i].data());
        if (nq[i] < 0) return fail(QSV_ERR_UNSUPPORTED, "gate kind cannot be fused");
        for (int j = 0; j < nq[i]; j++) queue[rq[i][j]].push_back((int)i);
    }
    uint64_t st = seed * 0x9E3779B97F4A7C15ull + 0x2545F4914F6CDD1Dull;
    auto next = [&]() {
        st ^= st << 13;
        st ^= st >> 7;
        st ^= st << 17;
        return st;
    };
    std::vector<size_t> head(64);
    std::vector<int> ready;
    std::vector<char> in_ready(m);
    for (int t = 0; t < tries; t++) {
        std::fill(head.begin(), head.end(), 0);
        std::fill(in_ready.begin(), in_ready.end(), 0);
        ready.clear();
        auto front = [&](int i) {
            for (int j = 0; j < nq[i]; j++) {
                const std::vector<int> &qq = queue[rq[i][j]];
                if (head[rq[i][j]] >= qq.size() || qq[head[rq[i][j]]] != i) return false;
            }
            return true;
        };
        for (size_t i = 0; i < m; i++)
            if (front((int)i)) {
                ready.push_back((int)i);
                in_ready[i] = 1;
            }
        cand.clear();
        while (!ready.empty()) {
            const size_t pick = next() % ready.size();
            const int i = ready[pick];
            ready[pick] = ready.back();
            ready.pop_back();
            cand.push_back(prog[i]);
            for (int j = 0; j < nq[i]; j++) {
                const int q = rq[i][j];
                head[q]++;
                if (head[q] < queue[q].size()) {
                    const int f = queue[q][head[q]];
                    if (!in_ready[f] && front(f)) {
                        ready.push_back(f);
                        in_ready[f] = 1;
                    }
                }
            }
        }
        if (cand.size() != m) return fail(QSV_ERR_ARG, "internal: dependency order left records behind");
        if ((rc = consider())) return rc;
    }
    return QSV_OK;
}

int fuse_window(const qsv_gate *gates, const std::vector<int64_t> &taken, const double *mats, int max_k, int exact,
                bool shape_only, std::vector<PlanItem> &blocks, int tries, uint64_t seed)
{
    std::vector<int64_t> order;
    int rc = best_fusion_order(gates, taken, mats, max_k, exact, tries, seed, order, blocks);
    if (rc || shape_only) return rc;
    return fuse_order(gates, order, mats, max_k, exact, false, blocks);
}

int plan_window(const qsv_gate *gates, int64_t ng, const double *mats, int max_k, int exact, int tb, int n,
                std::vector<PlanItem> &plan, std::vector<int> &window, bool shape_only,
                const std::vector<uint64_t> *masks_in, std::vector<uint64_t> *masks_out)
{
    size_t mi = 0;
    if (masks_out) masks_out->clear();
    if (max_k < 4 || max_k > 5) return fail(QSV_ERR_ARG, "window planning fuses blocks of 4 or 5 qubits");
    if (n < 4 || n > 62) return fail(QSV_ERR_ARG, "window planning needs 4..62 qubits");
    tb = std::min(tb, n);
    plan.clear();
    window.clear();
    int wid = 0;
    int64_t i = 0;
    while (i < ng) {
        if (gates[i].kind == QSV_GATE_MEASURE || gates[i].kind == QSV_GATE_RESET) {
            PlanItem it;
            it.barrier = i;
            plan.push_back(std::move(it));
            window.push_back(wid++);
            i++;
            continue;
        }
        int64_t j = i;
        while (j < ng && gates[j].kind != QSV_GATE_MEASURE && gates[j].kind != QSV_GATE_RESET) j++;
        Segment sg;
        int rc0 = build_segment(gates, i, j, n, sg);
        if (rc0) return rc0;
        int64_t left = sg.count;
        std::vector<int64_t> taken;
        std::vector<PlanItem> blocks;
        while (left > 0) {
            size_t hc[64];
            uint64_t w;
            if (masks_in) {
                if (mi >= masks_in->size()) return fail(QSV_ERR_ARG, "fewer window masks than windows");
                w = (*masks_in)[mi++];
            } else {
                w = sg.grow(tb, kLookahead);
            }
            if (masks_out) masks_out->push_back(w);
            taken.clear();
            std::copy(sg.head.begin(), sg.head.end(), hc);
            const int64_t got = sg.closure(w, hc, &taken);
            std::copy(hc, hc + n, sg.head.begin());
            if (got == 0) return fail(QSV_ERR_ARG, "window planner made no progress");
            left -= got;
            for (int64_t &g : taken) g += i;
            int rc = fuse_window(gates, taken, mats, max_k, exact, shape_only, blocks);
            if (rc) return rc;
            for (PlanItem &b : blocks) {
                plan.push_back(std::move(b));
                window.push_back(wid);
            }
            wid++;
        }
        i = j;
    }
    return QSV_OK;
}

// Window stream: the windows of the plan in order, one at a time, so that the
// caller can run a window while later ones are planned.  With width 1 these
// are plan_window's windows (the lookahead growth from the positions the
// committed windows left).  With width > 1 a beam search looks for windows
// that serve later windows better than the most gates now: every state (the
// queue positions after a sequence of windows) expands into the `branch`
// windows that run the most records among the lookahead growth's own window
// and the completions from its `first` best first candidates; the `width`
// states that have run the most records survive each depth (duplicates of
// the same positions once).  A window is committed once the best state is
// `delay` windows past it (fewer for the first windows of a segment, so the
// first sweep starts early); states that disagree with it are dropped.  The
// expansions of a depth run on a few worker threads.  On the 30-qubit
// depth-200 benchmark circuit: 234 windows with width 1, 221 with width 4,
// branch 3, first 6, delay 4, 216 with width 12, branch 4, first 8, delay 6
// (another seed: 238, 223, 219); relabelled (each window grows from the four
// qubits the previous one left at bits 0..3): 188 with width 1, 180 with
// width 4, 176 with width 8 or 12 (another seed: 177).
struct WindowStream::Impl {
    const qsv_gate *gates = nullptr;
    int64_t ng = 0;
    int n = 0, tb = 0;
    BeamOpts b;
    int64_t next_rec = 0;              // first record not yet in a segment
    bool active = false;
    int64_t seg0 = 0;                  // the active segment's first record
    Segment sg;
    std::vector<size_t> committed;     // queue positions after the committed windows
    int64_t ncommit = 0;               // windows committed in the active segment
    int64_t total = 0;                 // windows committed in all
    uint64_t clow = 0;                 // qubits at bits 0..3 after the committed windows
    struct Step {
        uint64_t w, next_low;
    };
    struct State {
        std::vector<size_t> head;
        int64_t done = 0;
        uint64_t low = 0;              // qubits at bits 0..3 after the path
        std::vector<Step> path;        // windows not yet committed
    };
    std::vector<State> beam;
    std::deque<WindowUnit> out;
    std::unique_ptr<JobPool> pool;

    void commit(const Step &st)
    {
        const uint64_t w = st.w;
        WindowUnit u;
        u.mask = w;
        u.low = clow;
        u.next_low = st.next_low ? st.next_low : clow;
        clow = u.next_low;
        std::vector<int64_t> t;
        sg.closure(w, committed.data(), &t);
        u.taken.reserve(t.size());
        for (int64_t g : t) u.taken.push_back(g + seg0);
        out.push_back(std::move(u));
        ncommit++;
        total++;
    }

    int start_segment()
    {
        int64_t j = next_rec;
        while (j < ng && !is_barrier(gates[j])) j++;
        int rc = build_segment(gates, next_rec, j, n, sg);
        if (rc) return rc;
        seg0 = next_rec;
        next_rec = j;
        committed.assign(n, 0);
        ncommit = 0;
        clow = sg.base_window();
        beam.assign(1, State());
        beam[0].head.assign(n, 0);
        beam[0].low = clow;
        active = true;
        return QSV_OK;
    }

    int step()
    {
        std::vector<State> next;
        std::unordered_set<uint64_t> seen;
        // every state's starting windows, grown with lookahead on the pool
        struct Job {
            size_t st;
            uint64_t w0, w = 0;
            int64_t got = 0;
        };
        std::vector<Job> jobs;
        std::vector<uint64_t> ws;
        for (size_t i = 0; i < beam.size(); i++) {
            const uint64_t w0 = b.relabel ? beam[i].low : sg.base_window();
            if (b.width <= 1) ws.assign(1, w0);
            else sg.starts(beam[i].head.data(), w0, tb, b.first, ws);
            for (uint64_t w0s : ws) jobs.push_back({i, w0s});
        }
        if (b.width > 1 && !pool) {
            const int hw = (int)std::thread::hardware_concurrency();
            pool.reset(new JobPool(std::max(0, std::min(6, hw - 2))));
        }
        auto grow = [&](size_t j) {
            Job &jb = jobs[j];
            const size_t *from = beam[jb.st].head.data();
            jb.w = sg.grow_from(jb.w0, from, tb, kLookahead);
            size_t h[64];
            std::copy(from, from + n, h);
            jb.got = sg.closure(jb.w, h, nullptr);
        };
        if (pool) pool->run(jobs.size(), grow);
        else
            for (size_t j = 0; j < jobs.size(); j++) grow(j);
        // each state's `branch` best distinct windows
        for (size_t i = 0, j = 0; i < beam.size(); i++) {
            std::vector<std::pair<int64_t, uint64_t>> fin;
            for (; j < jobs.size() && jobs[j].st == i; j++) {
                bool dup = false;
                for (const auto &x : fin) dup |= x.second == jobs[j].w;
                if (!dup) fin.push_back({jobs[j].got, jobs[j].w});
            }
            std::stable_sort(fin.begin(), fin.end(), [](const auto &x, const auto &y) { return x.first > y.first; });
            const State &st = beam[i];
            for (int f = 0; f < (int)fin.size() && f < std::max(1, b.branch); f++) {
                const uint64_t w = fin[f].second;
                State c = st;
                const int64_t got = sg.closure(w, c.head.data(), nullptr);
                if (got == 0) continue;
                c.done += got;
                c.path.push_back({w, 0});
                uint64_t hsh = 1469598103934665603ull;
                for (size_t x : c.head) hsh = (hsh ^ x) * 1099511628211ull;
                if (!seen.insert(hsh).second) continue;
                next.push_back(std::move(c));
            }
        }
        if (next.empty()) return fail(QSV_ERR_ARG, "window planner made no progress");
        std::stable_sort(next.begin(), next.end(), [](const State &x, const State &y) { return x.done > y.done; });
        if (next[0].done == sg.count) {
            for (const Step &st : next[0].path) commit(st);
            out.back().last = true;
            beam.clear();
            active = false;
            return QSV_OK;
        }
        if ((int)next.size() > b.width) next.resize(b.width);
        if (b.relabel) {
            // the survivors' next starting qubits, on the pool
            auto choose = [&](size_t j) {
                State &c = next[j];
                // c.low is still the low the window grew from (its parent's)
                c.low = sg.next_low(c.head.data(), c.path.back().w, c.low, tb, kLookahead);
                c.path.back().next_low = c.low;
            };
            if (pool) pool->run(next.size(), choose);
            else
                for (size_t j = 0; j < next.size(); j++) choose(j);
        }
        const size_t delay = b.width <= 1 ? 0 : (size_t)std::min<int64_t>(b.delay, ncommit + 1);
        while (next[0].path.size() > delay) {
            const Step s0 = next[0].path[0];
            commit(s0);
            std::vector<State> keep;
            for (State &st : next)
                if (st.path[0].w == s0.w && st.path[0].next_low == s0.next_low) {
                    st.path.erase(st.path.begin());
                    keep.push_back(std::move(st));
                }
            next.swap(keep);
        }
        beam.swap(next);
        return QSV_OK;
    }
};

WindowStream::WindowStream(const qsv_gate *gates, int64_t ng, int n, int tb, const BeamOpts &b) : impl_(new Impl)
{
    impl_->gates = gates;
    impl_->ng = ng;
    impl_->n = n;
    impl_->tb = std::min(tb, n);
    impl_->b = b;
}

WindowStream::~WindowStream() { delete impl_; }

int64_t WindowStream::windows() const { return impl_->total; }

int WindowStream::next(WindowUnit &u, bool *end)
{
    Impl &m = *impl_;
    *end = false;
    if (m.n < 4 || m.n > 62) return fail(QSV_ERR_ARG, "window planning needs 4..62 qubits");
    while (m.out.empty()) {
        if (!m.active) {
            if (m.next_rec >= m.ng) {
                *end = true;
                return QSV_OK;
            }
            if (is_barrier(m.gates[m.next_rec])) {
                u = WindowUnit();
                u.barrier = m.next_rec++;
                m.total++;
                return QSV_OK;
            }
            int rc = m.start_segment();
            if (rc) return rc;
        }
        int rc = m.step();
        if (rc) return rc;
    }
    u = std::move(m.out.front());
    m.out.pop_front();
    return QSV_OK;
}

// A block's rows are its qubits' tile bits and its columns the others; the 32
// lanes of a half-wave take consecutive columns.  Widening with the lowest
// free qubits (0..3) would make the tile's lowest bits rows, so the lanes'
// LDS addresses would step by 2^(rows below them) and collide in banks;
// widening with the window's highest qubits keeps the low tile bits as
// columns.  Every qubit chosen is already in the window's tile.
std::vector<int> window_widening(const std::vector<PlanItem> &plan, const std::vector<int> &window, size_t i,
                                 int kq)
{
    std::vector<int> add;
    const PlanItem &it = plan[i];
    if (it.barrier >= 0 || (int)it.qubits.size() >= kq) return add;
    uint64_t u = 0xF;
    for (size_t j = i; j > 0 && window[j - 1] == window[i]; j--)
        for (int q : plan[j - 1].qubits) u |= uint64_t(1) << q;
    for (size_t j = i; j < plan.size() && window[j] == window[i]; j++)
        for (int q : plan[j].qubits) u |= uint64_t(1) << q;
    for (int q : it.qubits) u &= ~(uint64_t(1) << q);
    for (int b = 63; b >= 0 && (int)(it.qubits.size() + add.size()) < kq; b--)
        if (u >> b & 1) add.push_back(b);
    return add;
}

std::vector<std::pair<int, int>> route_within(const std::vector<int> &dest, uint64_t in)
{
    // every position whose data belongs at another position of the set sends
    // it there; the last position of each chain (whose data belongs outside)
    // sends its data to the chain's first position
    const int n = (int)dest.size();
    std::vector<int> src(n);
    for (int p = 0; p < n; p++) src[dest[p]] = p;
    auto inside = [&](int p) { return (in >> p & 1) != 0; };
    std::vector<std::pair<int, int>> moves;
    for (int p = 0; p < n; p++) {
        if (!inside(p) || dest[p] == p) continue;
        int to = dest[p];
        if (!inside(to)) {
            to = p;
            while (inside(src[to]) && src[to] != p) to = src[to];
        }
        if (to != p) moves.push_back({p, to});
    }
    return moves;
}

std::vector<std::vector<std::pair<int, int>>> qubit_routes(std::vector<int> dest, int tb)
{
    // each pass takes the positions 0..3 and walks the cycles of p -> dest[p]
    // (those through 0..3 first) until the tile is full; inside it every
    // position whose data belongs at another position of the tile sends it
    // there, and the last position of each chain (whose data belongs outside)
    // sends its data to the chain's first position.  A chain of m positions
    // puts m - 1 of them in place, so a cycle of c positions outside 0..3
    // takes ceil((c - 1) / (tb - 1)) passes.
    const int n = (int)dest.size();
    std::vector<std::vector<std::pair<int, int>>> out;
    std::vector<char> in(n), seen(n);
    for (;;) {
        bool done = true;
        for (int p = 0; p < n; p++) done &= dest[p] == p;
        if (done) return out;
        std::fill(in.begin(), in.end(), 0);
        std::fill(seen.begin(), seen.end(), 0);
        int cnt = 0;
        auto add = [&](int p) {
            if (in[p]) return true;
            if (cnt >= tb) return false;
            in[p] = 1;
            cnt++;
            return true;
        };
        for (int p = 0; p < std::min(4, n); p++) add(p);
        auto walk = [&](int p0) {
            int p = p0;
            do {
                seen[p] = 1;
                if (!add(p)) break;
                p = dest[p];
            } while (p != p0 && !seen[p]);
        };
        for (int p = 0; p < std::min(4, n); p++)
            if (!seen[p] && dest[p] != p) walk(p);
        for (int p = 4; p < n && cnt < tb; p++)
            if (!seen[p] && dest[p] != p) walk(p);
        uint64_t mask = 0;
        for (int p = 0; p < n; p++)
            if (in[p]) mask |= uint64_t(1) << p;
        std::vector<std::pair<int, int>> moves = route_within(dest, mask);
        if (moves.empty()) return {};       // cannot happen: a walk adds a position and its successor
        std::vector<int> nd(dest);
        for (const auto &m : moves) nd[m.second] = dest[m.first];
        dest.swap(nd);
        out.push_back(std::move(moves));
    }
}

}  // namespace qsv

using namespace qsv;

// Both host planners' output format (k_out .. nblocks as qsv_fuse_plan, plus
// window_out).
static int emit_plan(const std::vector<PlanItem> &plan, const std::vector<int> &window, int64_t max_blocks, int *k_out,
                     int *qubits_out, double *mats_out, int64_t *ngates_out, int *window_out, int64_t *nblocks)
{
    *nblocks = (int64_t)plan.size();
    if ((int64_t)plan.size() > max_blocks) return fail(QSV_ERR_ARG, "plan larger than the output buffers");
    for (size_t b = 0; b < plan.size(); b++) {
        const PlanItem &it = plan[b];
        const int k = it.barrier >= 0 ? 0 : (int)it.qubits.size();
        if (k_out) k_out[b] = k;
        if (ngates_out) ngates_out[b] = it.barrier >= 0 ? it.barrier : it.ngates;
        if (window_out) window_out[b] = window[b];
        if (qubits_out)
            for (int j = 0; j < 5; j++) qubits_out[b * 5 + j] = j < k ? it.qubits[j] : -1;
        if (mats_out && k > 0)
            for (size_t j = 0; j < it.u.size(); j++) mats_out[b * 2048 + j] = it.u[j];
    }
    return QSV_OK;
}

extern "C" int qsv_window_plan_beam(const qsv_gate *gates, int64_t ngates, const double *mats, int max_k, int exact,
                                    int tb, int n, int width, int branch, int first, int delay, int fuse_tries,
                                    int relabel, int64_t max_blocks, int *k_out, int *qubits_out, double *mats_out,
                                    int64_t *ngates_out, int *window_out, int64_t *nblocks)
{
    if ((!gates && ngates) || !nblocks) return fail(QSV_ERR_ARG, "null argument");
    if (tb < 4 || tb > kMaxPassTileBits) return fail(QSV_ERR_ARG, "window tile bits must be 4..13");
    if (tb < 7 && tb < n) return fail(QSV_ERR_ARG, "window tile bits must be >= 7 below the state size");
    if (max_k < 4 || max_k > 5) return fail(QSV_ERR_ARG, "window planning fuses blocks of 4 or 5 qubits");
    if (width < 1 || width > 64 || branch < 1 || branch > 16 || first < 0 || first > 32 || delay < 0 || delay > 64 ||
        fuse_tries < 0 || fuse_tries > 1024)
        return fail(QSV_ERR_ARG, "beam width 1..64, branch 1..16, first 0..32, delay 0..64, fusion tries 0..1024");
    BeamOpts bo;
    bo.width = width;
    bo.branch = branch;
    bo.first = first;
    bo.delay = delay;
    bo.relabel = relabel != 0;
    WindowStream ws(gates, ngates, n, tb, bo);
    std::vector<PlanItem> plan, blocks;
    std::vector<int> window;
    int wid = 0;
    for (;;) {
        WindowUnit u;
        bool end = false;
        int rc = ws.next(u, &end);
        if (rc) return rc;
        if (end) break;
        if (u.barrier >= 0) {
            PlanItem it;
            it.barrier = u.barrier;
            plan.push_back(std::move(it));
            window.push_back(wid++);
            continue;
        }
        if ((rc = fuse_window(gates, u.taken, mats, max_k, exact, false, blocks, fuse_tries, (uint64_t)u.taken[0])))
            return rc;
        for (PlanItem &b : blocks) {
            plan.push_back(std::move(b));
            window.push_back(wid);
        }
        wid++;
    }
    return emit_plan(plan, window, max_blocks, k_out, qubits_out, mats_out, ngates_out, window_out, nblocks);
}

extern "C" int qsv_window_plan(const qsv_gate *gates, int64_t ngates, const double *mats, int max_k, int exact,
                               int tb, int n, int64_t max_blocks, int *k_out, int *qubits_out, double *mats_out,
                               int64_t *ngates_out, int *window_out, int64_t *nblocks)
{
    if ((!gates && ngates) || !nblocks) return fail(QSV_ERR_ARG, "null argument");
    if (tb < 4 || tb > kMaxPassTileBits) return fail(QSV_ERR_ARG, "window tile bits must be 4..13");
    // a window holds qubits 0..3 and grows from there, so it needs room for
    // the widest record (3 qubits: Toffoli, Fredkin) beside them, unless it
    // holds the whole state
    if (tb < 7 && tb < n) return fail(QSV_ERR_ARG, "window tile bits must be >= 7 below the state size");
    std::vector<PlanItem> plan;
    std::vector<int> window;
    int rc = plan_window(gates, ngates, mats, max_k, exact, tb, n, plan, window);
    if (rc) return rc;
    return emit_plan(plan, window, max_blocks, k_out, qubits_out, mats_out, ngates_out, window_out, nblocks);
}
