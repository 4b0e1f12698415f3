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
#include <unordered_map>

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

    // Growth with lookahead: at each step, complete the growth greedily from
    // each of the `look` best candidates and take the one whose completed
    // window runs the most records (ties: the better first step).
    uint64_t grow(int tb, int look) const
    {
        uint64_t w = 0;
        for (int q = 0; q < std::min(4, n); q++) w |= uint64_t(1) << q;
        size_t h[64], h2[64];
        std::copy(head.begin(), head.end(), h);
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
                        std::copy(head.begin(), head.end(), h2);
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
};

}  // namespace

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
        int64_t left = sg.count;
        std::vector<int64_t> taken;
        std::vector<qsv_gate> sub, sub_prog;
        std::vector<PlanItem> blocks, blocks_prog;
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
            // greedy fusion depends on the order it sees the gates in: fuse
            // the window's gates in the closure's order and in program order
            // (as shapes) and keep the order that gives fewer blocks
            sub.clear();
            for (int64_t g : taken) sub.push_back(sg.gates[g]);
            std::sort(taken.begin(), taken.end());
            sub_prog.clear();
            for (int64_t g : taken) sub_prog.push_back(sg.gates[g]);
            int rc = plan_greedy(sub.data(), (int64_t)sub.size(), mats, max_k, exact, blocks, true);
            if (!rc) rc = plan_greedy(sub_prog.data(), (int64_t)sub_prog.size(), mats, max_k, exact, blocks_prog, true);
            if (rc) return rc;
            if (blocks_prog.size() < blocks.size()) sub.swap(sub_prog);
            if (!shape_only || blocks_prog.size() < blocks.size())
                rc = plan_greedy(sub.data(), (int64_t)sub.size(), mats, max_k, exact, blocks, shape_only);
            if (rc) return rc;
            // every record is counted by some block (no-op records as the
            // identity on their qubit), one draw each
            int64_t counted = 0;
            for (const PlanItem &b : blocks) counted += b.ngates;
            if (counted != (int64_t)sub.size()) return fail(QSV_ERR_ARG, "internal: window records not all counted");
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

}  // namespace qsv

using namespace qsv;

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
