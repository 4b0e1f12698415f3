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
//   1. grow the window from {0, 1, 2, 3}: each step adds the candidate -- a
//      single qubit, or the qubits a gate at the front of some qubit's queue
//      still misses -- that makes the most further gates runnable per added
//      qubit, until the window holds TB qubits or nothing gains;
//   2. take the runnable gates in dependency order (front-of-queue closure)
//      and fuse them into blocks of <= max_k qubits with plan_greedy, so every
//      block lies inside the window (plan_greedy ends by merging narrow blocks
//      into commuting neighbours, absorb_blocks);
//   3. emit the blocks as one window.
// On 30-qubit random circuits of depth 200 (TB = 13) this applies ~31 gates
// per sweep against ~21 for block-first grouping, with more (but cheaper per
// gate) blocks per sweep.  The gates a window takes commute with every gate
// left behind that precedes them in the list (they act on other qubits), so
// the state is the circuit's own up to that reordering.
#include "qsv_sched.h"

#include <algorithm>

namespace qsv {

namespace {

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
    int64_t closure(uint64_t w, size_t *h, std::vector<int64_t> *out) const
    {
        int64_t cnt = 0;
        bool changed = true;
        while (changed) {
            changed = false;
            for (int q = 0; q < n; q++) {
                if (!(w >> q & 1)) continue;
                while (h[q] < queue[q].size()) {
                    const int64_t g = queue[q][h[q]];
                    const int *gq = &qs[3 * g];
                    bool ok = true;
                    for (int j = 0; j < nq[g] && ok; j++) {
                        const int x = gq[j];
                        ok = (w >> x & 1) && h[x] < queue[x].size() && queue[x][h[x]] == g;
                    }
                    if (!ok) break;
                    for (int j = 0; j < nq[g]; j++) h[gq[j]]++;
                    if (out) out->push_back(g);
                    cnt++;
                    changed = true;
                }
            }
        }
        return cnt;
    }
};

}  // namespace

int plan_window(const qsv_gate *gates, int64_t ng, const double *mats, int max_k, int exact, int tb, int n,
                std::vector<PlanItem> &plan, std::vector<int> &window, bool shape_only)
{
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
        std::vector<qsv_gate> sub;
        std::vector<PlanItem> blocks;
        while (left > 0) {
            uint64_t w = 0;
            for (int q = 0; q < std::min(4, n); q++) w |= uint64_t(1) << q;
            size_t hw[64], hc[64];
            std::copy(sg.head.begin(), sg.head.end(), hw);
            sg.closure(w, hw, nullptr);
            while (__builtin_popcountll(w) < tb) {
                const int room = tb - __builtin_popcountll(w);
                double best = 0;
                uint64_t best_add = 0;
                auto consider = [&](uint64_t add) {
                    const int k = __builtin_popcountll(add);
                    if (k == 0 || k > room) return;
                    std::copy(hw, hw + n, hc);
                    const double gain = (double)sg.closure(w | add, hc, nullptr) / k;
                    if (gain > best) {
                        best = gain;
                        best_add = add;
                    }
                };
                // a candidate gains only if the record at the front of one of
                // its qubits needs nothing outside w | add
                for (int q = 0; q < n; q++) {
                    if ((w >> q & 1) || hw[q] >= sg.queue[q].size()) continue;
                    const int64_t g = sg.queue[q][hw[q]];
                    uint64_t add = 0;
                    for (int x = 0; x < sg.nq[g]; x++) add |= uint64_t(1) << sg.qs[3 * g + x];
                    add &= ~w;
                    consider(add);
                }
                if (best_add == 0) break;
                w |= best_add;
                sg.closure(w, hw, nullptr);
            }
            taken.clear();
            std::copy(sg.head.begin(), sg.head.end(), hc);
            const int64_t got = sg.closure(w, hc, &taken);
            std::copy(hc, hc + n, sg.head.begin());
            if (got == 0) return fail(QSV_ERR_ARG, "window planner made no progress");
            left -= got;
            sub.clear();
            for (int64_t g : taken) sub.push_back(sg.gates[g]);
            int rc = plan_greedy(sub.data(), (int64_t)sub.size(), mats, max_k, exact, blocks, shape_only);
            if (rc) return rc;
            // no-op records fuse to nothing; keep them counted (one draw each)
            int64_t counted = 0;
            for (const PlanItem &b : blocks) counted += b.ngates;
            if (blocks.empty()) {
                PlanItem id;                       // only no-op records: an identity block
                id.qubits = {0};
                id.u = {1, 0, 0, 0, 0, 0, 1, 0};
                blocks.push_back(std::move(id));
            }
            blocks.back().ngates += (int64_t)sub.size() - counted;
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
