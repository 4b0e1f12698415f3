// qsv_fusion.cpp -- fusing a gate list into dense k-qubit blocks.
//
// Greedy planner.  Every qubit has at most one "open" block (the latest block
// touching it).  For each gate with qubit set S:
//   * the open blocks touching S are candidates for merging; the planner
//     keeps the subset that fits in max_k qubits together with S and holds the
//     most gates, closes the others, merges the kept ones with the gate;
//   * a block that changes is moved to the end of the emission order.  That is
//     always legal: a block is open on all of its qubits, so every block
//     emitted after it acts on other qubits and commutes with it.
// MEASURE and RESET records close every block and run on their own.
//
// Block matrices are accumulated in FP64 on the host (gate matrices follow
// the CPU simulator's conventions, CPUQuantumGates.cpp); the device sees one
// FP32 MFMA sweep per block (qsv_fused.hip).  Small blocks are widened to 4
// qubits with identity on the lowest free qubits.
#include "qsv_fusion.h"
#include "qsv_sched.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <limits>
#include <mutex>
#include <string>
#include <thread>

namespace qsv {

using cd = std::complex<double>;

namespace {

const double kSq2Ref = 0.707106781186548;
const double kSq2Exact = 0.70710678118654752440;

// Local matrix of a gate: qs[j] is the qubit of row bit j.
int gate_matrix(const qsv_gate &r, const double *mats, int exact, int *qs, std::vector<cd> &m)
{
    auto set2 = [&](cd a, cd b, cd c, cd d) { m = {a, b, c, d}; };
    const double h = exact ? kSq2Exact : kSq2Ref;
    switch (r.kind) {
    case QSV_GATE_H: qs[0] = r.q0; set2(h, h, h, -h); return 1;
    case QSV_GATE_X: qs[0] = r.q0; set2(0, 1, 1, 0); return 1;
    case QSV_GATE_RX: {
        qs[0] = r.q0;
        const double c = std::cos(r.theta / 2), s = std::sin(r.theta / 2);
        const cd k(0, r.dagger ? s : -s);
        set2(c, k, k, c);
        return 1;
    }
    case QSV_GATE_RY: {
        qs[0] = r.q0;
        const double c = std::cos(r.theta / 2), s = std::sin(r.theta / 2), g = r.dagger ? -1 : 1;
        set2(c, -g * s, g * s, c);
        return 1;
    }
    case QSV_GATE_RZ: {
        qs[0] = r.q0;
        const double c = std::cos(r.theta / 2), s = std::sin(r.theta / 2), s1 = r.dagger ? -s : s;
        set2(cd(c, -s1), 0, 0, cd(c, s1));
        return 1;
    }
    case QSV_GATE_U1: {
        qs[0] = r.q0;
        const double *u = mats + 2 * (size_t)r.mat;
        set2(cd(u[0], u[1]), cd(u[2], u[3]), cd(u[4], u[5]), cd(u[6], u[7]));
        return 1;
    }
    case QSV_GATE_CNOT:
        if (r.q0 == r.q1) return 0;
        qs[0] = r.q1; qs[1] = r.q0;        // bit0 = target, bit1 = control
        m.assign(16, 0);
        m[0 * 4 + 0] = m[1 * 4 + 1] = 1;
        m[2 * 4 + 3] = m[3 * 4 + 2] = 1;
        return 2;
    case QSV_GATE_CZ: case QSV_GATE_CR: {
        if (r.q0 == r.q1) return 0;
        qs[0] = r.q0; qs[1] = r.q1;
        m.assign(16, 0);
        m[0] = m[5] = m[10] = 1;
        if (r.kind == QSV_GATE_CZ) m[15] = -1;
        else m[15] = cd(std::cos(r.theta), r.dagger ? -std::sin(r.theta) : std::sin(r.theta));
        return 2;
    }
    case QSV_GATE_ISWAP: case QSV_GATE_SQISWAP: {
        if (r.q0 == r.q1) return 0;
        qs[0] = r.q0; qs[1] = r.q1;
        m.assign(16, 0);
        m[0] = m[15] = 1;
        const cd k(0, r.dagger ? -1 : 1);
        if (r.kind == QSV_GATE_ISWAP) {
            m[1 * 4 + 2] = m[2 * 4 + 1] = k;
        } else {
            m[1 * 4 + 1] = m[2 * 4 + 2] = h;
            m[1 * 4 + 2] = m[2 * 4 + 1] = k * h;
        }
        return 2;
    }
    case QSV_GATE_U2: {
        qs[0] = r.q1; qs[1] = r.q0;        // row = 2*bit(q0) + bit(q1)
        const double *u = mats + 2 * (size_t)r.mat;
        m.resize(16);
        for (int i = 0; i < 16; i++) m[i] = cd(u[2 * i], u[2 * i + 1]);
        return 2;
    }
    case QSV_GATE_TOFFOLI:
        qs[0] = r.q2; qs[1] = r.q0; qs[2] = r.q1;   // bit0 = target
        m.assign(64, 0);
        for (int i = 0; i < 6; i++) m[i * 8 + i] = 1;
        m[6 * 8 + 7] = m[7 * 8 + 6] = 1;
        return 3;
    case QSV_GATE_CSWAP:
        qs[0] = r.q0; qs[1] = r.q1; qs[2] = r.q2;   // bit0 = control
        m.assign(64, 0);
        for (int i = 0; i < 8; i++) m[i * 8 + i] = 1;
        m[3 * 8 + 3] = m[5 * 8 + 5] = 0;
        m[3 * 8 + 5] = m[5 * 8 + 3] = 1;
        return 3;
    default:
        return -1;
    }
}

// Apply an m-qubit matrix g (on block positions pos[0..m-1]) to every column
// of the block matrix B (dimension 2^nq): B <- G B.
// Blocks have at most 5 qubits (D <= 32).  The rows of B are combined whole,
// with the complex products written out in real arithmetic: std::complex's
// operator* goes through the C99 Annex G helper (__muldc3), which made this
// loop most of the planning time of k = 5 plans.
void apply_small(std::vector<cd> &B, int nq, const int *pos, int m, const std::vector<cd> &g)
{
    const int D = 1 << nq, d = 1 << m;
    int mask = 0;
    for (int j = 0; j < m; j++) mask |= 1 << pos[j];
    int off[32];
    for (int r = 0; r < d; r++) {
        int o = 0;
        for (int j = 0; j < m; j++)
            if (r >> j & 1) o |= 1 << pos[j];
        off[r] = o;
    }
    double *b = reinterpret_cast<double *>(B.data());
    const double *gm = reinterpret_cast<const double *>(g.data());
    double acc[2 * 32 * 32];
    for (int base = 0; base < D; base++) {
        if (base & mask) continue;
        for (int r = 0; r < d; r++) {
            double *a = acc + 2 * r * D;
            for (int col = 0; col < 2 * D; col++) a[col] = 0;
            for (int c = 0; c < d; c++) {
                const double gr = gm[2 * (r * d + c)], gi = gm[2 * (r * d + c) + 1];
                if (gr == 0 && gi == 0) continue;
                const double *x = b + 2 * (size_t)(base | off[c]) * D;
                for (int col = 0; col < D; col++) {
                    a[2 * col] += gr * x[2 * col] - gi * x[2 * col + 1];
                    a[2 * col + 1] += gr * x[2 * col + 1] + gi * x[2 * col];
                }
            }
        }
        for (int r = 0; r < d; r++)
            std::memcpy(b + 2 * (size_t)(base | off[r]) * D, acc + 2 * r * D, sizeof(double) * 2 * D);
    }
}

struct Block {
    std::vector<int> qs;
    std::vector<cd> m;
    int64_t seq = 0;
    int64_t ngates = 0;
    bool open = true;
    bool alive = true;
    int64_t barrier = -1;      // record index for MEASURE/RESET items
};

int index_of(const std::vector<int> &v, int q)
{
    for (size_t i = 0; i < v.size(); i++)
        if (v[i] == q) return (int)i;
    return -1;
}

// Re-express block b on the qubit list `qs` (a superset of b.qs).
void widen(Block &b, const std::vector<int> &qs)
{
    if (qs == b.qs) return;
    const int D = 1 << qs.size();
    std::vector<cd> I((size_t)D * D, 0);
    for (int i = 0; i < D; i++) I[(size_t)i * D + i] = 1;
    if (!b.qs.empty()) {
        std::vector<int> pos(b.qs.size());
        for (size_t j = 0; j < b.qs.size(); j++) pos[j] = index_of(qs, b.qs[j]);
        apply_small(I, (int)qs.size(), pos.data(), (int)b.qs.size(), b.m);
    }
    b.qs = qs;
    b.m.swap(I);
}

}  // namespace

int plan_greedy(const qsv_gate *gates, int64_t ng, const double *mats, int max_k, int exact,
                std::vector<PlanItem> &plan, bool shape_only)
{
    if (max_k < 1 || max_k > 5) return fail(QSV_ERR_ARG, "fusion size must be 1..5");
    std::vector<Block> blocks;
    int cur[64];
    for (int &c : cur) c = -1;
    int64_t clock = 0;
    auto close = [&](int bi) {
        Block &b = blocks[bi];
        if (!b.open) return;
        b.open = false;
        for (int q : b.qs)
            if (cur[q] == bi) cur[q] = -1;
    };
    for (int64_t gi = 0; gi < ng; gi++) {
        const qsv_gate &r = gates[gi];
        if (r.kind == QSV_GATE_MEASURE || r.kind == QSV_GATE_RESET) {
            for (size_t bi = 0; bi < blocks.size(); bi++) close((int)bi);
            Block b;
            b.barrier = gi;
            b.seq = ++clock;
            b.open = false;
            blocks.push_back(b);
            continue;
        }
        int qs[3];
        std::vector<cd> gm;
        int nq = gate_matrix(r, mats, exact, qs, gm);
        if (nq < 0) return fail(QSV_ERR_UNSUPPORTED, "gate kind cannot be fused");
        if (nq == 0) {
            // a no-op record (CNOT(q, q) and the like) changes no amplitude but
            // the reference executor still draws once for it: plan it as the
            // identity on q0 so that its block counts it
            qs[0] = r.q0;
            nq = 1;
            gm = {1, 0, 0, 1};
        }
        // candidate blocks: open blocks on the gate's qubits
        std::vector<int> cand;
        for (int j = 0; j < nq; j++)
            if (cur[qs[j]] >= 0 && std::find(cand.begin(), cand.end(), cur[qs[j]]) == cand.end())
                cand.push_back(cur[qs[j]]);
        // best subset that fits with the gate
        int best = -1;
        int64_t best_gates = -1;
        size_t best_width = 99;
        for (int sub = 0; sub < (1 << cand.size()); sub++) {
            std::vector<int> u(qs, qs + nq);
            int64_t gcount = 0;
            for (size_t c = 0; c < cand.size(); c++)
                if (sub >> c & 1) {
                    for (int q : blocks[cand[c]].qs)
                        if (index_of(u, q) < 0) u.push_back(q);
                    gcount += blocks[cand[c]].ngates;
                }
            if ((int)u.size() > std::max(max_k, nq)) continue;   // a wide gate may stand alone
            if (gcount > best_gates || (gcount == best_gates && u.size() < best_width)) {
                best = sub;
                best_gates = gcount;
                best_width = u.size();
            }
        }
        if (best < 0) return fail(QSV_ERR_ARG, "gate wider than the fusion size");
        std::vector<int> keep;
        for (size_t c = 0; c < cand.size(); c++) {
            if (best >> c & 1) keep.push_back(cand[c]);
            else close(cand[c]);
        }
        // target block: the first kept one, or a new block
        int ti;
        if (keep.empty()) {
            Block b;
            blocks.push_back(b);
            ti = (int)blocks.size() - 1;
        } else {
            ti = keep[0];
        }
        std::vector<int> u = blocks[ti].qs;
        for (size_t c = 1; c < keep.size(); c++)
            for (int q : blocks[keep[c]].qs)
                if (index_of(u, q) < 0) u.push_back(q);
        for (int j = 0; j < nq; j++)
            if (index_of(u, qs[j]) < 0) u.push_back(qs[j]);
        if (shape_only) blocks[ti].qs = u;
        else widen(blocks[ti], u);
        for (size_t c = 1; c < keep.size(); c++) {
            Block &o = blocks[keep[c]];
            std::vector<int> pos(o.qs.size());
            for (size_t j = 0; j < o.qs.size(); j++) pos[j] = index_of(u, o.qs[j]);
            if (!shape_only) apply_small(blocks[ti].m, (int)u.size(), pos.data(), (int)o.qs.size(), o.m);
            blocks[ti].ngates += o.ngates;
            o.alive = false;
            o.open = false;
        }
        int pos[3];
        for (int j = 0; j < nq; j++) pos[j] = index_of(u, qs[j]);
        if (!shape_only) apply_small(blocks[ti].m, (int)u.size(), pos, nq, gm);
        Block &t = blocks[ti];
        t.ngates += 1;
        t.seq = ++clock;
        t.open = true;
        t.alive = true;
        for (int q : t.qs) cur[q] = ti;
    }
    std::vector<int> order;
    for (size_t i = 0; i < blocks.size(); i++)
        if (blocks[i].alive) order.push_back((int)i);
    std::sort(order.begin(), order.end(), [&](int x, int y) { return blocks[x].seq < blocks[y].seq; });
    plan.clear();
    for (int i : order) {
        PlanItem it;
        it.barrier = blocks[i].barrier;
        it.qubits = blocks[i].qs;
        it.ngates = blocks[i].ngates;
        it.u.resize(2 * blocks[i].m.size());
        for (size_t j = 0; j < blocks[i].m.size(); j++) {
            it.u[2 * j] = blocks[i].m[j].real();
            it.u[2 * j + 1] = blocks[i].m[j].imag();
        }
        plan.push_back(std::move(it));
    }
    absorb_blocks(plan, max_k, shape_only);
    return QSV_OK;
}

// Widen a block to `kq` qubits with identity on the lowest qubits it does not
// use (U (x) I, the new qubits taking the high row bits).
void widen_to(PlanItem &it, int kq)
{
    uint64_t used = 0;
    for (int q : it.qubits) used |= uint64_t(1) << q;
    std::vector<int> add;
    for (int b = 0; (int)(it.qubits.size() + add.size()) < kq; b++)
        if (!(used >> b & 1)) add.push_back(b);
    widen_with(it, add);
}

// Make qubit q (if the block holds it) the block's row bit 0, permuting the
// matrix's rows and columns to match.  The sweep kernels move a block's data
// as pairs of amplitudes that differ in qubit 0, along rows when qubit 0 is
// row bit 0.
void move_qubit_first(PlanItem &it, int q)
{
    const int k = (int)it.qubits.size();
    int j = -1;
    for (int i = 0; i < k; i++)
        if (it.qubits[i] == q) j = i;
    if (j <= 0) return;
    std::vector<int> order{j};                    // new row bit b <- old row bit order[b]
    for (int i = 0; i < k; i++)
        if (i != j) order.push_back(i);
    std::vector<int> qs(k);
    for (int b = 0; b < k; b++) qs[b] = it.qubits[order[b]];
    if (it.u.empty()) {                           // a shape: qubits only
        it.qubits.swap(qs);
        return;
    }
    const int d = 1 << k;
    auto old_index = [&](int x) {
        int o = 0;
        for (int b = 0; b < k; b++)
            if (x >> b & 1) o |= 1 << order[b];
        return o;
    };
    std::vector<double> u(it.u.size());
    for (int r = 0; r < d; r++)
        for (int c = 0; c < d; c++) {
            const size_t src = (size_t)old_index(r) * d + old_index(c), dst = (size_t)r * d + c;
            u[2 * dst] = it.u[2 * src];
            u[2 * dst + 1] = it.u[2 * src + 1];
        }
    it.qubits.swap(qs);
    it.u.swap(u);
}

// Widen a block with identity on the qubits `add` (not in the block).
void widen_with(PlanItem &it, const std::vector<int> &add)
{
    const int k = (int)it.qubits.size();
    const int kq = k + (int)add.size();
    if (add.empty()) return;
    std::vector<int> qs = it.qubits;
    qs.insert(qs.end(), add.begin(), add.end());
    if (it.u.empty()) {                           // a shape: qubits only
        it.qubits = qs;
        return;
    }
    const int d = 1 << k, D = 1 << kq;
    std::vector<double> w(2 * (size_t)D * D, 0.0);
    for (int hi = 0; hi < D / d; hi++)
        for (int r = 0; r < d; r++)
            for (int c = 0; c < d; c++) {
                const size_t R = (size_t)hi * d + r, C = (size_t)hi * d + c;
                w[2 * (R * D + C)] = it.u[2 * ((size_t)r * d + c)];
                w[2 * (R * D + C) + 1] = it.u[2 * ((size_t)r * d + c) + 1];
            }
    it.qubits = qs;
    it.u.swap(w);
}

namespace {

uint64_t qubit_mask(const PlanItem &it)
{
    uint64_t m = 0;
    for (int q : it.qubits) m |= uint64_t(1) << q;
    return m;
}

std::vector<cd> complex_of(const std::vector<double> &u)
{
    std::vector<cd> m(u.size() / 2);
    for (size_t i = 0; i < m.size(); i++) m[i] = cd(u[2 * i], u[2 * i + 1]);
    return m;
}

// The block "first, then second" on the union of their qubits (second's
// qubits first, then first's others); only the qubits if shape_only.
PlanItem product_block(const PlanItem &first, const PlanItem &second, bool shape_only)
{
    std::vector<int> u = second.qubits;
    for (int q : first.qubits)
        if (index_of(u, q) < 0) u.push_back(q);
    if (shape_only) {
        PlanItem out;
        out.qubits = u;
        out.ngates = first.ngates + second.ngates;
        return out;
    }
    Block b;
    b.qs = first.qubits;
    b.m = complex_of(first.u);
    widen(b, u);
    std::vector<int> pos(second.qubits.size());
    for (size_t j = 0; j < pos.size(); j++) pos[j] = index_of(u, second.qubits[j]);
    apply_small(b.m, (int)u.size(), pos.data(), (int)pos.size(), complex_of(second.u));
    PlanItem out;
    out.qubits = u;
    out.u.resize(2 * b.m.size());
    for (size_t j = 0; j < b.m.size(); j++) {
        out.u[2 * j] = b.m[j].real();
        out.u[2 * j + 1] = b.m[j].imag();
    }
    out.ngates = first.ngates + second.ngates;
    return out;
}

}  // namespace

// Greedy fusion closes a block as soon as a gate cannot join it, which leaves
// narrow blocks behind (a window of a 30-qubit random circuit: ~20% single-
// and ~25% two-qubit blocks), and in a multi-block sweep every block costs the
// same pass over the tile in LDS whatever its width.  A block commutes with
// every block on disjoint qubits, so it can move forward or backward until it
// meets one that shares a qubit; merge it into the reachable block (the one it
// meets included) whose union with it is narrowest, if that is <= max_k
// qubits.  Barrier records are never crossed.
void absorb_blocks(std::vector<PlanItem> &plan, int max_k, bool shape_only)
{
    bool changed = true;
    while (changed) {
        changed = false;
        size_t b = 0;
        while (b < plan.size()) {
            if (plan[b].barrier >= 0) {
                b++;
                continue;
            }
            const uint64_t mb = qubit_mask(plan[b]);
            size_t best = b;
            int best_w = max_k + 1;
            for (size_t c = b + 1; c < plan.size() && plan[c].barrier < 0; c++) {
                const uint64_t mc = qubit_mask(plan[c]);
                const int w = __builtin_popcountll(mb | mc);
                if (w < best_w) {
                    best = c;
                    best_w = w;
                }
                if (mc & mb) break;
            }
            for (size_t c = b; c-- > 0 && plan[c].barrier < 0;) {
                const uint64_t mc = qubit_mask(plan[c]);
                const int w = __builtin_popcountll(mb | mc);
                if (w < best_w) {
                    best = c;
                    best_w = w;
                }
                if (mc & mb) break;
            }
            if (best == b) {
                b++;
                continue;
            }
            if (best > b) plan[best] = product_block(plan[b], plan[best], shape_only);
            else plan[best] = product_block(plan[best], plan[b], shape_only);
            plan.erase(plan.begin() + (std::ptrdiff_t)b);   // the next block moves into slot b
            changed = true;
        }
    }
}

namespace {

constexpr int kKqMin = 4;

// Widening and kernel choice of a plan's blocks: window plans widen their
// small blocks with the window's highest qubits (window_widening), the rest
// widen to >= 4 qubits with the lowest free ones where an MFMA tile fits;
// other blocks use the FP64 kernel.  mfma[i] = 1 for MFMA blocks.
void prepare_blocks(const qsv_state *s, std::vector<PlanItem> &plan, const std::vector<int> &window,
                    std::vector<char> &mfma)
{
    if (!window.empty() && s->n >= kKqMin + 6)
        for (size_t i = 0; i < plan.size(); i++)
            if (plan[i].barrier < 0) widen_with(plan[i], window_widening(plan, window, i, kKqMin));
    mfma.assign(plan.size(), 0);
    for (size_t i = 0; i < plan.size(); i++) {
        PlanItem &it = plan[i];
        if (it.barrier >= 0) continue;
        const int kq = std::max(kKqMin, (int)it.qubits.size());
        if (s->n < kq + 6) continue;
        widen_to(it, kKqMin);
        move_qubit_first(it, 0);
        mfma[i] = 1;
    }
}

// The window heuristic's launches: one sweep per window (several beyond the
// pass kernel's block limit).
void window_launches(const qsv_state *s, const std::vector<PlanItem> &plan, const std::vector<int> &window,
                     const std::vector<char> &mfma, const qsv_run_opts &o, std::vector<Launch> &launches)
{
    const int tb_min = o.pass_tb ? o.pass_tb : 11;
    for (size_t i = 0; i < plan.size();) {
        size_t j = i + 1;
        while (j < plan.size() && plan[i].barrier < 0 && plan[j].barrier < 0 && window[j] == window[i]) j++;
        for (size_t b = i; b < j; b += kMaxPassBlocks) {
            Launch L;
            L.first = b;
            L.count = (int)std::min<size_t>(kMaxPassBlocks, j - b);
            if (plan[b].barrier >= 0) {
                L.count = 1;
            } else if (L.count == 1) {
                const int kq = (int)plan[b].qubits.size();
                L.tc = mfma[b] ? std::max(6, std::min(kDefaultTileCols, s->n - kq)) : 0;
            } else {
                std::vector<const std::vector<int> *> q;
                for (int g = 0; g < L.count; g++) q.push_back(&plan[b + g].qubits);
                L.tb = std::max(pass_tile_bits(q, s->n), std::min(tb_min, s->n));
            }
            launches.push_back(L);
        }
        i = j;
    }
}

// Operands of every MFMA block (float offsets op_off[i] into ops), built on
// up to `threads` threads.
void build_operands(const std::vector<PlanItem> &plan, const std::vector<char> &mfma, int split, int threads,
                    std::vector<size_t> &op_off, std::vector<float> &ops)
{
    op_off.assign(plan.size(), 0);
    size_t total = 0;
    for (size_t i = 0; i < plan.size(); i++) {
        if (!mfma[i]) continue;
        op_off[i] = total;
        total += fused_operand_floats((int)plan[i].qubits.size(), split);
    }
    ops.assign(total, 0.0f);
    const size_t nth = std::max<size_t>(1, std::min<size_t>({(size_t)threads, std::thread::hardware_concurrency(),
                                                             plan.size() / 64 + 1}));
    auto work = [&](size_t t) {
        for (size_t i = t; i < plan.size(); i += nth)
            if (mfma[i]) fused_operands((int)plan[i].qubits.size(), plan[i].u.data(), split, ops.data() + op_off[i]);
    };
    if (nth == 1) {
        work(0);
        return;
    }
    std::vector<std::thread> pool;
    for (size_t t = 0; t < nth; t++) pool.emplace_back(work, t);
    for (std::thread &th : pool) th.join();
}

// Issues a plan's launches on the state's stream; operands of block i at
// dops + op_off[i].  Barrier records (MEASURE / RESET) draw and measure on
// the host's side of the stream.
int exec_launches(qsv_state *s, const qsv_gate *gates, const std::vector<PlanItem> &plan,
                  const std::vector<char> &mfma, const std::vector<Launch> &launches, const float *dops,
                  const std::vector<size_t> &op_off, const qsv_run_opts &o, int32_t *rng, int *outcomes,
                  Profiler *prof, int64_t *nblocks)
{
    const double N = (double)s->dim;
    int rc = QSV_OK;
    for (const Launch &L : launches) {
        if (rc) break;
        if (L.count == 0) {
            if ((rc = prof->begin())) break;
            rc = launch_pass_dev(s, {}, {}, 0, kMaxPassTileBits, 0, &L.moves);
            if (!rc) rc = prof->end(QSV_K_PASS, 2 * 8 * N, 0);
            continue;
        }
        const PlanItem &it0 = plan[L.first];
        if (it0.barrier >= 0) {
            const qsv_gate &r = gates[it0.barrier];
            // one draw for the measurement, one more for a Reset's NOT on outcome 1
            double draw = 1.0;
            if (rng) {
                *rng = qsv_pm_next(*rng);
                draw = (double)*rng / 2147483647;
            }
            int out = 0;
            if ((rc = prof->begin())) break;
            rc = qsv_measure(s, r.q0, draw, &out);
            if (!rc && r.kind == QSV_GATE_MEASURE && outcomes) outcomes[r.q1] = out;
            if (!rc && r.kind == QSV_GATE_RESET && out) {
                if (rng) *rng = qsv_pm_next(*rng);
                rc = qsv_x(s, r.q0, 0);
            }
            if (!rc) rc = prof->end(QSV_K_READOUT, 2 * 8 * N, 0);
            continue;
        }
        double flops = 0;
        int64_t gates_here = 0;
        std::vector<const std::vector<int> *> qsets;
        std::vector<const float *> mp;
        for (int j = 0; j < L.count; j++) {
            const PlanItem &it = plan[L.first + j];
            gates_here += it.ngates;
            qsets.push_back(&it.qubits);
            mp.push_back(dops + op_off[L.first + j]);
            flops += 8.0 * N * (1 << it.qubits.size());     // FP32 MFMA product (the BF16 correction is extra)
        }
        if (rng) *rng = qsv_pm_jump(*rng, (uint64_t)gates_here);
        if ((rc = prof->begin())) break;
        int cls = QSV_K_FUSED;
        bool single = L.count == 1 && L.moves.empty();
        int tb = L.tb;
        if (single && !mfma[L.first]) {
            rc = launch_block(s, it0.qubits.data(), (int)it0.qubits.size(), it0.u.data(), 0);
            cls = QSV_K_BLOCK64;
        } else if (single) {
            const int tc = L.tc ? L.tc : kDefaultTileCols;
            tb = (int)it0.qubits.size() + tc;
            rc = launch_fused_dev(s, it0.qubits.data(), (int)it0.qubits.size(), mp[0], o.split, tc);
        } else {
            rc = launch_pass_dev(s, qsets, mp, o.split, tb, o.pass_kernel, L.moves.empty() ? nullptr : &L.moves);
            cls = QSV_K_PASS;
        }
        int64_t tag = -1;
        if (!rc && prof->on() && (cls == QSV_K_FUSED || cls == QSV_K_PASS)) {
            float f[QSV_CM_FEATURES];
            launch_features(s->n, o.split, qsets, single, tb, f);
            tag = (int64_t)(s->tlog_feats.size() / QSV_CM_FEATURES);
            s->tlog_feats.insert(s->tlog_feats.end(), f, f + QSV_CM_FEATURES);
        }
        if (!rc) rc = prof->end(cls, 2 * 8 * N, flops, tag);
        *nblocks += L.count;
    }
    return rc;
}

// ---- streamed window schedules ---------------------------------------------
//
// The window schedules (the learned one and the window heuristic) are planned
// and run as a pipeline: a planning thread takes the windows from a
// WindowStream (for the learned schedule, the beam search), fuses each
// window's records, lets the cost model choose the window's block size (4 or
// 5) and launch split, and builds its operands; the calling thread copies
// each window's operands through a pinned ring into a device ring and issues
// its sweeps.  So all but the first window's planning overlaps the device's
// work on earlier windows.  The operand copies are ordered on the state's
// stream after every earlier sweep, so a ring slot is only overwritten once the
// sweeps that read it have run; a host slot is reused once the event recorded
// after its window's sweeps has completed.

// k = 5 only on a predicted gain above the model's per-launch error (at 5%
// the model picked k = 5 for windows that ran slower)
constexpr double kK5Margin = 0.10;
// smallest state the window schedules relabel by default (held-out circuits
// at 26 qubits ran 0.80-1.06x as fast relabelled, at 28 qubits 1.04-1.06x)
constexpr int kRelabelFrom = 28;
// random dependency orders the learned schedule's fusion tries per window
// (~12 us each at k = 4 and 5 on ~40 records): 24 where a window's sweep
// takes >= 2 ms (>= 29 qubits), 8 at 28 qubits (~1 ms), none below, so that
// the planning thread stays ahead of the sweeps
int fuse_tries(int n) { return n >= 29 ? 24 : n >= 28 ? 8 : 0; }

struct ReadyWindow {
    int64_t barrier = -1;
    bool end = false;
    int rc = QSV_OK;
    std::string err;
    std::vector<PlanItem> plan;
    std::vector<char> mfma;
    std::vector<Launch> launches;
    std::vector<size_t> op_off;
    std::vector<float> ops;
    int k = 4;
    double pred = 0;
};

class WindowQueue {
public:
    explicit WindowQueue(size_t cap) : cap_(cap) {}
    void push(ReadyWindow &&w)
    {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return q_.size() < cap_ || stop_; });
        if (stop_) return;
        q_.push_back(std::move(w));
        cv_.notify_all();
    }
    ReadyWindow pop()
    {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return !q_.empty(); });
        ReadyWindow w = std::move(q_.front());
        q_.pop_front();
        cv_.notify_all();
        return w;
    }
    void stop()
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
        cv_.notify_all();
    }
    bool stopped()
    {
        std::lock_guard<std::mutex> lk(mu_);
        return stop_;
    }

private:
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<ReadyWindow> q_;
    size_t cap_;
    bool stop_ = false;
};

void produce_windows(qsv_state *s, const qsv_gate *gates, int64_t ngates, const double *mats, const qsv_run_opts &o,
                     bool learned, WindowQueue *q)
{
    auto send_error = [&](int rc) {
        ReadyWindow w;
        w.rc = rc;
        w.err = qsv_last_error();
        q->push(std::move(w));
    };
    // a new thread's current device is device 0: plan on the state's
    if (hipSetDevice(s->device) != hipSuccess) {
        send_error(fail(QSV_ERR_DEVICE, "cannot select the device"));
        return;
    }
    // the beam search costs host time per window (~0.5 ms at width 4 on 29
    // qubits; with relabelling ~2.8 ms at width 8 on 30, ~4.4 ms at width 12
    // for the same windows): it is sized to stay ahead of the device, whose
    // window sweeps take ~2 ms at 29 qubits and ~4.7 ms at 30
    BeamOpts bo;
    if (learned && s->n >= 30) {
        bo.width = 8;
        bo.branch = 4;
        bo.first = 8;
        bo.delay = 6;
    } else if (learned && s->n >= 29) {
        bo.width = 4;
        bo.branch = 3;
        bo.first = 6;
        bo.delay = 4;
    }
    // relabelling: the window stream plans in logical qubits; phys[x] is the
    // bit position of qubit x during the current window and at[p] the qubit
    // at position p.  Each window's blocks are renamed to positions, and its
    // last sweep stores the tile with the window's next_low qubits moved to
    // positions 0..3 (exchanged with the qubits there).  Before a MEASURE /
    // RESET and at the end, sweeps that only move qubits bring every qubit
    // back to its own position.
    const int n = s->n;
    // (by default from 28 qubits: below, a window sweep takes < 1 ms and the
    // planning thread, which relabelling gives one more growth per window,
    // falls behind the device; fixed_low = -1 relabels from 14 qubits)
    const int relabel_from = o.fixed_low < 0 ? kMaxPassTileBits + 1 : kRelabelFrom;
    bo.relabel = o.fixed_low <= 0 && n >= relabel_from && (o.pass_tb == 0 || o.pass_tb == kMaxPassTileBits);
    std::vector<int> phys(n), at(n);
    for (int x = 0; x < n; x++) phys[x] = at[x] = x;
    auto restore = [&]() {
        bool moved = false;
        for (int p = 0; p < n; p++) moved |= at[p] != p;
        if (!moved) return;
        ReadyWindow w;
        for (auto &mv : qubit_routes(at, kMaxPassTileBits)) {
            Launch L;
            L.count = 0;
            L.tb = kMaxPassTileBits;
            L.moves = std::move(mv);
            w.launches.push_back(std::move(L));
        }
        for (int x = 0; x < n; x++) phys[x] = at[x] = x;
        q->push(std::move(w));
    };
    auto rename = [&](std::vector<PlanItem> &items) {
        for (PlanItem &it : items)
            for (int &x : it.qubits) x = phys[x];
    };
    const int fixed_k = std::max(4, std::min(o.fuse_k, 5));
    WindowStream ws(gates, ngates, s->n, o.pass_tb ? o.pass_tb : kMaxPassTileBits, bo);
    std::vector<PlanItem> shape, shape5;
    std::vector<char> shape_mfma;
    std::vector<Launch> shape_launches, launches5;
    std::vector<int64_t> order4, order5;
    while (!q->stopped()) {
        WindowUnit u;
        bool end = false;
        int rc = ws.next(u, &end);
        if (rc) return send_error(rc);
        ReadyWindow w;
        if (end) {
            restore();
            w.end = true;
            q->push(std::move(w));
            return;
        }
        if (u.barrier >= 0) {
            restore();
            w.barrier = u.barrier;
            PlanItem it;
            it.barrier = u.barrier;
            w.plan.push_back(std::move(it));
            w.mfma.assign(1, 0);
            Launch L;
            L.first = 0;
            w.launches.push_back(L);
            w.op_off.assign(1, 0);
            q->push(std::move(w));
            continue;
        }
        // the learned schedule predicts the window at k = 4 and k = 5 (as
        // shapes, both in one batched inference) and takes k = 5 only for a
        // predicted gain above the model's resolution; then the blocks get
        // their matrices
        if (bo.relabel) {
            uint64_t low = 0;
            for (int p = 0; p < 4; p++) low |= uint64_t(1) << at[p];
            if (low != u.low || (u.low & ~u.mask) || (u.next_low & ~u.mask) || __builtin_popcountll(u.next_low) != 4)
                return send_error(fail(QSV_ERR_ARG, "internal: relabelled window does not match the qubit order"));
        }
        int k = fixed_k;
        double best = 0;
        if (learned) {
            // fusion searches random dependency orders too (seeded by the
            // window's first record, so plans repeat run to run)
            const uint64_t seed = (uint64_t)u.taken[0];
            const int tries = fuse_tries(s->n);
            if ((rc = best_fusion_order(gates, u.taken, mats, 4, o.exact, tries, seed, order4, shape)) ||
                (rc = best_fusion_order(gates, u.taken, mats, 5, o.exact, tries, seed, order5, shape5)))
                return send_error(rc);
            const size_t n4 = shape.size();
            std::vector<int> win(n4, 0);
            win.resize(n4 + shape5.size(), 1);
            shape.insert(shape.end(), std::make_move_iterator(shape5.begin()), std::make_move_iterator(shape5.end()));
            rename(shape);
            prepare_blocks(s, shape, win, shape_mfma);
            double total = 0;
            if ((rc = group_launches(s, shape, o, launches5, &total, &win))) return send_error(rc);
            double p4 = 0, p5 = 0;
            shape_launches.clear();
            for (const Launch &L : launches5) (L.first < n4 ? p4 : p5) += L.pred_ms;
            k = 4;
            best = p4;
            if (p5 < p4 * (1 - kK5Margin)) {
                k = 5;
                best = p5;
            }
            for (Launch L : launches5) {
                if ((L.first < n4) != (k == 4)) continue;
                if (k == 5) L.first -= n4;
                shape_launches.push_back(L);
            }
            if (k == 5) shape.erase(shape.begin(), shape.begin() + (std::ptrdiff_t)n4);
            else shape.resize(n4);
        }
        rc = learned ? fuse_order(gates, k == 5 ? order5 : order4, mats, k, o.exact, false, w.plan)
                     : fuse_window(gates, u.taken, mats, k, o.exact, false, w.plan);
        if (rc) return send_error(rc);
        rename(w.plan);
        const std::vector<int> window(w.plan.size(), 0);
        prepare_blocks(s, w.plan, window, w.mfma);
        if (learned) {
            bool same = w.plan.size() == shape.size();
            for (size_t i = 0; same && i < shape.size(); i++) same = w.plan[i].qubits == shape[i].qubits;
            if (!same) return send_error(fail(QSV_ERR_UNSUPPORTED, "planner shapes differ with and without block matrices"));
            w.launches = shape_launches;
        } else {
            window_launches(s, w.plan, window, w.mfma, o, w.launches);
        }
        if (bo.relabel && u.last) {
            // the segment's last window starts putting qubits back: one pass
            // of the restoring routes over its own tile's positions
            uint64_t tile = 0;
            for (uint64_t m = u.mask; m; m &= m - 1) tile |= uint64_t(1) << phys[__builtin_ctzll(m)];
            Launch &L = w.launches.back();
            L.moves = route_within(at, tile);
            if (!L.moves.empty()) {
                std::vector<int> nat(at);
                for (const auto &mv : L.moves) nat[mv.second] = at[mv.first];
                at.swap(nat);
                for (int p = 0; p < n; p++) phys[at[p]] = p;
                L.tb = kMaxPassTileBits;
                L.tc = 0;
            }
        } else if (bo.relabel) {
            // next_low's qubits move into the positions 0..3 whose qubits
            // leave; then, among the tile's other positions, every qubit whose
            // own bit is one of them goes there (free in the same store, and
            // the order to restore at the end stays short)
            std::vector<int> nat(at);
            int slot[4], ns = 0;
            for (int p = 0; p < 4; p++)
                if (!(u.next_low >> at[p] & 1)) slot[ns++] = p;
            int si = 0;
            for (uint64_t m = u.next_low; m; m &= m - 1) {
                const int x = __builtin_ctzll(m);
                if (phys[x] < 4) continue;
                const int j = slot[si++], p = phys[x];
                std::swap(nat[j], nat[p]);
            }
            uint64_t tile = 0;
            for (uint64_t m = u.mask; m; m &= m - 1) tile |= uint64_t(1) << phys[__builtin_ctzll(m)];
            const std::vector<int> mid(nat);
            for (const auto &mv : route_within(mid, tile & ~uint64_t(15))) nat[mv.second] = mid[mv.first];
            Launch &L = w.launches.back();
            for (int p = 0; p < n; p++)
                if (nat[p] != at[p]) L.moves.push_back({phys[nat[p]], p});
            at.swap(nat);
            for (int p = 0; p < n; p++) phys[at[p]] = p;
            if (!L.moves.empty()) {
                L.tb = kMaxPassTileBits;
                L.tc = 0;
            }
        }
        w.k = k;
        w.pred = best;
        build_operands(w.plan, w.mfma, o.split, 1, w.op_off, w.ops);
        q->push(std::move(w));
    }
}

// Operand rings of the streamed schedules (kept with the state).
struct OperandRing {
    static constexpr size_t kBytes = size_t(32) << 20;
    float *dev = nullptr;
    float *host = nullptr;              // pinned
    size_t head = 0;                    // next free byte
    struct Rec {
        size_t off, bytes;
        hipEvent_t ev;
    };
    std::deque<Rec> live;
    std::vector<hipEvent_t> spare;
};

}  // namespace

void free_operand_ring(qsv_state *s)
{
    OperandRing *r = reinterpret_cast<OperandRing *>(s->op_ring);
    if (!r) return;
    for (auto &x : r->live) (void)hipEventDestroy(x.ev);
    for (hipEvent_t e : r->spare) (void)hipEventDestroy(e);
    if (r->dev) (void)hipFree(r->dev);
    if (r->host) (void)hipHostFree(r->host);
    delete r;
    s->op_ring = nullptr;
}

namespace {

// Room for `bytes` of operands in the rings: returns the byte offset, after
// waiting for the windows whose operands occupied that range.
int ring_alloc(qsv_state *s, size_t bytes, size_t *off)
{
    OperandRing *r = reinterpret_cast<OperandRing *>(s->op_ring);
    if (!r) {
        r = new OperandRing;
        s->op_ring = r;
        if (hipMalloc((void **)&r->dev, OperandRing::kBytes) != hipSuccess) {
            r->dev = nullptr;
            free_operand_ring(s);
            return fail(QSV_ERR_ALLOC, "cannot allocate the operand ring");
        }
        if (hipHostMalloc((void **)&r->host, OperandRing::kBytes, hipHostMallocDefault) != hipSuccess) {
            r->host = nullptr;
            free_operand_ring(s);
            return fail(QSV_ERR_ALLOC, "cannot allocate the pinned operand ring");
        }
    }
    if (bytes > OperandRing::kBytes) return fail(QSV_ERR_ARG, "window operands exceed the operand ring");
    bytes = (bytes + 255) & ~size_t(255);
    if (r->head + bytes > OperandRing::kBytes) r->head = 0;
    const size_t a = r->head, b = a + bytes;
    for (auto it = r->live.begin(); it != r->live.end();) {
        if (it->off < b && a < it->off + it->bytes) {
            QSV_HIP(hipEventSynchronize(it->ev));
            r->spare.push_back(it->ev);
            it = r->live.erase(it);
        } else {
            ++it;
        }
    }
    r->head = b;
    *off = a;
    return QSV_OK;
}

int ring_done(qsv_state *s, size_t off, size_t bytes)
{
    OperandRing *r = reinterpret_cast<OperandRing *>(s->op_ring);
    hipEvent_t ev;
    if (!r->spare.empty()) {
        ev = r->spare.back();
        r->spare.pop_back();
    } else {
        QSV_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    }
    const hipError_t e = hipEventRecord(ev, s->stream);
    if (e != hipSuccess) {
        (void)hipEventDestroy(ev);
        return hip_fail(e, "hipEventRecord");
    }
    r->live.push_back({off, (bytes + 255) & ~size_t(255), ev});
    return QSV_OK;
}

int run_streamed(qsv_state *s, const qsv_gate *gates, int64_t ngates, const double *mats, const qsv_run_opts &o,
                 bool learned, int32_t *rng, int *outcomes, int64_t *blocks, double *pred_ms, Profiler *prof,
                 hipEvent_t start, std::chrono::steady_clock::time_point t_plan, double *plan_ms, int *k_used,
                 int *planner)
{
    WindowQueue q(64);
    std::thread producer(produce_windows, s, gates, ngates, mats, o, learned, &q);
    // the FP16x3 state scale is measured while the first windows are planned
    int rc = o.split == 3 ? f16_state_scale(s) : QSV_OK;
    bool first = true;
    int64_t nblocks = 0;
    double pred = 0;
    int kmax = 0;
    if (!rc && prof->on()) {
        s->tlog_feats.clear();
        s->tlog_ms.clear();
    }
    while (!rc) {
        ReadyWindow w = q.pop();
        if (w.rc) {
            rc = fail(w.rc, w.err);
            break;
        }
        if (w.end) break;
        if (first) {
            *plan_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_plan).count();
            const hipError_t e = hipEventRecord(start, s->stream);    // device time starts with the first window
            if (e != hipSuccess) {
                rc = hip_fail(e, "hipEventRecord");
                break;
            }
            first = false;
        }
        if (w.barrier >= 0) {
            rc = exec_launches(s, gates, w.plan, w.mfma, w.launches, nullptr, w.op_off, o, rng, outcomes, prof,
                               &nblocks);
            continue;
        }
        const size_t bytes = w.ops.size() * sizeof(float);
        size_t off = 0;
        const float *dops = nullptr;
        if (bytes) {
            if ((rc = ring_alloc(s, bytes, &off))) break;
            OperandRing *r = reinterpret_cast<OperandRing *>(s->op_ring);
            std::memcpy(reinterpret_cast<char *>(r->host) + off, w.ops.data(), bytes);
            const hipError_t e = hipMemcpyAsync(reinterpret_cast<char *>(r->dev) + off,
                                                reinterpret_cast<char *>(r->host) + off, bytes,
                                                hipMemcpyHostToDevice, s->stream);
            if (e != hipSuccess) {
                rc = hip_fail(e, "operand upload");
                break;
            }
            dops = reinterpret_cast<const float *>(reinterpret_cast<const char *>(r->dev) + off);
        }
        rc = exec_launches(s, gates, w.plan, w.mfma, w.launches, dops, w.op_off, o, rng, outcomes, prof, &nblocks);
        if (!rc && bytes) rc = ring_done(s, off, bytes);
        pred += w.pred;
        kmax = std::max(kmax, w.k);
    }
    if (rc) q.stop();
    producer.join();
    if (first) *plan_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_plan).count();
    *blocks = nblocks;
    *pred_ms = pred;
    *k_used = kmax ? kmax : std::max(4, std::min(o.fuse_k, 5));
    *planner = 1;
    return rc;
}

}  // namespace

int run_fused(qsv_state *s, const qsv_gate *gates, int64_t ngates, const double *mats,
              const qsv_run_opts &o, int32_t *rng, int *outcomes, int64_t *blocks, double *pred_ms,
              Profiler *prof, hipEvent_t start, double *plan_ms, int *k_used, int *planner)
{
    if (s->prec != QSV_C64) return fail(QSV_ERR_PRECISION, "the fused path runs on c64 states");
    const auto t_plan = std::chrono::steady_clock::now();
    if (o.tile != 0 && (o.tile < 6 || o.tile > 8)) return fail(QSV_ERR_ARG, "tile column bits must be 6, 7 or 8");
    if (o.pass_tb != 0 && (o.pass_tb < 11 || o.pass_tb > kMaxPassTileBits))
        return fail(QSV_ERR_ARG, "pass tile bits must be 0 or 11..13");
    const bool windowed = o.scheduler == 3 && s->n >= 11;
    // window schedules of states that hold a 13-bit tile window run as a
    // planning / sweep pipeline (the learned schedule with its beam search)
    const bool learned_windows = o.scheduler == 1 && s->cm && s->n >= 11;
    if (windowed || learned_windows)
        return run_streamed(s, gates, ngates, mats, o, learned_windows, rng, outcomes, blocks, pred_ms, prof, start,
                            t_plan, plan_ms, k_used, planner);
    std::vector<PlanItem> plan;
    std::vector<int> window;
    int rc;
    SchedChoice choice;
    if (o.scheduler == 1) {
        rc = plan_learned(s, gates, ngates, mats, o, plan, &choice, &window);
    } else {
        rc = plan_greedy(gates, ngates, mats, std::min(o.fuse_k, 5), o.exact, plan);
        choice.fuse_k = o.fuse_k;
    }
    if (rc) return rc;
    std::vector<char> mfma;
    prepare_blocks(s, plan, window, mfma);
    std::vector<Launch> launches;
    double pred = 0;
    if (o.scheduler == 1 && s->cm) {
        rc = group_launches(s, plan, o, launches, &pred, window.empty() ? nullptr : &window);
    } else {
        qsv_run_opts g = o;
        if (o.scheduler != 2) g.group = 1;
        rc = group_launches(s, plan, g, launches, &pred);
    }
    if (rc) return rc;
    std::vector<size_t> op_off;
    std::vector<float> ops;
    build_operands(plan, mfma, o.split, 8, op_off, ops);
    float *dops = nullptr;
    if (!ops.empty()) {
        if (hipMalloc((void **)&dops, ops.size() * sizeof(float)) != hipSuccess)
            return fail(QSV_ERR_ALLOC, "cannot allocate block operands");
        hipError_t e = hipMemcpyAsync(dops, ops.data(), ops.size() * sizeof(float), hipMemcpyHostToDevice, s->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        if (e != hipSuccess) {
            (void)hipFree(dops);
            return hip_fail(e, "operand upload");
        }
    }
    if (o.split == 3 && (rc = f16_state_scale(s))) {
        if (dops) (void)hipFree(dops);
        return rc;
    }
    int64_t nblocks = 0;
    *plan_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_plan).count();
    *k_used = choice.fuse_k;
    *planner = choice.window;
    QSV_HIP(hipEventRecord(start, s->stream));       // device time starts after planning
    if (prof->on()) {
        s->tlog_feats.clear();
        s->tlog_ms.clear();
    }
    rc = exec_launches(s, gates, plan, mfma, launches, dops, op_off, o, rng, outcomes, prof, &nblocks);
    if (dops) {
        (void)hipStreamSynchronize(s->stream);
        (void)hipFree(dops);
    }
    *blocks = nblocks;
    *pred_ms = pred;
    return rc;
}

}  // namespace qsv

using namespace qsv;

extern "C" int qsv_fuse_plan(const qsv_gate *gates, int64_t ngates, const double *mats, int max_k, int exact,
                             int64_t max_blocks, int *k_out, int *qubits_out, double *mats_out,
                             int64_t *ngates_out, int64_t *nblocks)
{
    if ((!gates && ngates) || !nblocks) return fail(QSV_ERR_ARG, "null argument");
    std::vector<PlanItem> plan;
    int rc = plan_greedy(gates, ngates, mats, max_k, exact, plan);
    if (rc) return rc;
    *nblocks = (int64_t)plan.size();
    if ((int64_t)plan.size() > max_blocks) return fail(QSV_ERR_ARG, "plan larger than the output buffers");
    for (size_t b = 0; b < plan.size(); b++) {
        const PlanItem &it = plan[b];
        const int k = it.barrier >= 0 ? 0 : (int)it.qubits.size();
        if (k_out) k_out[b] = k;
        if (ngates_out) ngates_out[b] = it.barrier >= 0 ? it.barrier : it.ngates;
        if (qubits_out)
            for (int j = 0; j < 5; j++) qubits_out[b * 5 + j] = j < k ? it.qubits[j] : -1;
        if (mats_out && k > 0) std::memcpy(mats_out + b * 2048, it.u.data(), it.u.size() * sizeof(double));
    }
    return QSV_OK;
}
