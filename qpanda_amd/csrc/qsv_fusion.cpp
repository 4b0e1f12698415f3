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
#include <cstring>
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
    std::vector<int> qs(k);
    for (int b = 0; b < k; b++) qs[b] = it.qubits[order[b]];
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

int run_fused(qsv_state *s, const qsv_gate *gates, int64_t ngates, const double *mats,
              const qsv_run_opts &o, int32_t *rng, int *outcomes, int64_t *blocks, double *pred_ms,
              Profiler *prof, hipEvent_t start, double *plan_ms, int *k_used, int *planner)
{
    if (s->prec != QSV_C64) return fail(QSV_ERR_PRECISION, "the fused path runs on c64 states");
    const auto t_plan = std::chrono::steady_clock::now();
    if (o.tile != 0 && (o.tile < 6 || o.tile > 8)) return fail(QSV_ERR_ARG, "tile column bits must be 6, 7 or 8");
    std::vector<PlanItem> plan;
    std::vector<int> window;
    int rc;
    SchedChoice choice;
    const bool windowed = o.scheduler == 3 && s->n >= 11;
    if (o.pass_tb != 0 && (o.pass_tb < 11 || o.pass_tb > kMaxPassTileBits))
        return fail(QSV_ERR_ARG, "pass tile bits must be 0 or 11..13");
    if (o.scheduler == 1) {
        rc = plan_learned(s, gates, ngates, mats, o, plan, &choice, &window);
    } else if (windowed) {
        choice.fuse_k = std::max(4, std::min(o.fuse_k, 5));
        choice.window = 1;
        rc = plan_window(gates, ngates, mats, choice.fuse_k, o.exact, o.pass_tb ? o.pass_tb : kMaxPassTileBits,
                         s->n, plan, window);
    } else {
        rc = plan_greedy(gates, ngates, mats, std::min(o.fuse_k, 5), o.exact, plan);
        choice.fuse_k = o.fuse_k;
    }
    if (rc) return rc;
    const int kq_min = 4;
    // window plans widen their small blocks with the window's highest qubits
    // (window_widening); the rest widen to >= 4 qubits with the lowest free
    // ones where an MFMA tile fits; other blocks use the FP64 kernel
    if (!window.empty() && s->n >= kq_min + 6)
        for (size_t i = 0; i < plan.size(); i++)
            if (plan[i].barrier < 0) widen_with(plan[i], window_widening(plan, window, i, kq_min));
    std::vector<char> mfma(plan.size(), 0);
    for (size_t i = 0; i < plan.size(); i++) {
        PlanItem &it = plan[i];
        if (it.barrier >= 0) continue;
        const int kq = std::max(kq_min, (int)it.qubits.size());
        if (s->n < kq + 6) continue;
        widen_to(it, kq_min);
        move_qubit_first(it, 0);
        mfma[i] = 1;
    }
    std::vector<Launch> launches;
    double pred = 0;
    if (o.scheduler == 1 && s->cm) {
        rc = group_launches(s, plan, o, launches, &pred, window.empty() ? nullptr : &window);
    } else if (windowed) {
        // one sweep per window (several beyond the pass kernel's block limit)
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
    } else {
        qsv_run_opts g = o;
        if (o.scheduler != 2) g.group = 1;
        rc = group_launches(s, plan, g, launches, &pred);
    }
    if (rc) return rc;
    // operands for every MFMA block in one upload, built on a few threads
    std::vector<size_t> op_off(plan.size(), 0);
    size_t op_total = 0;
    for (size_t i = 0; i < plan.size(); i++) {
        if (!mfma[i]) continue;
        op_off[i] = op_total;
        op_total += fused_operand_floats((int)plan[i].qubits.size(), o.split);
    }
    std::vector<float> ops(op_total);
    {
        const size_t nth = std::max<size_t>(1, std::min<size_t>({8, std::thread::hardware_concurrency(),
                                                                  plan.size() / 64 + 1}));
        std::vector<std::thread> pool;
        for (size_t t = 0; t < nth; t++)
            pool.emplace_back([&, t]() {
                for (size_t i = t; i < plan.size(); i += nth)
                    if (mfma[i])
                        fused_operands((int)plan[i].qubits.size(), plan[i].u.data(), o.split, ops.data() + op_off[i]);
            });
        for (std::thread &th : pool) th.join();
    }
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
    const double N = (double)s->dim;
    int64_t nblocks = 0;
    *plan_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_plan).count();
    *k_used = choice.fuse_k;
    *planner = choice.window;
    QSV_HIP(hipEventRecord(start, s->stream));       // device time starts after planning
    rc = QSV_OK;
    if (prof->on()) {
        s->tlog_feats.clear();
        s->tlog_ms.clear();
    }
    for (const Launch &L : launches) {
        if (rc) break;
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
        bool single = L.count == 1;
        int tb = L.tb;
        if (single && !mfma[L.first]) {
            rc = launch_block(s, it0.qubits.data(), (int)it0.qubits.size(), it0.u.data(), 0);
            cls = QSV_K_BLOCK64;
        } else if (single) {
            const int tc = L.tc ? L.tc : kDefaultTileCols;
            tb = (int)it0.qubits.size() + tc;
            rc = launch_fused_dev(s, it0.qubits.data(), (int)it0.qubits.size(), mp[0], o.split, tc);
        } else {
            rc = launch_pass_dev(s, qsets, mp, o.split, tb, o.pass_kernel);
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
        nblocks += L.count;
    }
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
