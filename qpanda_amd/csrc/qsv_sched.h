// qsv_sched.h -- fusion plans, fused-block operands and the learned scheduler.
#pragma once

#include "qsv_internal.h"

#include <vector>

namespace qsv {

class Profiler;

// One item of a fusion plan: a dense block (qubits, matrix) or a barrier
// record (MEASURE / RESET) that runs on its own.
struct PlanItem {
    int64_t barrier = -1;          // record index, or -1 for a block
    std::vector<int> qubits;       // row bit j = bit(qubits[j])
    std::vector<double> u;         // 2^k x 2^k complex, row-major (re, im)
    int64_t ngates = 0;            // logical gates in the block
    int tile = 0;                  // tile-shape choice (0 = default)
    double pred_ms = 0;            // predicted device time
};

struct SchedChoice {
    int fuse_k = 4;
    int window = 0;                // 1: the plan came from the window planner
    double pred_ms = 0;
};

// shape_only: qubit lists and gate counts only (PlanItem::u left empty); the
// fusion decisions depend on the qubits alone, so the blocks are the same.
int plan_greedy(const qsv_gate *gates, int64_t ng, const double *mats, int max_k, int exact,
                std::vector<PlanItem> &plan, bool shape_only = false);
void widen_to(PlanItem &it, int kq);
void move_qubit_first(PlanItem &it, int q);
void widen_with(PlanItem &it, const std::vector<int> &add);
// Merge narrow blocks into commuting neighbours while the union stays within
// max_k qubits (the plan's product is unchanged).
void absorb_blocks(std::vector<PlanItem> &plan, int max_k, bool shape_only = false);

// Window planner (qsv_window.cpp): blocks of <= max_k (4, 5) qubits grouped
// into windows of tb tile bits; window[i] numbers the window of plan[i]
// (barrier records get their own).
// masks_out receives each window's qubit mask; masks_in replays such masks
// (same gates, tb and n) instead of growing the windows again.
int plan_window(const qsv_gate *gates, int64_t ng, const double *mats, int max_k, int exact, int tb, int n,
                std::vector<PlanItem> &plan, std::vector<int> &window, bool shape_only = false,
                const std::vector<uint64_t> *masks_in = nullptr, std::vector<uint64_t> *masks_out = nullptr);
// One window's records fused into blocks of <= max_k qubits: taken = the
// records (indices into gates) in a dependency order; the closure's order,
// program order and `tries` random dependency orders (seeded by `seed`) are
// fused and the first order with the fewest blocks is kept.
int fuse_window(const qsv_gate *gates, const std::vector<int64_t> &taken, const double *mats, int max_k, int exact,
                bool shape_only, std::vector<PlanItem> &blocks, int tries = 0, uint64_t seed = 0);
// The order fuse_window keeps, with its blocks as shapes.
int best_fusion_order(const qsv_gate *gates, const std::vector<int64_t> &taken, const double *mats, int max_k,
                      int exact, int tries, uint64_t seed, std::vector<int64_t> &order, std::vector<PlanItem> &shapes);
// Greedy fusion of records in exactly this order.
int fuse_order(const qsv_gate *gates, const std::vector<int64_t> &order, const double *mats, int max_k, int exact,
               bool shape_only, std::vector<PlanItem> &blocks);

// The windows of a gate list one at a time (qsv_window.cpp): a window's
// qubit mask and records, or a MEASURE / RESET record on its own.
struct WindowUnit {
    int64_t barrier = -1;              // record index of a MEASURE / RESET, else -1
    uint64_t mask = 0;                 // the window's qubits
    std::vector<int64_t> taken;        // its records, in a dependency order
    uint64_t low = 0;                  // the qubits at bits 0..3 during the sweep
    uint64_t next_low = 0;             // the qubits the sweep leaves at bits 0..3
    bool last = false;                 // the last window before a MEASURE / RESET or the end
};
struct BeamOpts {
    int width = 1;                     // states kept per depth (1: plan_window's windows)
    int branch = 3;                    // windows expanded per state
    int first = 6;                     // first growth candidates completed per state
    int delay = 4;                     // windows a commitment trails the search
    bool relabel = false;              // windows grow from the qubits the previous
                                       // one left at bits 0..3 (else from 0..3)
};
class WindowStream {
public:
    WindowStream(const qsv_gate *gates, int64_t ng, int n, int tb, const BeamOpts &b);
    ~WindowStream();
    WindowStream(const WindowStream &) = delete;
    WindowStream &operator=(const WindowStream &) = delete;
    int next(WindowUnit &u, bool *end);          // *end: no units left
    int64_t windows() const;                     // units handed out so far
private:
    struct Impl;
    Impl *impl_;
};

// Qubits that widen block i of a window plan to kq qubits: the highest of
// its window's qubits (the union of the window's blocks and qubits 0..3)
// that it does not use.
std::vector<int> window_widening(const std::vector<PlanItem> &plan, const std::vector<int> &window, size_t i,
                                 int kq);

// Learned scheduler (qsv_costmodel.hip): plans at every k, predicts each
// block's time with the cost model, keeps the cheapest plan.
int plan_learned(qsv_state *s, const qsv_gate *gates, int64_t ng, const double *mats,
                 const qsv_run_opts &o, std::vector<PlanItem> &plan, SchedChoice *choice,
                 std::vector<int> *window);

// Fused MFMA kernel helpers (qsv_fused.hip).
void fused_operands(int kq, const double *u, int split, float *out);
size_t fused_operand_floats(int kq, int split);
constexpr int kDefaultTileCols = 7;
int launch_fused_dev(qsv_state *s, const int *qubits, int kq, const float *mops, int split, int tc);

// Passes that move the data at position p to position dest[p] for every p
// (dest a permutation of the positions 0..n-1): each pass exchanges the qubits
// of one tile of tb positions that holds 0..3, given as (from, to) pairs
// (qsv_window.cpp).
std::vector<std::vector<std::pair<int, int>>> qubit_routes(std::vector<int> dest, int tb);
// One such pass over the positions in mask `in` (which holds 0..3): the moves
// that put the most of them in place.
std::vector<std::pair<int, int>> route_within(const std::vector<int> &dest, uint64_t in);

// Multi-block passes (qsv_pass.hip).
constexpr int kMaxPassBlocks = 16;
constexpr int kMaxPassTileBits = 13;
int pass_tile_bits(const std::vector<const std::vector<int> *> &qsets, int n);
int launch_pass_dev(qsv_state *s, const std::vector<const std::vector<int> *> &qsets,
                    const std::vector<const float *> &mops, int split, int tb, int kernel = 0,
                    const std::vector<std::pair<int, int>> *moves = nullptr);

// One device launch of a fused run: plan items [first, first + count) (all
// blocks) in one sweep.  count == 1 uses the single-block kernel with tile
// column bits `tc`; count > 1 uses the pass kernel with `tb` tile bits.
struct Launch {
    size_t first = 0;
    int count = 1;                     // 0: a sweep that only moves qubits
    int tc = 0;
    int tb = 0;
    double pred_ms = 0;
    std::vector<std::pair<int, int>> moves;   // (from, to) qubit moves of the sweep's store
};

// Cost-model features of one launch (single block: tb = kq + tc).
void launch_features(int n, int split, const std::vector<const std::vector<int> *> &qsets, bool single,
                     int tb, float *f);

// Group a block plan into launches.  scheduler 0: one block per launch;
// 2: greedy maximal groups (opts.group, opts.pass_tb); 1: cost-model DP.
int group_launches(qsv_state *s, const std::vector<PlanItem> &plan, const qsv_run_opts &o,
                   std::vector<Launch> &launches, double *pred_total, const std::vector<int> *window = nullptr);

}  // namespace qsv
