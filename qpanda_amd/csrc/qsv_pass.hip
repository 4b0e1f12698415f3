// qsv_pass.hip -- several consecutive fused blocks in one HBM sweep.
//
// A pass takes G <= 8 consecutive blocks of a fusion plan whose qubit sets
// together with a low run of qubits fit in TB tile bits (11..14; 2^TB complex64
// amplitudes = 16..128 KiB of LDS).  L = (union of the blocks' qubits) u
// {lowest other qubits until |L| = TB}; a tile is the 2^TB amplitudes whose
// non-L bits are fixed.  One workgroup per tile:
//   load   16-byte loads into LDS re / im planes (compact position over L);
//   apply  for each block in order: its 2^(TB-k) groups are the columns; each
//          wave takes 32-column chunks and runs the same MFMA product as the
//          single-block kernel (FP32 v_mfma_f32_32x32x2f32 for M_hi, BF16
//          v_mfma_f32_32x32x16_bf16 for the M_lo correction), writing the
//          results back in place; a barrier separates blocks;
//   store  16-byte coalesced stores.
// HBM traffic is one read and one write of the state per pass instead of per
// block; MFMA work is the sum over the blocks.  Whether a grouping pays is a
// cost-model decision (qsv_costmodel.hip).
#include "qsv_internal.h"
#include "qsv_sched.h"

namespace qsv {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

struct PassBlock {
    const float *mops;      // operand-ready matrix (hi floats, then bf16 lo)
    int kq;                 // 4 or 5
    int ncolbits;           // TB - kq
    int erow[32];           // compact tile position of row (amplitude) r
    int col_rank[10];       // rank in L of the i-th column bit
};

struct PassArgs {
    InsertSet tile_ins;     // L positions, ascending
    int run;                // L contains qubits 0..run-1
    int hpos[14];           // L positions above the run
    int nblocks;
    PassBlock b[kMaxPassBlocks];
};

template <int KQ, bool SPLIT>
__device__ __forceinline__ void pass_apply(float *pre, float *pim, const PassBlock &pb, int lane, int wave)
{
    constexpr int D = 1 << KQ, HALVES = (2 * D) / 32, KSTEPS = D, LOBLK = D / 8;
    float ahi[HALVES][KSTEPS];
    bf16x8 alo[SPLIT ? HALVES : 1][SPLIT ? LOBLK : 1];
#pragma unroll
    for (int h = 0; h < HALVES; h++)
#pragma unroll
        for (int s = 0; s < KSTEPS; s++) ahi[h][s] = pb.mops[(h * KSTEPS + s) * 64 + lane];
    if (SPLIT) {
        const bf16x8 *lo = reinterpret_cast<const bf16x8 *>(pb.mops + HALVES * KSTEPS * 64);
#pragma unroll
        for (int h = 0; h < HALVES; h++)
#pragma unroll
            for (int sb = 0; sb < LOBLK; sb++) alo[SPLIT ? h : 0][SPLIT ? sb : 0] = lo[(h * LOBLK + sb) * 64 + lane];
    }
    const int chunks = (1 << pb.ncolbits) >> 5;
    const float *const bplane = (lane >> 5) ? pim : pre;
    for (int chunk = wave; chunk < chunks; chunk += 4) {
        const int col = chunk * 32 + (lane & 31);
        int ecol = 0;
        for (int i = 0; i < pb.ncolbits; i++)
            if (col >> i & 1) ecol |= 1 << pb.col_rank[i];
        floatx16 acc[HALVES];
#pragma unroll
        for (int h = 0; h < HALVES; h++)
#pragma unroll
            for (int v = 0; v < 16; v++) acc[h][v] = 0.0f;
        float bv[KSTEPS];
#pragma unroll
        for (int s = 0; s < KSTEPS; s++) bv[s] = bplane[pb.erow[s] | ecol];
#pragma unroll
        for (int s = 0; s < KSTEPS; s++)
#pragma unroll
            for (int h = 0; h < HALVES; h++) acc[h] = __builtin_amdgcn_mfma_f32_32x32x2f32(ahi[h][s], bv[s], acc[h], 0, 0, 0);
        if (SPLIT) {
            // M_lo in the permuted K order of the operand layout: the B values
            // are this lane's FP32 B values of K-steps 8 sb .. 8 sb + 7
#pragma unroll
            for (int sb = 0; sb < LOBLK; sb++) {
                bf16x8 bl;
#pragma unroll
                for (int e = 0; e < 8; e++) bl[e] = (__bf16)bv[8 * sb + e];
#pragma unroll
                for (int h = 0; h < HALVES; h++)
                    acc[h] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo[SPLIT ? h : 0][SPLIT ? sb : 0], bl, acc[h], 0, 0, 0);
            }
        }
#pragma unroll
        for (int h = 0; h < HALVES; h++)
#pragma unroll
            for (int v = 0; v < 16; v++) {
                const int m = 32 * h + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
                const int e = pb.erow[m >> 1] | ecol;
                if (m & 1) pim[e] = acc[h][v];
                else pre[e] = acc[h][v];
            }
    }
}

template <int TB, bool SPLIT>
__global__ __launch_bounds__(256) void k_fused_pass(float2 *__restrict__ a, uint64_t ntiles, PassArgs pa)
{
    constexpr int TSIZE = 1 << TB;
    __shared__ float planes[2][TSIZE];
    float *const pre = planes[0];
    float *const pim = planes[1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint64_t tile = blockIdx.x;
    if (tile >= ntiles) return;
    const uint64_t base = expand_index(tile, pa.tile_ins);
    const int runmask = (1 << pa.run) - 1, nh = TB - pa.run;
    auto addr_of = [&](int e) -> uint64_t {
        uint64_t o = (uint64_t)(e & runmask);
        const int hi = e >> pa.run;
        for (int j = 0; j < nh; j++)
            if (hi >> j & 1) o |= uint64_t(1) << pa.hpos[j];
        return base | o;
    };
    for (int i = 0; i < TSIZE / 512; i++) {
        const int e = 2 * (tid + 256 * i);
        const float4 v = *reinterpret_cast<const float4 *>(a + addr_of(e));
        pre[e] = v.x; pim[e] = v.y;
        pre[e + 1] = v.z; pim[e + 1] = v.w;
    }
    __syncthreads();
    for (int g = 0; g < pa.nblocks; g++) {
        if (pa.b[g].kq == 4) pass_apply<4, SPLIT>(pre, pim, pa.b[g], lane, wave);
        else pass_apply<5, SPLIT>(pre, pim, pa.b[g], lane, wave);
        __syncthreads();
    }
    for (int i = 0; i < TSIZE / 512; i++) {
        const int e = 2 * (tid + 256 * i);
        float4 v;
        v.x = pre[e]; v.y = pim[e];
        v.z = pre[e + 1]; v.w = pim[e + 1];
        *reinterpret_cast<float4 *>(a + addr_of(e)) = v;
    }
}

// Tile bits a pass over these (widened) blocks needs: the union of their
// qubits plus at least 5 fill qubits (>= 32 columns for a 5-qubit block and a
// >= 256 B contiguous run), and at least 11.
int pass_tile_bits(const std::vector<const std::vector<int> *> &qsets, int n)
{
    uint64_t u = 0;
    for (auto *q : qsets)
        for (int x : *q) u |= uint64_t(1) << x;
    const int tb = std::max(11, __builtin_popcountll(u) + 5);
    return tb <= std::min(14, n) ? tb : -1;
}

template <int TB>
static void launch_tb(qsv_state *s, uint64_t ntiles, const PassArgs &pa, int split)
{
    float2 *a = (float2 *)s->amps;
    if (split) hipLaunchKernelGGL((k_fused_pass<TB, true>), dim3((unsigned)ntiles), dim3(256), 0, s->stream, a, ntiles, pa);
    else hipLaunchKernelGGL((k_fused_pass<TB, false>), dim3((unsigned)ntiles), dim3(256), 0, s->stream, a, ntiles, pa);
}

// blocks[g]: widened qubit list (4 or 5 qubits) and device operand pointer.
int launch_pass_dev(qsv_state *s, const std::vector<const std::vector<int> *> &qsets,
                    const std::vector<const float *> &mops, int split, int tb)
{
    if (s->prec != QSV_C64) return fail(QSV_ERR_PRECISION, "fused passes need a c64 state");
    const int G = (int)qsets.size();
    if (G < 1 || G > kMaxPassBlocks) return fail(QSV_ERR_ARG, "1..8 blocks per pass");
    if (tb < 11 || tb > 14 || tb > s->n) return fail(QSV_ERR_ARG, "pass tile bits must be 11..14");
    uint64_t u = 0;
    for (auto *q : qsets) {
        if (q->size() != 4 && q->size() != 5) return fail(QSV_ERR_ARG, "pass blocks must have 4 or 5 qubits");
        for (int x : *q) u |= uint64_t(1) << x;
    }
    if (__builtin_popcountll(u) + 5 > tb) return fail(QSV_ERR_ARG, "blocks do not fit the pass tile");
    uint64_t lmask = u;
    for (int b = 0; __builtin_popcountll(lmask) < tb; b++) lmask |= uint64_t(1) << b;
    PassArgs pa{};
    pa.tile_ins = insert_set_from_mask(lmask);
    const int *lpos = pa.tile_ins.pos;
    int run = 0;
    while (run < tb && lpos[run] == run) run++;
    pa.run = run;
    for (int j = 0; j < tb - run; j++) pa.hpos[j] = lpos[run + j];
    auto rank = [&](int bit) {
        for (int i = 0; i < tb; i++)
            if (lpos[i] == bit) return i;
        return -1;
    };
    pa.nblocks = G;
    for (int g = 0; g < G; g++) {
        PassBlock &pb = pa.b[g];
        const std::vector<int> &q = *qsets[g];
        pb.mops = mops[g];
        pb.kq = (int)q.size();
        pb.ncolbits = tb - pb.kq;
        for (int r = 0; r < (1 << pb.kq); r++) {
            int e = 0;
            for (int j = 0; j < pb.kq; j++)
                if (r >> j & 1) e |= 1 << rank(q[j]);
            pb.erow[r] = e;
        }
        int c = 0;
        for (int i = 0; i < tb; i++) {
            bool inq = false;
            for (int x : q) inq |= (x == lpos[i]);
            if (!inq) pb.col_rank[c++] = i;
        }
    }
    const uint64_t ntiles = s->dim >> tb;
    switch (tb) {
    case 11: launch_tb<11>(s, ntiles, pa, split); break;
    case 12: launch_tb<12>(s, ntiles, pa, split); break;
    case 13: launch_tb<13>(s, ntiles, pa, split); break;
    default: launch_tb<14>(s, ntiles, pa, split); break;
    }
    s->launches++;
    QSV_HIP(hipGetLastError());
    return QSV_OK;
}

}  // namespace qsv
