// qsv_pass.hip -- several consecutive fused blocks in one HBM sweep.
//
// A pass takes G <= 16 consecutive blocks of a fusion plan.  L = qubits 0..3 u
// (the blocks' qubits) u {lowest other qubits until |L| = TB}, TB = 11..13
// (2^TB complex64 amplitudes = 16..64 KiB of LDS, so several workgroups share
// a CU); a tile is the 2^TB amplitudes whose non-L bits are fixed, made of
// contiguous runs of >= 16 amplitudes (128 B).  One workgroup of NT threads
// (pass_threads) per tile:
//   load   16-byte loads issued first; element e of the tile sits at
//          pdep(e, L), split as pdep(2 tid, L) + pdep(2 NT i, L) so each
//          thread computes its address part once;
//   apply  for each block in order: its 2^(TB-k) groups are the columns; each
//          wave takes 32-column chunks and runs the same MFMA product as the
//          single-block kernel (FP16x3 on the FP16 MFMA, BF16x6 on the BF16
//          MFMA, or FP32 v_mfma_f32_32x32x2f32 with an optional BF16 lo
//          correction), writing the results back in place; a barrier
//          separates blocks;
//   store  16-byte stores to the same addresses.
// LDS layout: tile element e sits at float index P(e) = e ^ S(e) of its
// plane, S XOR-ing bank bits 1..4 with a fixed linear function of bits 5..8
// (kLdsSwizzle).  Without it, the 32 columns a half-wave reads or writes land
// 16 floats apart when the block holds the tile's lowest qubits (a 16-way bank
// conflict) and 2-8 ways for most other blocks; the constants were chosen over
// the blocks of random, QFT and HEA circuits (average bank cycles per access
// 2.7-3.6 without, 1.2-1.4 with).  P is linear over GF(2) and leaves bit 0
// alone, so a position is still the XOR of precomputed row and column parts
// and (re, im) pair writes stay 8-byte aligned.
// HBM traffic is one read and one write of the state per pass instead of per
// block; MFMA work is the sum over the blocks.  (A persistent grid-stride form
// that prefetches the next tile into registers measured ~10% slower at 30
// qubits than one workgroup per tile with the FP32 product, and 4% slower with
// BF16x6 on 64 KiB tiles, where it costs no occupancy; capping registers for
// five resident 32 KiB tiles per CU measured 4% slower as well.  On the 30-qubit
// benchmark plan with FP16x3, 4.21 ms per window sweep here against: one
// persistent 1024-thread workgroup per CU with the next tile prefetched into
// registers and the operands in LDS, 4.79 ms (bit-identical; without blocks it
// streams at 3.65 ms against 3.46 ms for this form); two persistent
// 512-thread workgroups per CU with the same prefetch, 4.49 ms.)  Whether a
// grouping pays is a scheduling decision (greedy grouping, or the cost model in
// qsv_costmodel.hip).
#include "qsv_internal.h"
#include "qsv_mfma.h"
#include "qsv_sched.h"
#include "qsv_stream.h"

namespace qsv {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// Threads per workgroup: 512 (8 waves) for 64 KiB tiles of k = 4 blocks, so
// that the two workgroups LDS admits per CU still give each SIMD 4 waves (123
// VGPRs); else 256 (the k = 5 product needs up to 210 VGPRs).
__host__ __device__ constexpr int pass_threads(int tb, bool any5) { return tb == 13 && !any5 ? 512 : 256; }

// Bank bits 1..4 XOR-ed into the LDS float index by bits 5..12.  The 32
// columns of a block's half-wave access are the 5 lowest tile bits outside
// the block (all below bit 9), so bits 9..12 only rotate a block's banks; they
// spread the lanes of a permuted store (k_fused_pass with PERM), whose lanes
// may read elements that differ in those bits alone.
constexpr int kLdsSwizzle[8] = {0x1E, 0x12, 0x14, 0x1C, 0x06, 0x0C, 0x18, 0x0A};

__host__ __device__ constexpr int lds_swizzle(int e)
{
    int s = 0;
#pragma unroll
    for (int i = 0; i < 8; i++)
        if (e >> (5 + i) & 1) s ^= kLdsSwizzle[i];
    return e ^ s;
}

// How a k = 4 block moves its tile data (see pass_apply4_rows / _cols).
enum PairMode { PAIR_NONE = 0, PAIR_ROWS = 1, PAIR_COLS = 2 };

struct PassBlock {
    const float *mops;      // operand-ready matrix (layout of fused_operands for the product)
    int kq;                 // 4 or 5
    int ncolbits;           // TB - kq
    int pair;               // PairMode
    int erow[32];           // swizzled LDS index of row (amplitude) r
    int colp[10];           // swizzled LDS index of the i-th column bit alone
};

struct PassArgs {
    InsertSet tile_ins;     // L positions, ascending
    int lpos[16];           // compact tile bit j -> qubit
    uint64_t hioff[16];     // pdep(2 NT i, L): address part of load i (NT = threads that move the tile)
    // permuted store (PERM): the amplitudes leave with some of the tile's
    // qubits exchanged.  Thread tid's store i writes the output elements
    // o and o + 1 at HBM offset (OR of st_hbit[j] over tid's set bits j) +
    // st_hioff[i], read from the LDS floats x and x ^ st_pair, x = (XOR of
    // st_lbit[j] over tid's set bits j) ^ st_li[i]
    uint64_t st_hbit[9];
    uint64_t st_hioff[16];
    int st_lbit[9];
    int st_li[16];
    int st_pair;
    int nblocks;
    float bscale, binv;     // FP16x3: state scale 2^s and 2^-s
    float one;              // 1.0f (see fp16x3_product)
    PassBlock b[kMaxPassBlocks];
};

typedef float float2v __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// Exact three-piece bf16 split of two FP32 values, one packed conversion per
// piece pair: b = b0 + b1 + b2 (each residual is exact in FP32 by Sterbenz's
// lemma and the last has at most 8 significant bits).  The bf16 pairs come
// back as the 32-bit words the MFMA operand registers hold.
__device__ __forceinline__ void split3_pair(float x0, float x1, uint32_t &p0, uint32_t &p1, uint32_t &p2)
{
    auto pk = [](float a, float b) {
        const bf16x2 v = __builtin_convertvector((float2v){a, b}, bf16x2);
        return __builtin_bit_cast(uint32_t, v);
    };
    auto lo = [](uint32_t w) { return __builtin_bit_cast(float, w << 16); };
    auto hi = [](uint32_t w) { return __builtin_bit_cast(float, w & 0xFFFF0000u); };
    p0 = pk(x0, x1);
    const float r0 = x0 - lo(p0), r1 = x1 - hi(p0);
    p1 = pk(r0, r1);
    p2 = pk(r0 - lo(p1), r1 - hi(p1));
}

// acc[h] += M B with the BF16x6 product (see qsv_fused.hip): B split exactly
// into three bf16 pieces, the six piece products with i + j <= 2, smallest first.
template <int HALVES, int LOBLK>
__device__ __forceinline__ void pass_bf16x6(floatx16 (&acc)[HALVES], const bf16x8 (&ap)[3][HALVES][LOBLK],
                                            const float *bv)
{
#pragma unroll
    for (int sb = 0; sb < LOBLK; sb++) {
        uint32_t w0[4], w1[4], w2[4];
#pragma unroll
        for (int e = 0; e < 4; e++) split3_pair(bv[8 * sb + 2 * e], bv[8 * sb + 2 * e + 1], w0[e], w1[e], w2[e]);
        const bf16x8 b0 = __builtin_bit_cast(bf16x8, w0);
        const bf16x8 b1 = __builtin_bit_cast(bf16x8, w1);
        const bf16x8 b2 = __builtin_bit_cast(bf16x8, w2);
#pragma unroll
        for (int h = 0; h < HALVES; h++) {
            acc[h] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ap[2][h][sb], b0, acc[h], 0, 0, 0);
            acc[h] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ap[1][h][sb], b1, acc[h], 0, 0, 0);
            acc[h] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ap[0][h][sb], b2, acc[h], 0, 0, 0);
            acc[h] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ap[1][h][sb], b0, acc[h], 0, 0, 0);
            acc[h] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ap[0][h][sb], b1, acc[h], 0, 0, 0);
            acc[h] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ap[0][h][sb], b0, acc[h], 0, 0, 0);
        }
    }
}

// MODE: 0 FP32 product, 1 FP32 + BF16 lo correction, 2 BF16x6, 3 FP16x3.
// NW: waves per workgroup.
template <int KQ, int MODE, int NW>
__device__ __forceinline__ void pass_apply(float *pre, float *pim, const PassBlock &pb, int lane, int wave,
                                           float bscale, float binv)
{
    constexpr int D = 1 << KQ, HALVES = (2 * D) / 32, KSTEPS = D, LOBLK = D / 8;
    constexpr bool SPLIT = MODE == 1, X6 = MODE == 2, H3 = MODE == 3, NOHI = X6 || H3;
    float ahi[NOHI ? 1 : HALVES][NOHI ? 1 : KSTEPS];
    bf16x8 alo[SPLIT ? HALVES : 1][SPLIT ? LOBLK : 1];
    bf16x8 ap[X6 ? 3 : 1][X6 ? HALVES : 1][X6 ? LOBLK : 1];
    f16x8 ah[H3 ? 2 : 1][H3 ? HALVES : 1][H3 ? LOBLK : 1];
    if (H3) {
        const f16x8 *p = reinterpret_cast<const f16x8 *>(pb.mops);
#pragma unroll
        for (int q = 0; q < 2; q++)
#pragma unroll
            for (int h = 0; h < HALVES; h++)
#pragma unroll
                for (int sb = 0; sb < LOBLK; sb++)
                    ah[H3 ? q : 0][H3 ? h : 0][H3 ? sb : 0] = p[((q * HALVES + h) * LOBLK + sb) * 64 + lane];
    } else if (X6) {
        const bf16x8 *p = reinterpret_cast<const bf16x8 *>(pb.mops);
#pragma unroll
        for (int q = 0; q < 3; q++)
#pragma unroll
            for (int h = 0; h < HALVES; h++)
#pragma unroll
                for (int sb = 0; sb < LOBLK; sb++)
                    ap[X6 ? q : 0][X6 ? h : 0][X6 ? sb : 0] = p[((q * HALVES + h) * LOBLK + sb) * 64 + lane];
    } else {
#pragma unroll
        for (int h = 0; h < HALVES; h++)
#pragma unroll
            for (int s = 0; s < KSTEPS; s++) ahi[NOHI ? 0 : h][NOHI ? 0 : s] = pb.mops[(h * KSTEPS + s) * 64 + lane];
    }
    if (SPLIT) {
        const bf16x8 *lo = reinterpret_cast<const bf16x8 *>(pb.mops + HALVES * KSTEPS * 64);
#pragma unroll
        for (int h = 0; h < HALVES; h++)
#pragma unroll
            for (int sb = 0; sb < LOBLK; sb++) alo[SPLIT ? h : 0][SPLIT ? sb : 0] = lo[(h * LOBLK + sb) * 64 + lane];
    }
    const int chunks = (1 << pb.ncolbits) >> 5;
    const float *const bplane = (lane >> 5) ? pim : pre;
    for (int chunk = wave; chunk < chunks; chunk += NW) {
        const int col = chunk * 32 + (lane & 31);
        int ecol = 0;
        for (int i = 0; i < pb.ncolbits; i++)
            if (col >> i & 1) ecol ^= pb.colp[i];
        floatx16 acc[HALVES];
#pragma unroll
        for (int h = 0; h < HALVES; h++)
#pragma unroll
            for (int v = 0; v < 16; v++) acc[h][v] = 0.0f;
        float bv[KSTEPS];
#pragma unroll
        for (int s = 0; s < KSTEPS; s++) bv[s] = bplane[pb.erow[s] ^ ecol];
        if constexpr (X6) {
            pass_bf16x6<HALVES, LOBLK>(acc, ap, bv);
        } else if constexpr (H3) {
            fp16x3_product<HALVES, LOBLK>(acc, ah, bv, bscale);
#pragma unroll
            for (int h = 0; h < HALVES; h++) acc[h] *= binv;
        } else {
#pragma unroll
            for (int s = 0; s < KSTEPS; s++)
#pragma unroll
                for (int h = 0; h < HALVES; h++)
                    acc[h] = __builtin_amdgcn_mfma_f32_32x32x2f32(ahi[NOHI ? 0 : h][NOHI ? 0 : s], bv[s], acc[h], 0, 0, 0);
        }
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
                const int e = pb.erow[out_row(m)] ^ ecol;
                if (out_plane(m)) pim[e] = acc[h][v];
                else pre[e] = acc[h][v];
            }
    }
}

// k = 4 operands of one block in registers (A of the FP32 product and, with
// the split, A of the BF16 correction).
template <int MODE>
struct Ops4 {
    float hi[MODE >= 2 ? 1 : 16];
    bf16x8 lo[MODE == 1 ? 2 : 1];
    bf16x8 p[MODE == 2 ? 3 : 1][1][MODE == 2 ? 2 : 1];   // BF16x6 pieces [piece][half][K-block]
    f16x8 h[MODE == 3 ? 2 : 1][1][MODE == 3 ? 2 : 1];    // FP16x3 pieces [piece][half][K-block]
};

template <int MODE>
__device__ __forceinline__ void load_ops4(Ops4<MODE> &o, const float *mops, int lane)
{
    if (MODE == 3) {
        const f16x8 *p = reinterpret_cast<const f16x8 *>(mops);
#pragma unroll
        for (int q = 0; q < 2; q++)
#pragma unroll
            for (int sb = 0; sb < 2; sb++) o.h[MODE == 3 ? q : 0][0][MODE == 3 ? sb : 0] = p[(q * 2 + sb) * 64 + lane];
        return;
    }
    if (MODE == 2) {
        const bf16x8 *p = reinterpret_cast<const bf16x8 *>(mops);
#pragma unroll
        for (int q = 0; q < 3; q++)
#pragma unroll
            for (int sb = 0; sb < 2; sb++) o.p[MODE == 2 ? q : 0][0][MODE == 2 ? sb : 0] = p[(q * 2 + sb) * 64 + lane];
        return;
    }
#pragma unroll
    for (int s = 0; s < 16; s++) o.hi[MODE >= 2 ? 0 : s] = mops[s * 64 + lane];
    if (MODE == 1) {
        const bf16x8 *lo = reinterpret_cast<const bf16x8 *>(mops + 16 * 64);
#pragma unroll
        for (int sb = 0; sb < 2; sb++) o.lo[MODE == 1 ? sb : 0] = lo[sb * 64 + lane];
    }
}

// One k = 4 block over the tile: 2^(TB-4) columns in 32-column chunks, chunk
// c on wave c % 4.  (Interleaving two chunks' MFMA chains and prefetching the
// next block's operands during this one measured within 3% of this form.)
// Addressing: column bits 0..4 come from the lane (their tile position is
// computed once per block), bits 5.. from the wave-uniform chunk index (scalar
// instructions); the 16 row addresses of a chunk serve the B reads and the
// write-back (a lane writes rows 4j' + 2 hf + {0, 1} of both planes).
template <int TB, int MODE, int NW>
__device__ __forceinline__ void pass_apply4(float *pre, const PassBlock &pb, const Ops4<MODE> &o, int lane,
                                            int wave, float bscale, float binv)
{
    constexpr bool SPLIT = MODE == 1;
    constexpr int NCB = TB - 4;                      // column bits
    constexpr int PER = ((1 << NCB) >> 5) / NW;      // chunks per wave: 1, 2, 4
    static_assert(PER >= 1, "fewer 32-column chunks than waves");
    constexpr int PLANE = 1 << TB;                   // im plane offset (floats)
    const int hf = lane >> 5;
    // byte offsets: the swizzled index is linear, so an element's offset is
    // the XOR of its row part, its column bits' parts and the plane bit.  The
    // lane's masks (0 or -1 per lane bit) do not depend on the block, so the
    // block loop keeps them; per block each column bit then costs one
    // three-input bit operation.
    int elane = hf << (TB + 2);
#pragma unroll
    for (int i = 0; i < 5; i++) elane ^= (-((lane >> i) & 1)) & (4 * pb.colp[i]);
    // rows this lane writes: 4 (j >> 1) + (j & 1) + 2 hf, i.e. the wave-uniform
    // offset wa0[j] of row r0 = 4 (j >> 1) + (j & 1), XOR-ed with wdm[j] (the
    // difference to row r0 + 2 in the lanes of the upper half); an indexed read
    // of erow would index the array dynamically, and hf * (b - a) costs a
    // quarter-rate multiply per row
    const int hfmask = -hf;
    int wa0[8], wdm[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const int r0 = 4 * (j >> 1) + (j & 1);
        wa0[j] = 4 * pb.erow[r0];
        wdm[j] = (wa0[j] ^ (4 * pb.erow[r0 + 2])) & hfmask;
    }
    const int wv = __builtin_amdgcn_readfirstlane(wave);
    char *const base = reinterpret_cast<char *>(pre);
    for (int c = 0; c < PER; c++) {
        const int chunk = wv + NW * c;               // wave-uniform
        int ehi = 0;
#pragma unroll
        for (int i = 5; i < NCB; i++)
            if (chunk >> (i - 5) & 1) ehi ^= 4 * pb.colp[i];
        const int ecol = elane ^ ehi;                // includes this lane's plane
        floatx16 acc;
#pragma unroll
        for (int v = 0; v < 16; v++) acc[v] = 0.0f;
        float bv[16];
#pragma unroll
        for (int s = 0; s < 16; s++) bv[s] = *reinterpret_cast<const float *>(base + ((4 * pb.erow[s]) ^ ecol));
        if constexpr (MODE == 2) {
            floatx16 a1[1] = {acc};
            pass_bf16x6<1, 2>(a1, o.p, bv);
            acc = a1[0];
        } else if constexpr (MODE == 3) {
            floatx16 a1[1] = {acc};
            fp16x3_product<1, 2>(a1, o.h, bv, bscale);
            acc = a1[0] * binv;
        } else {
#pragma unroll
            for (int s = 0; s < 16; s++)
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(o.hi[MODE >= 2 ? 0 : s], bv[s], acc, 0, 0, 0);
        }
        if (SPLIT) {
#pragma unroll
            for (int sb = 0; sb < 2; sb++) {
                bf16x8 bl;
#pragma unroll
                for (int e = 0; e < 8; e++) bl[e] = (__bf16)bv[8 * sb + e];
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(o.lo[SPLIT ? sb : 0], bl, acc, 0, 0, 0);
            }
        }
        // output register v holds real row m = 8 (v >> 2) + 4 hf + (v & 3):
        // complex row 4 (v >> 2) + 2 hf + (v & 1), plane (v >> 1) & 1 (out_row,
        // out_plane); the plane bit of ecol belongs to the B reads, so it is
        // cleared here
        const int wcol = ecol & ~(1 << (TB + 2));
#pragma unroll
        for (int v = 0; v < 16; v++) {
            const int j = 2 * (v >> 2) + (v & 1);
            *reinterpret_cast<float *>(base + (wa0[j] ^ wdm[j] ^ wcol) + ((v >> 1) & 1) * 4 * PLANE) = acc[v];
        }
    }
}

// Paired forms of pass_apply4 for the BF16x6 / FP16x3 products.  Qubit 0 is
// tile position 0 in every pass, and the LDS swizzle keeps bit 0, so the two
// amplitudes that differ in qubit 0 sit in adjacent floats of each plane.
// Every block then moves its data as 8-byte pairs: half the LDS instructions
// and half the address arithmetic of the dword form (a wave's ds_read_b32
// moves ~60 B/clk per CU on gfx950, ds_read_b64 ~110).
//
// PAIR_ROWS: qubit 0 is the block's row bit 0 (the planner puts it first).
// A lane reads its column's rows (2t, 2t + 1) with one 8-byte read and writes
// its output rows (2u, 2u + 1) of both planes with one ds_write2st64_b64.
template <int TB, int MODE, int NW>
__device__ __forceinline__ void pass_apply4_rows(float *pre, const PassBlock &pb, const Ops4<MODE> &o, int lane,
                                                 int wave, float one)
{
    static_assert(MODE == 2 || MODE == 3, "paired forms serve the half-precision products");
    constexpr int NCB = TB - 4;
    constexpr int PER = ((1 << NCB) >> 5) / NW;
    static_assert(PER >= 1, "fewer 32-column chunks than waves");
    constexpr int PLANE = 1 << TB;
    const int hf = lane >> 5;
    int elane = hf << (TB + 2);
#pragma unroll
    for (int i = 0; i < 5; i++) elane ^= (-((lane >> i) & 1)) & (4 * pb.colp[i]);
    // output rows 4 j + 2 hf (+1): offset of row 4 j, XOR-ed in the upper half
    // with the difference to row 4 j + 2
    const int hfmask = -hf;
    int wa0[4], wdm[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        wa0[j] = 4 * pb.erow[4 * j];
        wdm[j] = (wa0[j] ^ (4 * pb.erow[4 * j + 2])) & hfmask;
    }
    const int wv = __builtin_amdgcn_readfirstlane(wave);
    char *const base = reinterpret_cast<char *>(pre);
    for (int c = 0; c < PER; c++) {
        const int chunk = wv + NW * c;
        int ehi = 0;
#pragma unroll
        for (int i = 5; i < NCB; i++)
            if (chunk >> (i - 5) & 1) ehi ^= 4 * pb.colp[i];
        const int ecol = elane ^ ehi;
        float bv[16];
#pragma unroll
        for (int t = 0; t < 8; t++) {
            const float2 v = *reinterpret_cast<const float2 *>(base + ((4 * pb.erow[2 * t]) ^ ecol));
            bv[2 * t] = v.x;
            bv[2 * t + 1] = v.y;
        }
        floatx16 acc[1];
#pragma unroll
        for (int v = 0; v < 16; v++) acc[0][v] = 0.0f;
        if constexpr (MODE == 2) pass_bf16x6<1, 2>(acc, o.p, bv);
        else fp16x3_product<1, 2>(acc, o.h, bv, one);
        // acc[4 j + u]: row 4 j + 2 hf + (u & 1), plane u >> 1 (out_row,
        // out_plane), so each register pair is one plane's 8-byte pair
        const int wcol = ecol & ~(1 << (TB + 2));
        // both planes in one ds_write2st64_b64, straight from the accumulators
#pragma unroll
        for (int j = 0; j < 4; j++) {
            char *const a = base + (wa0[j] ^ wdm[j] ^ wcol);
            *reinterpret_cast<float2 *>(a) = make_float2(acc[0][4 * j], acc[0][4 * j + 1]);
            *reinterpret_cast<float2 *>(a + 4 * PLANE) = make_float2(acc[0][4 * j + 2], acc[0][4 * j + 3]);
        }
    }
}

// PAIR_COLS: qubit 0 is a column of the block (column bit 0).  A wave's two
// 32-column chunks are the two halves of 64 columns that differ in that bit,
// so one 8-byte read gives a lane the same row of both of its columns; lane
// bits take column bits 1..5 and the wave index the bits above.  The two
// products are independent chains on the same operands.
template <int TB, int MODE, int NW>
__device__ __forceinline__ void pass_apply4_cols(float *pre, const PassBlock &pb, const Ops4<MODE> &o, int lane,
                                                 int wave, float one)
{
    static_assert(MODE == 2 || MODE == 3, "paired forms serve the half-precision products");
    constexpr int NCB = TB - 4;
    static_assert((1 << NCB) == 64 * NW, "one pair of 32-column chunks per wave");
    constexpr int PLANE = 1 << TB;
    const int hf = lane >> 5;
    int elane = hf << (TB + 2);
#pragma unroll
    for (int i = 0; i < 5; i++) elane ^= (-((lane >> i) & 1)) & (4 * pb.colp[i + 1]);
    const int wv = __builtin_amdgcn_readfirstlane(wave);
    int ehi = 0;
#pragma unroll
    for (int i = 6; i < NCB; i++)
        if (wv >> (i - 6) & 1) ehi ^= 4 * pb.colp[i];
    const int ecol = elane ^ ehi;
    char *const base = reinterpret_cast<char *>(pre);
    float bv0[16], bv1[16];
#pragma unroll
    for (int s = 0; s < 16; s++) {
        const float2 v = *reinterpret_cast<const float2 *>(base + ((4 * pb.erow[s]) ^ ecol));
        bv0[s] = v.x;
        bv1[s] = v.y;
    }
    floatx16 acc0[1], acc1[1];
#pragma unroll
    for (int v = 0; v < 16; v++) {
        acc0[0][v] = 0.0f;
        acc1[0][v] = 0.0f;
    }
    if constexpr (MODE == 2) {
        pass_bf16x6<1, 2>(acc0, o.p, bv0);
        pass_bf16x6<1, 2>(acc1, o.p, bv1);
    } else {
        fp16x3_product<1, 2>(acc0, o.h, bv0, one);
        fp16x3_product<1, 2>(acc1, o.h, bv1, one);
    }
    // output register v: row 4 (v >> 2) + 2 hf + (v & 1), plane (v >> 1) & 1
    const int wcol = ecol & ~(1 << (TB + 2));
    const int hfmask = -hf;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const int r0 = 4 * (j >> 1) + (j & 1);
        const int a0 = 4 * pb.erow[r0], d = (a0 ^ (4 * pb.erow[r0 + 2])) & hfmask;
        char *const a = base + (a0 ^ d ^ wcol);
        const int v = 4 * (j >> 1) + (j & 1);
        *reinterpret_cast<float2 *>(a) = make_float2(acc0[0][v], acc1[0][v]);
        *reinterpret_cast<float2 *>(a + 4 * PLANE) = make_float2(acc0[0][v + 2], acc1[0][v + 2]);
    }
}

// ANY5: some block of the pass has 5 qubits (otherwise only the k = 4 product
// is compiled in, which keeps the register count at the k = 4 level).
template <int TB, int MODE, bool ANY5, bool PERM>
__device__ __forceinline__ void fused_pass_body(float2 *__restrict__ a, uint64_t ntiles, const PassArgs &pa)
{
    constexpr int TSIZE = 1 << TB;
    constexpr int NT = pass_threads(TB, ANY5), NW = NT / 64, TBITS = NT == 512 ? 9 : 8;
    constexpr int NLD = TSIZE / (2 * NT);    // 16-byte loads per thread
    __shared__ float planes[2][TSIZE];
    float *const pre = planes[0];
    float *const pim = planes[1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // XCD mapping (benchmark: 2.2% faster than tile = b; runs of 8, 16, 32,
    // 128, 256, 1024 tiles and a contiguous eighth per XCD measured 0.4-1.5%
    // slower than runs of 64; bit-identical)
    const uint64_t tile = xcd_block(blockIdx.x, gridDim.x);
    if (tile >= ntiles) return;
    const uint64_t base = expand_index(tile, pa.tile_ins);
    // element 2 tid + 2 NT i sits at base + pdep(2 tid, L) + pdep(2 NT i, L);
    // bit 0 of L is qubit 0, so element + 1 is the next amplitude in memory
    uint64_t toff = 0;
#pragma unroll
    for (int j = 1; j <= TBITS; j++)
        if ((tid >> (j - 1)) & 1) toff |= uint64_t(1) << pa.lpos[j];
    float4 buf[NLD];
    // nontemporal loads and stores: each amplitude is read and written once
    // per sweep (with the XCD tile mapping, 2% faster per benchmark circuit
    // than nontemporal loads with cached stores)
#pragma unroll
    for (int i = 0; i < NLD; i++) buf[i] = ld_nt(reinterpret_cast<const float4 *>(a + base + toff + pa.hioff[i]));
    // elements 2 tid + 2 NT i: the swizzle is linear, so it is one XOR for
    // tid's bits and a constant per (unrolled) i; it keeps bit 0, so element +
    // 1 follows at index + 1
    const int sw = lds_swizzle(2 * tid) ^ (2 * tid);
    // FP16x3: the tile sits in LDS scaled by 2^s (exact) from staging to the
    // store, so the blocks split it without a multiply on either side
    const float ls = MODE == 3 ? pa.bscale : 1.0f, li = MODE == 3 ? pa.binv : 1.0f;
    // (e, e + 1) move as one aligned 8-byte access per plane: as two dword
    // accesses a half-wave's even addresses would collide 2-way in the banks
#pragma unroll
    for (int i = 0; i < NLD; i++) {
        const int e = (2 * (tid + NT * i)) ^ sw ^ (lds_swizzle(2 * NT * i) ^ (2 * NT * i));
        *reinterpret_cast<float2 *>(pre + e) = make_float2(buf[i].x * ls, buf[i].z * ls);
        *reinterpret_cast<float2 *>(pim + e) = make_float2(buf[i].y * ls, buf[i].w * ls);
    }
    __syncthreads();
    if constexpr (ANY5) {
        for (int g = 0; g < pa.nblocks; g++) {
            if (pa.b[g].kq == 4) pass_apply<4, MODE, NW>(pre, pim, pa.b[g], lane, wave, pa.one, 1.0f);
            else pass_apply<5, MODE, NW>(pre, pim, pa.b[g], lane, wave, pa.one, 1.0f);
            __syncthreads();
        }
    } else {
        for (int g = 0; g < pa.nblocks; g++) {
            Ops4<MODE> ops;
            load_ops4<MODE>(ops, pa.b[g].mops, lane);
            if constexpr ((MODE == 2 || MODE == 3) && (TB == 12 || TB == 13)) {
                if (pa.b[g].pair == PAIR_ROWS) pass_apply4_rows<TB, MODE, NW>(pre, pa.b[g], ops, lane, wave, pa.one);
                else if (pa.b[g].pair == PAIR_COLS) pass_apply4_cols<TB, MODE, NW>(pre, pa.b[g], ops, lane, wave, pa.one);
                else pass_apply4<TB, MODE, NW>(pre, pa.b[g], ops, lane, wave, pa.one, 1.0f);
            } else {
                pass_apply4<TB, MODE, NW>(pre, pa.b[g], ops, lane, wave, pa.one, 1.0f);
            }
            __syncthreads();
        }
    }
    if constexpr (PERM) {
        uint64_t soff = 0;
        int sl = 0;
#pragma unroll
        for (int j = 0; j < TBITS; j++)
            if ((tid >> j) & 1) {
                soff |= pa.st_hbit[j];
                sl ^= pa.st_lbit[j];
            }
        if (pa.st_pair == 1) {
            // qubit 0 stays: the pair is adjacent and 8-byte aligned
#pragma unroll
            for (int i = 0; i < NLD; i++) {
                const int e = sl ^ pa.st_li[i];
                const float2 r = *reinterpret_cast<const float2 *>(pre + e), m = *reinterpret_cast<const float2 *>(pim + e);
                float4 v;
                v.x = r.x * li; v.y = m.x * li;
                v.z = r.y * li; v.w = m.y * li;
                st_nt(reinterpret_cast<float4 *>(a + base + soff + pa.st_hioff[i]), v);
            }
        } else {
#pragma unroll
            for (int i = 0; i < NLD; i++) {
                const int e0 = sl ^ pa.st_li[i], e1 = e0 ^ pa.st_pair;
                float4 v;
                v.x = pre[e0] * li; v.y = pim[e0] * li;
                v.z = pre[e1] * li; v.w = pim[e1] * li;
                st_nt(reinterpret_cast<float4 *>(a + base + soff + pa.st_hioff[i]), v);
            }
        }
    } else {
#pragma unroll
        for (int i = 0; i < NLD; i++) {
            const int e = (2 * (tid + NT * i)) ^ sw ^ (lds_swizzle(2 * NT * i) ^ (2 * NT * i));
            const float2 r = *reinterpret_cast<const float2 *>(pre + e), m = *reinterpret_cast<const float2 *>(pim + e);
            float4 v;
            v.x = r.x * li; v.y = m.x * li;
            v.z = r.y * li; v.w = m.y * li;
            st_nt(reinterpret_cast<float4 *>(a + base + toff + pa.hioff[i]), v);
        }
    }
}

static_assert(sizeof(PassArgs) <= 4096, "pass arguments must fit the 4 KiB kernel-argument budget");

template <int TB, int MODE, bool ANY5, bool PERM>
__global__ __launch_bounds__(pass_threads(TB, ANY5)) void k_fused_pass(float2 *__restrict__ a, uint64_t ntiles, PassArgs pa)
{
    fused_pass_body<TB, MODE, ANY5, PERM>(a, ntiles, pa);
}

// Tile bits a pass over these (widened) blocks needs: qubits 0..3 (the
// contiguous run) plus every block qubit above them, at least 11 and at most
// 13 (64 KiB of LDS).  -1 when they do not fit.
int pass_tile_bits(const std::vector<const std::vector<int> *> &qsets, int n)
{
    uint64_t u = 0;
    for (auto *q : qsets)
        for (int x : *q) u |= uint64_t(1) << x;
    const int tb = std::max(11, 4 + __builtin_popcountll(u & ~uint64_t(15)));
    return tb <= std::min(kMaxPassTileBits, n) ? tb : -1;
}

template <int TB, bool ANY5, bool PERM>
static void launch_tb5(qsv_state *s, uint64_t ntiles, const PassArgs &pa, int split)
{
    float2 *a = (float2 *)s->amps;
    const dim3 g((unsigned)ntiles), t(pass_threads(TB, ANY5));
    if (split == 3) hipLaunchKernelGGL((k_fused_pass<TB, 3, ANY5, PERM>), g, t, 0, s->stream, a, ntiles, pa);
    else if (split == 2) hipLaunchKernelGGL((k_fused_pass<TB, 2, ANY5, PERM>), g, t, 0, s->stream, a, ntiles, pa);
    else if (split) hipLaunchKernelGGL((k_fused_pass<TB, 1, ANY5, PERM>), g, t, 0, s->stream, a, ntiles, pa);
    else hipLaunchKernelGGL((k_fused_pass<TB, 0, ANY5, PERM>), g, t, 0, s->stream, a, ntiles, pa);
}

// perm: the pass exchanges qubits on the way out (13-bit tiles only)
template <int TB>
static void launch_tb(qsv_state *s, uint64_t ntiles, const PassArgs &pa, int split, bool perm)
{
    bool any5 = false;
    for (int g = 0; g < pa.nblocks; g++) any5 |= pa.b[g].kq == 5;
    if constexpr (TB == kMaxPassTileBits) {
        if (perm) {
            if (any5) launch_tb5<TB, true, true>(s, ntiles, pa, split);
            else launch_tb5<TB, false, true>(s, ntiles, pa, split);
            return;
        }
    }
    if (any5) launch_tb5<TB, true, false>(s, ntiles, pa, split);
    else launch_tb5<TB, false, false>(s, ntiles, pa, split);
}

// The permuted store's thread layout.  tau[j] = the input tile bit whose
// qubit leaves at output tile bit j.  Output bits 1..3 are thread bits 0..2
// (eight lanes store one 128-byte run); thread bits 3..5 take the three of
// output bits 4..8 that spread a wave's LDS reads over the most banks (32
// banks per half-wave and 64 per wave, both scored), so that a wave's stores
// stay within 4 KiB; the remaining bits follow in ascending order.  (Lane bits
// taken from the whole tile, which may hold qubits 2^26 amplitudes apart, made
// a pass that exchanges qubits 0..3 with 26..29 take 4.65 ms at 30 qubits
// against 3.1-3.4 ms for other exchanges.)
static void permuted_store(PassArgs &pa, int tb, int nt, const int *tau)
{
    const int tbits = __builtin_ctz(nt);
    auto q = [&](int j) { return lds_swizzle(1 << tau[j]); };
    int ob[16], nob = 0;
    for (int j = 1; j <= 3; j++) ob[nob++] = j;
    int best[3] = {4, 5, 6}, best_score = 1 << 30;
    const int top = std::min(tb, 9);
    for (int x = 4; x < top; x++)
        for (int y = x + 1; y < top; y++)
            for (int z = 4; z < top; z++) {
                if (z == x || z == y) continue;
                const int lb[6] = {1, 2, 3, x, y, z};
                int c32[2][32] = {}, c64[64] = {}, m32 = 0, m64 = 0;
                for (int l = 0; l < 64; l++) {
                    int e = 0;
                    for (int b = 0; b < 6; b++)
                        if (l >> b & 1) e ^= q(lb[b]);
                    m32 = std::max(m32, ++c32[l >> 5][e & 31]);
                    m64 = std::max(m64, ++c64[e & 63]);
                }
                const int score = 64 * m32 + m64;
                if (score < best_score) {
                    best_score = score;
                    best[0] = x;
                    best[1] = y;
                    best[2] = z;
                }
            }
    for (int b = 0; b < 3; b++) ob[nob++] = best[b];
    for (int j = 4; j < tb; j++)
        if (j != best[0] && j != best[1] && j != best[2]) ob[nob++] = j;
    // ob[0 .. tbits) are thread bits, the rest store-index bits
    for (int b = 0; b < tbits; b++) {
        pa.st_hbit[b] = uint64_t(1) << pa.lpos[ob[b]];
        pa.st_lbit[b] = q(ob[b]);
    }
    const int nld = (1 << tb) / (2 * nt);
    for (int i = 0; i < nld; i++) {
        uint64_t h = 0;
        int l = 0;
        for (int c = 0; tbits + c < tb - 1; c++)
            if (i >> c & 1) {
                h |= uint64_t(1) << pa.lpos[ob[tbits + c]];
                l ^= q(ob[tbits + c]);
            }
        pa.st_hioff[i] = h;
        pa.st_li[i] = l;
    }
    pa.st_pair = q(0);
}

// blocks[g]: widened qubit list (4 or 5 qubits) and device operand pointer.
// kernel: 0 or 1 (one workgroup per tile, the only sweep kernel).
// moves: (from, to) qubit pairs: the amplitudes leave the pass with the
// qubit at `from` moved to `to` (a permutation of some of the tile's qubits;
// 13-bit tiles only; a pass that only moves qubits has no blocks).
int launch_pass_dev(qsv_state *s, const std::vector<const std::vector<int> *> &qsets,
                    const std::vector<const float *> &mops, int split, int tb, int kernel,
                    const std::vector<std::pair<int, int>> *moves)
{
    if (s->prec != QSV_C64) return fail(QSV_ERR_PRECISION, "fused passes need a c64 state");
    const int G = (int)qsets.size();
    const bool perm = moves && !moves->empty();
    if (G < (perm ? 0 : 1) || G > kMaxPassBlocks) return fail(QSV_ERR_ARG, "1..16 blocks per pass");
    if (perm && tb != kMaxPassTileBits) return fail(QSV_ERR_ARG, "passes that move qubits have 13-bit tiles");
    if (tb < 11 || tb > kMaxPassTileBits || tb > s->n) return fail(QSV_ERR_ARG, "pass tile bits must be 11..13");
    if (split < 0 || split > 3) return fail(QSV_ERR_ARG, "split must be 0, 1, 2 or 3");
    if (kernel < 0 || kernel > 1) return fail(QSV_ERR_UNSUPPORTED, "pass kernel must be 0 or 1");
    uint64_t u = 0;
    for (auto *q : qsets) {
        if (q->size() != 4 && q->size() != 5) return fail(QSV_ERR_ARG, "pass blocks must have 4 or 5 qubits");
        for (int x : *q) {
            if (x < 0 || x >= s->n) return fail(QSV_ERR_ARG, "qubit out of range");
            u |= uint64_t(1) << x;
        }
    }
    uint64_t from = 0, to = 0;
    if (perm)
        for (const auto &m : *moves) {
            if (m.first < 0 || m.first >= s->n || m.second < 0 || m.second >= s->n)
                return fail(QSV_ERR_ARG, "moved qubit out of range");
            if ((from >> m.first & 1) || (to >> m.second & 1)) return fail(QSV_ERR_ARG, "qubit moved twice");
            from |= uint64_t(1) << m.first;
            to |= uint64_t(1) << m.second;
        }
    if (from != to) return fail(QSV_ERR_ARG, "qubit moves are not a permutation");
    u |= from;
    if (4 + __builtin_popcountll(u & ~uint64_t(15)) > tb) return fail(QSV_ERR_ARG, "blocks do not fit the pass tile");
    uint64_t lmask = u;
    for (int b = 0; __builtin_popcountll(lmask) < tb; b++) lmask |= uint64_t(1) << b;
    PassArgs pa{};
    pa.bscale = s->f16_scale;
    pa.binv = 1.0f / s->f16_scale;
    pa.one = 1.0f;
    pa.tile_ins = insert_set_from_mask(lmask);
    const int *lpos = pa.tile_ins.pos;
    for (int j = 0; j < tb; j++) pa.lpos[j] = lpos[j];
    if (lpos[0] != 0) return fail(QSV_ERR_ARG, "pass tile without qubit 0");
    auto pdep = [&](uint64_t e) {
        uint64_t o = 0;
        for (int j = 0; j < tb; j++)
            if (e >> j & 1) o |= uint64_t(1) << lpos[j];
        return o;
    };
    bool any5 = false;
    for (auto *q : qsets) any5 |= q->size() == 5;
    const int nt = pass_threads(tb, any5);
    for (int i = 0; i < (1 << tb) / (2 * nt); i++) pa.hioff[i] = pdep((uint64_t)(2 * nt) * i);
    auto rank = [&](int bit) {
        for (int i = 0; i < tb; i++)
            if (lpos[i] == bit) return i;
        return -1;
    };
    if (perm) {
        // tau[output tile bit] = input tile bit
        int tau[16];
        for (int i = 0; i < tb; i++) {
            int dst = lpos[i];
            for (const auto &m : *moves)
                if (m.first == lpos[i]) dst = m.second;
            tau[rank(dst)] = i;
        }
        permuted_store(pa, tb, nt, tau);
    }
    pa.nblocks = G;
    for (int g = 0; g < G; g++) {
        PassBlock &pb = pa.b[g];
        const std::vector<int> &q = *qsets[g];
        pb.mops = mops[g];
        pb.kq = (int)q.size();
        pb.ncolbits = tb - pb.kq;
        // paired data movement (k = 4 blocks, half-precision products, 12- or
        // 13-bit tiles): rows when qubit 0 is the block's row bit 0, columns
        // when the block does not hold qubit 0
        pb.pair = PAIR_NONE;
        if (!any5 && (split == 2 || split == 3) && (tb == 12 || tb == 13)) {
            bool has0 = false;
            for (int x : q) has0 |= x == 0;
            pb.pair = !has0 ? PAIR_COLS : (q[0] == 0 ? PAIR_ROWS : PAIR_NONE);
        }
        for (int r = 0; r < (1 << pb.kq); r++) {
            int e = 0;
            for (int j = 0; j < pb.kq; j++)
                if (r >> j & 1) e |= 1 << rank(q[j]);
            pb.erow[r] = lds_swizzle(e);
        }
        int c = 0;
        for (int i = 0; i < tb; i++) {
            bool inq = false;
            for (int x : q) inq |= (x == lpos[i]);
            if (!inq) pb.colp[c++] = lds_swizzle(1 << i);
        }
        // the pairs must be adjacent floats and every 8-byte access aligned
        if (pb.pair == PAIR_ROWS) {
            bool ok = pb.erow[1] == 1;
            for (int r = 0; r < 16; r += 2) ok &= (pb.erow[r] & 1) == 0 && pb.erow[r + 1] == (pb.erow[r] ^ 1);
            for (int i = 0; i < pb.ncolbits; i++) ok &= (pb.colp[i] & 1) == 0;
            if (!ok) pb.pair = PAIR_NONE;
        } else if (pb.pair == PAIR_COLS) {
            bool ok = pb.colp[0] == 1;
            for (int r = 0; r < 16; r++) ok &= (pb.erow[r] & 1) == 0;
            for (int i = 1; i < pb.ncolbits; i++) ok &= (pb.colp[i] & 1) == 0;
            if (!ok) pb.pair = PAIR_NONE;
        }
    }
    const uint64_t ntiles = s->dim >> tb;
    switch (tb) {
    case 11: launch_tb<11>(s, ntiles, pa, split, false); break;
    case 12: launch_tb<12>(s, ntiles, pa, split, false); break;
    default: launch_tb<13>(s, ntiles, pa, split, perm); break;
    }
    s->launches++;
    QSV_HIP(hipGetLastError());
    return QSV_OK;
}

}  // namespace qsv

using namespace qsv;

extern "C" int qsv_permute_qubits(qsv_state *s, const int *perm)
{
    QSV_CHECK_STATE(s);
    if (!perm) return fail(QSV_ERR_ARG, "null permutation");
    if (s->prec != QSV_C64) return fail(QSV_ERR_PRECISION, "qubit permutation runs on c64 states");
    if (s->n < kMaxPassTileBits) return fail(QSV_ERR_UNSUPPORTED, "qubit permutation needs >= 13 qubits");
    std::vector<int> dest(s->n);
    uint64_t hit = 0;
    for (int q = 0; q < s->n; q++) {
        if (perm[q] < 0 || perm[q] >= s->n || (hit >> perm[q] & 1)) return fail(QSV_ERR_ARG, "not a permutation");
        hit |= uint64_t(1) << perm[q];
        dest[q] = perm[q];
    }
    for (const auto &moves : qubit_routes(dest, kMaxPassTileBits)) {
        const int rc = launch_pass_dev(s, {}, {}, 0, kMaxPassTileBits, 0, &moves);
        if (rc) return rc;
    }
    return QSV_OK;
}
