// qsv_fused.hip -- fused k-qubit blocks (k = 4, 5) as FP32 complex GEMMs on MFMA.
//
// A block applies a dense 2^k x 2^k complex matrix U to every group of 2^k
// amplitudes that differ only in the block's qubits Q.  Written as a real GEMM,
// Y = M X with M the 2d x 2d real form of U (d = 2^k):
//     M[2r][2c] = Re U[r][c],  M[2r][2c+1] = -Im U[r][c],
//     M[2r+1][2c] = Im U[r][c], M[2r+1][2c+1] = Re U[r][c],
// and X's column j holding group j's amplitudes as (re, im) row pairs.
//
// Tiling:
//   L  = Q u {the TC lowest qubits not in Q}; L always contains qubits 0..TC-1,
//        so a tile -- the 2^(k+TC) amplitudes whose non-L bits are fixed --
//        consists of contiguous runs of >= 2^TC amplitudes (>= 512 B).
//   One workgroup per tile.  (Persistent workgroups that prefetch the next
//   tile into registers measured 5-15% slower at 30 qubits: the one-shot
//   kernel already streams at the in-place read-modify-write rate of a plain
//   float4 copy kernel, ~5.2 TB/s on MI355X.)
//   load:  16-byte loads of amplitude pairs (issued before the operand loads),
//          staged in LDS as separate re / im planes indexed by the compact
//          position e over L;
//   math:  wave w owns the 32 groups (columns) w*32..w*32+31 of the tile; per
//          K-step s it feeds amplitude s of its 32 columns (re in lanes 0-31,
//          im in lanes 32-63) as the B operand of v_mfma_f32_32x32x2f32, with M's
//          columns 2s, 2s+1 as A;
//          split = 1 adds the correction M_lo X with M_lo = M - fp32(M): its
//          terms are ~2^-24 of the result, so it runs on the BF16 MFMA
//          (v_mfma_f32_32x32x16_bf16) with both operands rounded to bf16
//          (relative error of the correction <= 2^-8, i.e. ~2^-32 of the
//          result), accumulated into the same FP32 registers.  Its K order is
//          permuted (K-block sb, lane half hf, element e <-> real column
//          2(8 sb + e) + hf) so that each lane's B values are exactly the FP32
//          B values it already holds from the K-steps 8 sb .. 8 sb + 7;
//          split = 2 (BF16x6) runs the whole product on the BF16 MFMA instead:
//          M = a0 + a1 + a2 and each B value b = b0 + b1 + b2 are split into
//          three bf16 pieces (b's split is exact: 3 x 8 significand bits
//          cover FP32's 24), and the six pieces products a_i b_j with
//          i + j <= 2 are accumulated in FP32, smallest first -- the
//          BF16x6 scheme of Henry, Tang and Heinecke (ARITH 2019).  The
//          dropped products are ~2^-24 of the result, the FP32 rounding
//          level.  Per K of 16 real rows it costs 6 v_mfma_f32_32x32x16_bf16
//          (6 x 32 cycles) against 8 v_mfma_f32_32x32x2f32 (8 x 64 cycles),
//          2.7x less matrix-core time, plus the VALU work of splitting B;
//          split = 3 (FP16x3, qsv_mfma.h) splits matrix and scaled state into
//          two fp16 pieces each and accumulates three piece products on the
//          FP16 MFMA: 22 of FP32's 24 significand bits for half of BF16x6's
//          matrix-core work;
//   store: results go back to the planes, then 16-byte coalesced stores.
// Per tile the kernel moves 2 * 8 * 2^(k+TC) bytes and performs
// 2 * (2d)^2 * 2^TC flops in the FP32 product, so k = 4 is HBM-bound (8 flop/B,
// below the FP32 MFMA ridge of ~20 flop/B) and k = 5 sits near the ridge.
//
// MFMA 32x32x2 f32 operand layout (one block): A[i][k] in lane i + 32k,
// B[k][j] in lane j + 32k, D[i][j] in lane j + 32*((i/4)%2), register
// (i/8)*4 + i%4.  32x32x16 bf16: A[i][k] in lane i + 32*(k/8), element k%8;
// B[k][j] in lane j + 32*(k/8), element k%8; same D layout.
#include "qsv_internal.h"
#include "qsv_stream.h"
#include "qsv_mfma.h"
#include "qsv_sched.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace qsv {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

struct FusedArgs {
    uint64_t off_high[32];   // address offset of the tile bits above the low run
    InsertSet tile_ins;      // L positions, to expand a tile index into its base
    int erow[32];            // compact tile position of row (amplitude) r
    int col_rank[8];         // rank in L of the i-th free tile (column) bit
    float bscale, binv;      // FP16x3: state scale 2^s and 2^-s
    float one;               // 1.0f (see fp16x3_product)
};

// acc[h] += M B over one K-block pair of bf16 pieces.  ap[p][h][sb]: piece p
// of the A operand (permuted K order); bv: this lane's FP32 B values.
template <int HALVES, int LOBLK>
__device__ __forceinline__ void bf16x6_product(floatx16 (&acc)[HALVES], const bf16x8 (&ap)[3][HALVES][LOBLK],
                                               const float *bv)
{
#pragma unroll
    for (int sb = 0; sb < LOBLK; sb++) {
        bf16x8 b0, b1, b2;
#pragma unroll
        for (int e = 0; e < 8; e++) {
            // b = b0 + b1 + b2 exactly: each residual is exact in FP32
            // (Sterbenz) and the last has at most 8 significant bits
            const float x = bv[8 * sb + e];
            const __bf16 t0 = (__bf16)x;
            const float r1 = x - (float)t0;
            const __bf16 t1 = (__bf16)r1;
            b0[e] = t0;
            b1[e] = t1;
            b2[e] = (__bf16)(r1 - (float)t1);
        }
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

// KQ: block qubits (4, 5).  TC: column bits per tile (6, 7, 8 -> 2, 4, 8 waves).
// MODE: 0 FP32 product, 1 FP32 + BF16 lo correction, 2 BF16x6, 3 FP16x3.
template <int KQ, int TC, int MODE>
__global__ __launch_bounds__(2 << TC) void k_fused_mfma(float2 *__restrict__ a, uint64_t ntiles, FusedArgs fa,
                                                         const float *__restrict__ mops)
{
    constexpr int D = 1 << KQ;               // amplitudes per group
    constexpr int THREADS = 2 << TC;         // one wave per 32 columns
    constexpr int TSIZE = 1 << (KQ + TC);    // amplitudes per tile
    constexpr int RUN = 1 << TC;             // contiguous low run (qubits 0..TC-1)
    constexpr int HALVES = (2 * D) / 32;     // 32-row output halves: 1 (k=4) or 2 (k=5)
    constexpr int KSTEPS = D;                // K = 2D real rows, 2 per step
    constexpr int NLD = TSIZE / (2 * THREADS);   // 16-byte loads per thread
    constexpr int LOBLK = D / 8;             // bf16 K-blocks (16 real rows = 8 amplitudes)
    constexpr bool SPLIT = MODE == 1;
    constexpr bool X6 = MODE == 2;
    constexpr bool H3 = MODE == 3;
    constexpr bool NOHI = X6 || H3;           // no FP32 A operand
    __shared__ float planes[2][TSIZE];     // [0] = re, [1] = im
    float *const pre = planes[0];
    float *const pim = planes[1];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint64_t tile = xcd_block(blockIdx.x, gridDim.x);
    if (tile >= ntiles) return;
    const uint64_t base = expand_index(tile, fa.tile_ins);

    // issue the tile's loads first; their latency covers the operand loads
    uint64_t off[NLD];
    float4 buf[NLD];
#pragma unroll
    for (int i = 0; i < NLD; i++) {
        const int e = 2 * (tid + THREADS * i);
        off[i] = base + (uint64_t)(e & (RUN - 1)) + fa.off_high[e >> TC];
        buf[i] = ld_nt(reinterpret_cast<const float4 *>(a + off[i]));
    }

    // A operands (the matrix), operand-ready layout: mops[(h*KSTEPS + s)*64 + lane]
    float ahi[NOHI ? 1 : HALVES][NOHI ? 1 : KSTEPS];
    bf16x8 alo[SPLIT ? HALVES : 1][SPLIT ? LOBLK : 1];
    bf16x8 ap[X6 ? 3 : 1][X6 ? HALVES : 1][X6 ? LOBLK : 1];
    f16x8 ah[H3 ? 2 : 1][H3 ? HALVES : 1][H3 ? LOBLK : 1];
    if (X6) {
        const bf16x8 *p = reinterpret_cast<const bf16x8 *>(mops);
#pragma unroll
        for (int q = 0; q < 3; q++)
#pragma unroll
            for (int h = 0; h < HALVES; h++)
#pragma unroll
                for (int sb = 0; sb < LOBLK; sb++)
                    ap[X6 ? q : 0][X6 ? h : 0][X6 ? sb : 0] = p[((q * HALVES + h) * LOBLK + sb) * 64 + lane];
    } else if (H3) {
        const f16x8 *p = reinterpret_cast<const f16x8 *>(mops);
#pragma unroll
        for (int q = 0; q < 2; q++)
#pragma unroll
            for (int h = 0; h < HALVES; h++)
#pragma unroll
                for (int sb = 0; sb < LOBLK; sb++)
                    ah[H3 ? q : 0][H3 ? h : 0][H3 ? sb : 0] = p[((q * HALVES + h) * LOBLK + sb) * 64 + lane];
    } else {
#pragma unroll
        for (int h = 0; h < HALVES; h++)
#pragma unroll
            for (int s = 0; s < KSTEPS; s++) ahi[NOHI ? 0 : h][NOHI ? 0 : s] = mops[(h * KSTEPS + s) * 64 + lane];
    }
    if (SPLIT) {
        const bf16x8 *lo = reinterpret_cast<const bf16x8 *>(mops + HALVES * KSTEPS * 64);
#pragma unroll
        for (int h = 0; h < HALVES; h++)
#pragma unroll
            for (int sb = 0; sb < LOBLK; sb++) alo[SPLIT ? h : 0][SPLIT ? sb : 0] = lo[(h * LOBLK + sb) * 64 + lane];
    }

    // FP16x3: the tile sits in LDS scaled by 2^s (exact), so the split needs
    // no multiply; the store scales back
    const float ls = H3 ? fa.bscale : 1.0f, li = H3 ? fa.binv : 1.0f;
    // (e, e + 1) as one aligned 8-byte access per plane (two dword accesses
    // at stride 2 would collide 2-way in the banks)
#pragma unroll
    for (int i = 0; i < NLD; i++) {
        const int e = 2 * (tid + THREADS * i);
        *reinterpret_cast<float2 *>(pre + e) = make_float2(buf[i].x * ls, buf[i].z * ls);
        *reinterpret_cast<float2 *>(pim + e) = make_float2(buf[i].y * ls, buf[i].w * ls);
    }
    __syncthreads();

    // this lane's column and its compact position bits
    const int col = wave * 32 + (lane & 31);
    int ecol = 0;
#pragma unroll
    for (int i = 0; i < TC; i++)
        if (col >> i & 1) ecol |= 1 << fa.col_rank[i];
    const float *const bplane = planes[lane >> 5];   // lanes 32-63 feed the im rows

    floatx16 acc[HALVES];
#pragma unroll
    for (int h = 0; h < HALVES; h++)
#pragma unroll
        for (int v = 0; v < 16; v++) acc[h][v] = 0.0f;
    float bv[KSTEPS];
#pragma unroll
    for (int s = 0; s < KSTEPS; s++) bv[s] = bplane[fa.erow[s] | ecol];
    if constexpr (X6) {
        bf16x6_product<HALVES, LOBLK>(acc, ap, bv);
    } else if constexpr (H3) {
        fp16x3_product<HALVES, LOBLK>(acc, ah, bv, fa.one);
    } else {
#pragma unroll
        for (int s = 0; s < KSTEPS; s++)
#pragma unroll
            for (int h = 0; h < HALVES; h++)
                acc[h] = __builtin_amdgcn_mfma_f32_32x32x2f32(ahi[NOHI ? 0 : h][NOHI ? 0 : s], bv[s], acc[h], 0, 0, 0);
    }
    if (SPLIT) {
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
    // all B reads of this wave precede its writes (data dependence); other
    // waves touch other columns, so no barrier is needed before the write-back
#pragma unroll
    for (int h = 0; h < HALVES; h++)
#pragma unroll
        for (int v = 0; v < 16; v++) {
            const int m = 32 * h + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);   // real output row
            const int e = fa.erow[out_row(m)] | ecol;
            if (out_plane(m)) pim[e] = acc[h][v];
            else pre[e] = acc[h][v];
        }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NLD; i++) {
        const int e = 2 * (tid + THREADS * i);
        const float2 r = *reinterpret_cast<const float2 *>(pre + e), m = *reinterpret_cast<const float2 *>(pim + e);
        float4 v;
        v.x = r.x * li; v.y = m.x * li;
        v.z = r.y * li; v.w = m.y * li;
        st_nt(reinterpret_cast<float4 *>(a + off[i]), v);
    }
}

// Host-side: operand-ready matrices for a block.  u: d x d complex doubles
// (row bit j = bit(qubits[j])).  Layout: HALVES * D * 64 floats of M_hi
// (value M[32h + l%32][2s + l/32] at (h*D + s)*64 + l), then, with split = 1,
// HALVES * (D/8) * 64 * 8 bf16 of M_lo in the permuted K order (value
// M_lo[32h + l%32][2(8sb + e) + l/32] at ((h*(D/8) + sb)*64 + l)*8 + e).
// split = 2: only bf16, pieces p = 0, 1, 2 of M (a0 = bf16(M), a1 = bf16(M -
// a0), a2 = bf16(M - a0 - a1), residuals in FP64) in the permuted K order,
// value a_p[32h + l%32][2(8sb + e) + l/32] at (((p*HALVES + h)*(D/8) + sb)*64
// + l)*8 + e: 1.5x the floats of M_hi, the size of the split = 1 layout.
// split = 3: fp16 pieces p = 0, 1 (a0 = fp16(M), a1 = fp16(M - a0), each
// rounded once from FP64) in the same permuted K order and index formula
// with 2 pieces: the floats of M_hi.
static uint16_t to_bf16(float f)
{
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t lsb = (x >> 16) & 1;          // round to nearest even
    x += 0x7FFFu + lsb;
    return (uint16_t)(x >> 16);
}

static double from_bf16(uint16_t b)
{
    const uint32_t x = (uint32_t)b << 16;
    float f;
    std::memcpy(&f, &x, 4);
    return f;
}

void fused_operands(int kq, const double *u, int split, float *out)
{
    const int d = 1 << kq, R = 2 * d, halves = R / 32;
    // real row i: complex row out_row(i), plane out_plane(i); real column j:
    // complex column j >> 1, plane j & 1 (the K order of the B operand)
    auto M = [&](int i, int j) -> double {
        const int r = out_row(i), c = j >> 1;
        const double ur = u[2 * (r * d + c)], ui = u[2 * (r * d + c) + 1];
        if (out_plane(i) == 0) return (j & 1) == 0 ? ur : -ui;
        return (j & 1) == 0 ? ui : ur;
    };
    if (split == 3) {
        _Float16 *p = reinterpret_cast<_Float16 *>(out);
        const int lob = d / 8;
        for (int h = 0; h < halves; h++)
            for (int sb = 0; sb < lob; sb++)
                for (int l = 0; l < 64; l++)
                    for (int e = 0; e < 8; e++) {
                        double r = M(32 * h + (l & 31), 2 * (8 * sb + e) + (l >> 5));
                        for (int q = 0; q < 2; q++) {
                            const _Float16 b = (_Float16)r;
                            p[(((q * halves + h) * lob + sb) * 64 + l) * 8 + e] = b;
                            r -= (double)b;
                        }
                    }
        return;
    }
    if (split == 2) {
        uint16_t *p = reinterpret_cast<uint16_t *>(out);
        const int lob = d / 8;
        for (int h = 0; h < halves; h++)
            for (int sb = 0; sb < lob; sb++)
                for (int l = 0; l < 64; l++)
                    for (int e = 0; e < 8; e++) {
                        double r = M(32 * h + (l & 31), 2 * (8 * sb + e) + (l >> 5));
                        for (int q = 0; q < 3; q++) {
                            const uint16_t b = to_bf16((float)r);
                            p[(((q * halves + h) * lob + sb) * 64 + l) * 8 + e] = b;
                            r -= from_bf16(b);
                        }
                    }
        return;
    }
    for (int h = 0; h < halves; h++)
        for (int s = 0; s < d; s++)
            for (int l = 0; l < 64; l++)
                out[(h * d + s) * 64 + l] = (float)M(32 * h + (l & 31), 2 * s + (l >> 5));
    if (!split) return;
    uint16_t *lo = reinterpret_cast<uint16_t *>(out + halves * d * 64);
    const int lob = d / 8;
    for (int h = 0; h < halves; h++)
        for (int sb = 0; sb < lob; sb++)
            for (int l = 0; l < 64; l++)
                for (int e = 0; e < 8; e++) {
                    const double v = M(32 * h + (l & 31), 2 * (8 * sb + e) + (l >> 5));
                    lo[((h * lob + sb) * 64 + l) * 8 + e] = to_bf16((float)(v - (double)(float)v));
                }
}

size_t fused_operand_floats(int kq, int split)
{
    const int d = 1 << kq, halves = (2 * d) / 32;
    const size_t hi = (size_t)halves * d * 64;
    // split 1: the bf16 lo part is half the bytes of the hi part; split 2:
    // three bf16 pieces, the same 1.5x; split 3: two fp16 pieces, 1x
    if (split == 3) return hi;
    return hi + (split ? hi / 2 : 0);
}

template <int KQ, int TC>
static void launch_kq_tc(qsv_state *s, uint64_t ntiles, const FusedArgs &fa, const float *mops, int split)
{
    float2 *a = (float2 *)s->amps;
    const dim3 g((unsigned)ntiles), t(2 << TC);
    if (split == 3) hipLaunchKernelGGL((k_fused_mfma<KQ, TC, 3>), g, t, 0, s->stream, a, ntiles, fa, mops);
    else if (split == 2) hipLaunchKernelGGL((k_fused_mfma<KQ, TC, 2>), g, t, 0, s->stream, a, ntiles, fa, mops);
    else if (split) hipLaunchKernelGGL((k_fused_mfma<KQ, TC, 1>), g, t, 0, s->stream, a, ntiles, fa, mops);
    else hipLaunchKernelGGL((k_fused_mfma<KQ, TC, 0>), g, t, 0, s->stream, a, ntiles, fa, mops);
}

// Launch one fused block (kq = 4 or 5, qubits[0..kq-1] distinct) with
// operands already on the device.  tc = tile column bits (6, 7, 8; 0 = 7).
int launch_fused_dev(qsv_state *s, const int *qubits, int kq, const float *mops, int split, int tc)
{
    if (tc == 0) tc = kDefaultTileCols;
    if (s->prec != QSV_C64) return fail(QSV_ERR_PRECISION, "fused MFMA blocks need a c64 state");
    if (kq != 4 && kq != 5) return fail(QSV_ERR_ARG, "fused block size must be 4 or 5");
    if (tc < 6 || tc > 8) return fail(QSV_ERR_ARG, "tile column bits must be 6, 7 or 8");
    if (split < 0 || split > 3) return fail(QSV_ERR_ARG, "split must be 0, 1, 2 or 3");
    if (s->n < kq + tc) return fail(QSV_ERR_ARG, "state too small for a fused tile");
    FusedArgs fa{};
    fa.bscale = s->f16_scale;
    fa.binv = 1.0f / s->f16_scale;
    fa.one = 1.0f;
    uint64_t qmask = 0;
    for (int j = 0; j < kq; j++) qmask |= uint64_t(1) << qubits[j];
    // the tc lowest free bits are the tile's columns
    int free_bits[8], nf = 0;
    for (int b = 0; b < s->n && nf < tc; b++)
        if (!(qmask >> b & 1)) free_bits[nf++] = b;
    uint64_t lmask = qmask;
    for (int i = 0; i < tc; i++) lmask |= uint64_t(1) << free_bits[i];
    fa.tile_ins = insert_set_from_mask(lmask);
    int lpos[16];
    for (int i = 0; i < fa.tile_ins.n; i++) lpos[i] = fa.tile_ins.pos[i];
    auto rank = [&](int bit) {
        for (int i = 0; i < kq + tc; i++)
            if (lpos[i] == bit) return i;
        return -1;
    };
    for (int i = 0; i < tc; i++) fa.col_rank[i] = rank(free_bits[i]);
    const int d = 1 << kq;
    for (int r = 0; r < d; r++) {
        int e = 0;
        for (int j = 0; j < kq; j++)
            if (r >> j & 1) e |= 1 << rank(qubits[j]);
        fa.erow[r] = e;
    }
    // L positions 0..tc-1 are qubits 0..tc-1; ranks >= tc map the high tile bits
    for (int hbits = 0; hbits < d; hbits++) {
        uint64_t o = 0;
        for (int j = 0; j < kq; j++)
            if (hbits >> j & 1) o |= uint64_t(1) << lpos[tc + j];
        fa.off_high[hbits] = o;
    }
    const uint64_t ntiles = s->dim >> (kq + tc);
    if (kq == 4) {
        if (tc == 6) launch_kq_tc<4, 6>(s, ntiles, fa, mops, split);
        else if (tc == 7) launch_kq_tc<4, 7>(s, ntiles, fa, mops, split);
        else launch_kq_tc<4, 8>(s, ntiles, fa, mops, split);
    } else {
        if (tc == 6) launch_kq_tc<5, 6>(s, ntiles, fa, mops, split);
        else if (tc == 7) launch_kq_tc<5, 7>(s, ntiles, fa, mops, split);
        else launch_kq_tc<5, 8>(s, ntiles, fa, mops, split);
    }
    s->launches++;
    QSV_HIP(hipGetLastError());
    return QSV_OK;
}

// FP16x3 state scale from the state's FP64 norm: s = 2^(15 - e) with 2^e >=
// the largest possible amplitude sqrt(norm), so every |x s| stays below 2^15.
int f16_state_scale(qsv_state *s)
{
    double nrm = 0;
    int rc = reduce_norm2(s, &nrm);
    if (rc) return rc;
    int e = 0;
    if (nrm > 0 && std::isfinite(nrm)) e = (int)std::ceil(0.5 * std::log2(nrm));
    e = std::max(-60, std::min(60, e));
    s->f16_scale = std::ldexp(1.0f, 15 - e);
    return QSV_OK;
}

// qsv_apply_block on a c64 state: blocks of 4 or 5 qubits go to the MFMA
// kernel; smaller blocks are widened with identity on the lowest free qubits
// (same sweep, same kernel); states too small for a tile use the FP64 block
// kernel.
int launch_block_mfma(qsv_state *s, const int *qubits, int k, const double *u, int split)
{
    int kq = k < 4 ? 4 : k;
    if (s->n < kq + kDefaultTileCols) return launch_block(s, qubits, k, u, 0);
    std::vector<int> qs(qubits, qubits + k);
    std::vector<double> w;
    const double *uu = u;
    if (kq != k) {
        uint64_t used = 0;
        for (int q : qs) used |= uint64_t(1) << q;
        for (int b = 0; (int)qs.size() < kq; b++)
            if (!(used >> b & 1)) qs.push_back(b);
        const int d = 1 << k, D = 1 << kq;
        w.assign(2 * (size_t)D * D, 0.0);
        // U (x) I: rows/cols with equal high (padding) bits carry U
        for (int hi = 0; hi < D / d; hi++)
            for (int r = 0; r < d; r++)
                for (int c = 0; c < d; c++) {
                    const size_t R = (size_t)hi * d + r, Cc = (size_t)hi * d + c;
                    w[2 * (R * D + Cc)] = u[2 * (r * d + c)];
                    w[2 * (R * D + Cc) + 1] = u[2 * (r * d + c) + 1];
                }
        uu = w.data();
    }
    if (split == 3) {
        int rc = f16_state_scale(s);
        if (rc) return rc;
    }
    std::vector<float> ops(fused_operand_floats(kq, split));
    fused_operands(kq, uu, split, ops.data());
    const size_t bytes = ops.size() * sizeof(float);
    if (bytes > sizeof(double) * (kRedSlots - 8)) return fail(QSV_ERR_ARG, "operand block too large");
    QSV_HIP(hipMemcpyAsync(s->red, ops.data(), bytes, hipMemcpyHostToDevice, s->stream));
    int rc = launch_fused_dev(s, qs.data(), kq, (const float *)s->red, split, 0);
    if (!rc) QSV_HIP(hipStreamSynchronize(s->stream));   // ops is a host temporary
    return rc;
}

}  // namespace qsv
