// qsv_mfma.h -- the FP16x3 split product shared by the fused-block kernels.
//
// FP16x3 runs a real FP32 GEMM on the FP16 matrix cores
// (v_mfma_f32_32x32x16_f16, FP32 accumulation):
//   * the matrix M (|entries| <= 1) is split on the host into two fp16
//     pieces, M = a0 + a1 + r, a1 = fp16(M - a0), |r| <= 2^-22 (or 2^-25
//     absolute where a1 is subnormal);
//   * each state value x is scaled by a power of two s (exact), then split
//     the same way, x s = b0 + b1 + r' (x s - b0 is exact in FP32);
//   * the three products a1 b0 + a0 b1 + a0 b0 (smallest first) accumulate in
//     FP32 and the result is multiplied by 1/s (exact).
// The kernels keep a tile scaled by s in LDS from its load to its store, so
// the blocks in between split it without a multiply.
// The dropped terms (a1 b1, r, r') are ~2^-22 of each product: the operands
// keep 22 of FP32's 24 significand bits, against BF16x6's 24, for half of its
// matrix-core work (3 MFMAs per K-block instead of 6).  s = 2^15 for a state
// of norm <= 1 keeps |x s| < 2^15 below fp16's maximum 65504 while the
// typical amplitude of a 30-qubit state (2^-15) lands near 1; values below
// 2^-29 become fp16 subnormals, whose absolute error 2^-25 / s = 2^-40 is far
// below the FP32 rounding of the state itself.  The host derives s from the
// state's FP64 norm (f16_state_scale), so unnormalised states stay in range.
#pragma once

#include <hip/hip_runtime.h>

namespace qsv {

typedef float floatx16 __attribute__((ext_vector_type(16)));

// Output rows of the real product.  Real row m of the 2^(k+1) x 2^(k+1) real
// matrix holds complex row out_row(m), plane out_plane(m) (0 = re, 1 = im):
// m = 8 a + 4 b + 2 c + d -> row 4 a + 2 b + d, plane c.  The 32x32 MFMA
// leaves rows 8 (v >> 2) + 4 hf + (v & 3) in accumulator register v of lane
// half hf, so registers (v, v + 1), v even, are two adjacent complex rows of
// one plane: an 8-byte pair of the planar LDS tile when the rows differ in
// tile bit 0, written without register moves.
__host__ __device__ constexpr int out_row(int m) { return 4 * (m >> 3) + 2 * ((m >> 2) & 1) + (m & 1); }
__host__ __device__ constexpr int out_plane(int m) { return (m >> 1) & 1; }
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x2v __attribute__((ext_vector_type(2)));
typedef _Float16 f16x2v __attribute__((ext_vector_type(2)));

// acc[h] += M B over LOBLK K-blocks.  ap[p][h][sb]: fp16 piece p of the A
// operand in the permuted K order of the BF16 products (K-block sb, lane half
// hf, element e <-> real column 2 (8 sb + e) + hf), so the B values of K-block
// sb are this lane's FP32 B values bv[8 sb .. 8 sb + 7], already scaled.
// `one` must be 1.0f but unknown to the compiler (a kernel argument), so the
// residual x * one - fp16(x) is formed by fma-mix instructions that read the
// fp16 half in place (with a literal 1 it costs a conversion and a
// subtraction).  Written-out v_fma_mixlo/hi_f16 (one instruction per value
// instead of the compiler's mix-to-FP32 plus a packing conversion) measured
// 1% slower per window sweep: the VALU count is not what bounds the blocks.
template <int HALVES, int LOBLK>
__device__ __forceinline__ void fp16x3_product(floatx16 (&acc)[HALVES], const f16x8 (&ap)[2][HALVES][LOBLK],
                                               const float *bv, float one)
{
#pragma unroll
    for (int sb = 0; sb < LOBLK; sb++) {
        f16x8 b0, b1;
#pragma unroll
        for (int e = 0; e < 8; e += 2) {
            const float x0 = bv[8 * sb + e], x1 = bv[8 * sb + e + 1];
            const f16x2v h = __builtin_convertvector((f32x2v){x0, x1}, f16x2v);
            b0[e] = h[0];
            b0[e + 1] = h[1];
            // the residual is exact in FP32 and rounds to fp16 (v_fma_mix*)
            b1[e] = (_Float16)__builtin_fmaf(x0, one, -(float)h[0]);
            b1[e + 1] = (_Float16)__builtin_fmaf(x1, one, -(float)h[1]);
        }
#pragma unroll
        for (int h = 0; h < HALVES; h++) {
            acc[h] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ap[1][h][sb], b0, acc[h], 0, 0, 0);
            acc[h] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ap[0][h][sb], b1, acc[h], 0, 0, 0);
            acc[h] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ap[0][h][sb], b0, acc[h], 0, 0, 0);
        }
    }
}

}  // namespace qsv
