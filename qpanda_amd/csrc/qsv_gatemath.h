// qsv_gatemath.h -- the per-gate amplitude arithmetic, shared by the per-gate
// sweeps (qsv_gates.hip) and the LDS-tiled multi-gate sweep (qsv_tiles.hip),
// so that both compute bit-identical results: every function evaluates its
// products and sums as written, with no fused multiply-add contraction, so
// the result cannot depend on how the surrounding kernel was compiled.
//   pair ops  (H, X, 2x2, diagonal) on the two members of a pair;
//   phase     one amplitude times a complex factor (CR, CZ);
//   block     a dense 2^k x 2^k FP64 matrix on a group of 2^k amplitudes.
// The formulas are those of QPanda's CPU simulator (CPUQuantumGates.cpp:
// Hadamard :92-117, NOT :1277-1304, RX/RY/RZ :271-570, CR :773-825, iSWAP /
// sqiSWAP :838-1039, Fredkin :934-965), written in the engine's pair form.
#pragma once

#include "qsv_internal.h"

namespace qsv {

// Pair kernel operations (launch_pair)
enum PairOp { OP_H = 0, OP_X = 1, OP_U = 2, OP_DIAG = 3 };

template <typename T> struct Real;
template <> struct Real<c128> { using type = double; };
template <> struct Real<c64> { using type = float; };

template <typename R> struct Coef {
    R u[8];
};

template <typename T>
__device__ __forceinline__ T cmadd(T acc, typename Real<T>::type ur, typename Real<T>::type ui, T x)
{
#pragma clang fp contract(off)
    acc.x += ur * x.x - ui * x.y;
    acc.y += ur * x.y + ui * x.x;
    return acc;
}

// (x0, x1) -> (y0, y1) for one pair; c: OP_H c.u[0] = the Hadamard constant,
// OP_DIAG c.u[0..3] = d0, d1, OP_U c.u[0..7] = the row-major 2x2 matrix.
template <typename T, int OP>
__device__ __forceinline__ void pair_math(T x0, T x1, const Coef<typename Real<T>::type> &c, T &y0, T &y1)
{
#pragma clang fp contract(off)
    using R = typename Real<T>::type;
    if (OP == OP_H) {
        const R h = c.u[0];
        y0.x = (x0.x + x1.x) * h; y0.y = (x0.y + x1.y) * h;
        y1.x = (x0.x - x1.x) * h; y1.y = (x0.y - x1.y) * h;
    } else if (OP == OP_X) {
        y0 = x1; y1 = x0;
    } else if (OP == OP_DIAG) {
        y0.x = c.u[0] * x0.x - c.u[1] * x0.y; y0.y = c.u[0] * x0.y + c.u[1] * x0.x;
        y1.x = c.u[2] * x1.x - c.u[3] * x1.y; y1.y = c.u[2] * x1.y + c.u[3] * x1.x;
    } else {
        y0.x = 0; y0.y = 0; y1.x = 0; y1.y = 0;
        y0 = cmadd(y0, c.u[0], c.u[1], x0); y0 = cmadd(y0, c.u[2], c.u[3], x1);
        y1 = cmadd(y1, c.u[4], c.u[5], x0); y1 = cmadd(y1, c.u[6], c.u[7], x1);
    }
}

template <typename T>
__device__ __forceinline__ T phase_math(T x, typename Real<T>::type re, typename Real<T>::type im)
{
#pragma clang fp contract(off)
    T y;
    y.x = re * x.x - im * x.y;
    y.y = re * x.y + im * x.x;
    return y;
}

// Row r of a dense d x d FP64 block (u: row-major complex doubles) applied to
// the group's amplitudes (xr, xi), in the result type of T.
template <typename T, int MAXD>
__device__ __forceinline__ T block_row_math(const double *u, int d, int r, const double (&xr)[MAXD],
                                            const double (&xi)[MAXD])
{
#pragma clang fp contract(off)
    double sr = 0, si = 0;
#pragma unroll
    for (int c = 0; c < MAXD; c++) {
        if (c < d) {
            const double ur = u[2 * (r * d + c)], ui = u[2 * (r * d + c) + 1];
            sr += ur * xr[c] - ui * xi[c];
            si += ur * xi[c] + ui * xr[c];
        }
    }
    T y;
    y.x = (typename Real<T>::type)sr;
    y.y = (typename Real<T>::type)si;
    return y;
}

}  // namespace qsv
