// qsv_gatemath.h -- the per-gate amplitude arithmetic, shared by the per-gate
// sweeps (qsv_gates.hip) and the LDS-tiled multi-gate sweep (qsv_tiles.hip),
// so that both compute bit-identical results: every function spells out its
// fused multiply-adds (and evaluates the rest as written, with no
// contraction), so the result cannot depend on how the surrounding kernel
// was compiled.
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

__device__ __forceinline__ double fmadd(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fmadd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// acc + u x for complex u = ur + i ui: two fused multiply-adds per component
template <typename T>
__device__ __forceinline__ T cmadd(T acc, typename Real<T>::type ur, typename Real<T>::type ui, T x)
{
    T y;
    y.x = fmadd(ur, x.x, fmadd(-ui, x.y, acc.x));
    y.y = fmadd(ur, x.y, fmadd(ui, x.x, acc.y));
    return y;
}

// u x for complex u = re + i im: one product and one fused multiply-add per
// component
template <typename T>
__device__ __forceinline__ T cmul(T x, typename Real<T>::type re, typename Real<T>::type im)
{
    T y;
    y.x = fmadd(re, x.x, -(im * x.y));
    y.y = fmadd(re, x.y, im * x.x);
    return y;
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
        y0 = cmul(x0, c.u[0], c.u[1]);
        y1 = cmul(x1, c.u[2], c.u[3]);
    } else {
        y0.x = 0; y0.y = 0; y1.x = 0; y1.y = 0;
        y0 = cmadd(y0, c.u[0], c.u[1], x0); y0 = cmadd(y0, c.u[2], c.u[3], x1);
        y1 = cmadd(y1, c.u[4], c.u[5], x0); y1 = cmadd(y1, c.u[6], c.u[7], x1);
    }
}

template <typename T>
__device__ __forceinline__ T phase_math(T x, typename Real<T>::type re, typename Real<T>::type im)
{
    return cmul(x, re, im);
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
            sr = fmadd(ur, xr[c], fmadd(-ui, xi[c], sr));
            si = fmadd(ur, xi[c], fmadd(ui, xr[c], si));
        }
    }
    T y;
    y.x = (typename Real<T>::type)sr;
    y.y = (typename Real<T>::type)si;
    return y;
}

}  // namespace qsv
