/*
 * qsv_oracle.c -- CPU restatement of QPanda's CPU state-vector simulator.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle: only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark may load
 * it, and only as the checker.  The engine never calls it.
 *
 * It restates, in plain C99 on interleaved (re, im) doubles, the loops of
 * QPandaSDK/QuantumInstructionHandle/CPUQuantumGates.cpp (cited per function)
 * and the Park-Miller generator of QuantumGates.cpp:46-85.  `exact` selects the
 * ground-truth mode (1/sqrt(2) instead of the reference's truncated
 * 0.707106781186548, exact -1 for CZ); exact = 0 follows the reference bit for
 * bit in formula (operation order may differ; parity is 1e-12 per amplitude).
 *
 * Parity pins (see tests/test_oracle.py): the reference's own known-answer
 * tests (testMain/testCPUQuantumGates.cpp), golden vectors produced by the
 * compiled reference sources (tests/golden/), closed forms (GHZ, QFT = DFT via
 * numpy.fft), brute-force Kronecker products, and the Park-Miller KAT.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/qsv.h"

#define SQ2_REF 0.707106781186548          /* CPUQuantumGates.cpp:36 */
#define SQ2_EXACT 0.70710678118654752440

#define RE(p, i) ((p)[2 * (size_t)(i)])
#define IM(p, i) ((p)[2 * (size_t)(i) + 1])

/* ---------------------------------------------------------------- RNG ---- */
/* Schrage step of the 16807 generator, QuantumGates.cpp:46-85. */
int32_t orc_pm_next(int32_t s)
{
    const int32_t ia = 16807, im = 2147483647, iq = 127773, ir = 2836;
    int32_t t = ia * (s % iq) - ir * (s / iq);
    if (t < 0) t += im;
    return t;
}

static uint64_t mulmod31(uint64_t a, uint64_t b) { return (a * b) % 2147483647ull; }

/* s * 16807^k mod (2^31 - 1), square-and-multiply. */
int32_t orc_pm_jump(int32_t s, uint64_t k)
{
    uint64_t base = 16807, acc = 1;
    while (k) {
        if (k & 1) acc = mulmod31(acc, base);
        base = mulmod31(base, base);
        k >>= 1;
    }
    return (int32_t)mulmod31((uint64_t)s, acc);
}

/* Seed expression of QuantumGates.cpp:68-72. */
int32_t orc_pm_seed_from_tm(int y, int mo, int d, int h, int mi, int se)
{
    return y + 70 * (mo + 1 + 12 * (d + 31 * (h + 23 * (mi + 59 * se))));
}

static double pm_draw(int32_t *rng)
{
    if (!rng) return 1.0;
    *rng = orc_pm_next(*rng);
    return (double)*rng / 2147483647;
}

/* ------------------------------------------------------------ helpers ---- */
static inline void cmul(double ar, double ai, double br, double bi, double *cr, double *ci)
{
    *cr = ar * br - ai * bi;
    *ci = ar * bi + ai * br;
}

void orc_init_zero(double *psi, int n)
{
    size_t dim = (size_t)1 << n;
    memset(psi, 0, 2 * dim * sizeof(double));
    psi[0] = 1.0;
}

/* Index of the i-th amplitude whose bits at the sorted positions pos[0..m-1]
 * are all zero: the zero-bit insertion that CPUQuantumGates.cpp:231-249 does
 * one bit at a time. */
static size_t insert_zeros(size_t i, const int *pos_sorted, int m)
{
    for (int j = 0; j < m; j++) {
        int p = pos_sorted[j];
        size_t lo = i & (((size_t)1 << p) - 1);
        i = ((i >> p) << (p + 1)) | lo;
    }
    return i;
}

static int sort_list(const int *in, int m, int *out)
{
    for (int i = 0; i < m; i++) out[i] = in[i];
    for (int i = 1; i < m; i++) {
        int v = out[i], j = i - 1;
        while (j >= 0 && out[j] > v) { out[j + 1] = out[j]; j--; }
        out[j + 1] = v;
    }
    return m;
}

/* -------------------------------------------------- single-qubit gates ---- */
/* Hadamard, CPUQuantumGates.cpp:92-117. */
void orc_h(double *psi, int n, int q, int exact)
{
    const double c = exact ? SQ2_EXACT : SQ2_REF;
    size_t dim = (size_t)1 << n, st = (size_t)1 << q;
    for (size_t i = 0; i < dim; i += 2 * st)
        for (size_t j = i; j < i + st; j++) {
            double ar = RE(psi, j), ai = IM(psi, j), br = RE(psi, j + st), bi = IM(psi, j + st);
            RE(psi, j) = (ar + br) * c;      IM(psi, j) = (ai + bi) * c;
            RE(psi, j + st) = (ar - br) * c; IM(psi, j + st) = (ai - bi) * c;
        }
}

/* NOT, CPUQuantumGates.cpp:1277-1304. */
void orc_x(double *psi, int n, int q)
{
    size_t dim = (size_t)1 << n, st = (size_t)1 << q;
    for (size_t i = 0; i < dim; i += 2 * st)
        for (size_t j = i; j < i + st; j++) {
            double tr = RE(psi, j), ti = IM(psi, j);
            RE(psi, j) = RE(psi, j + st); IM(psi, j) = IM(psi, j + st);
            RE(psi, j + st) = tr;          IM(psi, j + st) = ti;
        }
}

/* RX, CPUQuantumGates.cpp:271-309: a' = a c + k s b, b' = b c + k s a with
 * k = -i (dagger: +i). */
void orc_rx(double *psi, int n, int q, double theta, int dagger)
{
    double c = cos(theta / 2), s = sin(theta / 2);
    double ki = dagger ? 1.0 : -1.0;                 /* k = (0, ki) */
    size_t dim = (size_t)1 << n, st = (size_t)1 << q;
    for (size_t i = 0; i < dim; i += 2 * st)
        for (size_t j = i; j < i + st; j++) {
            double ar = RE(psi, j), ai = IM(psi, j), br = RE(psi, j + st), bi = IM(psi, j + st);
            /* k*s*b = (0, ki*s) * (br, bi) = (-ki*s*bi, ki*s*br) */
            RE(psi, j) = ar * c - ki * s * bi;      IM(psi, j) = ai * c + ki * s * br;
            RE(psi, j + st) = br * c - ki * s * ai; IM(psi, j + st) = bi * c + ki * s * ar;
        }
}

/* RY, CPUQuantumGates.cpp:473-511. */
void orc_ry(double *psi, int n, int q, double theta, int dagger)
{
    double c = cos(theta / 2), s = sin(theta / 2);
    size_t dim = (size_t)1 << n, st = (size_t)1 << q;
    for (size_t i = 0; i < dim; i += 2 * st)
        for (size_t j = i; j < i + st; j++) {
            double ar = RE(psi, j), ai = IM(psi, j), br = RE(psi, j + st), bi = IM(psi, j + st);
            if (dagger) {
                RE(psi, j) = ar * c + s * br;      IM(psi, j) = ai * c + s * bi;
                RE(psi, j + st) = br * c - s * ar; IM(psi, j + st) = bi * c - s * ai;
            } else {
                RE(psi, j) = ar * c - s * br;      IM(psi, j) = ai * c - s * bi;
                RE(psi, j + st) = br * c + s * ar; IM(psi, j + st) = bi * c + s * ai;
            }
        }
}

/* RZ, CPUQuantumGates.cpp:524-570: |0> *= (c, -s), |1> *= (c, s); dagger
 * conjugates both. */
void orc_rz(double *psi, int n, int q, double theta, int dagger)
{
    double c = cos(theta / 2), s = sin(theta / 2);
    double s1 = dagger ? -s : s;                     /* phase on |1> */
    size_t dim = (size_t)1 << n, st = (size_t)1 << q;
    for (size_t i = 0; i < dim; i += 2 * st)
        for (size_t j = i; j < i + st; j++) {
            double r, m;
            cmul(c, -s1, RE(psi, j), IM(psi, j), &r, &m);
            RE(psi, j) = r; IM(psi, j) = m;
            cmul(c, s1, RE(psi, j + st), IM(psi, j + st), &r, &m);
            RE(psi, j + st) = r; IM(psi, j + st) = m;
        }
}

/* General 2x2 (a correct unitarySingleQubitGate; the reference's addresses
 * qubit qn-1, CPUQuantumGates.cpp:1188, which this does not reproduce). */
void orc_u1(double *psi, int n, int q, const double *u)
{
    size_t dim = (size_t)1 << n, st = (size_t)1 << q;
    for (size_t i = 0; i < dim; i += 2 * st)
        for (size_t j = i; j < i + st; j++) {
            double ar = RE(psi, j), ai = IM(psi, j), br = RE(psi, j + st), bi = IM(psi, j + st);
            double r0, i0, r1, i1;
            cmul(u[0], u[1], ar, ai, &r0, &i0); cmul(u[2], u[3], br, bi, &r1, &i1);
            RE(psi, j) = r0 + r1; IM(psi, j) = i0 + i1;
            cmul(u[4], u[5], ar, ai, &r0, &i0); cmul(u[6], u[7], br, bi, &r1, &i1);
            RE(psi, j + st) = r0 + r1; IM(psi, j + st) = i0 + i1;
        }
}

/* ----------------------------------------- controlled single-qubit gates ---
 * ctrl = controls followed by the target (Instruction.cpp:62 pushes the
 * target last); m = list length.  The loop enumerates the 2^(n-m) indices with
 * every listed bit 0, then adds `block` (all listed bits 1) minus the target
 * step, as CPUQuantumGates.cpp:215-257 does.  op: 0 H, 1 X, 2 RX, 3 RY, 4 RZ,
 * 5 U1.  The target is ctrl[m-1] (RY and NOT read it from the list,
 * CPUQuantumGates.cpp:407, :1323). */
void orc_ctrl_1q(double *psi, int n, const int *ctrl, int m, int op,
                 double theta, int dagger, const double *u, int exact)
{
    int sorted[64];
    int q = ctrl[m - 1];
    sort_list(ctrl, m, sorted);
    size_t block = 0;
    for (int j = 0; j < m; j++) block += (size_t)1 << sorted[j];
    size_t st = (size_t)1 << q, M = ((size_t)1 << n) >> m;
    double c = cos(theta / 2), s = sin(theta / 2);
    const double h = exact ? SQ2_EXACT : SQ2_REF;
    for (size_t i = 0; i < M; i++) {
        size_t i0 = insert_zeros(i, sorted, m) + block - st, i1 = i0 + st;
        double ar = RE(psi, i0), ai = IM(psi, i0), br = RE(psi, i1), bi = IM(psi, i1);
        double a2r, a2i, b2r, b2i;
        switch (op) {
        case 0: /* CPUQuantumGates.cpp:252-253 */
            a2r = (ar + br) * h; a2i = (ai + bi) * h; b2r = (ar - br) * h; b2i = (ai - bi) * h;
            break;
        case 1: /* CPUQuantumGates.cpp:1360-1362 */
            a2r = br; a2i = bi; b2r = ar; b2i = ai;
            break;
        case 2: { /* CPUQuantumGates.cpp:340-380: a' = a c - k s b, k = +i (dagger -i) */
            double ki = dagger ? -1.0 : 1.0;
            a2r = ar * c + ki * s * bi; a2i = ai * c - ki * s * br;
            b2r = br * c + ki * s * ai; b2i = bi * c - ki * s * ar;
            break; }
        case 3: /* CPUQuantumGates.cpp:444-455 */
            if (dagger) { a2r = ar * c + s * br; a2i = ai * c + s * bi; b2r = br * c - s * ar; b2i = bi * c - s * ai; }
            else        { a2r = ar * c - s * br; a2i = ai * c - s * bi; b2r = br * c + s * ar; b2i = bi * c + s * ai; }
            break;
        case 4: { /* CPUQuantumGates.cpp:610-650 */
            double s1 = dagger ? -s : s;
            cmul(c, -s1, ar, ai, &a2r, &a2i);
            cmul(c, s1, br, bi, &b2r, &b2i);
            break; }
        default: {
            double r0, j0, r1, j1;
            cmul(u[0], u[1], ar, ai, &r0, &j0); cmul(u[2], u[3], br, bi, &r1, &j1);
            a2r = r0 + r1; a2i = j0 + j1;
            cmul(u[4], u[5], ar, ai, &r0, &j0); cmul(u[6], u[7], br, bi, &r1, &j1);
            b2r = r0 + r1; b2i = j0 + j1;
            break; }
        }
        RE(psi, i0) = a2r; IM(psi, i0) = a2i; RE(psi, i1) = b2r; IM(psi, i1) = b2i;
    }
}

/* --------------------------------------------------- two/three-qubit ---- */
/* CNOT, CPUQuantumGates.cpp:668-737 (swap target pair where control = 1). */
void orc_cnot(double *psi, int n, int c, int t)
{
    size_t dim = (size_t)1 << n, sc = (size_t)1 << c, stt = (size_t)1 << t;
    for (size_t i = 0; i < dim; i++)
        if ((i & sc) && !(i & stt)) {
            double tr = RE(psi, i), ti = IM(psi, i);
            RE(psi, i) = RE(psi, i + stt); IM(psi, i) = IM(psi, i + stt);
            RE(psi, i + stt) = tr; IM(psi, i + stt) = ti;
        }
}

/* CR, CPUQuantumGates.cpp:773-825: |11> *= (cos theta, +-sin theta). */
void orc_cr(double *psi, int n, int q1, int q2, double theta, int dagger)
{
    double c = cos(theta), s = dagger ? -sin(theta) : sin(theta);
    size_t dim = (size_t)1 << n, s1 = (size_t)1 << q1, s2 = (size_t)1 << q2;
    if (q1 == q2) return;                            /* empty loop in the reference */
    for (size_t i = 0; i < dim; i++)
        if ((i & s1) && (i & s2)) {
            double r, m;
            cmul(RE(psi, i), IM(psi, i), c, s, &r, &m);
            RE(psi, i) = r; IM(psi, i) = m;
        }
}

/* |11> *= -1 (exact CZ, ground-truth mode). */
void orc_cz_exact(double *psi, int n, int q1, int q2)
{
    size_t dim = (size_t)1 << n, s1 = (size_t)1 << q1, s2 = (size_t)1 << q2;
    for (size_t i = 0; i < dim; i++)
        if ((i & s1) && (i & s2)) { RE(psi, i) = -RE(psi, i); IM(psi, i) = -IM(psi, i); }
}

/* iSWAP, CPUQuantumGates.cpp:838-899: |01> <-> |10> times k = +i (dagger -i).
 * sqiSWAP, :978-1039: a' = a h + k b h, b' = k a h + b h. */
static void iswap_like(double *psi, int n, int q1, int q2, int dagger, int sq, double h)
{
    int hi = q1 > q2 ? q1 : q2, lo = q1 > q2 ? q2 : q1;
    size_t dim = (size_t)1 << n, sh = (size_t)1 << hi, sl = (size_t)1 << lo;
    double ki = dagger ? -1.0 : 1.0;
    if (q1 == q2) return;
    for (size_t k = 0; k < dim; k++) {
        if ((k & sh) || !(k & sl)) continue;          /* k = |hi=0, lo=1> */
        size_t k2 = k + sh - sl;                       /* |hi=1, lo=0>     */
        double ar = RE(psi, k), ai = IM(psi, k), br = RE(psi, k2), bi = IM(psi, k2);
        if (!sq) {
            RE(psi, k) = -ki * bi; IM(psi, k) = ki * br;
            RE(psi, k2) = -ki * ai; IM(psi, k2) = ki * ar;
        } else {
            RE(psi, k) = ar * h + (-ki * bi) * h;  IM(psi, k) = ai * h + (ki * br) * h;
            RE(psi, k2) = (-ki * ai) * h + br * h; IM(psi, k2) = (ki * ar) * h + bi * h;
        }
    }
}
void orc_iswap(double *psi, int n, int q1, int q2, int dagger) { iswap_like(psi, n, q1, q2, dagger, 0, 0); }
void orc_sqiswap(double *psi, int n, int q1, int q2, int dagger, int exact)
{
    iswap_like(psi, n, q1, q2, dagger, 1, exact ? SQ2_EXACT : SQ2_REF);
}

/* toffoli, CPUQuantumGates.cpp:1378-1435: swap the target pair where both
 * controls are 1. */
void orc_toffoli(double *psi, int n, int c1, int c2, int t)
{
    size_t dim = (size_t)1 << n, a = (size_t)1 << c1, b = (size_t)1 << c2, st = (size_t)1 << t;
    for (size_t i = 0; i < dim; i++)
        if ((i & a) && (i & b) && (i & st)) {
            double tr = RE(psi, i), ti = IM(psi, i);
            RE(psi, i) = RE(psi, i - st); IM(psi, i) = IM(psi, i - st);
            RE(psi, i - st) = tr; IM(psi, i - st) = ti;
        }
}

/* controlSwap, CPUQuantumGates.cpp:934-965: where bit c = 1, bit a = 1 and
 * bit b = 0, exchange with the index that has a = 0, b = 1. */
void orc_cswap(double *psi, int n, int c, int a, int b)
{
    size_t dim = (size_t)1 << n, sc = (size_t)1 << c, sa = (size_t)1 << a, sb = (size_t)1 << b;
    for (size_t j = 0; j < dim; j++)
        if ((j & sc) && (j & sa) && !(j & sb)) {
            size_t k = j - sa + sb;
            double tr = RE(psi, j), ti = IM(psi, j);
            RE(psi, j) = RE(psi, k); IM(psi, j) = IM(psi, k);
            RE(psi, k) = tr; IM(psi, k) = ti;
        }
}

/* Dense k-qubit block: row/column bit j = bit(qubits[j]).  No reference
 * counterpart; pinned by Kronecker-product brute force and gate products. */
void orc_block(double *psi, int n, const int *qubits, int k, const double *u)
{
    int sorted[8];
    size_t d = (size_t)1 << k, dim = (size_t)1 << n;
    size_t off[32];
    double xr[32], xi[32];
    sort_list(qubits, k, sorted);
    for (size_t r = 0; r < d; r++) {
        off[r] = 0;
        for (int j = 0; j < k; j++)
            if (r >> j & 1) off[r] |= (size_t)1 << qubits[j];
    }
    for (size_t g = 0; g < (dim >> k); g++) {
        size_t base = insert_zeros(g, sorted, k);
        for (size_t c = 0; c < d; c++) { xr[c] = RE(psi, base + off[c]); xi[c] = IM(psi, base + off[c]); }
        for (size_t r = 0; r < d; r++) {
            double sr = 0, si = 0;
            for (size_t c = 0; c < d; c++) {
                double ur = u[2 * (r * d + c)], ui = u[2 * (r * d + c) + 1];
                sr += ur * xr[c] - ui * xi[c];
                si += ur * xi[c] + ui * xr[c];
            }
            RE(psi, base + off[r]) = sr; IM(psi, base + off[r]) = si;
        }
    }
}

/* Two-qubit dense matrix, row = 2*bit(qa) + bit(qb). */
void orc_u2(double *psi, int n, int qa, int qb, const double *u)
{
    int qs[2] = {qb, qa};                 /* block bit 0 = qb, bit 1 = qa */
    orc_block(psi, n, qs, 2, u);
}

/* ------------------------------------------------------------- readout ---- */
/* p0 of qubitMeasure, CPUQuantumGates.cpp:1074-1083: serial index order,
 * abs(a)*abs(a) with abs = hypot. */
double orc_prob0(const double *psi, int n, int q)
{
    size_t dim = (size_t)1 << n, st = (size_t)1 << q;
    double p = 0;
    for (size_t i = 0; i < dim; i += 2 * st)
        for (size_t j = i; j < i + st; j++) {
            double a = hypot(RE(psi, j), IM(psi, j));
            p += a * a;
        }
    return p;
}

/* qubitMeasure, CPUQuantumGates.cpp:1071-1126, with the draw r given. */
int orc_measure(double *psi, int n, int q, double r)
{
    size_t dim = (size_t)1 << n, st = (size_t)1 << q;
    double p = orc_prob0(psi, n, q);
    float fr = (float)r;
    int out = (fr > p) ? 1 : 0;
    double sc = out ? 1 / sqrt(1 - p) : 1 / sqrt(p);
    for (size_t i = 0; i < dim; i += 2 * st)
        for (size_t j = i; j < i + st; j++) {
            size_t keep = out ? j + st : j, drop = out ? j : j + st;
            RE(psi, keep) *= sc; IM(psi, keep) *= sc;
            RE(psi, drop) = 0;   IM(psi, drop) = 0;
        }
    return out;
}

/* Marginal distribution, CPUQuantumGates.cpp:159-176 (before the top-10 sort):
 * outcome bit j = bit(qnum[j]). */
void orc_marginals(const double *psi, int n, const int *qnum, int m, double *out)
{
    size_t dim = (size_t)1 << n;
    memset(out, 0, ((size_t)1 << m) * sizeof(double));
    for (size_t i = 0; i < dim; i++) {
        size_t idx = 0;
        for (int j = 0; j < m; j++) idx += ((i >> qnum[j]) & 1) << j;
        double a = hypot(RE(psi, i), IM(psi, i));
        out[idx] += a * a;
    }
}

double orc_norm2(const double *psi, int n)
{
    size_t dim = (size_t)1 << n;
    double s = 0;
    for (size_t i = 0; i < dim; i++) s += RE(psi, i) * RE(psi, i) + IM(psi, i) * IM(psi, i);
    return s;
}

/* sum_{i=0}^{n-2} <Z_i Z_{i+1}> = sum_x |a_x|^2 sum_i (-1)^(x_i xor x_{i+1}). */
double orc_expect_zz_chain(const double *psi, int n)
{
    size_t dim = (size_t)1 << n;
    double s = 0;
    for (size_t x = 0; x < dim; x++) {
        double p = RE(psi, x) * RE(psi, x) + IM(psi, x) * IM(psi, x);
        int e = 0;
        for (int i = 0; i + 1 < n; i++) e += (((x >> i) ^ (x >> (i + 1))) & 1) ? -1 : 1;
        s += p * e;
    }
    return s;
}

/* ------------------------------------------------------- circuit runner ---
 * Apply gate records in order.  Draw consumption follows the reference
 * executor: one randGenerator() per gate call (e.g. CPUQuantumGates.cpp:97),
 * one per measure (:1086); Reset = measure, then NOT (one more draw) when the
 * outcome is 1 (:126-136).  CZ runs as CR(pi) unless exact (the reference has
 * no CZ; QRunes programs write CR "pi").  Returns 0, or -1 on a bad record. */
int orc_apply_gate(double *psi, int n, const qsv_gate *g, const double *mats, int exact,
                   int32_t *rng, int *outcomes)
{
    switch (g->kind) {
    case QSV_GATE_H: pm_draw(rng); orc_h(psi, n, g->q0, exact); break;
    case QSV_GATE_X: pm_draw(rng); orc_x(psi, n, g->q0); break;
    case QSV_GATE_RX: pm_draw(rng); orc_rx(psi, n, g->q0, g->theta, g->dagger); break;
    case QSV_GATE_RY: pm_draw(rng); orc_ry(psi, n, g->q0, g->theta, g->dagger); break;
    case QSV_GATE_RZ: pm_draw(rng); orc_rz(psi, n, g->q0, g->theta, g->dagger); break;
    case QSV_GATE_CNOT: pm_draw(rng); orc_cnot(psi, n, g->q0, g->q1); break;
    case QSV_GATE_CZ:
        pm_draw(rng);
        if (exact) orc_cz_exact(psi, n, g->q0, g->q1);
        else orc_cr(psi, n, g->q0, g->q1, 3.14159265358979323846, 0);
        break;
    case QSV_GATE_CR: pm_draw(rng); orc_cr(psi, n, g->q0, g->q1, g->theta, g->dagger); break;
    case QSV_GATE_ISWAP: pm_draw(rng); orc_iswap(psi, n, g->q0, g->q1, g->dagger); break;
    case QSV_GATE_SQISWAP: pm_draw(rng); orc_sqiswap(psi, n, g->q0, g->q1, g->dagger, exact); break;
    case QSV_GATE_TOFFOLI: pm_draw(rng); orc_toffoli(psi, n, g->q0, g->q1, g->q2); break;
    case QSV_GATE_CSWAP: pm_draw(rng); orc_cswap(psi, n, g->q0, g->q1, g->q2); break;
    case QSV_GATE_U1: pm_draw(rng); orc_u1(psi, n, g->q0, mats + 2 * (size_t)g->mat); break;
    case QSV_GATE_U2: pm_draw(rng); orc_u2(psi, n, g->q0, g->q1, mats + 2 * (size_t)g->mat); break;
    case QSV_GATE_MEASURE: {
        double r = pm_draw(rng);
        int o = orc_measure(psi, n, g->q0, r);
        if (outcomes) outcomes[g->q1] = o;
        break; }
    case QSV_GATE_RESET: {
        double r = pm_draw(rng);
        if (orc_measure(psi, n, g->q0, r)) { pm_draw(rng); orc_x(psi, n, g->q0); }
        break; }
    default: return -1;
    }
    return 0;
}

int orc_run(double *psi, int n, const qsv_gate *g, int64_t ng, const double *mats, int exact,
            int32_t *rng, int *outcomes)
{
    for (int64_t i = 0; i < ng; i++)
        if (orc_apply_gate(psi, n, &g[i], mats, exact, rng, outcomes)) return -1;
    return 0;
}

/* Per-shot measurement chain of the reference shot loop (QPandaAPI.cpp:173-182,
 * QuantumInstructionHandleAPI.cpp:96-114): every shot re-simulates the program
 * from |0..0> and consumes `gate_draws` draws before its measurements, so the
 * state before the measurements is the same each shot.  `final` holds that
 * state; each shot measures qubits mq[0..m-1] in order into bits of
 * shots_out[shot] (bit j = outcome of mq[j]).  scratch: 2^n complex doubles. */
void orc_sample_shots(const double *final, int n, const int *mq, int m, int64_t gate_draws,
                      int32_t *rng, int64_t shots, uint64_t *shots_out, double *scratch)
{
    size_t bytes = ((size_t)2 << n) * sizeof(double);
    for (int64_t s = 0; s < shots; s++) {
        memcpy(scratch, final, bytes);
        *rng = orc_pm_jump(*rng, (uint64_t)gate_draws);
        uint64_t bits = 0;
        for (int j = 0; j < m; j++) {
            double r = pm_draw(rng);
            bits |= (uint64_t)orc_measure(scratch, n, mq[j], r) << j;
        }
        shots_out[s] = bits;
    }
}
