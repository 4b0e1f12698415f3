/*
 * mlp_oracle.c -- CPU restatement of the engine's cost-model network, in FP64.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP forward / backward kernels
 * (qpanda_amd/csrc/qsv_costmodel.hip).  Same architecture and parameter layout:
 * 12 -> 32 -> 32 -> 1, ReLU; W1[32][12], b1[32], W2[32][32], b2[32], W3[32], b3.
 * Loss 0.5 * mean((y - t)^2).  No reference counterpart (QPanda has no cost
 * model); pinned by central finite differences in tests/test_costmodel_cpu.py.
 */
#include <stdint.h>
#include <string.h>

#define F 12
#define H 32
#define OW1 0
#define OB1 (OW1 + H * F)
#define OW2 (OB1 + H)
#define OB2 (OW2 + H * H)
#define OW3 (OB2 + H)
#define OB3 (OW3 + H)
#define NP (OB3 + 1)

int orc_mlp_nparams(void) { return NP; }

static double fwd1(const double *p, const double *x, double *h1, double *h2)
{
    for (int j = 0; j < H; j++) {
        double s = p[OB1 + j];
        for (int i = 0; i < F; i++) s += p[OW1 + j * F + i] * x[i];
        h1[j] = s > 0 ? s : 0;
    }
    double out = p[OB3];
    for (int j = 0; j < H; j++) {
        double s = p[OB2 + j];
        for (int i = 0; i < H; i++) s += p[OW2 + j * H + i] * h1[i];
        h2[j] = s > 0 ? s : 0;
        out += p[OW3 + j] * h2[j];
    }
    return out;
}

void orc_mlp_forward(const double *p, const double *x, int64_t batch, double *y)
{
    double h1[H], h2[H];
    for (int64_t b = 0; b < batch; b++) y[b] = fwd1(p, x + b * F, h1, h2);
}

double orc_mlp_loss(const double *p, const double *x, const double *t, int64_t batch)
{
    double h1[H], h2[H], s = 0;
    for (int64_t b = 0; b < batch; b++) {
        const double d = fwd1(p, x + b * F, h1, h2) - t[b];
        s += 0.5 * d * d;
    }
    return s / (double)batch;
}

/* Gradient of orc_mlp_loss by the chain rule (ReLU derivative 0 at 0). */
double orc_mlp_grad(const double *p, const double *x, const double *t, int64_t batch, double *g)
{
    double h1[H], h2[H], dh2[H], loss = 0;
    memset(g, 0, sizeof(double) * NP);
    for (int64_t b = 0; b < batch; b++) {
        const double *xb = x + b * F;
        const double y = fwd1(p, xb, h1, h2);
        const double d = y - t[b];
        loss += 0.5 * d * d;
        const double dy = d / (double)batch;
        g[OB3] += dy;
        for (int j = 0; j < H; j++) {
            g[OW3 + j] += dy * h2[j];
            dh2[j] = h2[j] > 0 ? dy * p[OW3 + j] : 0;
            g[OB2 + j] += dh2[j];
            for (int i = 0; i < H; i++) g[OW2 + j * H + i] += dh2[j] * h1[i];
        }
        for (int i = 0; i < H; i++) {
            double s = 0;
            for (int j = 0; j < H; j++) s += p[OW2 + j * H + i] * dh2[j];
            const double dh1 = h1[i] > 0 ? s : 0;
            g[OB1 + i] += dh1;
            for (int k = 0; k < F; k++) g[OW1 + i * F + k] += dh1 * xb[k];
        }
    }
    return loss / (double)batch;
}
