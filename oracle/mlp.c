/*
 * mlp.c -- FP64 oracle of the learned short-range force correction
 * (no ct2 counterpart: the reference evaluates its far field by L2P,
 * kick.cpp:29, with no learned part).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Network: x (KIN = 128 inputs, the first 111 used) -> H = 256, GELU -> 256,
 * GELU -> 32 outputs (the first 3 used).  Weights in the flat padded layout
 *   W1[128][256] b1[256] W2[256][256] b2[256] W3[256][32] b3[32]
 * (row = input, column = output).  GELU is the tanh form
 *   gelu(z) = z/2 (1 + tanh(c (z + 0.044715 z^3))),  c = sqrt(2/pi).
 * With bf16 != 0 the inputs of every layer (x, h1, h2) and the weights are
 * rounded to bfloat16 (round to nearest even) first, as on the GPU, and all
 * sums are FP64.  The loss is L = sum_r w_r sum_{o<3} (y_ro - t_ro)^2 / m with
 * a weight w_r per sample (T[r][3]); the training sets use w = 1/|F|^2 so that
 * L measures relative force error.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define KIN 128
#define HID 256
#define KOUT 32
#define NOUT 3

static const double GELU_C = 0.7978845608028654; /* sqrt(2/pi) */

size_t orc_mlp_num_params(void) { return (size_t)KIN * HID + HID + (size_t)HID * HID + HID + (size_t)HID * KOUT + KOUT; }

double orc_bf16_round(double v) {
    float f = (float)v;
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return v; /* inf / nan */
    const uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    u &= 0xffff0000u;
    memcpy(&f, &u, 4);
    return (double)f;
}

static double gelu(double z) { return 0.5 * z * (1.0 + tanh(GELU_C * (z + 0.044715 * z * z * z))); }
static double gelu_d(double z) {
    const double t = tanh(GELU_C * (z + 0.044715 * z * z * z));
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * GELU_C * (1.0 + 3.0 * 0.044715 * z * z);
}

void orc_gelu(size_t n, const double* z, double* g, double* dg) {
    for (size_t i = 0; i < n; i++) {
        if (g) g[i] = gelu(z[i]);
        if (dg) dg[i] = gelu_d(z[i]);
    }
}

typedef struct {
    const double *W1, *b1, *W2, *b2, *W3, *b3;
} net;

static net split(const double* w) {
    net p;
    p.W1 = w;
    p.b1 = p.W1 + KIN * HID;
    p.W2 = p.b1 + HID;
    p.b2 = p.W2 + HID * HID;
    p.W3 = p.b2 + HID;
    p.b3 = p.W3 + HID * KOUT;
    return p;
}

static double* rounded_weights(const double* w, int bf16) {
    const size_t P = orc_mlp_num_params();
    double* r = (double*)malloc(sizeof(double) * P);
    memcpy(r, w, sizeof(double) * P);
    if (bf16) { /* biases stay FP32 on the GPU: round only the matrices */
        net p = split(r);
        for (size_t i = 0; i < (size_t)KIN * HID; i++) ((double*)p.W1)[i] = orc_bf16_round(p.W1[i]);
        for (size_t i = 0; i < (size_t)HID * HID; i++) ((double*)p.W2)[i] = orc_bf16_round(p.W2[i]);
        for (size_t i = 0; i < (size_t)HID * KOUT; i++) ((double*)p.W3)[i] = orc_bf16_round(p.W3[i]);
    }
    return r;
}

/* one sample forward; keeps pre-activations z1, z2 and layer inputs */
static void fwd1(const net* p, const double* x_in, int bf16, double* x, double* z1, double* h1, double* z2,
                 double* h2, double* y) {
    for (int k = 0; k < KIN; k++) x[k] = bf16 ? orc_bf16_round(x_in[k]) : x_in[k];
    for (int j = 0; j < HID; j++) {
        double s = p->b1[j];
        for (int k = 0; k < KIN; k++) s += x[k] * p->W1[k * HID + j];
        z1[j] = s;
        h1[j] = bf16 ? orc_bf16_round(gelu(s)) : gelu(s);
    }
    for (int j = 0; j < HID; j++) {
        double s = p->b2[j];
        for (int k = 0; k < HID; k++) s += h1[k] * p->W2[k * HID + j];
        z2[j] = s;
        h2[j] = bf16 ? orc_bf16_round(gelu(s)) : gelu(s);
    }
    for (int j = 0; j < KOUT; j++) {
        double s = p->b3[j];
        for (int k = 0; k < HID; k++) s += h2[k] * p->W3[k * KOUT + j];
        y[j] = s;
    }
}

/* Forward over m samples: X[m][128] -> Y[m][32]. */
void orc_mlp_forward(const double* w, size_t m, const double* X, int bf16, double* Y) {
    double* r = rounded_weights(w, bf16);
    net p = split(r);
#pragma omp parallel
    {
        double x[KIN], z1[HID], h1[HID], z2[HID], h2[HID];
#pragma omp for schedule(static)
        for (ptrdiff_t s = 0; s < (ptrdiff_t)m; s++) fwd1(&p, X + (size_t)s * KIN, bf16, x, z1, h1, z2, h2, Y + (size_t)s * KOUT);
    }
    free(r);
}

/* Loss and its gradient with respect to the flat weights (same layout).
 * T[m][4]: targets of outputs 0..2 and the sample weight; outputs 3..31
 * carry no loss. */
double orc_mlp_grad(const double* w, size_t m, const double* X, const double* T, int bf16, double* grad) {
    const size_t P = orc_mlp_num_params();
    double* r = rounded_weights(w, bf16);
    net p = split(r);
    memset(grad, 0, sizeof(double) * P);
    double* gW1 = grad;
    double* gb1 = gW1 + KIN * HID;
    double* gW2 = gb1 + HID;
    double* gb2 = gW2 + HID * HID;
    double* gW3 = gb2 + HID;
    double* gb3 = gW3 + HID * KOUT;
    double loss = 0.0;
    double x[KIN], z1[HID], h1[HID], z2[HID], h2[HID], y[KOUT], dz2[HID], dz1[HID];
    for (size_t s = 0; s < m; s++) {
        fwd1(&p, X + s * KIN, bf16, x, z1, h1, z2, h2, y);
        double dy[KOUT] = {0};
        const double wr = T[s * 4 + 3];
        for (int o = 0; o < NOUT; o++) {
            const double e = y[o] - T[s * 4 + o];
            loss += wr * e * e / (double)m;
            dy[o] = 2.0 * wr * e / (double)m;
        }
        for (int o = 0; o < KOUT; o++) {
            gb3[o] += dy[o];
            for (int k = 0; k < HID; k++) gW3[k * KOUT + o] += h2[k] * dy[o];
        }
        for (int k = 0; k < HID; k++) {
            double d = 0.0;
            for (int o = 0; o < KOUT; o++) d += dy[o] * p.W3[k * KOUT + o];
            dz2[k] = d * gelu_d(z2[k]);
        }
        for (int j = 0; j < HID; j++) {
            gb2[j] += dz2[j];
            for (int k = 0; k < HID; k++) gW2[k * HID + j] += h1[k] * dz2[j];
        }
        for (int k = 0; k < HID; k++) {
            double d = 0.0;
            for (int j = 0; j < HID; j++) d += dz2[j] * p.W2[k * HID + j];
            dz1[k] = d * gelu_d(z1[k]);
        }
        for (int j = 0; j < HID; j++) {
            gb1[j] += dz1[j];
            for (int k = 0; k < KIN; k++) gW1[k * HID + j] += x[k] * dz1[j];
        }
    }
    free(r);
    return loss;
}
