/*
 * FP64 CPU oracle of the constitutive-surrogate MLP (test infrastructure
 * only; no reference counterpart). Forward: y = W3 relu(W2 relu(W1 x + b1)
 * + b2) + b3 on normalised inputs x = (raw - in_mu) * in_isd. Backward:
 * textbook reverse-mode differentiation of the mean squared error.
 *
 * Pinned by: central finite differences of the loss with bf16 = 0 (every
 * parameter block), closed forms (zero weights give y = b3; a single ReLU
 * path reduces to a product of scalars), and, with bf16 = 1, the GPU
 * kernels within FP32-accumulation tolerance.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "mpm_oracle.h"

enum { NI = 16, NIU = 13, NH = 64, NO = 16, NOU = 7 };
enum { OW1 = 0, OB1 = OW1 + NH * NI, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH, OB3 = OW3 + NO * NH };

/* float -> bfloat16 -> float, round to nearest even */
static double bf(double x) {
    float f = (float)x;
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    u &= 0xFFFF0000u;
    memcpy(&f, &u, 4);
    return (double)f;
}

static double rf(double x, int on) { return on ? bf(x) : x; }
static double f32(double x, int on) { return on ? (double)(float)x : x; }

/* one sample's forward pass; keeps the activations for the backward pass */
typedef struct {
    double x[NI], h1[NH], h2[NH], y[NO];
} acts;

static void weights(const double *p, int b, double *w) {
    for (int i = 0; i < ORC_MLP_NPARAMS; i++) {
        const int is_w = (i < OB1) || (i >= OW2 && i < OB2) || (i >= OW3 && i < OB3);
        w[i] = is_w ? rf(p[i], b) : f32(p[i], b);
    }
}

static void fwd(const double *w, const double *sc, const float *raw, int b, acts *a) {
    for (int f = 0; f < NI; f++) {
        double v = 0.0;
        if (f < NIU) {
            if (b) v = (double)(((float)raw[f] - (float)sc[f]) * (float)sc[16 + f]);
            else v = ((double)raw[f] - sc[f]) * sc[16 + f];
        }
        a->x[f] = rf(v, b);
    }
    for (int j = 0; j < NH; j++) {
        double s = 0.0;
        for (int k = 0; k < NI; k++) s += w[OW1 + j * NI + k] * a->x[k];
        s = f32(s, b) + w[OB1 + j];
        a->h1[j] = rf(s > 0.0 ? s : 0.0, b);
    }
    for (int j = 0; j < NH; j++) {
        double s = 0.0;
        for (int k = 0; k < NH; k++) s += w[OW2 + j * NH + k] * a->h1[k];
        s = f32(s, b) + w[OB2 + j];
        a->h2[j] = rf(s > 0.0 ? s : 0.0, b);
    }
    for (int o = 0; o < NO; o++) {
        double s = 0.0;
        for (int k = 0; k < NH; k++) s += w[OW3 + o * NH + k] * a->h2[k];
        a->y[o] = f32(f32(s, b) + w[OB3 + o], b);
    }
}

void orc_mlp_forward(const double *params, const double *scaling, const float *in, long n, int bf16, double *out) {
    double *w = malloc(sizeof(double) * ORC_MLP_NPARAMS);
    weights(params, bf16, w);
    acts a;
    for (long i = 0; i < n; i++) {
        fwd(w, scaling, in + i * NI, bf16, &a);
        for (int o = 0; o < NO; o++) out[i * NO + o] = (o < NOU) ? a.y[o] * scaling[48 + o] + scaling[32 + o] : 0.0;
    }
    free(w);
}

double orc_mlp_grad(const double *params, const double *scaling, const float *in, const float *tgt, long n, int bf16,
                    double *grad) {
    double *w = malloc(sizeof(double) * ORC_MLP_NPARAMS);
    weights(params, bf16, w);
    memset(grad, 0, sizeof(double) * ORC_MLP_NPARAMS);
    const double inv_n = bf16 ? (double)(1.0f / (float)n) : 1.0 / (double)n;
    double loss = 0.0;
    acts a;
    double dy[NO], d2[NH], d1[NH];
    for (long i = 0; i < n; i++) {
        fwd(w, scaling, in + i * NI, bf16, &a);
        for (int o = 0; o < NO; o++) {
            double d = 0.0;
            if (o < NOU) {
                double tn;
                if (bf16) tn = (double)(((float)tgt[i * NO + o] - (float)scaling[32 + o]) / (float)scaling[48 + o]);
                else tn = ((double)tgt[i * NO + o] - scaling[32 + o]) / scaling[48 + o];
                d = a.y[o] - tn;
                loss += 0.5 * d * d * inv_n;
                d = f32(d * inv_n, bf16);
            }
            dy[o] = rf(d, bf16);
        }
        /* layer 3 */
        for (int o = 0; o < NO; o++) {
            grad[OB3 + o] += dy[o];
            for (int k = 0; k < NH; k++) grad[OW3 + o * NH + k] += dy[o] * a.h2[k];
        }
        for (int k = 0; k < NH; k++) {
            double s = 0.0;
            for (int o = 0; o < NO; o++) s += dy[o] * w[OW3 + o * NH + k];
            d2[k] = rf(a.h2[k] > 0.0 ? f32(s, bf16) : 0.0, bf16);
        }
        /* layer 2 */
        for (int j = 0; j < NH; j++) {
            grad[OB2 + j] += d2[j];
            for (int k = 0; k < NH; k++) grad[OW2 + j * NH + k] += d2[j] * a.h1[k];
        }
        for (int k = 0; k < NH; k++) {
            double s = 0.0;
            for (int j = 0; j < NH; j++) s += d2[j] * w[OW2 + j * NH + k];
            d1[k] = rf(a.h1[k] > 0.0 ? f32(s, bf16) : 0.0, bf16);
        }
        /* layer 1 */
        for (int j = 0; j < NH; j++) {
            grad[OB1 + j] += d1[j];
            for (int k = 0; k < NI; k++) grad[OW1 + j * NI + k] += d1[j] * a.x[k];
        }
    }
    free(w);
    return loss;
}
