/*
 * oracle_mlp.c — FP64 restatement of the surrogate network (forward, loss, backward).
 *
 * TEST INFRASTRUCTURE ONLY (see sirius_oracle.h). The reference has no network: the contract
 * is sirius_hip.h's (5 -> 4 x 256 GELU -> 6, layer-by-layer W_l then b_l, cross-entropy plus
 * weighted residual error). With round_bf16 the GEMM operands of the hidden layers are rounded
 * to BF16 where the device rounds them: their input activations, their weights and the
 * pre-activation gradients; biases, pre-activations and bias gradients stay unrounded, and so
 * does the output head (the last hidden activation, the head's weights and the output
 * gradient), which the device keeps in FP32.
 * Pins: tests/test_oracle_mlp.py (finite differences of the FP64 loss).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "sirius_oracle.h"

#define NL 5
static const int DIM_IN[NL] = {SIRIUS_MLP_IN, SIRIUS_MLP_HIDDEN, SIRIUS_MLP_HIDDEN, SIRIUS_MLP_HIDDEN,
                               SIRIUS_MLP_HIDDEN};
static const int DIM_OUT[NL] = {SIRIUS_MLP_HIDDEN, SIRIUS_MLP_HIDDEN, SIRIUS_MLP_HIDDEN, SIRIUS_MLP_HIDDEN,
                                SIRIUS_MLP_OUT};

size_t oracle_mlp_param_count(void) {
    size_t n = 0;
    for (int l = 0; l < NL; ++l) n += (size_t)DIM_OUT[l] * DIM_IN[l] + DIM_OUT[l];
    return n;
}

/* Round to the nearest BF16 (ties to even), through FP32 as the device stores values. */
static double bf(double v, int on) {
    if (!on) return v;
    float f = (float)v;
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return f;
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    memcpy(&f, &u, 4);
    return f;
}

/* GELU, tanh form, and its derivative */
static const double GC = 0.7978845608028654, GA = 0.044715;
static double gelu(double z) { return 0.5 * z * (1.0 + tanh(GC * (z + GA * z * z * z))); }
static double gelu_d(double z) {
    double t = tanh(GC * (z + GA * z * z * z));
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * GC * (1.0 + 3.0 * GA * z * z);
}

typedef struct {
    const float* W[NL];
    const float* b[NL];
    size_t offW[NL], offb[NL];
} layout;

static void make_layout(const float* params, layout* L) {
    size_t o = 0;
    for (int l = 0; l < NL; ++l) {
        L->offW[l] = o;
        L->W[l] = params + o;
        o += (size_t)DIM_OUT[l] * DIM_IN[l];
        L->offb[l] = o;
        L->b[l] = params + o;
        o += DIM_OUT[l];
    }
}

/* Forward for one sample; h[l] (layer inputs, rounded as the device holds them) and z[l]
 * (pre-activations) are kept for the backward pass. */
static void forward1(const layout* L, const float* x, double h[NL][SIRIUS_MLP_HIDDEN],
                     double z[NL][SIRIUS_MLP_HIDDEN], double y[SIRIUS_MLP_OUT], int rb) {
    for (int k = 0; k < SIRIUS_MLP_IN; ++k) h[0][k] = bf(x[k], rb);
    for (int l = 0; l < NL; ++l) {
        for (int o = 0; o < DIM_OUT[l]; ++o) {
            double acc = 0;
            const float* w = L->W[l] + (size_t)o * DIM_IN[l];
            const int rl = rb && l < NL - 1;
            for (int k = 0; k < DIM_IN[l]; ++k) acc += bf(w[k], rl) * h[l][k];
            acc += L->b[l][o];
            if (l < NL - 1) {
                z[l][o] = acc;
                h[l + 1][o] = bf(gelu(acc), rb && l + 1 < NL - 1);
            } else {
                y[o] = acc;
            }
        }
    }
}

void oracle_mlp_forward(const float* params, const float* x, int n, double* out, int round_bf16) {
    layout L;
    make_layout(params, &L);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        double h[NL][SIRIUS_MLP_HIDDEN], z[NL][SIRIUS_MLP_HIDDEN];
        forward1(&L, x + (size_t)i * SIRIUS_MLP_IN, h, z, out + (size_t)i * SIRIUS_MLP_OUT, round_bf16);
    }
}

/* Per-sample loss and the output gradient dL/dy (before the 1/n of the batch mean). */
static double loss_out(const double y[SIRIUS_MLP_OUT], int cls, const float* dir, double gy[SIRIUS_MLP_OUT]) {
    double m = y[0];
    for (int k = 1; k < 3; ++k) m = y[k] > m ? y[k] : m;
    double e[3], s = 0;
    for (int k = 0; k < 3; ++k) { e[k] = exp(y[k] - m); s += e[k]; }
    double loss = -(y[cls] - m - log(s));
    for (int k = 0; k < 3; ++k) gy[k] = e[k] / s - (k == cls ? 1.0 : 0.0);
    for (int k = 0; k < 3; ++k) {
        double d = (cls == 0) ? y[3 + k] - dir[k] : 0.0;
        loss += SIRIUS_MLP_DIR_WEIGHT * d * d;
        gy[3 + k] = 2.0 * SIRIUS_MLP_DIR_WEIGHT * d;
    }
    return loss;
}

double oracle_mlp_loss_grad(const float* params, const float* x, const int32_t* cls,
                            const float* dir, int n, double* grad, int round_bf16) {
    layout L;
    make_layout(params, &L);
    size_t P = oracle_mlp_param_count();
    memset(grad, 0, P * sizeof(double));
    double total = 0;
    const int rb = round_bf16;
#pragma omp parallel
    {
        double* g = (double*)calloc(P, sizeof(double));
        double local = 0;
#pragma omp for schedule(static)
        for (int i = 0; i < n; ++i) {
            double h[NL][SIRIUS_MLP_HIDDEN], z[NL][SIRIUS_MLP_HIDDEN], y[SIRIUS_MLP_OUT], gy[SIRIUS_MLP_OUT];
            forward1(&L, x + (size_t)i * SIRIUS_MLP_IN, h, z, y, rb);
            local += loss_out(y, cls[i], dir + (size_t)i * 3, gy);
            /* delta: gradient w.r.t. the current layer's output pre-activation, 1/n included */
            double delta[SIRIUS_MLP_HIDDEN], dr[SIRIUS_MLP_HIDDEN], gh[SIRIUS_MLP_HIDDEN];
            for (int o = 0; o < SIRIUS_MLP_OUT; ++o) delta[o] = gy[o] / n;
            for (int l = NL - 1; l >= 0; --l) {
                int no = DIM_OUT[l], ni = DIM_IN[l];
                const int rl = rb && l < NL - 1;
                for (int o = 0; o < no; ++o) dr[o] = bf(delta[o], rl);
                for (int o = 0; o < no; ++o) {
                    double* gw = g + L.offW[l] + (size_t)o * ni;
                    for (int k = 0; k < ni; ++k) gw[k] += dr[o] * h[l][k];
                    g[L.offb[l] + o] += delta[o];
                }
                if (l == 0) break;
                for (int k = 0; k < ni; ++k) {
                    double acc = 0;
                    for (int o = 0; o < no; ++o) acc += dr[o] * bf(L.W[l][(size_t)o * ni + k], rl);
                    gh[k] = acc;
                }
                for (int k = 0; k < ni; ++k) delta[k] = gh[k] * gelu_d(z[l - 1][k]);
            }
        }
#pragma omp critical
        {
            total += local;
            for (size_t p = 0; p < P; ++p) grad[p] += g[p];
        }
        free(g);
    }
    return total / n;
}
