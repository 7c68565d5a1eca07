/* mlp.c -- CPU oracle of the learned insertion density's network (no reference counterpart).
 *
 * Test infrastructure only: the checker of the device MLP, never part of the product path.
 *
 * Network: 64 Fourier features of a fractional position f (sin and cos of 2 pi k.f for the 32
 * integer vectors k of orc_mlp_frequencies) -> 4 hidden layers of width 256 with SiLU -> 1
 * output.  Parameters packed as W1[64][256] b1[256] W2[256][256] b2 W3 b3 W4 b4 w5[256] b5.
 * Loss: mean squared error over the rows.
 *
 * Two modes.  bf16 = 0: plain FP64 arithmetic (the mathematical network, pinned by central
 * finite differences of the loss).  bf16 = 1: the same sums in FP64 but with every GEMM operand
 * rounded to bfloat16 (round to nearest even) where the device rounds it: features, W1..W4, the
 * hidden activations silu(z), and the back-propagated dL/dz; the output layer, biases and
 * pre-activations stay unrounded.  The device accumulates in FP32, so its results agree with
 * bf16 = 1 to accumulation-order rounding, not bitwise.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define IN ORC_MLP_IN
#define HID ORC_MLP_HIDDEN
#define ORC_PI 3.14159265358979323846

static const size_t kOff[10] = {
    0,
    IN * HID,
    IN * HID + HID,
    IN * HID + HID + HID * HID,
    IN * HID + 2 * HID + HID * HID,
    IN * HID + 2 * HID + 2 * HID * HID,
    IN * HID + 3 * HID + 2 * HID * HID,
    IN * HID + 3 * HID + 3 * HID * HID,
    IN * HID + 4 * HID + 3 * HID * HID,
    IN * HID + 5 * HID + 3 * HID * HID,
};

void orc_mlp_frequencies(int8_t out[3 * ORC_MLP_FREQS]) {
    int n = 0;
    for (int j = 0; j < 6; ++j)
        for (int a = 0; a < 3; ++a) {
            for (int c = 0; c < 3; ++c) out[3 * n + c] = (int8_t)(c == a ? (1 << j) : 0);
            ++n;
        }
    static const int pairs[3][2] = {{0, 1}, {1, 2}, {0, 2}};
    for (int j = 0; j < 4; ++j)
        for (int p = 0; p < 3; ++p) {
            for (int c = 0; c < 3; ++c) out[3 * n + c] = (int8_t)((c == pairs[p][0] || c == pairs[p][1]) ? (1 << j) : 0);
            ++n;
        }
    for (int j = 0; j < 2; ++j) {
        for (int c = 0; c < 3; ++c) out[3 * n + c] = (int8_t)(1 << j);
        ++n;
    }
}

/* float -> bfloat16 (nearest even) -> value */
static double bf(double v) {
    float f = (float)v;
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) != 0x7f800000u) u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    memcpy(&f, &u, 4);
    return (double)f;
}

static double silu(double z) { return z / (1.0 + exp(-z)); }
static double silu_grad(double z) {
    const double s = 1.0 / (1.0 + exp(-z));
    return s * (1.0 + z * (1.0 - s));
}

typedef struct {
    double x[IN];
    double z[4][HID];
    double h[4][HID];
    double y;
} row_state;

static void features(const double f[3], int bf16, double x[IN]) {
    int8_t kv[3 * ORC_MLP_FREQS];
    orc_mlp_frequencies(kv);
    for (int c = 0; c < ORC_MLP_FREQS; ++c) {
        double ph = (double)kv[3 * c] * f[0] + (double)kv[3 * c + 1] * f[1] + (double)kv[3 * c + 2] * f[2];
        ph -= floor(ph); /* sin, cos of 2 pi ph are periodic in ph */
        const double s = sin(2.0 * ORC_PI * ph), co = cos(2.0 * ORC_PI * ph);
        x[2 * c] = bf16 ? bf(s) : s;
        x[2 * c + 1] = bf16 ? bf(co) : co;
    }
}

static void forward_row(const double* p, const double f[3], int bf16, row_state* st) {
    features(f, bf16, st->x);
    const double* in = st->x;
    int kin = IN;
    for (int l = 0; l < 4; ++l) {
        const double* W = p + kOff[2 * l];
        const double* b = p + kOff[2 * l + 1];
        for (int j = 0; j < HID; ++j) {
            double acc = 0.0;
            for (int i = 0; i < kin; ++i) acc += in[i] * (bf16 ? bf(W[(size_t)i * HID + j]) : W[(size_t)i * HID + j]);
            st->z[l][j] = acc + b[j];
            const double h = silu(st->z[l][j]);
            st->h[l][j] = bf16 ? bf(h) : h;
        }
        in = st->h[l];
        kin = HID;
    }
    double y = 0.0;
    for (int k = 0; k < HID; ++k) y += st->h[3][k] * p[kOff[8] + k];
    st->y = y + p[kOff[9]];
}

void orc_mlp_forward(const double* params, const double* frac, size_t n, int bf16, double* y) {
    row_state* st = (row_state*)malloc(sizeof(row_state));
    for (size_t r = 0; r < n; ++r) {
        forward_row(params, frac + 3 * r, bf16, st);
        y[r] = st->y;
    }
    free(st);
}

double orc_mlp_loss_grad(const double* params, const double* frac, const double* target, size_t n, int bf16,
                         double* grad) {
    memset(grad, 0, ORC_MLP_PARAMS * sizeof(double));
    row_state* st = (row_state*)malloc(sizeof(row_state));
    double dz[HID], dz_prev[HID];
    double loss = 0.0;
    for (size_t r = 0; r < n; ++r) {
        forward_row(params, frac + 3 * r, bf16, st);
        const double d = st->y - target[r];
        loss += d * d;
        const double dy = 2.0 * d / (double)n;
        /* output layer */
        for (int k = 0; k < HID; ++k) grad[kOff[8] + k] += st->h[3][k] * dy;
        grad[kOff[9]] += dy;
        for (int k = 0; k < HID; ++k) {
            const double v = dy * params[kOff[8] + k] * silu_grad(st->z[3][k]);
            dz[k] = bf16 ? bf(v) : v;
        }
        for (int l = 3; l >= 0; --l) {
            const double* in = l == 0 ? st->x : st->h[l - 1];
            const int kin = l == 0 ? IN : HID;
            double* gW = grad + kOff[2 * l];
            double* gb = grad + kOff[2 * l + 1];
            for (int i = 0; i < kin; ++i)
                for (int j = 0; j < HID; ++j) gW[(size_t)i * HID + j] += in[i] * dz[j];
            for (int j = 0; j < HID; ++j) gb[j] += dz[j];
            if (l == 0) break;
            const double* W = params + kOff[2 * l];
            for (int i = 0; i < HID; ++i) {
                double acc = 0.0;
                for (int j = 0; j < HID; ++j) acc += dz[j] * (bf16 ? bf(W[(size_t)i * HID + j]) : W[(size_t)i * HID + j]);
                const double v = acc * silu_grad(st->z[l - 1][i]);
                dz_prev[i] = bf16 ? bf(v) : v;
            }
            memcpy(dz, dz_prev, sizeof(dz));
        }
    }
    free(st);
    return loss / (double)n;
}
