/*
 * plis_mlp.c -- CPU reference of the learned tau proposal (TEST
 * INFRASTRUCTURE ONLY: the checker of the gfx950 MLP kernels).
 *
 * Network 12 -> 64 -> 64 -> 2, GELU or ReLU activations, inputs standardised
 * per feature.  GELU(z) = z Phi(z) is evaluated as
 *     max(z, 0) - |z| max(0, p(|z|)),  p(a) = 0.5 + a (C1 + a (C2 + a C3)),
 * a cubic fitted (minimax over the real line, p(0) = 1/2 so that the slope
 * at 0 is GELU's 1/2) to the correction |z| Phi(-|z|) = relu(z) - GELU(z);
 * its largest error against the erf form is 0.0058.  Parameter blob (float, 5378 values, the first layer padded to
 * 16 inputs whose last 4 rows stay zero):
 *   W1[16][64] | b1[64] | W2[64][64] | b2[64] | W3[64][2] | b3[2]
 * W[i][o] multiplies input i into output o.  With bf16 != 0 every GEMM
 * operand (weights, activations, back-propagated errors) is rounded to
 * bfloat16 (round to nearest even) as on the device; sums are in double.
 *
 * Loss per sample: (o0 - y)^2 + softplus(o1) - c*o1, averaged over the
 * batch; y is the standardised log tau target, c in {0,1} the critical
 * flag: the reject label of the record's trial leap.
 *
 * Features of a subvolume (12, float):
 *   f[s]     = log(1 + x_s)                         s < 5 (0 if s >= S)
 *   f[5 + s] = log(1 + sum_{j: r_sj>0} a_j + d_s x_s W_v)
 *   f[10]    = highest order among reactions with a_j > 0
 *   f[11]    = log of the realization's previous tau
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "plis_oracle.h"

#define NIN 16
#define NH 64
#define NOUT 2
#define P_W1 0
#define P_B1 (P_W1 + NIN * NH)
#define P_W2 (P_B1 + NH)
#define P_B2 (P_W2 + NH * NH)
#define P_W3 (P_B2 + NH)
#define P_B3 (P_W3 + NH * NOUT)
#define P_TOTAL (P_B3 + NOUT)

int orc_mlp_param_count(void) { return P_TOTAL; }

static double bf(double x, int on)
{
    if (!on) return x;
    float f = (float)x;
    uint32_t b;
    memcpy(&b, &f, 4);
    b = (b + 0x7FFFu + ((b >> 16) & 1u)) & 0xFFFF0000u;
    memcpy(&f, &b, 4);
    return (double)f;
}

/* fp32 rounding of accumulator outputs, only when emulating the device */
static double f32(double x, int on) { return on ? (double)(float)x : x; }
static double relu(double z) { return z > 0.0 ? z : 0.0; }
static double relu_d(double z) { return z > 0.0 ? 1.0 : 0.0; }

#define G_C1 (-0.47735822f)
#define G_C2 0.15120307f
#define G_C3 (-0.01585033f)

/* the device's single-precision evaluation, every fma explicit */
float orc_gelu(float z)
{
    float a = fabsf(z);
    float p = fmaf(fmaf(fmaf(G_C3, a, G_C2), a, G_C1), a, 0.5f);
    return fmaxf(z, 0.0f) - a * fmaxf(p, 0.0f);
}

static float gelu_d(float z)
{
    float a = fabsf(z);
    float p = fmaf(fmaf(fmaf(G_C3, a, G_C2), a, G_C1), a, 0.5f);
    float dp = fmaf(fmaf(3.0f * G_C3, a, 2.0f * G_C2), a, G_C1);
    float q = p > 0.0f ? fmaf(a, dp, p) : 0.0f;
    return z > 0.0f ? 1.0f - q : q;
}

/* the same function in double (finite-difference checks of the gradient) */
static double gelu_dbl(double z, double *deriv)
{
    double a = fabs(z);
    double p = fma(fma(fma(G_C3, a, G_C2), a, G_C1), a, 0.5);
    double dp = fma(fma(3.0 * G_C3, a, 2.0 * G_C2), a, G_C1);
    double q = p > 0.0 ? fma(a, dp, p) : 0.0;
    *deriv = z > 0.0 ? 1.0 - q : q;
    return (z > 0.0 ? z : 0.0) - a * (p > 0.0 ? p : 0.0);
}

/* kind 1 = GELU, 0 = ReLU; single precision where the device is emulated */
static double act(double z, int kind, int bfon)
{
    double d;
    if (!kind) return relu(z);
    return bfon ? (double)orc_gelu((float)z) : gelu_dbl(z, &d);
}

static double act_d(double z, int kind, int bfon)
{
    double d;
    if (!kind) return relu_d(z);
    if (bfon) return (double)gelu_d((float)z);
    gelu_dbl(z, &d);
    return d;
}

/* norm: mu[12], sig[12]; returns per-sample outputs and (optionally) grads */
static double mlp_pass(const float *w, const float *norm, const float *feats, int32_t n, int bfon,
                       int kind, const float *ylog, const float *ycrit, double *grad, float *out)
{
    double loss = 0.0;
    for (int32_t i = 0; i < n; i++) {
        double x[NIN], z1[NH], h1[NH], z2[NH], h2[NH], o[NOUT];
        for (int k = 0; k < NIN; k++) {
            double f = k < 12 ? ((double)feats[(size_t)i * 12 + k] - norm[k]) / norm[12 + k] : 0.0;
            x[k] = bf(f32(f, bfon), bfon);
        }
        for (int h = 0; h < NH; h++) {
            double s = 0.0;
            for (int k = 0; k < NIN; k++) s += x[k] * bf(w[P_W1 + k * NH + h], bfon);
            z1[h] = f32(s + w[P_B1 + h], bfon);
            h1[h] = bf(f32(act(z1[h], kind, bfon), bfon), bfon);
        }
        for (int h = 0; h < NH; h++) {
            double s = 0.0;
            for (int k = 0; k < NH; k++) s += h1[k] * bf(w[P_W2 + k * NH + h], bfon);
            z2[h] = f32(s + w[P_B2 + h], bfon);
            h2[h] = bf(f32(act(z2[h], kind, bfon), bfon), bfon);
        }
        for (int q = 0; q < NOUT; q++) {
            double s = 0.0;
            for (int k = 0; k < NH; k++) s += h2[k] * bf(w[P_W3 + k * NOUT + q], bfon);
            o[q] = f32(s + w[P_B3 + q], bfon);
        }
        if (out) { out[(size_t)i * 2] = (float)o[0]; out[(size_t)i * 2 + 1] = (float)o[1]; }
        if (!grad) continue;
        double y = ylog[i], c = ycrit[i];
        double sp = (o[1] > 0 ? o[1] : 0.0) + log1p(exp(-fabs(o[1])));
        loss += (o[0] - y) * (o[0] - y) + sp - c * o[1];
        double dO[NOUT], dz2[NH], dz1[NH];
        dO[0] = bf(f32(2.0 * (o[0] - y) / n, bfon), bfon);
        dO[1] = bf(f32((1.0 / (1.0 + exp(-o[1])) - c) / n, bfon), bfon);
        for (int q = 0; q < NOUT; q++) {
            grad[P_B3 + q] += dO[q];
            for (int k = 0; k < NH; k++) grad[P_W3 + k * NOUT + q] += h2[k] * dO[q];
        }
        for (int k = 0; k < NH; k++) {
            double s = 0.0;
            for (int q = 0; q < NOUT; q++) s += dO[q] * bf(w[P_W3 + k * NOUT + q], bfon);
            dz2[k] = bf(f32(s * bf(act_d(z2[k], kind, bfon), bfon), bfon), bfon);
        }
        for (int h = 0; h < NH; h++) {
            grad[P_B2 + h] += dz2[h];
            for (int k = 0; k < NH; k++) grad[P_W2 + k * NH + h] += h1[k] * dz2[h];
        }
        for (int k = 0; k < NH; k++) {
            double s = 0.0;
            for (int h = 0; h < NH; h++) s += dz2[h] * bf(w[P_W2 + k * NH + h], bfon);
            dz1[k] = bf(f32(s * bf(act_d(z1[k], kind, bfon), bfon), bfon), bfon);
        }
        for (int h = 0; h < NH; h++) {
            grad[P_B1 + h] += dz1[h];
            for (int k = 0; k < NIN; k++) grad[P_W1 + k * NH + h] += x[k] * dz1[h];
        }
    }
    return loss / (n > 0 ? n : 1);
}

void orc_mlp_forward(const float *w, const float *norm, const float *feats, int32_t n, int32_t bfon,
                     int32_t act_kind, float *out)
{
    mlp_pass(w, norm, feats, n, bfon, act_kind, NULL, NULL, NULL, out);
}

double orc_mlp_loss_grad(const float *w, const float *norm, const float *feats, int32_t n,
                         int32_t bfon, int32_t act_kind, const float *ylog, const float *ycrit,
                         double *grad)
{
    memset(grad, 0, sizeof(double) * P_TOTAL);
    return mlp_pass(w, norm, feats, n, bfon, act_kind, ylog, ycrit, grad, NULL);
}

/* Adam (Kingma & Ba 2015) with bias correction; step t counts from 1 */
void orc_adam_step(float *w, float *m, float *v, const float *g, int32_t count, int32_t t,
                   float lr, float b1, float b2, float eps)
{
    float c1 = 1.0f - powf(b1, (float)t), c2 = 1.0f - powf(b2, (float)t);
    for (int32_t i = 0; i < count; i++) {
        m[i] = b1 * m[i] + (1.0f - b1) * g[i];
        v[i] = b2 * v[i] + (1.0f - b2) * g[i] * g[i];
        float mh = m[i] / c1, vh = v[i] / c2;
        w[i] -= lr * mh / (sqrtf(vh) + eps);
    }
}
