/*
 * oracle/ar_oracle.c — CPU oracle (FP64) of the learned proposal network.
 *
 * TEST INFRASTRUCTURE ONLY: the checker used by tests/ and bench.py.
 *
 * The reference has no learned proposal (its autoregressive ansatz kind is
 * rejected by nqs_ansatz_create, src/nqs/nqs_ansatz.c:250-251), so this file
 * restates this library's own definition of the network: N -> 64 -> 64 -> N
 * masked autoregressive MLP (MADE degrees 1 + floor(k (N-1) / 64)), ReLU, logit l_j
 * of p(s_j = -1 | s_<j), parameters [W1 | b1 | W2 | b2 | W3 | b3] in FP32.
 * With quantize = 1 the weights and the activations a1, a2 are rounded to
 * BF16 (round to nearest even) where the device rounds them, sums are FP64,
 * and log q uses the device's conditional: p = sigmoid(l) evaluated in FP32
 * with l clamped to [-16, 16], log p or log(1 - p) in FP64.
 * With quantize = 0 everything is FP64 (used for finite-difference checks of
 * the backward pass).  A magnetisation constraint forces a site once the
 * remaining sites must all be up or all down; forced sites contribute
 * neither to log q nor to the loss.
 */
#include "nqs_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ARH 64

typedef unsigned long long u64;

static inline u64 ar_xs_next(u64 *st) {
    u64 x = *st;
    x ^= x << 13;
    x ^= x >> 7;
    x ^= x << 17;
    *st = x;
    return x;
}
static inline double ar_xs_uniform(u64 *st) {
    return (double)(ar_xs_next(st) >> 11) * (1.0 / 9007199254740992.0);
}
static u64 ar_splitmix64(u64 x) {
    u64 z = x + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static uint16_t f2bf(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
static float bf2f(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}
static double qw(float x, int quantize) { return quantize ? (double)bf2f(f2bf(x)) : (double)x; }
static double qa(double x, int quantize) {
    double r = x > 0.0 ? x : 0.0;
    return quantize ? (double)bf2f(f2bf((float)r)) : r;
}

static int degree(int k, int N) { return N > 1 ? 1 + (k * (N - 1)) / ARH : 1; }
static int mask1(int k, int i, int N) { return degree(k, N) >= i + 1; }
static int mask2(int kp, int k, int N) { return degree(kp, N) >= degree(k, N); }
static int mask3(int j, int kp, int N) { return j + 1 > degree(kp, N); }

typedef struct {
    long oW1, ob1, oW2, ob2, oW3, ob3, P;
} ar_dims;
static ar_dims dims(int N) {
    ar_dims d;
    d.oW1 = 0;
    d.ob1 = (long)ARH * N;
    d.oW2 = d.ob1 + ARH;
    d.ob2 = d.oW2 + (long)ARH * ARH;
    d.oW3 = d.ob2 + ARH;
    d.ob3 = d.oW3 + (long)N * ARH;
    d.P = d.ob3 + N;
    return d;
}
static int free_param(const ar_dims *d, int N, long p) {
    if (p < d->ob1) return mask1((int)(p / N), (int)(p % N), N);
    if (p < d->oW2) return 1;
    if (p < d->ob2) return mask2((int)((p - d->oW2) / ARH), (int)((p - d->oW2) % ARH), N);
    if (p < d->oW3) return 1;
    if (p < d->ob3) return mask3((int)((p - d->oW3) / ARH), (int)((p - d->oW3) % ARH), N);
    return 1;
}

long orc_ar_num_params(int N) { return dims(N).P; }

/* Initial weights: xorshift64 seeded with splitmix64(seed ^ 0xA17E5EED)
 * (1 if zero), one draw per parameter in layout order; W1 entries
 * scale (2u-1)/sqrt(N), W2 and W3 entries scale (2u-1)/8, biases 0,
 * masked entries 0. */
void orc_ar_init(int N, unsigned long long seed, double scale, float *w) {
    const ar_dims d = dims(N);
    u64 st = ar_splitmix64((u64)seed ^ 0xA17E5EEDULL);
    if (!st) st = 1;
    for (long p = 0; p < d.P; p++) {
        const double u = ar_xs_uniform(&st);
        double x = 0.0;
        if (p < d.ob1) x = scale * (2.0 * u - 1.0) / sqrt((double)N);
        else if ((p >= d.oW2 && p < d.ob2) || (p >= d.oW3 && p < d.ob3))
            x = scale * (2.0 * u - 1.0) / 8.0;
        w[p] = free_param(&d, N, p) ? (float)x : 0.0f;
    }
}

static double log_sigmoid(double x) { return x >= 0.0 ? -log1p(exp(-x)) : x - log1p(exp(x)); }

/* forward of one configuration: z1, a1, z2, a2 [64], logits l [N], and the
 * free-site flags */
static void forward(int N, const float *w, const int *s, int constrained, int quantize,
                    double *z1, double *a1, double *z2, double *a2, double *l, int *fr) {
    const ar_dims d = dims(N);
    for (int k = 0; k < ARH; k++) {
        double z = (double)w[d.ob1 + k];
        for (int i = 0; i < N; i++)
            if (mask1(k, i, N)) z += qw(w[d.oW1 + (long)k * N + i], quantize) * (double)s[i];
        z1[k] = z;
        a1[k] = qa(z, quantize);
    }
    for (int kp = 0; kp < ARH; kp++) {
        double z = (double)w[d.ob2 + kp];
        for (int k = 0; k < ARH; k++)
            if (mask2(kp, k, N)) z += qw(w[d.oW2 + (long)kp * ARH + k], quantize) * a1[k];
        z2[kp] = z;
        a2[kp] = qa(z, quantize);
    }
    int D = 0;
    for (int i = 0; i < N; i++) D += s[i] < 0;
    int before = 0;
    for (int j = 0; j < N; j++) {
        double z = (double)w[d.ob3 + j];
        for (int kp = 0; kp < ARH; kp++)
            if (mask3(j, kp, N)) z += qw(w[d.oW3 + (long)j * ARH + kp], quantize) * a2[kp];
        l[j] = z;
        const int rem_d = D - before, rem = N - j;
        fr[j] = !(constrained && (rem_d <= 0 || rem_d >= rem));
        before += s[j] < 0;
    }
}

int orc_ar_log_prob(int N, const float *w, const int *spins, int B, int constrained,
                    int quantize, double *out) {
    if (N < 2 || N > 64 || !w || !spins || !out || B < 0) return -1;
    double z1[ARH], a1[ARH], z2[ARH], a2[ARH], l[64];
    int fr[64];
    for (int b = 0; b < B; b++) {
        const int *s = spins + (long)b * N;
        forward(N, w, s, constrained, quantize, z1, a1, z2, a2, l, fr);
        double lq = 0.0;
        for (int j = 0; j < N; j++) {
            if (!fr[j]) continue;
            if (quantize) {
                /* the device conditional: FP32 sigmoid of l clamped to
                 * [-16, 16]; probabilities p and 1 - p in FP64 */
                const float lc = fminf(fmaxf((float)l[j], -16.0f), 16.0f);
                const double p = (double)(1.0f / (1.0f + expf(-lc)));   /* device: v_exp_f32, v_rcp_f32 */
                lq += s[j] < 0 ? log(p) : log(1.0 - p);
            } else {
                const double ls = log_sigmoid(l[j]);
                lq += s[j] < 0 ? ls : ls - l[j];
            }
        }
        out[b] = lq;
    }
    return 0;
}

/* Gradient of the mean negative log-likelihood over B configurations (the
 * BF16 roundings of the quantised forward are treated as identities). */
int orc_ar_gradient(int N, const float *w, const int *spins, int B, int constrained,
                    int quantize, double *grad, double *mean_nll) {
    if (N < 2 || N > 64 || !w || !spins || !grad || B <= 0) return -1;
    const ar_dims d = dims(N);
    memset(grad, 0, sizeof(double) * d.P);
    double z1[ARH], a1[ARH], z2[ARH], a2[ARH], l[64], g3[64], g2[ARH], g1[ARH];
    int fr[64];
    double nll = 0.0;
    for (int b = 0; b < B; b++) {
        const int *s = spins + (long)b * N;
        forward(N, w, s, constrained, quantize, z1, a1, z2, a2, l, fr);
        for (int j = 0; j < N; j++) {
            g3[j] = 0.0;
            if (!fr[j]) continue;
            const double x = s[j] < 0 ? 1.0 : 0.0;
            const double ls = log_sigmoid(l[j]);
            nll -= x > 0.0 ? ls : ls - l[j];
            g3[j] = (exp(ls) - x) / (double)B;
        }
        for (int kp = 0; kp < ARH; kp++) {
            double acc = 0.0;
            for (int j = 0; j < N; j++)
                if (mask3(j, kp, N)) acc += qw(w[d.oW3 + (long)j * ARH + kp], quantize) * g3[j];
            g2[kp] = z2[kp] > 0.0 ? acc : 0.0;
        }
        for (int k = 0; k < ARH; k++) {
            double acc = 0.0;
            for (int kp = 0; kp < ARH; kp++)
                if (mask2(kp, k, N)) acc += qw(w[d.oW2 + (long)kp * ARH + k], quantize) * g2[kp];
            g1[k] = z1[k] > 0.0 ? acc : 0.0;
        }
        for (int j = 0; j < N; j++) {
            grad[d.ob3 + j] += g3[j];
            for (int kp = 0; kp < ARH; kp++)
                if (mask3(j, kp, N)) grad[d.oW3 + (long)j * ARH + kp] += g3[j] * a2[kp];
        }
        for (int kp = 0; kp < ARH; kp++) {
            grad[d.ob2 + kp] += g2[kp];
            for (int k = 0; k < ARH; k++)
                if (mask2(kp, k, N)) grad[d.oW2 + (long)kp * ARH + k] += g2[kp] * a1[k];
        }
        for (int k = 0; k < ARH; k++) {
            grad[d.ob1 + k] += g1[k];
            for (int i = 0; i < N; i++)
                if (mask1(k, i, N)) grad[d.oW1 + (long)k * N + i] += g1[k] * (double)s[i];
        }
    }
    if (mean_nll) *mean_nll = nll / (double)B;
    return 0;
}

/* One Adam step (beta 0.9 / 0.999, eps 1e-8, bias-corrected at step t >= 1)
 * on the free parameters, FP64 moments. */
void orc_ar_adam_step(int N, float *w, double *m, double *v, const double *grad, long t, double lr) {
    const ar_dims d = dims(N);
    const double bc1 = 1.0 - pow(0.9, (double)t), bc2 = 1.0 - pow(0.999, (double)t);
    for (long p = 0; p < d.P; p++) {
        if (!free_param(&d, N, p)) continue;
        m[p] = 0.9 * m[p] + 0.1 * grad[p];
        v[p] = 0.999 * v[p] + 0.001 * grad[p] * grad[p];
        w[p] = (float)((double)w[p] - lr * (m[p] / bc1) / (sqrt(v[p] / bc2) + 1e-8));
    }
}
