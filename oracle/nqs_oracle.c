/*
 * oracle/nqs_oracle.c — CPU oracle (FP64) for the NQS VMC path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker used by tests/, smoke() and the
 * cpu_baseline leg of bench.py.  The product never links or calls it.
 *
 * Restates, function by function, the reference algorithm of
 * tsotchke/spin_based_neural_network (files cited per function).  Loop
 * structure follows the reference where it fixes a random stream, a bond
 * order or a stopping rule; the local energy re-evaluates log psi from
 * scratch for every connected configuration exactly as the reference does
 * (src/nqs/nqs_gradient.c:24-45), which also makes this file the
 * "reference algorithm" CPU timing proxy.
 */
#define _GNU_SOURCE
#include "nqs_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef unsigned long long u64;

/* xorshift64 (<<13, >>7, <<17): src/nqs/nqs_ansatz.c:50-56,
 * src/nqs/nqs_sampler.c:31-38. */
static inline u64 xs_next(u64 *st) {
    u64 x = *st;
    x ^= x << 13;
    x ^= x >> 7;
    x ^= x << 17;
    *st = x;
    return x;
}
/* 53-bit uniform in [0,1): nqs_ansatz.c:50-56, nqs_sampler.c:40-44 */
static inline double xs_uniform(u64 *st) {
    return (double)(xs_next(st) >> 11) * (1.0 / 9007199254740992.0);
}
/* Box-Muller, one normal per call, u1 floored at 1e-15: nqs_ansatz.c:58-63 */
static double xs_gauss(u64 *st) {
    double u1 = xs_uniform(st);
    double u2 = xs_uniform(st);
    if (u1 < 1e-15) u1 = 1e-15;
    return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

static u64 splitmix64(u64 x) {
    u64 z = x + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* ===================================================================== */
/* ansatz — src/nqs/nqs_ansatz.c                                           */
/* ===================================================================== */
struct orc_ansatz {
    int kind;       /* nqs_ansatz_kind_t                                   */
    int N, M;
    long P;         /* total real parameters                               */
    double *p;
};

/* nqs_ansatz.c:65-252 (mean-field :76-85, RBM :88-111, complex RBM
 * :219-248).  Other kinds return NULL (:250-251). */
orc_ansatz_t *orc_ansatz_create(const nqs_config_t *cfg, int N) {
    if (!cfg || N <= 0) return NULL;
    int kind = cfg->ansatz;
    if (kind != NQS_ANSATZ_LEGACY_MLP && kind != NQS_ANSATZ_RBM &&
        kind != NQS_ANSATZ_COMPLEX_RBM)
        return NULL;
    orc_ansatz_t *a = calloc(1, sizeof(*a));
    if (!a) return NULL;
    a->kind = kind;
    a->N = N;
    u64 seed = cfg->rng_seed ? (u64)cfg->rng_seed : 0xDEADBEEFCAFEBABEULL;
    if (kind == NQS_ANSATZ_LEGACY_MLP) {
        a->M = 0;
        a->P = N;
        a->p = calloc((size_t)N, sizeof(double));
        for (int i = 0; i < N; i++) a->p[i] = 0.01 * (xs_uniform(&seed) - 0.5);
        return a;
    }
    int M = cfg->rbm_hidden_units > 0 ? cfg->rbm_hidden_units : 2 * N;
    long Pr = (long)N + M + (long)M * N;
    a->M = M;
    a->P = (kind == NQS_ANSATZ_COMPLEX_RBM) ? 2 * Pr : Pr;
    a->p = calloc((size_t)a->P, sizeof(double));
    if (!a->p) { free(a); return NULL; }
    double s = cfg->rbm_init_scale > 0.0 ? cfg->rbm_init_scale : 0.01;
    double ws = s / sqrt((double)N);
    /* real block: a (uniform), b (uniform), W (Gaussian / sqrt N) */
    for (int i = 0; i < N; i++) a->p[i] = s * (xs_uniform(&seed) - 0.5);
    for (int h = 0; h < M; h++) a->p[N + h] = s * (xs_uniform(&seed) - 0.5);
    for (long k = 0; k < (long)M * N; k++) a->p[N + M + k] = ws * xs_gauss(&seed);
    if (kind == NQS_ANSATZ_COMPLEX_RBM) {
        double si = 0.1 * s, wsi = si / sqrt((double)N);
        double *q = a->p + Pr;
        for (int i = 0; i < N; i++) q[i] = si * (xs_uniform(&seed) - 0.5);
        for (int h = 0; h < M; h++) q[N + h] = si * (xs_uniform(&seed) - 0.5);
        for (long k = 0; k < (long)M * N; k++) q[N + M + k] = wsi * xs_gauss(&seed);
    }
    return a;
}

void orc_ansatz_free(orc_ansatz_t *a) {
    if (!a) return;
    free(a->p);
    free(a);
}
long orc_ansatz_num_params(const orc_ansatz_t *a) { return a ? a->P : 0; }
int orc_ansatz_num_hidden(const orc_ansatz_t *a) { return a ? a->M : 0; }
int orc_ansatz_is_complex(const orc_ansatz_t *a) {
    return (a && a->kind == NQS_ANSATZ_COMPLEX_RBM) ? 1 : 0;
}
double *orc_ansatz_params_raw(orc_ansatz_t *a) { return a ? a->p : NULL; }

/* log(2 cosh z) for z = x + iy in the overflow-safe form of
 * nqs_ansatz.c:309-318: 2cosh z = e^|x| [(1+e^-2|x|) cos y
 * + i sgn(x)(1-e^-2|x|) sin y]; zero magnitude maps to -1e300. */
static void logcosh2_c(double x, double y, double *la, double *ar) {
    double ax = fabs(x), e = exp(-2.0 * ax);
    double re = (1.0 + e) * cos(y);
    double im = (x >= 0.0 ? 1.0 : -1.0) * (1.0 - e) * sin(y);
    double mag = hypot(re, im);
    *la = ax + (mag > 0.0 ? log(mag) : -1e300);
    *ar = atan2(im, re);
}

/* nqs_ansatz.c:846-873 dispatch with mf (:264-270), rbm (:277-294),
 * crbm (:320-355).  (0,0) on NULL / size mismatch. */
void orc_ansatz_log_amp(const int *s, int N, void *user, double *ola, double *oar) {
    orc_ansatz_t *a = (orc_ansatz_t *)user;
    double la = 0.0, ar = 0.0;
    if (a && s && N == a->N) {
        const double *p = a->p;
        if (a->kind == NQS_ANSATZ_LEGACY_MLP) {
            for (int i = 0; i < N; i++) la += p[i] * (double)s[i];
        } else if (a->kind == NQS_ANSATZ_RBM) {
            int M = a->M;
            for (int i = 0; i < N; i++) la += p[i] * (double)s[i];
            for (int h = 0; h < M; h++) {
                const double *w = p + N + M + (long)h * N;
                double x = p[N + h];
                for (int i = 0; i < N; i++) x += w[i] * (double)s[i];
                double ax = fabs(x);
                la += ax + log1p(exp(-2.0 * ax));
            }
        } else {
            int M = a->M;
            long Pr = (long)N + M + (long)M * N;
            const double *q = p + Pr;
            for (int i = 0; i < N; i++) {
                la += p[i] * (double)s[i];
                ar += q[i] * (double)s[i];
            }
            for (int h = 0; h < M; h++) {
                const double *wr = p + N + M + (long)h * N;
                const double *wi = q + N + M + (long)h * N;
                double xr = p[N + h], xi = q[N + h];
                for (int i = 0; i < N; i++) {
                    xr += wr[i] * (double)s[i];
                    xi += wi[i] * (double)s[i];
                }
                double l1, a1;
                logcosh2_c(xr, xi, &l1, &a1);
                la += l1;
                ar += a1;
            }
        }
    }
    if (ola) *ola = la;
    if (oar) *oar = ar;
}

/* tanh(xr + i xi) through sinh/cosh with the 1e-300 floor of
 * nqs_ansatz.c:961-975. */
static void ctanh_ref(double xr, double xi, double *tr, double *ti) {
    double ch = cosh(xr), sh = sinh(xr), c = cos(xi), sn = sin(xi);
    double nr = sh * c, ni = ch * sn, dr = ch * c, di = sh * sn;
    double den = dr * dr + di * di;
    if (den < 1e-300) den = 1e-300;
    *tr = (nr * dr + ni * di) / den;
    *ti = (ni * dr - nr * di) / den;
}

static void hidden_pre(const orc_ansatz_t *a, const int *s, int h,
                       double *xr, double *xi) {
    int N = a->N, M = a->M;
    long Pr = (long)N + M + (long)M * N;
    const double *wr = a->p + N + M + (long)h * N;
    double r = a->p[N + h];
    for (int i = 0; i < N; i++) r += wr[i] * (double)s[i];
    *xr = r;
    if (xi) {
        const double *wi = a->p + Pr + N + M + (long)h * N;
        double q = a->p[Pr + N + h];
        for (int i = 0; i < N; i++) q += wi[i] * (double)s[i];
        *xi = q;
    }
}

/* Real-projected d log|psi|/d theta: mf nqs_ansatz.c:875-880,
 * rbm :885-911, crbm :933-995. */
int orc_ansatz_logpsi_gradient(orc_ansatz_t *a, const int *s, int N, double *g) {
    if (!a || !s || !g || N != a->N) return -1;
    int M = a->M;
    if (a->kind == NQS_ANSATZ_LEGACY_MLP) {
        for (int i = 0; i < N; i++) g[i] = (double)s[i];
        return 0;
    }
    for (int i = 0; i < N; i++) g[i] = (double)s[i];
    if (a->kind == NQS_ANSATZ_RBM) {
        for (int h = 0; h < M; h++) {
            double x;
            hidden_pre(a, s, h, &x, NULL);
            double t = tanh(x);
            g[N + h] = t;
            double *gw = g + N + M + (long)h * N;
            for (int i = 0; i < N; i++) gw[i] = t * (double)s[i];
        }
        return 0;
    }
    long Pr = (long)N + M + (long)M * N;
    for (int i = 0; i < N; i++) g[Pr + i] = 0.0;
    for (int h = 0; h < M; h++) {
        double xr, xi, tr, ti;
        hidden_pre(a, s, h, &xr, &xi);
        ctanh_ref(xr, xi, &tr, &ti);
        g[N + h] = tr;
        g[Pr + N + h] = -ti;
        double *gr = g + N + M + (long)h * N, *gi = g + Pr + N + M + (long)h * N;
        for (int i = 0; i < N; i++) {
            gr[i] = tr * (double)s[i];
            gi[i] = -ti * (double)s[i];
        }
    }
    return 0;
}

/* Holomorphic gradient: crbm nqs_ansatz.c:1025-1091; real kinds give the
 * real-projected gradient and zero imaginary part (:1258-1273). */
int orc_ansatz_logpsi_gradient_complex(orc_ansatz_t *a, const int *s, int N,
                                       double *re, double *im) {
    if (!a || !s || !re || !im || N != a->N) return -1;
    if (a->kind != NQS_ANSATZ_COMPLEX_RBM) {
        int rc = orc_ansatz_logpsi_gradient(a, s, N, re);
        if (rc == 0) memset(im, 0, (size_t)a->P * sizeof(double));
        return rc;
    }
    int M = a->M;
    long Pr = (long)N + M + (long)M * N;
    for (int i = 0; i < N; i++) {
        re[i] = (double)s[i];  im[i] = 0.0;
        re[Pr + i] = 0.0;      im[Pr + i] = (double)s[i];
    }
    for (int h = 0; h < M; h++) {
        double xr, xi, tr, ti;
        hidden_pre(a, s, h, &xr, &xi);
        ctanh_ref(xr, xi, &tr, &ti);
        re[N + h] = tr;        im[N + h] = ti;
        re[Pr + N + h] = -ti;  im[Pr + N + h] = tr;
        long o = N + M + (long)h * N;
        for (int i = 0; i < N; i++) {
            double si = (double)s[i];
            re[o + i] = tr * si;        im[o + i] = ti * si;
            re[Pr + o + i] = -ti * si;  im[Pr + o + i] = tr * si;
        }
    }
    return 0;
}

/* nqs_ansatz.c:1275-1282 */
int orc_ansatz_apply_update(orc_ansatz_t *a, const double *d, double step) {
    if (!a || !d) return -1;
    for (long k = 0; k < a->P; k++) a->p[k] += step * d[k];
    return 0;
}

/* ===================================================================== */
/* Marshall — src/nqs/nqs_marshall.c                                      */
/* ===================================================================== */
/* nqs_marshall.c:16-27 */
int orc_marshall_parity(const int *s, int Lx, int Ly) {
    int n = 0;
    for (int x = 0; x < Lx; x++)
        for (int y = 0; y < Ly; y++)
            if (((x + y) & 1) == 0 && s[x * Ly + y] < 0) n++;
    return n & 1;
}
/* nqs_marshall.c:29-40 */
void orc_marshall_log_amp(const int *s, int N, void *user, double *la, double *ar) {
    orc_marshall_wrapper_t *w = (orc_marshall_wrapper_t *)user;
    double l = 0.0, a = 0.0;
    w->base_log_amp(s, N, w->base_user, &l, &a);
    int par = orc_marshall_parity(s, w->size_x, w->size_y);
    if (la) *la = l;
    if (ar) *ar = a + (par ? M_PI : 0.0);
}

/* ===================================================================== */
/* sampler — src/nqs/nqs_sampler.c (+ chains / exchange / Sz=0 start)     */
/* ===================================================================== */
struct orc_sampler {
    int N, C;
    nqs_config_t cfg;
    nqs_hip_options_t opt;
    orc_log_amp_fn fn;
    void *user;
    int *cur;        /* C x N */
    int *prop;       /* N     */
    double *cur_la;  /* C     */
    double *cur_ar;  /* C     */
    u64 *rng;        /* C     */
    long acc, tot;
};

unsigned long long orc_chain_seed(unsigned rng_seed, int c, int C) {
    u64 s0 = rng_seed ? (u64)rng_seed : 0x9E3779B97F4A7C15ULL;
    (void)C;
    if (c == 0) return s0;
    u64 s = splitmix64(s0 ^ ((u64)c * 0xD1B54A32D192ED03ULL));
    return s ? s : 1ULL;
}

orc_sampler_t *orc_sampler_create_ex(int N, const nqs_config_t *cfg,
                                     const nqs_hip_options_t *opt,
                                     orc_log_amp_fn fn, void *user) {
    if (N <= 0 || !cfg || !fn) return NULL;
    orc_sampler_t *s = calloc(1, sizeof(*s));
    if (!s) return NULL;
    s->N = N;
    s->cfg = *cfg;
    if (opt) s->opt = *opt;
    else {
        memset(&s->opt, 0, sizeof(s->opt));
        s->opt.num_chains = 1;
    }
    s->C = s->opt.num_chains > 0 ? s->opt.num_chains : 1;
    s->fn = fn;
    s->user = user;
    int C = s->C;
    s->cur = malloc((size_t)C * N * sizeof(int));
    s->prop = malloc((size_t)N * sizeof(int));
    s->cur_la = malloc((size_t)C * sizeof(double));
    s->cur_ar = malloc((size_t)C * sizeof(double));
    s->rng = malloc((size_t)C * sizeof(u64));
    for (int c = 0; c < C; c++) {
        u64 *st = &s->rng[c];
        int *x = s->cur + (size_t)c * N;
        *st = orc_chain_seed(cfg->rng_seed, c, C);
        if (s->opt.sz_zero_start) {
            /* N/2 up then down, Fisher-Yates from the top index down */
            for (int i = 0; i < N; i++) x[i] = (i < N / 2) ? 1 : -1;
            for (int i = N - 1; i > 0; i--) {
                int j = (int)(xs_next(st) % (u64)(i + 1));
                int t = x[i]; x[i] = x[j]; x[j] = t;
            }
        } else {
            /* nqs_sampler.c:69-72 */
            for (int i = 0; i < N; i++) x[i] = (xs_next(st) & 1) ? 1 : -1;
        }
        fn(x, N, user, &s->cur_la[c], &s->cur_ar[c]);
    }
    return s;
}

/* nqs_sampler.c:50-76 */
orc_sampler_t *orc_sampler_create(int N, const nqs_config_t *cfg,
                                  orc_log_amp_fn fn, void *user) {
    return orc_sampler_create_ex(N, cfg, NULL, fn, user);
}

void orc_sampler_free(orc_sampler_t *s) {
    if (!s) return;
    free(s->cur); free(s->prop); free(s->cur_la); free(s->cur_ar); free(s->rng);
    free(s);
}

/* One move of chain c.  Flip: nqs_sampler.c:86-117.  Exchange: draw the
 * k-th up site (k = x % n_up) and l-th down site (l = x % n_down) in index
 * order and flip both; symmetric at fixed magnetisation.  A state with no
 * up or no down site proposes nothing and counts as a rejected move. */
static int orc_step(orc_sampler_t *s, int c) {
    int N = s->N;
    int *x = s->cur + (size_t)c * N;
    u64 *st = &s->rng[c];
    memcpy(s->prop, x, (size_t)N * sizeof(int));
    if (s->opt.move == NQS_HIP_MOVE_EXCHANGE) {
        int nu = 0;
        for (int i = 0; i < N; i++) nu += (x[i] > 0);
        int nd = N - nu;
        if (nu == 0 || nd == 0) { s->tot++; return 0; }
        int k = (int)(xs_next(st) % (u64)nu);
        int l = (int)(xs_next(st) % (u64)nd);
        int iu = -1, id = -1;
        for (int i = 0, cu = 0, cd = 0; i < N; i++) {
            if (x[i] > 0) { if (cu++ == k) iu = i; }
            else          { if (cd++ == l) id = i; }
        }
        s->prop[iu] = -s->prop[iu];
        s->prop[id] = -s->prop[id];
    } else {
        int site = (int)(xs_next(st) % (u64)N);
        s->prop[site] = -s->prop[site];
    }
    double pla, par;
    s->fn(s->prop, N, s->user, &pla, &par);
    double lr = 2.0 * (pla - s->cur_la[c]);
    int acc;
    if (lr >= 0.0) acc = 1;
    else acc = (log(xs_uniform(st)) < lr);
    s->tot++;
    if (acc) {
        s->acc++;
        memcpy(x, s->prop, (size_t)N * sizeof(int));
        s->cur_la[c] = pla;
        s->cur_ar[c] = par;
    }
    return acc;
}

/* nqs_sampler.c:119-124, per chain */
void orc_sampler_thermalize(orc_sampler_t *s) {
    if (!s) return;
    for (int c = 0; c < s->C; c++)
        for (int i = 0; i < s->cfg.num_thermalize; i++) orc_step(s, c);
}

/* nqs_sampler.c:126-133: every chain advances; chain 0 is returned */
const int *orc_sampler_next(orc_sampler_t *s) {
    if (!s) return NULL;
    int k = s->cfg.num_decorrelate > 0 ? s->cfg.num_decorrelate : 1;
    for (int c = 0; c < s->C; c++)
        for (int i = 0; i < k; i++) orc_step(s, c);
    return s->cur;
}

/* nqs_sampler.c:135-144 with round-robin over chains */
int orc_sampler_batch(orc_sampler_t *s, int n, int *out) {
    if (!s || !out || n <= 0) return -1;
    int C = s->C, N = s->N;
    int rounds = (n + C - 1) / C;
    for (int r = 0; r < rounds; r++) {
        orc_sampler_next(s);
        for (int c = 0; c < C; c++) {
            long b = (long)r * C + c;
            if (b < n) memcpy(out + b * N, s->cur + (size_t)c * N, (size_t)N * sizeof(int));
        }
    }
    return 0;
}

double orc_sampler_acceptance_ratio(const orc_sampler_t *s) {
    if (!s || s->tot == 0) return 0.0;
    return (double)s->acc / (double)s->tot;
}
const int *orc_sampler_current(const orc_sampler_t *s) { return s ? s->cur : NULL; }
const int *orc_sampler_chain_state(const orc_sampler_t *s, int c) {
    return (s && c >= 0 && c < s->C) ? s->cur + (size_t)c * s->N : NULL;
}
int orc_sampler_num_sites(const orc_sampler_t *s) { return s ? s->N : 0; }
unsigned long long orc_sampler_rng_state(const orc_sampler_t *s, int c) {
    return (s && c >= 0 && c < s->C) ? s->rng[c] : 0ULL;
}

/* ===================================================================== */
/* local energy — src/nqs/nqs_gradient.c                                   */
/* ===================================================================== */
typedef struct {
    const int *spins;
    int *scr;
    int N;
    double la0, ar0;
    orc_log_amp_fn fn;
    void *user;
} le_ctx;

/* psi(s')/psi(s) from a full evaluation of s': nqs_gradient.c:24-45 */
static void ratio_c(le_ctx *c, double *re, double *im) {
    double la, ar;
    c->fn(c->scr, c->N, c->user, &la, &ar);
    double m = exp(la - c->la0), d = ar - c->ar0;
    *re = m * cos(d);
    *im = m * sin(d);
}

/* (J/4 diag, J/2 off) pair term: nqs_gradient.c:110-123, 184-193, 397-409 */
static void pair_term(le_ctx *c, int a, int b, double dcoef, double ocoef,
                      double *diag, double *ore, double *oim) {
    int sa = c->spins[a], sb = c->spins[b];
    *diag += dcoef * (double)(sa * sb);
    if (sa != sb) {
        double r, i;
        c->scr[a] = -sa; c->scr[b] = -sb;
        ratio_c(c, &r, &i);
        c->scr[a] = sa; c->scr[b] = sb;
        *ore += ocoef * r;
        *oim += ocoef * i;
    }
}

/* neighbour along one periodic / open dimension; -1 when absent */
static int nb_next(int x, int L, int periodic) {
    if (x + 1 < L) return x + 1;
    if (periodic && L >= 3) return 0;
    return -1;
}

static void le_tfim(const nqs_config_t *cfg, int Lx, int Ly, int per, le_ctx *c,
                    double *re, double *im) {
    /* nqs_gradient.c:50-82 (open) with optional wrap */
    double J = cfg->j_coupling, G = cfg->transverse_field, diag = 0.0;
    for (int x = 0; x < Lx; x++)
        for (int y = 0; y < Ly; y++) {
            int s = c->spins[x * Ly + y];
            int xn = nb_next(x, Lx, per), yn = nb_next(y, Ly, per);
            if (xn >= 0) diag += -J * (double)(s * c->spins[xn * Ly + y]);
            if (yn >= 0) diag += -J * (double)(s * c->spins[x * Ly + yn]);
        }
    double ore = 0.0, oim = 0.0;
    for (int i = 0; i < c->N; i++) {
        double r, q;
        c->scr[i] = -c->scr[i];
        ratio_c(c, &r, &q);
        ore += -G * r;
        oim += -G * q;
        c->scr[i] = -c->scr[i];
    }
    *re = diag + ore;
    *im = oim;
}

static void le_xxz(int Lx, int Ly, int per, double Jxy, double Jz, le_ctx *c,
                   double *re, double *im) {
    /* nqs_gradient.c:87-129 */
    double diag = 0.0, ore = 0.0, oim = 0.0;
    for (int x = 0; x < Lx; x++)
        for (int y = 0; y < Ly; y++) {
            int a = x * Ly + y;
            int xn = nb_next(x, Lx, per), yn = nb_next(y, Ly, per);
            if (xn >= 0) pair_term(c, a, xn * Ly + y, 0.25 * Jz, 0.5 * Jxy, &diag, &ore, &oim);
            if (yn >= 0) pair_term(c, a, x * Ly + yn, 0.25 * Jz, 0.5 * Jxy, &diag, &ore, &oim);
        }
    *re = diag + ore;
    *im = oim;
}

static void le_j1j2(const nqs_config_t *cfg, int Lx, int Ly, int per, le_ctx *c,
                    double *re, double *im) {
    /* nqs_gradient.c:146-205: Heisenberg(J1) then both diagonals */
    double h1r, h1i;
    le_xxz(Lx, Ly, per, cfg->j_coupling, cfg->j_coupling, c, &h1r, &h1i);
    double J2 = cfg->j2_coupling, d2 = 0.0, o2r = 0.0, o2i = 0.0;
    int xe = per && Lx >= 3 ? Lx : Lx - 1;
    int ye = per && Ly >= 3 ? Ly : Ly - 1;
    for (int x = 0; x < xe; x++)
        for (int y = 0; y < ye; y++) {
            int x1 = (x + 1) % Lx, y1 = (y + 1) % Ly;
            pair_term(c, x * Ly + y, x1 * Ly + y1, 0.25 * J2, 0.5 * J2, &d2, &o2r, &o2i);
            pair_term(c, x * Ly + y1, x1 * Ly + y, 0.25 * J2, 0.5 * J2, &d2, &o2r, &o2i);
        }
    *re = h1r + d2 + o2r;
    *im = h1i + o2i;
}

static void le_kagome(const nqs_config_t *cfg, int Lx, int Ly, le_ctx *c,
                      double *re, double *im) {
    /* nqs_gradient.c:374-445 */
    double J = cfg->j_coupling, diag = 0.0, ore = 0.0, oim = 0.0;
    int pbc = cfg->kagome_pbc;
    for (int cx = 0; cx < Lx; cx++)
        for (int cy = 0; cy < Ly; cy++) {
            int A = 3 * (cx * Ly + cy), B = A + 1, C = A + 2;
            pair_term(c, A, B, 0.25 * J, 0.5 * J, &diag, &ore, &oim);
            pair_term(c, A, C, 0.25 * J, 0.5 * J, &diag, &ore, &oim);
            pair_term(c, B, C, 0.25 * J, 0.5 * J, &diag, &ore, &oim);
            int cxm, cym;
            if (pbc) {
                cxm = (cx - 1 + Lx) % Lx;
                cym = (cy - 1 + Ly) % Ly;
            } else {
                if (cx == 0 || cy == 0) continue;
                cxm = cx - 1;
                cym = cy - 1;
            }
            int Bm = 3 * (cxm * Ly + cy) + 1, Cm = 3 * (cx * Ly + cym) + 2;
            pair_term(c, A, Bm, 0.25 * J, 0.5 * J, &diag, &ore, &oim);
            pair_term(c, A, Cm, 0.25 * J, 0.5 * J, &diag, &ore, &oim);
            pair_term(c, Bm, Cm, 0.25 * J, 0.5 * J, &diag, &ore, &oim);
        }
    *re = diag + ore;
    *im = oim;
}

static int ham_sites(const nqs_config_t *cfg, int Lx, int Ly) {
    return cfg->hamiltonian == NQS_HAM_KAGOME_HEISENBERG ? 3 * Lx * Ly : Lx * Ly;
}

/* nqs_gradient.c:447-491 / :807-875 (one routine: the real path equals
 * the real part of the complex path term by term). */
static void le_one(const nqs_config_t *cfg, int Lx, int Ly, int per,
                   const int *spins, orc_log_amp_fn fn, void *user,
                   double *re, double *im) {
    int N = ham_sites(cfg, Lx, Ly);
    le_ctx c;
    c.spins = spins;
    c.N = N;
    c.fn = fn;
    c.user = user;
    c.scr = malloc((size_t)N * sizeof(int));
    memcpy(c.scr, spins, (size_t)N * sizeof(int));
    fn(spins, N, user, &c.la0, &c.ar0);
    switch (cfg->hamiltonian) {
    case NQS_HAM_HEISENBERG:
        le_xxz(Lx, Ly, per, cfg->j_coupling, cfg->j_coupling, &c, re, im); break;
    case NQS_HAM_XXZ:
        le_xxz(Lx, Ly, per, cfg->j_coupling, cfg->j_z_coupling, &c, re, im); break;
    case NQS_HAM_J1_J2:
        le_j1j2(cfg, Lx, Ly, per, &c, re, im); break;
    case NQS_HAM_KAGOME_HEISENBERG:
        le_kagome(cfg, Lx, Ly, &c, re, im); break;
    default:
        le_tfim(cfg, Lx, Ly, per, &c, re, im); break;
    }
    free(c.scr);
}

double orc_local_energy(const nqs_config_t *cfg, int Lx, int Ly, const int *spins,
                        orc_log_amp_fn fn, void *user) {
    if (!cfg || !spins || !fn) return 0.0;
    double re, im;
    le_one(cfg, Lx, Ly, 0, spins, fn, user, &re, &im);
    return re;
}
void orc_local_energy_complex(const nqs_config_t *cfg, int Lx, int Ly,
                              const int *spins, orc_log_amp_fn fn, void *user,
                              double *re, double *im) {
    if (!cfg || !spins || !fn || !re || !im) {
        if (re) *re = 0.0;
        if (im) *im = 0.0;
        return;
    }
    le_one(cfg, Lx, Ly, 0, spins, fn, user, re, im);
}
void orc_local_energy_batch(const nqs_config_t *cfg, int Lx, int Ly,
                            const int *spins, int B, orc_log_amp_fn fn,
                            void *user, double *out) {
    if (!cfg || !spins || !out || B <= 0) return;
    int N = ham_sites(cfg, Lx, Ly);
    for (int b = 0; b < B; b++)
        out[b] = orc_local_energy(cfg, Lx, Ly, spins + (size_t)b * N, fn, user);
}
void orc_local_energy_batch_complex(const nqs_config_t *cfg, int Lx, int Ly,
                                    const int *spins, int B, orc_log_amp_fn fn,
                                    void *user, double *re, double *im) {
    int N = ham_sites(cfg, Lx, Ly);
    for (int b = 0; b < B; b++)
        orc_local_energy_complex(cfg, Lx, Ly, spins + (size_t)b * N, fn, user,
                                 &re[b], &im[b]);
}
void orc_local_energy_batch_ex(const nqs_config_t *cfg, int Lx, int Ly, int per,
                               const int *spins, int B, orc_log_amp_fn fn,
                               void *user, double *re, double *im) {
    int N = ham_sites(cfg, Lx, Ly);
    for (int b = 0; b < B; b++) {
        double r, i;
        le_one(cfg, Lx, Ly, per, spins + (size_t)b * N, fn, user, &r, &i);
        if (re) re[b] = r;
        if (im) im[b] = i;
    }
}

/* ===================================================================== */
/* optimiser — src/nqs/nqs_optimizer.c                                     */
/* ===================================================================== */
static double dotv(const double *a, const double *b, long n) {
    double s = 0.0;
    for (long i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

/* (S + eps I) v on uncentred O: nqs_optimizer.c:44-80 */
static void qgt_mv(const double *O, long Ns, long P, const double *Om, double eps,
                   const double *v, double *out, double *u) {
    for (long i = 0; i < Ns; i++) u[i] = dotv(O + i * P, v, P);
    double um = 0.0;
    for (long i = 0; i < Ns; i++) um += u[i];
    um /= (double)Ns;
    for (long k = 0; k < P; k++) {
        double s = 0.0;
        for (long i = 0; i < Ns; i++) s += O[i * P + k] * u[i];
        out[k] = s / (double)Ns - Om[k] * um + eps * v[k];
    }
}

/* G v = 2[(R^T R v + I^T I v)/Ns - Rm <Rv> - Im <Iv>] + eps v:
 * nqs_optimizer.c:317-356 */
static void qgt_mv_c(const double *R, const double *I, long Ns, long P,
                     const double *Rm, const double *Im, double eps,
                     const double *v, double *out, double *uR, double *uI) {
    for (long i = 0; i < Ns; i++) {
        uR[i] = dotv(R + i * P, v, P);
        uI[i] = dotv(I + i * P, v, P);
    }
    double mR = 0.0, mI = 0.0;
    for (long i = 0; i < Ns; i++) { mR += uR[i]; mI += uI[i]; }
    mR /= (double)Ns;
    mI /= (double)Ns;
    for (long k = 0; k < P; k++) {
        double aR = 0.0, aI = 0.0;
        for (long i = 0; i < Ns; i++) {
            aR += R[i * P + k] * uR[i];
            aI += I[i * P + k] * uI[i];
        }
        out[k] = 2.0 * ((aR + aI) / (double)Ns - Rm[k] * mR - Im[k] * mI) + eps * v[k];
    }
}

/* CG from x0 = 0: nqs_optimizer.c:84-139 (real, breakdown message) and
 * :358-401 (complex, silent breakdown).  Stops when |r|^2 <= tol^2 |r0|^2
 * (returns iter+1, converged) or |pAp| < 1e-30 or max_iters. */
static int cg_core(int cplx, const double *R, const double *I, long Ns, long P,
                   const double *Rm, const double *Imn, double eps, const double *F,
                   int max_iters, double tol, double *x, int *conv) {
    memset(x, 0, (size_t)P * sizeof(double));
    double *r = malloc(P * sizeof(double)), *p = malloc(P * sizeof(double));
    double *Ap = malloc(P * sizeof(double));
    double *u1 = malloc(Ns * sizeof(double)), *u2 = malloc(Ns * sizeof(double));
    memcpy(r, F, P * sizeof(double));
    memcpy(p, F, P * sizeof(double));
    double rs = dotv(r, r, P), rs0 = rs;
    int it, ok = 0;
    for (it = 0; it < max_iters; it++) {
        if (cplx) qgt_mv_c(R, I, Ns, P, Rm, Imn, eps, p, Ap, u1, u2);
        else      qgt_mv(R, Ns, P, Rm, eps, p, Ap, u1);
        double pAp = dotv(p, Ap, P);
        if (!(fabs(pAp) >= 1e-30)) {
            if (!cplx)
                fprintf(stderr, "orc cg_solve: Krylov breakdown at iter=%d, pAp=%.3e\n",
                        it, pAp);
            break;
        }
        double al = rs / pAp;
        for (long k = 0; k < P; k++) x[k] += al * p[k];
        for (long k = 0; k < P; k++) r[k] += -al * Ap[k];
        double rn = dotv(r, r, P);
        if (rn <= tol * tol * rs0) { ok = 1; it++; break; }
        double be = rn / rs;
        for (long k = 0; k < P; k++) p[k] = r[k] + be * p[k];
        rs = rn;
    }
    free(r); free(p); free(Ap); free(u1); free(u2);
    if (conv) *conv = ok;
    return it;
}

int orc_cg_solve(const double *O, long Ns, long P, const double *Om, double eps,
                 const double *F, int max_iters, double tol, double *delta, int *conv) {
    return cg_core(0, O, NULL, Ns, P, Om, NULL, eps, F, max_iters, tol, delta, conv);
}
int orc_cg_solve_complex(const double *R, const double *I, long Ns, long P,
                         const double *Rm, const double *Im, double eps,
                         const double *F, int max_iters, double tol,
                         double *delta, int *conv) {
    return cg_core(1, R, I, Ns, P, Rm, Im, eps, F, max_iters, tol, delta, conv);
}

typedef struct {
    int N, B;
    long P;
    int *batch;
    double *E, *Ei;
    double *O, *Oi;
} sr_buf;

static int sr_collect(const nqs_config_t *cfg, int Lx, int Ly, orc_ansatz_t *a,
                      orc_sampler_t *s, orc_log_amp_fn fn, void *user,
                      int holo, orc_gradient_fn gfn, orc_holo_gradient_fn hfn,
                      void *guser, sr_buf *b) {
    b->N = orc_sampler_num_sites(s);
    b->P = orc_ansatz_num_params(a);
    b->B = cfg->num_samples;
    if (b->P <= 0 || b->B <= 0) return -1;
    size_t B = (size_t)b->B;
    b->batch = malloc(B * b->N * sizeof(int));
    b->E = malloc(B * sizeof(double));
    b->Ei = malloc(B * sizeof(double));
    b->O = malloc(B * b->P * sizeof(double));
    b->Oi = holo ? malloc(B * b->P * sizeof(double)) : NULL;
    orc_sampler_batch(s, b->B, b->batch);
    orc_local_energy_batch_ex(cfg, Lx, Ly, s->opt.periodic, b->batch, b->B, fn, user,
                              b->E, b->Ei);
    for (size_t i = 0; i < B; i++) {
        int *sp = b->batch + i * b->N;
        double *g = b->O + i * b->P;
        if (holo) {
            double *gi = b->Oi + i * b->P;
            if (hfn) hfn(guser, a, sp, b->N, g, gi);
            else orc_ansatz_logpsi_gradient_complex(a, sp, b->N, g, gi);
        } else {
            if (gfn) gfn(guser, a, sp, b->N, g);
            else orc_ansatz_logpsi_gradient(a, sp, b->N, g);
        }
    }
    return 0;
}
static void sr_free(sr_buf *b) {
    free(b->batch); free(b->E); free(b->Ei); free(b->O); free(b->Oi);
}

/* nqs_optimizer.c:141-241 */
int orc_sr_step_custom_full(const nqs_config_t *cfg, int Lx, int Ly,
                            orc_ansatz_t *a, orc_sampler_t *s,
                            orc_log_amp_fn fn, void *user,
                            orc_gradient_fn gfn, void *guser, orc_sr_info_t *info) {
    if (!cfg || !a || !s) return -1;
    if (!fn) { fn = orc_ansatz_log_amp; user = a; }
    sr_buf b;
    if (sr_collect(cfg, Lx, Ly, a, s, fn, user, 0, gfn, NULL, guser, &b)) return -1;
    long P = b.P;
    int B = b.B;
    double em = 0.0, e2 = 0.0;
    for (int i = 0; i < B; i++) { em += b.E[i]; e2 += b.E[i] * b.E[i]; }
    em /= (double)B;
    e2 /= (double)B;
    double *Om = calloc(P, sizeof(double)), *F = calloc(P, sizeof(double));
    double *d = calloc(P, sizeof(double));
    for (long k = 0; k < P; k++) {
        double sg = 0.0, sge = 0.0;
        for (int i = 0; i < B; i++) {
            double g = b.O[(size_t)i * P + k];
            sg += g;
            sge += g * b.E[i];
        }
        Om[k] = sg / (double)B;
        F[k] = sge / (double)B - Om[k] * em;
    }
    int conv = 0;
    int it = orc_cg_solve(b.O, B, P, Om, cfg->sr_diag_shift, F, cfg->sr_cg_max_iters,
                          cfg->sr_cg_tol, d, &conv);
    orc_ansatz_apply_update(a, d, -cfg->learning_rate);
    if (info) {
        info->mean_energy = em;
        info->variance_energy = e2 - em * em;
        info->update_norm = sqrt(dotv(d, d, P)) * cfg->learning_rate;
        info->acceptance_ratio = orc_sampler_acceptance_ratio(s);
        info->cg_iterations = it;
        info->converged = conv;
    }
    free(Om); free(F); free(d);
    sr_free(&b);
    return 0;
}

/* nqs_optimizer.c:264-284 */
int orc_sr_run_custom_full(const nqs_config_t *cfg, int Lx, int Ly,
                           orc_ansatz_t *a, orc_sampler_t *s,
                           orc_log_amp_fn fn, void *user,
                           orc_gradient_fn gfn, void *guser, double *trace) {
    if (!cfg || !a || !s) return -1;
    orc_sampler_thermalize(s);
    for (int it = 0; it < cfg->num_iterations; it++) {
        orc_sr_info_t inf;
        int rc = orc_sr_step_custom_full(cfg, Lx, Ly, a, s, fn, user, gfn, guser, &inf);
        if (rc) return rc;
        if (trace) trace[it] = inf.mean_energy;
    }
    return 0;
}

/* nqs_optimizer.c:415-524 */
int orc_sr_step_holomorphic_full(const nqs_config_t *cfg, int Lx, int Ly,
                                 orc_ansatz_t *a, orc_sampler_t *s,
                                 orc_log_amp_fn fn, void *user,
                                 orc_holo_gradient_fn gfn, void *guser,
                                 orc_sr_info_t *info) {
    if (!cfg || !a || !s) return -1;
    if (!fn) { fn = orc_ansatz_log_amp; user = a; }
    sr_buf b;
    if (sr_collect(cfg, Lx, Ly, a, s, fn, user, 1, NULL, gfn, guser, &b)) return -1;
    long P = b.P;
    int B = b.B;
    double erm = 0.0, eim = 0.0, er2 = 0.0;
    for (int i = 0; i < B; i++) {
        erm += b.E[i];
        eim += b.Ei[i];
        er2 += b.E[i] * b.E[i];
    }
    erm /= (double)B;
    eim /= (double)B;
    er2 /= (double)B;
    double *Rm = calloc(P, sizeof(double)), *Im = calloc(P, sizeof(double));
    double *F = calloc(P, sizeof(double)), *d = calloc(P, sizeof(double));
    for (long k = 0; k < P; k++) {
        double sR = 0.0, sI = 0.0, sRe = 0.0, sIe = 0.0;
        for (int i = 0; i < B; i++) {
            double rk = b.O[(size_t)i * P + k], ik = b.Oi[(size_t)i * P + k];
            sR += rk;
            sI += ik;
            sRe += rk * b.E[i];
            sIe += ik * b.Ei[i];
        }
        Rm[k] = sR / (double)B;
        Im[k] = sI / (double)B;
        F[k] = 2.0 * ((sRe + sIe) / (double)B - Rm[k] * erm - Im[k] * eim);
    }
    int conv = 0;
    int it = orc_cg_solve_complex(b.O, b.Oi, B, P, Rm, Im, cfg->sr_diag_shift, F,
                                  cfg->sr_cg_max_iters, cfg->sr_cg_tol, d, &conv);
    orc_ansatz_apply_update(a, d, -cfg->learning_rate);
    if (info) {
        info->mean_energy = erm;
        info->variance_energy = er2 - erm * erm;
        info->update_norm = sqrt(dotv(d, d, P)) * cfg->learning_rate;
        info->acceptance_ratio = orc_sampler_acceptance_ratio(s);
        info->cg_iterations = it;
        info->converged = conv;
    }
    free(Rm); free(Im); free(F); free(d);
    sr_free(&b);
    return 0;
}

/* nqs_optimizer.c:538-558 */
int orc_sr_run_holomorphic_full(const nqs_config_t *cfg, int Lx, int Ly,
                                orc_ansatz_t *a, orc_sampler_t *s,
                                orc_log_amp_fn fn, void *user,
                                orc_holo_gradient_fn gfn, void *guser,
                                double *trace) {
    if (!cfg || !a || !s) return -1;
    orc_sampler_thermalize(s);
    for (int it = 0; it < cfg->num_iterations; it++) {
        orc_sr_info_t inf;
        int rc = orc_sr_step_holomorphic_full(cfg, Lx, Ly, a, s, fn, user, gfn, guser, &inf);
        if (rc) return rc;
        if (trace) trace[it] = inf.mean_energy;
    }
    return 0;
}

/* Holomorphic force 2[(R^T ea + I^T eb)/Ns - Rm <ea> - Im <eb>] and the
 * column means of R and I (the sums of nqs_optimizer.c:496-509, 677-690,
 * 797-810, with (ea, eb) the two energy series each caller pairs). */
static void holo_force(const sr_buf *b, const double *ea, const double *eb, double sa,
                       const double *means, double *Rm, double *Im, double *F) {
    long P = b->P;
    int B = b->B;
    double ma = 0.0, mb = 0.0;
    if (means) {
        ma = means[0];
        mb = means[1];
    } else {
        for (int i = 0; i < B; i++) { ma += ea[i]; mb += eb[i]; }
        ma /= (double)B;
        mb /= (double)B;
    }
    for (long k = 0; k < P; k++) {
        double sR = 0.0, sI = 0.0, sRa = 0.0, sIb = 0.0;
        for (int i = 0; i < B; i++) {
            double rk = b->O[(size_t)i * P + k], ik = b->Oi[(size_t)i * P + k];
            sR += rk;
            sI += ik;
            sRa += rk * ea[i];
            sIb += ik * eb[i];
        }
        Rm[k] = sR / (double)B;
        Im[k] = sI / (double)B;
        F[k] = 2.0 * ((sRa + sa * sIb) / (double)B - Rm[k] * ma - sa * Im[k] * mb);
    }
}

/* Excited-state SR with the orthogonality penalty mu |<r>|^2,
 * r(s) = psi_ref(s) / psi(s) with the log-ratio clamped to +-10:
 * nqs_optimizer.c:562-716.  The local energy is augmented by
 * mu r(s) conj(<r>); the force is the holomorphic one of the augmented
 * energies; info reports the physical energy. */
int orc_sr_step_excited(const nqs_config_t *cfg, int Lx, int Ly, orc_ansatz_t *a,
                        orc_sampler_t *s, orc_log_amp_fn fn, void *user,
                        orc_log_amp_fn ref_fn, void *ref_user, double mu,
                        orc_sr_info_t *info) {
    if (!cfg || !a || !s || !ref_fn) return -1;
    if (!fn) { fn = orc_ansatz_log_amp; user = a; }
    sr_buf b;
    if (sr_collect(cfg, Lx, Ly, a, s, fn, user, 1, NULL, NULL, NULL, &b)) return -1;
    long P = b.P;
    int B = b.B, N = b.N;
    double *rr = malloc((size_t)B * sizeof(double)), *ri = malloc((size_t)B * sizeof(double));
    for (int i = 0; i < B; i++) {
        const int *sp = b.batch + (size_t)i * N;
        double lp, ap, lr, ar;
        fn(sp, N, user, &lp, &ap);
        ref_fn(sp, N, ref_user, &lr, &ar);
        double dl = lr - lp;
        if (dl > 10.0) dl = 10.0;
        if (dl < -10.0) dl = -10.0;
        const double m = exp(dl), dth = ar - ap;
        rr[i] = m * cos(dth);
        ri[i] = m * sin(dth);
    }
    double Rr = 0.0, Ri = 0.0, erm = 0.0, eim = 0.0, er2 = 0.0;
    for (int i = 0; i < B; i++) {
        Rr += rr[i];
        Ri += ri[i];
        erm += b.E[i];
        eim += b.Ei[i];
        er2 += b.E[i] * b.E[i];
    }
    Rr /= (double)B;
    Ri /= (double)B;
    erm /= (double)B;
    eim /= (double)B;
    er2 /= (double)B;
    for (int i = 0; i < B; i++) {
        b.E[i] += mu * (rr[i] * Rr + ri[i] * Ri);
        b.Ei[i] += mu * (ri[i] * Rr - rr[i] * Ri);
    }
    /* the augmented means in closed form, as nqs_optimizer.c:654-657 */
    const double means[2] = {erm + mu * (Rr * Rr + Ri * Ri), eim + mu * (Ri * Rr - Rr * Ri)};
    double *Rm = calloc(P, sizeof(double)), *Im = calloc(P, sizeof(double));
    double *F = calloc(P, sizeof(double)), *d = calloc(P, sizeof(double));
    holo_force(&b, b.E, b.Ei, 1.0, means, Rm, Im, F);
    int conv = 0;
    int it = orc_cg_solve_complex(b.O, b.Oi, B, P, Rm, Im, cfg->sr_diag_shift, F,
                                  cfg->sr_cg_max_iters, cfg->sr_cg_tol, d, &conv);
    orc_ansatz_apply_update(a, d, -cfg->learning_rate);
    if (info) {
        info->mean_energy = erm;
        info->variance_energy = er2 - erm * erm;
        info->update_norm = sqrt(dotv(d, d, P)) * cfg->learning_rate;
        info->acceptance_ratio = orc_sampler_acceptance_ratio(s);
        info->cg_iterations = it;
        info->converged = conv;
    }
    free(rr); free(ri); free(Rm); free(Im); free(F); free(d);
    sr_free(&b);
    return 0;
}

/* nqs_optimizer.c:718-742 */
int orc_sr_run_excited(const nqs_config_t *cfg, int Lx, int Ly, orc_ansatz_t *a,
                       orc_sampler_t *s, orc_log_amp_fn fn, void *user,
                       orc_log_amp_fn ref_fn, void *ref_user, double mu, double *trace) {
    if (!cfg || !a || !s || !ref_fn) return -1;
    orc_sampler_thermalize(s);
    for (int it = 0; it < cfg->num_iterations; it++) {
        orc_sr_info_t inf;
        int rc = orc_sr_step_excited(cfg, Lx, Ly, a, s, fn, user, ref_fn, ref_user, mu, &inf);
        if (rc) return rc;
        if (trace) trace[it] = inf.mean_energy;
    }
    return 0;
}

/* Real-time tVMC direction Re(S)^-1 Im(F) at the current parameters from a
 * fresh batch, F_k = 2[Cov(R_k, E_im) - Cov(I_k, E_re)]
 * (nqs_optimizer.c:747-828); the batch has the sampler's site count (the
 * reference's Heun helper sizes it Lx * Ly, which is the same number on
 * square lattices). */
static int tvmc_delta(const nqs_config_t *cfg, int Lx, int Ly, orc_ansatz_t *a,
                      orc_sampler_t *s, orc_log_amp_fn fn, void *user, double *d,
                      double *em_out, double *var_out, int *it_out, int *conv_out) {
    sr_buf b;
    if (sr_collect(cfg, Lx, Ly, a, s, fn, user, 1, NULL, NULL, NULL, &b)) return -1;
    long P = b.P;
    int B = b.B;
    double erm = 0.0, er2 = 0.0;
    for (int i = 0; i < B; i++) { erm += b.E[i]; er2 += b.E[i] * b.E[i]; }
    erm /= (double)B;
    er2 /= (double)B;
    double *Rm = calloc(P, sizeof(double)), *Im = calloc(P, sizeof(double));
    double *F = calloc(P, sizeof(double));
    /* 2[(R^T E_im - I^T E_re)/Ns - Rm <E_im> + Im <E_re>] */
    holo_force(&b, b.Ei, b.E, -1.0, NULL, Rm, Im, F);
    int conv = 0;
    int it = orc_cg_solve_complex(b.O, b.Oi, B, P, Rm, Im, cfg->sr_diag_shift, F,
                                  cfg->sr_cg_max_iters, cfg->sr_cg_tol, d, &conv);
    if (em_out) *em_out = erm;
    if (var_out) *var_out = er2 - erm * erm;
    if (it_out) *it_out = it;
    if (conv_out) *conv_out = conv;
    free(Rm); free(Im); free(F);
    sr_free(&b);
    return 0;
}

/* Forward Euler theta <- theta + dt delta: nqs_optimizer.c:904-993 */
int orc_tvmc_step_real_time(const nqs_config_t *cfg, double dt, int Lx, int Ly,
                            orc_ansatz_t *a, orc_sampler_t *s, orc_log_amp_fn fn,
                            void *user, orc_sr_info_t *info) {
    if (!cfg || !a || !s) return -1;
    if (!fn) { fn = orc_ansatz_log_amp; user = a; }
    long P = orc_ansatz_num_params(a);
    double *d = calloc(P, sizeof(double));
    double em, var;
    int it, conv;
    if (tvmc_delta(cfg, Lx, Ly, a, s, fn, user, d, &em, &var, &it, &conv)) { free(d); return -1; }
    orc_ansatz_apply_update(a, d, dt);
    if (info) {
        info->mean_energy = em;
        info->variance_energy = var;
        info->update_norm = sqrt(dotv(d, d, P)) * fabs(dt);
        info->acceptance_ratio = orc_sampler_acceptance_ratio(s);
        info->cg_iterations = it;
        info->converged = conv;
    }
    free(d);
    return 0;
}

/* Heun: k1 = delta(theta), k2 = delta(theta + dt k1) from a second batch,
 * theta <- theta + dt (k1 + k2) / 2: nqs_optimizer.c:830-902 */
int orc_tvmc_step_heun(const nqs_config_t *cfg, double dt, int Lx, int Ly,
                       orc_ansatz_t *a, orc_sampler_t *s, orc_log_amp_fn fn,
                       void *user, orc_sr_info_t *info) {
    if (!cfg || !a || !s) return -1;
    if (!fn) { fn = orc_ansatz_log_amp; user = a; }
    long P = orc_ansatz_num_params(a);
    double *k1 = calloc(P, sizeof(double)), *k2 = calloc(P, sizeof(double));
    double *th0 = malloc((size_t)P * sizeof(double));
    memcpy(th0, orc_ansatz_params_raw(a), (size_t)P * sizeof(double));
    double e1 = 0.0, v1 = 0.0;
    int it1 = 0, c1 = 0, it2 = 0, c2 = 0;
    int rc = tvmc_delta(cfg, Lx, Ly, a, s, fn, user, k1, &e1, &v1, &it1, &c1);
    if (!rc) {
        orc_ansatz_apply_update(a, k1, dt);
        rc = tvmc_delta(cfg, Lx, Ly, a, s, fn, user, k2, NULL, NULL, &it2, &c2);
    }
    memcpy(orc_ansatz_params_raw(a), th0, (size_t)P * sizeof(double));
    if (!rc) {
        for (long k = 0; k < P; k++) k1[k] = 0.5 * (k1[k] + k2[k]);
        orc_ansatz_apply_update(a, k1, dt);
        if (info) {
            info->mean_energy = e1;
            info->variance_energy = v1;
            info->update_norm = sqrt(dotv(k1, k1, P)) * fabs(dt);
            info->acceptance_ratio = orc_sampler_acceptance_ratio(s);
            info->cg_iterations = it1 + it2;
            info->converged = c1 && c2;
        }
    }
    free(k1); free(k2); free(th0);
    return rc;
}

/* Unblocked lower Cholesky in place, upper triangle zeroed on success:
 * nqs_optimizer.c:1026-1043.  Only the lower triangle is read. */
int orc_cholesky_lower(double *A, int n) {
    for (int j = 0; j < n; j++) {
        double d = A[(size_t)j * n + j];
        for (int k = 0; k < j; k++) d -= A[(size_t)j * n + k] * A[(size_t)j * n + k];
        if (!(d > 0.0)) return -1;
        double L = sqrt(d);
        A[(size_t)j * n + j] = L;
        double inv = 1.0 / L;
        for (int i = j + 1; i < n; i++) {
            double s = A[(size_t)i * n + j];
            for (int k = 0; k < j; k++) s -= A[(size_t)i * n + k] * A[(size_t)j * n + k];
            A[(size_t)i * n + j] = s * inv;
        }
    }
    for (int i = 0; i < n; i++)
        for (int j = i + 1; j < n; j++) A[(size_t)i * n + j] = 0.0;
    return 0;
}

/* L L^T x = b: nqs_optimizer.c:1046-1061 */
void orc_cholesky_solve(const double *L, int n, const double *b, double *x) {
    double *z = malloc((size_t)n * sizeof(double));
    for (int i = 0; i < n; i++) {
        double s = b[i];
        for (int k = 0; k < i; k++) s -= L[(size_t)i * n + k] * z[k];
        z[i] = s / L[(size_t)i * n + i];
    }
    for (int i = n - 1; i >= 0; i--) {
        double s = z[i];
        for (int k = i + 1; k < n; k++) s -= L[(size_t)k * n + i] * x[k];
        x[i] = s / L[(size_t)i * n + i];
    }
    free(z);
}

/* MinSR linear algebra: nqs_optimizer.c:1101-1177 */
int orc_minsr_solve(double *O, long Ns, long P, const double *E, double eps,
                    double *delta, double *T_out, double *y_out) {
    double em = 0.0;
    for (long i = 0; i < Ns; i++) em += E[i];
    em /= (double)Ns;
    double *Om = calloc(P, sizeof(double));
    for (long k = 0; k < P; k++) {
        double s = 0.0;
        for (long i = 0; i < Ns; i++) s += O[i * P + k];
        Om[k] = s / (double)Ns;
    }
    double *ep = malloc(Ns * sizeof(double));
    for (long i = 0; i < Ns; i++) ep[i] = E[i] - em;
    for (long i = 0; i < Ns; i++)
        for (long k = 0; k < P; k++) O[i * P + k] -= Om[k];
    double *T = malloc((size_t)Ns * Ns * sizeof(double));
    double inv = 1.0 / (double)Ns;
    for (long i = 0; i < Ns; i++) {
        for (long j = 0; j <= i; j++) {
            double t = dotv(O + i * P, O + j * P, P) * inv;
            T[i * Ns + j] = t;
            T[j * Ns + i] = t;
        }
        T[i * Ns + i] += eps;
    }
    if (T_out) memcpy(T_out, T, (size_t)Ns * Ns * sizeof(double));
    double *y = calloc(Ns, sizeof(double));
    int ok = 0;
    int rc = orc_cholesky_lower(T, (int)Ns);
    if (rc != 0) {
        /* retry on the partially factored matrix, as the reference does */
        for (long i = 0; i < Ns; i++) T[i * Ns + i] += 1e-6;
        rc = orc_cholesky_lower(T, (int)Ns);
    }
    if (rc == 0) {
        orc_cholesky_solve(T, (int)Ns, ep, y);
        ok = 1;
    } else {
        fprintf(stderr, "orc minsr: Cholesky breakdown after retry; update set to 0\n");
    }
    for (long k = 0; k < P; k++) {
        double s = 0.0;
        for (long i = 0; i < Ns; i++) s += O[i * P + k] * y[i];
        delta[k] = s * inv;
    }
    if (y_out) memcpy(y_out, y, Ns * sizeof(double));
    free(Om); free(ep); free(T); free(y);
    return ok;
}

/* nqs_optimizer.c:1063-1194 */
int orc_sr_step_minsr_full(const nqs_config_t *cfg, int Lx, int Ly,
                           orc_ansatz_t *a, orc_sampler_t *s,
                           orc_log_amp_fn fn, void *user,
                           orc_gradient_fn gfn, void *guser, orc_sr_info_t *info) {
    if (!cfg || !a || !s) return -1;
    if (!fn) { fn = orc_ansatz_log_amp; user = a; }
    sr_buf b;
    if (sr_collect(cfg, Lx, Ly, a, s, fn, user, 0, gfn, NULL, guser, &b)) return -1;
    long P = b.P;
    int B = b.B;
    double em = 0.0, e2 = 0.0;
    for (int i = 0; i < B; i++) { em += b.E[i]; e2 += b.E[i] * b.E[i]; }
    em /= (double)B;
    e2 /= (double)B;
    double *d = calloc(P, sizeof(double));
    int ok = orc_minsr_solve(b.O, B, P, b.E, cfg->sr_diag_shift, d, NULL, NULL);
    orc_ansatz_apply_update(a, d, -cfg->learning_rate);
    if (info) {
        info->mean_energy = em;
        info->variance_energy = e2 - em * em;
        info->update_norm = sqrt(dotv(d, d, P)) * cfg->learning_rate;
        info->acceptance_ratio = orc_sampler_acceptance_ratio(s);
        info->cg_iterations = 0;
        info->converged = ok;
    }
    free(d);
    sr_free(&b);
    return 0;
}

/* nqs_optimizer.c:1196-1214 */
int orc_sr_run_minsr(const nqs_config_t *cfg, int Lx, int Ly, orc_ansatz_t *a,
                     orc_sampler_t *s, orc_log_amp_fn fn, void *user, double *trace) {
    if (!cfg || !a || !s) return -1;
    orc_sampler_thermalize(s);
    for (int it = 0; it < cfg->num_iterations; it++) {
        orc_sr_info_t inf;
        int rc = orc_sr_step_minsr_full(cfg, Lx, Ly, a, s, fn, user, NULL, NULL, &inf);
        if (rc) return rc;
        if (trace) trace[it] = inf.mean_energy;
    }
    return 0;
}

/* ===================================================================== */
/* exact diagonalisation (own Lanczos over the same bond sets)            */
/* ===================================================================== */
typedef struct {
    int nd, nf, np;
    int *da, *db; double *dc;   /* diagonal zz terms           */
    int *fi; double *fc;        /* single flips (TFIM)         */
    int *pa, *pb; double *pc;   /* antiparallel pair exchanges */
} bonds_t;

static void bonds_push_pair(bonds_t *b, int a, int c, double dcoef, double ocoef) {
    b->da[b->nd] = a; b->db[b->nd] = c; b->dc[b->nd++] = dcoef;
    b->pa[b->np] = a; b->pb[b->np] = c; b->pc[b->np++] = ocoef;
}

/* Same bond sets, in the same order, as the local-energy routines above. */
static void bonds_build(const nqs_config_t *cfg, int Lx, int Ly, int per, bonds_t *b) {
    int cap = 16 * (3 * Lx * Ly + 4);
    b->da = malloc(cap * sizeof(int)); b->db = malloc(cap * sizeof(int));
    b->dc = malloc(cap * sizeof(double));
    b->fi = malloc(cap * sizeof(int)); b->fc = malloc(cap * sizeof(double));
    b->pa = malloc(cap * sizeof(int)); b->pb = malloc(cap * sizeof(int));
    b->pc = malloc(cap * sizeof(double));
    b->nd = b->nf = b->np = 0;
    int h = cfg->hamiltonian;
    if (h == NQS_HAM_KAGOME_HEISENBERG) {
        double J = cfg->j_coupling;
        for (int cx = 0; cx < Lx; cx++)
            for (int cy = 0; cy < Ly; cy++) {
                int A = 3 * (cx * Ly + cy);
                bonds_push_pair(b, A, A + 1, 0.25 * J, 0.5 * J);
                bonds_push_pair(b, A, A + 2, 0.25 * J, 0.5 * J);
                bonds_push_pair(b, A + 1, A + 2, 0.25 * J, 0.5 * J);
                int cxm, cym;
                if (cfg->kagome_pbc) { cxm = (cx - 1 + Lx) % Lx; cym = (cy - 1 + Ly) % Ly; }
                else { if (cx == 0 || cy == 0) continue; cxm = cx - 1; cym = cy - 1; }
                int Bm = 3 * (cxm * Ly + cy) + 1, Cm = 3 * (cx * Ly + cym) + 2;
                bonds_push_pair(b, A, Bm, 0.25 * J, 0.5 * J);
                bonds_push_pair(b, A, Cm, 0.25 * J, 0.5 * J);
                bonds_push_pair(b, Bm, Cm, 0.25 * J, 0.5 * J);
            }
        return;
    }
    if (h == NQS_HAM_TFIM || (h != NQS_HAM_HEISENBERG && h != NQS_HAM_XXZ && h != NQS_HAM_J1_J2)) {
        double J = cfg->j_coupling;
        for (int x = 0; x < Lx; x++)
            for (int y = 0; y < Ly; y++) {
                int xn = nb_next(x, Lx, per), yn = nb_next(y, Ly, per);
                if (xn >= 0) { b->da[b->nd] = x * Ly + y; b->db[b->nd] = xn * Ly + y; b->dc[b->nd++] = -J; }
                if (yn >= 0) { b->da[b->nd] = x * Ly + y; b->db[b->nd] = x * Ly + yn; b->dc[b->nd++] = -J; }
            }
        for (int i = 0; i < Lx * Ly; i++) { b->fi[b->nf] = i; b->fc[b->nf++] = -cfg->transverse_field; }
        return;
    }
    double Jxy = cfg->j_coupling, Jz = (h == NQS_HAM_XXZ) ? cfg->j_z_coupling : cfg->j_coupling;
    for (int x = 0; x < Lx; x++)
        for (int y = 0; y < Ly; y++) {
            int xn = nb_next(x, Lx, per), yn = nb_next(y, Ly, per);
            if (xn >= 0) bonds_push_pair(b, x * Ly + y, xn * Ly + y, 0.25 * Jz, 0.5 * Jxy);
            if (yn >= 0) bonds_push_pair(b, x * Ly + y, x * Ly + yn, 0.25 * Jz, 0.5 * Jxy);
        }
    if (h == NQS_HAM_J1_J2) {
        double J2 = cfg->j2_coupling;
        int xe = per && Lx >= 3 ? Lx : Lx - 1, ye = per && Ly >= 3 ? Ly : Ly - 1;
        for (int x = 0; x < xe; x++)
            for (int y = 0; y < ye; y++) {
                int x1 = (x + 1) % Lx, y1 = (y + 1) % Ly;
                bonds_push_pair(b, x * Ly + y, x1 * Ly + y1, 0.25 * J2, 0.5 * J2);
                bonds_push_pair(b, x * Ly + y1, x1 * Ly + y, 0.25 * J2, 0.5 * J2);
            }
    }
}
static void bonds_free(bonds_t *b) {
    free(b->da); free(b->db); free(b->dc); free(b->fi); free(b->fc);
    free(b->pa); free(b->pb); free(b->pc);
}

/* basis: bit i = 1 <-> s_i = -1 (src/nqs/nqs_lanczos.c:16-20) */
static double diag_e(const bonds_t *b, u64 st) {
    double e = 0.0;
    for (int k = 0; k < b->nd; k++) {
        int sa = ((st >> b->da[k]) & 1) ? -1 : 1, sb = ((st >> b->db[k]) & 1) ? -1 : 1;
        e += b->dc[k] * (double)(sa * sb);
    }
    return e;
}

/* combinatorial ranking of a fixed-popcount state (colex order) */
static double **binom_tab(int n) {
    double **C = malloc((n + 1) * sizeof(double *));
    for (int i = 0; i <= n; i++) {
        C[i] = calloc(n + 2, sizeof(double));
        C[i][0] = 1.0;
        for (int k = 1; k <= i; k++) C[i][k] = C[i - 1][k - 1] + (k <= i - 1 ? C[i - 1][k] : 0.0);
    }
    return C;
}
static long rank_state(u64 st, double **C) {
    long r = 0;
    int k = 1;
    while (st) {
        int p = __builtin_ctzll(st);
        r += (long)C[p][k];
        k++;
        st &= st - 1;
    }
    return r;
}

typedef struct {
    const bonds_t *b;
    int N, sector;
    long dim;
    u64 *states;   /* sector basis (NULL: full space)   */
    double **C;
    double *diag;
} edop_t;

static void ed_matvec(const edop_t *op, const double *x, double *y) {
    const bonds_t *b = op->b;
    for (long r = 0; r < op->dim; r++) {
        u64 st = op->states ? op->states[r] : (u64)r;
        double acc = op->diag[r] * x[r];
        for (int k = 0; k < b->nf; k++) {
            u64 t = st ^ (1ULL << b->fi[k]);
            acc += b->fc[k] * x[(long)t];
        }
        for (int k = 0; k < b->np; k++) {
            int ba = (st >> b->pa[k]) & 1, bb = (st >> b->pb[k]) & 1;
            if (ba != bb) {
                u64 t = st ^ (1ULL << b->pa[k]) ^ (1ULL << b->pb[k]);
                long j = op->states ? rank_state(t, op->C) : (long)t;
                acc += b->pc[k] * x[j];
            }
        }
        y[r] = acc;
    }
}

/* smallest eigenvalue of a symmetric tridiagonal matrix by bisection on the
 * Sturm count */
static double tridiag_min(const double *a, const double *bt, int m) {
    double lo = 1e300, hi = -1e300;
    for (int i = 0; i < m; i++) {
        double r = (i > 0 ? fabs(bt[i - 1]) : 0.0) + (i < m - 1 ? fabs(bt[i]) : 0.0);
        if (a[i] - r < lo) lo = a[i] - r;
        if (a[i] + r > hi) hi = a[i] + r;
    }
    for (int it = 0; it < 200; it++) {
        double mid = 0.5 * (lo + hi);
        int cnt = 0;
        double q = 1.0;
        for (int i = 0; i < m; i++) {
            double b2 = i > 0 ? bt[i - 1] * bt[i - 1] : 0.0;
            q = a[i] - mid - (i > 0 ? b2 / q : 0.0);
            if (q == 0.0) q = 1e-300;
            if (q < 0.0) cnt++;
        }
        if (cnt >= 1) hi = mid; else lo = mid;
        if (hi - lo < 1e-15 * (fabs(lo) + fabs(hi) + 1e-300)) break;
    }
    return 0.5 * (lo + hi);
}

double orc_ed_ground_energy(const nqs_config_t *cfg, int Lx, int Ly, int per,
                            int n_down, int max_iters, double tol) {
    int N = ham_sites(cfg, Lx, Ly);
    if (N > 30) return NAN;
    bonds_t b;
    bonds_build(cfg, Lx, Ly, per, &b);
    edop_t op;
    memset(&op, 0, sizeof(op));
    op.b = &b;
    op.N = N;
    if (n_down >= 0) {
        if (b.nf > 0) { bonds_free(&b); return NAN; }
        op.C = binom_tab(N);
        op.dim = (long)op.C[N][n_down];
        op.states = malloc(op.dim * sizeof(u64));
        u64 st = n_down ? ((1ULL << n_down) - 1) : 0;
        for (long r = 0; r < op.dim; r++) {
            op.states[r] = st;
            if (n_down == 0) break;
            u64 c = st & -st, rr = st + c;   /* next state with the same popcount */
            st = (((rr ^ st) >> 2) / c) | rr;
        }
    } else {
        op.dim = 1L << N;
    }
    op.diag = malloc(op.dim * sizeof(double));
    for (long r = 0; r < op.dim; r++) op.diag[r] = diag_e(&b, op.states ? op.states[r] : (u64)r);
    /* plain three-vector Lanczos from a deterministic start vector */
    double *v0 = calloc(op.dim, sizeof(double)), *v1 = malloc(op.dim * sizeof(double));
    double *w = malloc(op.dim * sizeof(double));
    double *al = malloc((max_iters + 1) * sizeof(double)), *be = malloc((max_iters + 1) * sizeof(double));
    u64 seed = 0xA5A5A5A5A5A5A5A5ULL ^ (u64)op.dim;
    double nrm = 0.0;
    for (long r = 0; r < op.dim; r++) { v1[r] = xs_uniform(&seed) - 0.5; nrm += v1[r] * v1[r]; }
    nrm = sqrt(nrm);
    for (long r = 0; r < op.dim; r++) v1[r] /= nrm;
    double e_old = 1e300, e = NAN, bprev = 0.0;
    for (int k = 0; k < max_iters; k++) {
        ed_matvec(&op, v1, w);
        double a = 0.0;
        for (long r = 0; r < op.dim; r++) a += w[r] * v1[r];
        for (long r = 0; r < op.dim; r++) w[r] -= a * v1[r] + bprev * v0[r];
        double bb = 0.0;
        for (long r = 0; r < op.dim; r++) bb += w[r] * w[r];
        bb = sqrt(bb);
        al[k] = a;
        be[k] = bb;
        e = tridiag_min(al, be, k + 1);
        if (k > 5 && fabs(e - e_old) < tol) break;
        if (bb < 1e-14) break;
        e_old = e;
        for (long r = 0; r < op.dim; r++) { v0[r] = v1[r]; v1[r] = w[r] / bb; }
        bprev = bb;
    }
    free(v0); free(v1); free(w); free(al); free(be); free(op.diag); free(op.states);
    if (op.C) { for (int i = 0; i <= N; i++) free(op.C[i]); free(op.C); }
    bonds_free(&b);
    return e;
}

double orc_exact_energy(const nqs_config_t *cfg, int Lx, int Ly, int per,
                        orc_log_amp_fn fn, void *user) {
    int N = ham_sites(cfg, Lx, Ly);
    if (N > 20) return NAN;
    long dim = 1L << N;
    double *la = malloc(dim * sizeof(double)), *ar = malloc(dim * sizeof(double));
    int *s = malloc(N * sizeof(int));
    double mx = -1e300;
    for (long st = 0; st < dim; st++) {
        for (int i = 0; i < N; i++) s[i] = ((st >> i) & 1) ? -1 : 1;
        fn(s, N, user, &la[st], &ar[st]);
        if (la[st] > mx) mx = la[st];
    }
    bonds_t b;
    bonds_build(cfg, Lx, Ly, per, &b);
    /* <psi|H|psi> with complex psi: sum_s conj(psi_s) (H psi)_s */
    double num = 0.0, den = 0.0;
    for (long st = 0; st < dim; st++) {
        double m = exp(la[st] - mx);
        double pr = m * cos(ar[st]), pi = m * sin(ar[st]);
        den += m * m;
        double hr = diag_e(&b, (u64)st) * pr, hi = diag_e(&b, (u64)st) * pi;
        for (int k = 0; k < b.nf; k++) {
            long t = st ^ (1L << b.fi[k]);
            double mt = exp(la[t] - mx);
            hr += b.fc[k] * mt * cos(ar[t]);
            hi += b.fc[k] * mt * sin(ar[t]);
        }
        for (int k = 0; k < b.np; k++) {
            if ((((st >> b.pa[k]) ^ (st >> b.pb[k])) & 1) == 0) continue;
            long t = st ^ (1L << b.pa[k]) ^ (1L << b.pb[k]);
            double mt = exp(la[t] - mx);
            hr += b.pc[k] * mt * cos(ar[t]);
            hi += b.pc[k] * mt * sin(ar[t]);
        }
        num += pr * hr + pi * hi;
    }
    bonds_free(&b);
    free(la); free(ar); free(s);
    return num / den;
}

/* Integrated autocorrelation time, tau = 1/2 + sum_{t=1}^{W} rho(t), with
 * the smallest window W >= 5 tau(W) (Sokal's automatic windowing). */
double orc_tau_int(const double *x, int n) {
    if (n < 4) return 0.5;
    double m = 0.0;
    for (int i = 0; i < n; i++) m += x[i];
    m /= n;
    double c0 = 0.0;
    for (int i = 0; i < n; i++) c0 += (x[i] - m) * (x[i] - m);
    c0 /= n;
    if (c0 <= 0.0) return 0.5;
    double tau = 0.5;
    for (int t = 1; t < n; t++) {
        double c = 0.0;
        for (int i = 0; i + t < n; i++) c += (x[i] - m) * (x[i + t] - m);
        c /= n;
        tau += c / c0;
        if (t >= 5.0 * tau) break;
    }
    return tau;
}

/* thin wrappers of the reference API surface: nqs_optimizer.c:243-305,
 * 403-413, 526-536 */
int orc_sr_step(const nqs_config_t *cfg, int Lx, int Ly, orc_ansatz_t *a,
                orc_sampler_t *s, orc_sr_info_t *info) {
    return orc_sr_step_custom_full(cfg, Lx, Ly, a, s, NULL, NULL, NULL, NULL, info);
}
int orc_sr_run(const nqs_config_t *cfg, int Lx, int Ly, orc_ansatz_t *a,
               orc_sampler_t *s, double *trace) {
    return orc_sr_run_custom_full(cfg, Lx, Ly, a, s, NULL, NULL, NULL, NULL, trace);
}
int orc_sr_step_custom(const nqs_config_t *cfg, int Lx, int Ly, orc_ansatz_t *a,
                       orc_sampler_t *s, orc_log_amp_fn fn, void *user,
                       orc_sr_info_t *info) {
    return orc_sr_step_custom_full(cfg, Lx, Ly, a, s, fn, user, NULL, NULL, info);
}
int orc_sr_run_custom(const nqs_config_t *cfg, int Lx, int Ly, orc_ansatz_t *a,
                      orc_sampler_t *s, orc_log_amp_fn fn, void *user, double *trace) {
    return orc_sr_run_custom_full(cfg, Lx, Ly, a, s, fn, user, NULL, NULL, trace);
}
int orc_sr_step_holomorphic(const nqs_config_t *cfg, int Lx, int Ly, orc_ansatz_t *a,
                            orc_sampler_t *s, orc_log_amp_fn fn, void *user,
                            orc_sr_info_t *info) {
    return orc_sr_step_holomorphic_full(cfg, Lx, Ly, a, s, fn, user, NULL, NULL, info);
}
int orc_sr_run_holomorphic(const nqs_config_t *cfg, int Lx, int Ly, orc_ansatz_t *a,
                           orc_sampler_t *s, orc_log_amp_fn fn, void *user, double *trace) {
    return orc_sr_run_holomorphic_full(cfg, Lx, Ly, a, s, fn, user, NULL, NULL, trace);
}

/* ====================================================================== */
/* sample diagnostics                                                      */
/* ====================================================================== */

/* nqs_diagnostics.c:14-92: a fresh batch, the complex gradient of every
 * sample, Tr(S) = sum_k <|O_k|^2> - |<O_k>|^2 with values in (-1e-12, 0)
 * taken as 0. */
int orc_compute_chi_F(const nqs_config_t *cfg, int size_x, int size_y,
                      orc_ansatz_t *a, orc_sampler_t *s,
                      double *out_trace_S, double *out_per_param) {
    (void)size_x; (void)size_y;
    if (!cfg || !a || !s || !out_trace_S) return -1;
    const long P = orc_ansatz_num_params(a);
    const int N = orc_sampler_num_sites(s), B = cfg->num_samples;
    if (P <= 0 || B <= 0) return -1;
    int *smp = malloc((size_t)B * N * sizeof(int));
    double *acc = calloc((size_t)3 * P, sizeof(double));   /* sum re, sum im, sum |O|^2 */
    double *g = malloc((size_t)2 * P * sizeof(double));
    int rc = (!smp || !acc || !g) ? -1 : orc_sampler_batch(s, B, smp);
    for (int b = 0; rc == 0 && b < B; b++) {
        if (orc_ansatz_logpsi_gradient_complex(a, smp + (size_t)b * N, N, g, g + P)) { rc = -1; break; }
        for (long k = 0; k < P; k++) {
            acc[k] += g[k];
            acc[P + k] += g[P + k];
            acc[2 * P + k] += g[k] * g[k] + g[P + k] * g[P + k];
        }
    }
    if (rc == 0) {
        const double inv = 1.0 / B;
        double tr = 0.0;
        for (long k = 0; k < P; k++) {
            const double mr = acc[k] * inv, mi = acc[P + k] * inv;
            double v = acc[2 * P + k] * inv - (mr * mr + mi * mi);
            if (v < 0.0 && v > -1e-12) v = 0.0;
            tr += v;
        }
        *out_trace_S = tr;
        if (out_per_param) *out_per_param = tr / (double)P;
    }
    free(smp); free(acc); free(g);
    return rc;
}

/* nqs_diagnostics.c:113-222: per cell (cx-major) the up-triangle bonds
 * A-B, A-C, B-C and, with periodic cells or cx, cy >= 1, the down-triangle
 * bonds A-Bm, A-Cm, Bm-Cm (Bm: B of cell (cx-1, cy), Cm: C of cell
 * (cx, cy-1)); for every opposite-spin bond of every sample the complex
 * ratio psi(s with both spins flipped) / psi(s) from two log-amplitude
 * evaluations, summed per sublattice class in that order. */
int orc_compute_kagome_bond_phase(const nqs_config_t *cfg, int Lx, int Ly,
                                  orc_ansatz_t *a, orc_sampler_t *s,
                                  double *out_re, double *out_im, long *out_counts) {
    if (!cfg || !a || !s || !out_re || !out_im) return -1;
    if (cfg->hamiltonian != NQS_HAM_KAGOME_HEISENBERG) return -2;
    if (Lx < 1 || Ly < 1) return -1;
    const int N = 3 * Lx * Ly, B = cfg->num_samples;
    if (orc_sampler_num_sites(s) != N || B <= 0) return -1;
    /* bond list of one sample, in the order of the walk */
    int (*bond)[2] = malloc((size_t)6 * Lx * Ly * sizeof *bond);
    int nb = 0;
    for (int cx = 0; cx < Lx; cx++)
        for (int cy = 0; cy < Ly; cy++) {
            const int A = 3 * (cx * Ly + cy);
            bond[nb][0] = A; bond[nb][1] = A + 1; nb++;
            bond[nb][0] = A; bond[nb][1] = A + 2; nb++;
            bond[nb][0] = A + 1; bond[nb][1] = A + 2; nb++;
            if (!cfg->kagome_pbc && (cx == 0 || cy == 0)) continue;
            const int bm = 3 * (((cx - 1 + Lx) % Lx) * Ly + cy) + 1;
            const int cm = 3 * (cx * Ly + (cy - 1 + Ly) % Ly) + 2;
            bond[nb][0] = A; bond[nb][1] = bm; nb++;
            bond[nb][0] = A; bond[nb][1] = cm; nb++;
            bond[nb][0] = bm; bond[nb][1] = cm; nb++;
        }
    int *smp = malloc((size_t)B * N * sizeof(int));
    int *w = malloc((size_t)N * sizeof(int));
    double sr[3] = {0, 0, 0}, si[3] = {0, 0, 0};
    long cnt[3] = {0, 0, 0};
    int rc = orc_sampler_batch(s, B, smp);
    for (int b = 0; rc == 0 && b < B; b++) {
        const int *x = smp + (size_t)b * N;
        memcpy(w, x, (size_t)N * sizeof(int));
        double l0, p0;
        orc_ansatz_log_amp(x, N, a, &l0, &p0);
        for (int t = 0; t < nb; t++) {
            const int u = bond[t][0], v = bond[t][1];
            if (x[u] == x[v]) continue;
            const int su = u % 3, sv = v % 3;
            const int lo = su < sv ? su : sv, hi = su < sv ? sv : su;
            const int c = lo == 0 ? (hi == 1 ? 0 : 1) : 2;
            w[u] = -x[u]; w[v] = -x[v];
            double l1, p1;
            orc_ansatz_log_amp(w, N, a, &l1, &p1);
            w[u] = x[u]; w[v] = x[v];
            const double m = exp(l1 - l0);
            sr[c] += m * cos(p1 - p0);
            si[c] += m * sin(p1 - p0);
            cnt[c]++;
        }
    }
    if (rc == 0)
        for (int c = 0; c < 3; c++) {
            out_re[c] = cnt[c] ? sr[c] / (double)cnt[c] : 0.0;
            out_im[c] = cnt[c] ? si[c] / (double)cnt[c] : 0.0;
            if (out_counts) out_counts[c] = cnt[c];
        }
    free(bond); free(smp); free(w);
    return rc;
}

/* ====================================================================== */
/* symmetry projection                                                     */
/* ====================================================================== */

/* The orbit of one configuration: (g.s)_i = s_{perm[g][i]}, base
 * log-amplitudes lp[g], arg[g], and m = max_g lp[g].  rows: G * N ints. */
static double orbit_amps(orc_log_amp_fn fn, void *user, const int *perm, int G, int N,
                         const int *s, int *rows, double *lp, double *arg) {
    double m = -1e300;
    for (int g = 0; g < G; g++) {
        int *r = rows + (size_t)g * N;
        for (int i = 0; i < N; i++) r[i] = s[perm[(size_t)g * N + i]];
        fn(r, N, user, &lp[g], &arg[g]);
        if (lp[g] > m) m = lp[g];
    }
    return m;
}

/* nqs_symproj.c:30-80 / nqs_translation.c:24-72: real combination
 * S = sum chi cos(arg) e^{lp - m}, log|S| + m - log(G)/2, arg 0 or pi */
static void orbit_log_amp(orc_log_amp_fn fn, void *user, const int *perm, const double *chi,
                          int G, int N, const int *s, double *ola, double *oar) {
    int *rows = malloc((size_t)G * N * sizeof(int));
    double *lp = malloc((size_t)G * sizeof(double)), *ar = malloc((size_t)G * sizeof(double));
    const double m = orbit_amps(fn, user, perm, G, N, s, rows, lp, ar);
    double S = 0.0;
    for (int g = 0; g < G; g++) S += chi[g] * cos(ar[g]) * exp(lp[g] - m);
    if (fabs(S) > 0.0) {
        if (ola) *ola = log(fabs(S)) + m - 0.5 * log((double)G);
        if (oar) *oar = S < 0.0 ? M_PI : 0.0;
    } else {
        if (ola) *ola = -1e300;
        if (oar) *oar = 0.0;
    }
    free(rows); free(lp); free(ar);
}

/* nqs_symproj.c:82-153 (real weights), 155-263 (complex weights): base
 * gradients of a at the orbit members, weighted by c_g / sum c, c_g =
 * chi e^{lp - m} cos(arg) (real) or chi e^{lp - m + i arg} (complex);
 * zero when the sum vanishes, members of zero weight skipped */
static int orbit_gradient(orc_log_amp_fn fn, void *user, const int *perm, const double *chi,
                          int G, int N, orc_ansatz_t *a, const int *s, int cplx,
                          double *ore, double *oim) {
    const long P = orc_ansatz_num_params(a);
    int *rows = malloc((size_t)G * N * sizeof(int));
    double *lp = malloc((size_t)G * sizeof(double)), *ar = malloc((size_t)G * sizeof(double));
    double *cr = malloc((size_t)G * sizeof(double)), *ci = calloc((size_t)G, sizeof(double));
    double *tr = malloc((size_t)P * sizeof(double)), *ti = malloc((size_t)P * sizeof(double));
    const double m = orbit_amps(fn, user, perm, G, N, s, rows, lp, ar);
    double dr = 0.0, di = 0.0;
    for (int g = 0; g < G; g++) {
        if (cplx) {
            const double mag = chi[g] * exp(lp[g] - m);
            cr[g] = mag * cos(ar[g]);
            ci[g] = mag * sin(ar[g]);
        } else {
            cr[g] = chi[g] * cos(ar[g]) * exp(lp[g] - m);
        }
        dr += cr[g];
        di += ci[g];
    }
    memset(ore, 0, (size_t)P * sizeof(double));
    if (cplx) memset(oim, 0, (size_t)P * sizeof(double));
    const double d2 = dr * dr + di * di;
    int rc = 0;
    if (!(cplx ? d2 == 0.0 : dr == 0.0)) {
        for (int g = 0; g < G && rc == 0; g++) {
            const double wr = cplx ? (cr[g] * dr + ci[g] * di) / d2 : cr[g] / dr;
            const double wi = cplx ? (ci[g] * dr - cr[g] * di) / d2 : 0.0;
            if (wr == 0.0 && wi == 0.0) continue;
            const int *r = rows + (size_t)g * N;
            if (!cplx) {
                rc = orc_ansatz_logpsi_gradient(a, r, N, tr);
                for (long k = 0; rc == 0 && k < P; k++) ore[k] += wr * tr[k];
            } else {
                rc = orc_ansatz_logpsi_gradient_complex(a, r, N, tr, ti);
                for (long k = 0; rc == 0 && k < P; k++) {
                    ore[k] += wr * tr[k] - wi * ti[k];
                    oim[k] += wr * ti[k] + wi * tr[k];
                }
            }
        }
    }
    free(rows); free(lp); free(ar); free(cr); free(ci); free(tr); free(ti);
    return rc;
}

void orc_symproj_log_amp(const int *spins, int num_sites, void *user, double *ola, double *oar) {
    const orc_symproj_wrapper_t *w = user;
    if (!w || !w->base_log_amp || !w->perm || !w->characters || num_sites != w->num_sites ||
        w->num_sites <= 0 || w->num_group_elements <= 0) {
        if (ola) *ola = 0.0;
        if (oar) *oar = 0.0;
        return;
    }
    orbit_log_amp(w->base_log_amp, w->base_user, w->perm, w->characters, w->num_group_elements,
                  w->num_sites, spins, ola, oar);
}

int orc_symproj_gradient(void *u, orc_ansatz_t *a, const int *s, int n, double *out) {
    const orc_symproj_wrapper_t *w = u;
    if (!w || !a || !s || !out || n != w->num_sites || n <= 0 || w->num_group_elements <= 0)
        return -1;
    return orbit_gradient(w->base_log_amp, w->base_user, w->perm, w->characters,
                          w->num_group_elements, n, a, s, 0, out, NULL);
}

int orc_symproj_gradient_complex(void *u, orc_ansatz_t *a, const int *s, int n, double *re,
                                 double *im) {
    const orc_symproj_wrapper_t *w = u;
    if (!w || !a || !s || !re || !im || n != w->num_sites || n <= 0 ||
        w->num_group_elements <= 0)
        return -1;
    return orbit_gradient(w->base_log_amp, w->base_user, w->perm, w->characters,
                          w->num_group_elements, n, a, s, 1, re, im);
}

/* nqs_translation.c:36-48: element (tx, ty), tx-major, output site x Ly + y
 * takes site ((x + tx) mod Lx) Ly + (y + ty) mod Ly */
static void translation_table(int Lx, int Ly, int *perm, double *chi) {
    const int N = Lx * Ly;
    int g = 0;
    for (int tx = 0; tx < Lx; tx++)
        for (int ty = 0; ty < Ly; ty++, g++) {
            chi[g] = 1.0;
            for (int x = 0; x < Lx; x++)
                for (int y = 0; y < Ly; y++)
                    perm[(size_t)g * N + x * Ly + y] = ((x + tx) % Lx) * Ly + (y + ty) % Ly;
        }
}

void orc_translation_log_amp(const int *spins, int num_sites, void *user, double *ola,
                             double *oar) {
    const orc_translation_wrapper_t *w = user;
    const int N = w ? w->size_x * w->size_y : 0;
    if (!w || N <= 0 || num_sites != N) {
        if (ola) *ola = 0.0;
        if (oar) *oar = 0.0;
        return;
    }
    int *perm = malloc((size_t)N * N * sizeof(int));
    double *chi = malloc((size_t)N * sizeof(double));
    translation_table(w->size_x, w->size_y, perm, chi);
    orbit_log_amp(w->base_log_amp, w->base_user, perm, chi, N, N, spins, ola, oar);
    free(perm); free(chi);
}

int orc_translation_gradient(void *u, orc_ansatz_t *a, const int *s, int n, double *out) {
    const orc_translation_wrapper_t *w = u;
    if (!w || !a || !s || !out) return -1;
    const int N = w->size_x * w->size_y;
    if (N <= 0 || n != N) return -1;
    int *perm = malloc((size_t)N * N * sizeof(int));
    double *chi = malloc((size_t)N * sizeof(double));
    translation_table(w->size_x, w->size_y, perm, chi);
    const int rc = orbit_gradient(w->base_log_amp, w->base_user, perm, chi, N, N, a, s, 0, out,
                                  NULL);
    free(perm); free(chi);
    return rc;
}
