/*
 * oracle/sanity_main.c — TEST INFRASTRUCTURE ONLY.
 *
 * A driver that walks every entry point of the CPU oracle (nqs_oracle.c,
 * ar_oracle.c) once on small systems, so that the oracle can be built with
 * AddressSanitizer and UndefinedBehaviorSanitizer (`make -C oracle
 * sanitize`) and any out-of-bounds access, leak or undefined operation in
 * the checker shows up as a sanitizer report.  It checks a few values that
 * the mathematics fixes, so that a driver that silently skips work fails:
 *   - TFIM 4-site open chain, uniform amplitude: E_loc(all up) = -3J - 4h
 *     (3 bonds, every flip ratio 1);
 *   - Heisenberg 2x2 open ground energy -2.0 by ED (a 4-site ring with J=1:
 *     the open 2x2 square has the 4 bonds of a ring);
 *   - sum of the learned network's q over all 2^N states is 1.
 * Exit status 0 when every check passes.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "nqs_oracle.h"

static int fails = 0;
#define CHECK(cond, ...)                  \
  do {                                    \
    if (!(cond)) {                        \
      fprintf(stderr, "FAIL: " __VA_ARGS__); \
      fprintf(stderr, "\n");              \
      fails++;                            \
    }                                     \
  } while (0)

static void uniform_amp(const int *s, int n, void *u, double *la, double *ar) {
  (void)s; (void)n; (void)u;
  *la = 0.0;
  *ar = 0.0;
}

static void run_sr_family(int ansatz, int ham, int Lx, int Ly, int periodic) {
  nqs_config_t cfg = nqs_config_defaults();
  cfg.ansatz = ansatz;
  cfg.rbm_hidden_units = 6;
  cfg.hamiltonian = ham;
  cfg.kagome_pbc = 1;
  cfg.rbm_init_scale = 0.3;
  cfg.num_samples = 48;
  cfg.num_thermalize = 16;
  cfg.num_decorrelate = 2;
  cfg.num_iterations = 2;
  cfg.sr_cg_max_iters = 30;
  cfg.sr_diag_shift = 1e-3;
  cfg.learning_rate = 0.02;
  const int N = (ham == NQS_HAM_KAGOME_HEISENBERG) ? 3 * Lx * Ly : Lx * Ly;
  orc_ansatz_t *a = orc_ansatz_create(&cfg, N);
  CHECK(a != NULL, "ansatz %d", ansatz);
  if (!a) return;
  nqs_hip_options_t opt;
  memset(&opt, 0, sizeof(opt));
  opt.num_chains = 3;
  opt.periodic = periodic;
  opt.move = (ham == NQS_HAM_TFIM) ? NQS_HIP_MOVE_FLIP : NQS_HIP_MOVE_EXCHANGE;
  opt.sz_zero_start = (ham != NQS_HAM_TFIM);
  orc_sampler_t *s = orc_sampler_create_ex(N, &cfg, &opt, orc_ansatz_log_amp, a);
  CHECK(s != NULL, "sampler");
  if (!s) { orc_ansatz_free(a); return; }
  orc_sampler_thermalize(s);
  const long P = orc_ansatz_num_params(a);
  int *batch = malloc(sizeof(int) * 16 * N);
  double *re = malloc(sizeof(double) * 16), *im = malloc(sizeof(double) * 16);
  double *g = malloc(sizeof(double) * P), *gi = malloc(sizeof(double) * P);
  CHECK(orc_sampler_batch(s, 16, batch) == 0, "batch");
  orc_local_energy_batch_ex(&cfg, Lx, Ly, periodic, batch, 16, orc_ansatz_log_amp, a, re, im);
  for (int b = 0; b < 16; b++) CHECK(isfinite(re[b]) && isfinite(im[b]), "E_loc finite");
  CHECK(orc_ansatz_logpsi_gradient(a, batch, N, g) == 0, "gradient");
  CHECK(orc_ansatz_logpsi_gradient_complex(a, batch, N, g, gi) == 0, "gradient complex");
  orc_sr_info_t info;
  double trace[2];
  if (!periodic) {
    CHECK(orc_sr_step(&cfg, Lx, Ly, a, s, &info) == 0, "sr step");
    CHECK(orc_sr_step_holomorphic(&cfg, Lx, Ly, a, s, orc_ansatz_log_amp, a, &info) == 0, "holo");
    CHECK(orc_sr_step_minsr_full(&cfg, Lx, Ly, a, s, orc_ansatz_log_amp, a, NULL, NULL, &info) == 0,
          "minsr");
    CHECK(orc_sr_run(&cfg, Lx, Ly, a, s, trace) == 0, "sr run");
    CHECK(orc_sr_run_minsr(&cfg, Lx, Ly, a, s, orc_ansatz_log_amp, a, trace) == 0, "minsr run");
    CHECK(isfinite(trace[1]), "trace finite");
  }
  orc_marshall_wrapper_t w = {orc_ansatz_log_amp, a, Lx, Ly};
  if (ham != NQS_HAM_KAGOME_HEISENBERG) {
    double e = orc_local_energy(&cfg, Lx, Ly, batch, orc_marshall_log_amp, &w);
    CHECK(isfinite(e), "Marshall E_loc");
  }
  if (N <= 12) {
    const double ex = orc_exact_energy(&cfg, Lx, Ly, periodic, orc_ansatz_log_amp, a);
    CHECK(isfinite(ex), "exact energy");
  }
  free(batch); free(re); free(im); free(g); free(gi);
  orc_sampler_free(s);
  orc_ansatz_free(a);
}

int main(void) {
  /* closed forms of the local energy and the ED */
  nqs_config_t cfg = nqs_config_defaults();
  cfg.hamiltonian = NQS_HAM_TFIM;
  cfg.j_coupling = 1.0;
  cfg.transverse_field = 0.5;
  int up[4] = {1, 1, 1, 1};
  const double e = orc_local_energy(&cfg, 4, 1, up, uniform_amp, NULL);
  CHECK(fabs(e - (-3.0 - 4.0 * 0.5)) < 1e-12, "TFIM uniform E_loc %.15g", e);
  cfg.hamiltonian = NQS_HAM_HEISENBERG;
  const double e0 = orc_ed_ground_energy(&cfg, 2, 2, 0, -1, 200, 1e-12);
  CHECK(fabs(e0 - (-2.0)) < 1e-9, "Heisenberg 2x2 E0 %.15g", e0);

  /* every SR family on every Hamiltonian kind, open and periodic */
  run_sr_family(NQS_ANSATZ_RBM, NQS_HAM_TFIM, 4, 1, 0);
  run_sr_family(NQS_ANSATZ_COMPLEX_RBM, NQS_HAM_HEISENBERG, 2, 3, 0);
  run_sr_family(NQS_ANSATZ_RBM, NQS_HAM_J1_J2, 3, 3, 1);
  run_sr_family(NQS_ANSATZ_COMPLEX_RBM, NQS_HAM_XXZ, 3, 3, 1);
  run_sr_family(NQS_ANSATZ_COMPLEX_RBM, NQS_HAM_KAGOME_HEISENBERG, 2, 1, 0);
  run_sr_family(NQS_ANSATZ_LEGACY_MLP, NQS_HAM_HEISENBERG, 2, 2, 0);

  /* solver building blocks */
  {
    const int n = 5;
    double A[25], b[5], x[5];
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) A[i * n + j] = (i == j) ? 4.0 : 1.0 / (1 + i + j);
    for (int i = 0; i < n; i++) b[i] = 1.0 + i;
    CHECK(orc_cholesky_lower(A, n) == 0, "cholesky");
    orc_cholesky_solve(A, n, b, x);
    double O[6 * 4], E[6], d[4], T[36], y[6];
    for (int i = 0; i < 24; i++) O[i] = sin(0.7 * i);
    for (int i = 0; i < 6; i++) E[i] = cos(1.3 * i);
    CHECK(orc_minsr_solve(O, 6, 4, E, 1e-3, d, T, y) >= 0, "minsr solve");
    double series[64];
    for (int i = 0; i < 64; i++) series[i] = sin(0.3 * i) + 0.1 * cos(2.1 * i);
    CHECK(orc_tau_int(series, 64) > 0.0, "tau_int");
  }

  /* learned network: normalisation over the full space (N = 6) */
  {
    const int N = 6;
    const long P = orc_ar_num_params(N);
    float *w = malloc(sizeof(float) * P);
    orc_ar_init(N, 7, 1.0, w);
    int *all = malloc(sizeof(int) * 64 * N);
    double *lq = malloc(sizeof(double) * 64);
    for (int k = 0; k < 64; k++)
      for (int i = 0; i < N; i++) all[k * N + i] = ((k >> i) & 1) ? -1 : 1;
    CHECK(orc_ar_log_prob(N, w, all, 64, 0, 0, lq) == 0, "ar log prob");
    double z = 0.0;
    for (int k = 0; k < 64; k++) z += exp(lq[k]);
    CHECK(fabs(z - 1.0) < 1e-9, "ar normalisation %.15g", z);
    double *grad = malloc(sizeof(double) * P), *m = calloc(P, sizeof(double)),
           *v = calloc(P, sizeof(double)), nll = 0.0;
    CHECK(orc_ar_gradient(N, w, all, 64, 0, 0, grad, &nll) == 0, "ar gradient");
    orc_ar_adam_step(N, w, m, v, grad, 1, 1e-3);
    free(grad); free(m); free(v); free(w); free(all); free(lq);
  }

  if (fails) {
    fprintf(stderr, "%d checks failed\n", fails);
    return 1;
  }
  printf("oracle sanity ok\n");
  return 0;
}
