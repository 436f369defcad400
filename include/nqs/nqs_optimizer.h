/*
 * nqs/nqs_optimizer.h — stochastic reconfiguration (CG-SR, holomorphic
 * CG-SR, MinSR), exported by libnqs_hip.so.
 *
 * Replaces include/nqs/nqs_optimizer.h (:35-186, :307-323) /
 * src/nqs/nqs_optimizer.c.  One step is: sample Ns = cfg->num_samples
 * configurations, local energies, the O_k = d log psi / d theta_k matrix,
 * means and force F_k = <O_k E> - <O_k><E>, solve (S + eps I) delta = F,
 * theta -= lr * delta.  Every stage runs on the GPU with FP64 reductions and
 * solves.  The CG stopping rule (relative residual tol^2 * |r0|^2, breakdown
 * |pAp| < 1e-30), MinSR's Cholesky retry (+1e-6 on the diagonal, zero
 * update on a second failure) and every nqs_sr_step_info_t field follow the
 * reference definitions.  The excited-state penalty step and real-time
 * tVMC (Euler and Heun) reuse the same stages.
 */
#ifndef NQS_OPTIMIZER_H
#define NQS_OPTIMIZER_H

#include "nqs_config.h"
#include "nqs_ansatz.h"
#include "nqs_sampler.h"

#ifdef __cplusplus
extern "C" {
#endif

/* nqs_optimizer.h:35-42 */
typedef struct {
    double mean_energy;
    double variance_energy;
    double update_norm;
    double acceptance_ratio;
    int    cg_iterations;
    int    converged;
} nqs_sr_step_info_t;

/* nqs_optimizer.h:69-72: custom real-projected gradient (host callback). */
typedef int (*nqs_gradient_fn_t)(void *grad_user, nqs_ansatz_t *ansatz,
                                 const int *spins, int num_sites,
                                 double *out_grad);

/* nqs_optimizer.h:155-159: custom holomorphic gradient (host callback). */
typedef int (*nqs_holomorphic_gradient_fn_t)(void *grad_user,
                                             nqs_ansatz_t *ansatz,
                                             const int *spins, int num_sites,
                                             double *out_re, double *out_im);

/* nqs_optimizer.h:48-52 / nqs_optimizer.c:255-262 */
int nqs_sr_step(const nqs_config_t *cfg, int size_x, int size_y,
                nqs_ansatz_t *ansatz, nqs_sampler_t *sampler,
                nqs_sr_step_info_t *out_info);

/* nqs_optimizer.h:58-62 / nqs_optimizer.c:298-305 */
int nqs_sr_run(const nqs_config_t *cfg, int size_x, int size_y,
               nqs_ansatz_t *ansatz, nqs_sampler_t *sampler,
               double *out_energy_trace);

/* nqs_optimizer.h:78-84 / nqs_optimizer.c:243-253 */
int nqs_sr_step_custom(const nqs_config_t *cfg, int size_x, int size_y,
                       nqs_ansatz_t *ansatz, nqs_sampler_t *sampler,
                       nqs_log_amp_fn_t log_amp_fn, void *log_amp_user,
                       nqs_sr_step_info_t *out_info);

/* nqs_optimizer.h:86-92 / nqs_optimizer.c:286-296 */
int nqs_sr_run_custom(const nqs_config_t *cfg, int size_x, int size_y,
                      nqs_ansatz_t *ansatz, nqs_sampler_t *sampler,
                      nqs_log_amp_fn_t log_amp_fn, void *log_amp_user,
                      double *out_energy_trace);

/* nqs_optimizer.h:97-105 / nqs_optimizer.c:141-241 */
int nqs_sr_step_custom_full(const nqs_config_t *cfg, int size_x, int size_y,
                            nqs_ansatz_t *ansatz, nqs_sampler_t *sampler,
                            nqs_log_amp_fn_t log_amp_fn, void *log_amp_user,
                            nqs_gradient_fn_t gradient_fn, void *grad_user,
                            nqs_sr_step_info_t *out_info);

/* nqs_optimizer.h:107-115 / nqs_optimizer.c:264-284 */
int nqs_sr_run_custom_full(const nqs_config_t *cfg, int size_x, int size_y,
                           nqs_ansatz_t *ansatz, nqs_sampler_t *sampler,
                           nqs_log_amp_fn_t log_amp_fn, void *log_amp_user,
                           nqs_gradient_fn_t gradient_fn, void *grad_user,
                           double *out_energy_trace);

/* nqs_optimizer.h:136-142 / nqs_optimizer.c:403-413 */
int nqs_sr_step_holomorphic(const nqs_config_t *cfg, int size_x, int size_y,
                            nqs_ansatz_t *ansatz, nqs_sampler_t *sampler,
                            nqs_log_amp_fn_t log_amp_fn, void *log_amp_user,
                            nqs_sr_step_info_t *out_info);

/* nqs_optimizer.h:144-150 / nqs_optimizer.c:526-536 */
int nqs_sr_run_holomorphic(const nqs_config_t *cfg, int size_x, int size_y,
                           nqs_ansatz_t *ansatz, nqs_sampler_t *sampler,
                           nqs_log_amp_fn_t log_amp_fn, void *log_amp_user,
                           double *out_energy_trace);

/* nqs_optimizer.h:169-177 / nqs_optimizer.c:415-524 */
int nqs_sr_step_holomorphic_full(const nqs_config_t *cfg, int size_x,
                                 int size_y, nqs_ansatz_t *ansatz,
                                 nqs_sampler_t *sampler,
                                 nqs_log_amp_fn_t log_amp_fn,
                                 void *log_amp_user,
                                 nqs_holomorphic_gradient_fn_t grad_fn,
                                 void *grad_user,
                                 nqs_sr_step_info_t *out_info);

/* nqs_optimizer.h:179-186 / nqs_optimizer.c:538-558 */
int nqs_sr_run_holomorphic_full(const nqs_config_t *cfg, int size_x,
                                int size_y, nqs_ansatz_t *ansatz,
                                nqs_sampler_t *sampler,
                                nqs_log_amp_fn_t log_amp_fn,
                                void *log_amp_user,
                                nqs_holomorphic_gradient_fn_t grad_fn,
                                void *grad_user, double *out_energy_trace);

/* nqs_optimizer.h:307-315 / nqs_optimizer.c:1063-1194.  MinSR: sample-space
 * Gram T = O_c O_c^T / Ns + eps I, Cholesky solve T y = E - <E>,
 * delta = O_c^T y / Ns.  cg_iterations = 0, converged = factorisation ok. */
int nqs_sr_step_minsr_full(const nqs_config_t *cfg, int size_x, int size_y,
                           nqs_ansatz_t *ansatz, nqs_sampler_t *sampler,
                           nqs_log_amp_fn_t log_amp_fn, void *log_amp_user,
                           nqs_gradient_fn_t gradient_fn, void *grad_user,
                           nqs_sr_step_info_t *out_info);

/* nqs_optimizer.h:317-323 / nqs_optimizer.c:1196-1214 */
int nqs_sr_run_minsr(const nqs_config_t *cfg, int size_x, int size_y,
                     nqs_ansatz_t *ansatz, nqs_sampler_t *sampler,
                     nqs_log_amp_fn_t log_amp_fn, void *log_amp_user,
                     double *out_energy_trace);

/* nqs_optimizer.h:206-213 / nqs_optimizer.c:904-993.  Real-time tVMC,
 * forward Euler: F_k = 2[Cov(R_k, E_im) - Cov(I_k, E_re)] over holomorphic
 * O = R + iI, (G + eps I) delta = F with the holomorphic metric G,
 * theta <- theta + dt delta; update_norm = |delta| |dt|. */
int nqs_tvmc_step_real_time(const nqs_config_t *cfg, double dt, int size_x,
                            int size_y, nqs_ansatz_t *ansatz,
                            nqs_sampler_t *sampler, nqs_log_amp_fn_t log_amp_fn,
                            void *log_amp_user, nqs_sr_step_info_t *out_info);

/* nqs_optimizer.h:242-251 / nqs_optimizer.c:562-716.  Excited-state SR:
 * local energies augmented by penalty_mu r(s) conj(<r>),
 * r = psi_ref / psi (log-ratio clamped to +-10), holomorphic SR on them;
 * mean_energy and variance_energy report the physical energy. */
int nqs_sr_step_excited(const nqs_config_t *cfg, int size_x, int size_y,
                        nqs_ansatz_t *ansatz, nqs_sampler_t *sampler,
                        nqs_log_amp_fn_t log_amp_fn, void *log_amp_user,
                        nqs_log_amp_fn_t ref_log_amp_fn, void *ref_log_amp_user,
                        double penalty_mu, nqs_sr_step_info_t *out_info);

/* nqs_optimizer.h:253-262 / nqs_optimizer.c:718-742: thermalise, then
 * num_iterations excited steps; trace[it] = physical mean energy. */
int nqs_sr_run_excited(const nqs_config_t *cfg, int size_x, int size_y,
                       nqs_ansatz_t *ansatz, nqs_sampler_t *sampler,
                       nqs_log_amp_fn_t log_amp_fn, void *log_amp_user,
                       nqs_log_amp_fn_t ref_log_amp_fn, void *ref_log_amp_user,
                       double penalty_mu, double *out_energy_trace);

/* nqs_optimizer.h:267-274 / nqs_optimizer.c:830-902.  Heun: k1 = delta(theta),
 * k2 = delta(theta + dt k1) from a second batch, theta <- theta +
 * dt (k1 + k2) / 2; info from the first stage, cg_iterations summed,
 * converged = both stages converged. */
int nqs_tvmc_step_heun(const nqs_config_t *cfg, double dt, int size_x, int size_y,
                       nqs_ansatz_t *ansatz, nqs_sampler_t *sampler,
                       nqs_log_amp_fn_t log_amp_fn, void *log_amp_user,
                       nqs_sr_step_info_t *out_info);

#ifdef __cplusplus
}
#endif

#endif /* NQS_OPTIMIZER_H */
