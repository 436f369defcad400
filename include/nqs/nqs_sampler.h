/*
 * nqs/nqs_sampler.h — Metropolis sampler over spin configurations drawn
 * from |psi|^2, exported by libnqs_hip.so.
 *
 * Replaces include/nqs/nqs_sampler.h (:35-75) / src/nqs/nqs_sampler.c.
 * With the default of one chain, a sampler created here walks exactly the
 * reference chain: xorshift64 seeded from rng_seed (0x9E3779B97F4A7C15 when
 * zero), uniform random start, single-site flips, accept when
 * 2*(log|psi'| - log|psi|) >= 0 or log(u) < that.  The chain runs on the GPU
 * when the callback is nqs_ansatz_log_amp or nqs_marshall_log_amp over it;
 * any other callback is a host function and is called from the host.
 *
 * More chains, Sz-conserving exchange moves and the learned proposal are
 * selected through nqs_hip/nqs_hip.h (nqs_hip_sampler_create_ex) or the
 * NQS_HIP_CHAINS environment variable.
 */
#ifndef NQS_SAMPLER_H
#define NQS_SAMPLER_H

#include <stddef.h>
#include "nqs_config.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nqs_sampler nqs_sampler_t;

/* nqs_sampler.h:35-37.  log|psi(s)| into out_re, arg psi(s) into out_im. */
typedef void (*nqs_log_amp_fn_t)(const int *spins, int num_sites, void *user,
                                 double *out_re, double *out_im);

/* nqs_sampler.h:42-45 / nqs_sampler.c:50-76.  NULL on bad args. */
nqs_sampler_t *nqs_sampler_create(int num_sites, const nqs_config_t *cfg,
                                  nqs_log_amp_fn_t log_amp, void *log_amp_user);

/* nqs_sampler.h:47 / nqs_sampler.c:78-83 */
void nqs_sampler_free(nqs_sampler_t *s);

/* nqs_sampler.h:51 / nqs_sampler.c:119-124: cfg.num_thermalize moves. */
void nqs_sampler_thermalize(nqs_sampler_t *s);

/* nqs_sampler.h:56 / nqs_sampler.c:126-133: max(num_decorrelate,1) moves,
 * returns the sampler-owned current configuration (chain 0). */
const int *nqs_sampler_next(nqs_sampler_t *s);

/* nqs_sampler.h:61 / nqs_sampler.c:135-144: n samples into out[n*N]. */
int nqs_sampler_batch(nqs_sampler_t *s, int n, int *out);

/* nqs_sampler.h:64 / nqs_sampler.c:146-149: lifetime acceptance ratio. */
double nqs_sampler_acceptance_ratio(const nqs_sampler_t *s);

/* nqs_sampler.h:67 / nqs_sampler.c:151-153 (chain 0). */
const int *nqs_sampler_current(const nqs_sampler_t *s);

/* nqs_sampler.h:75 / nqs_sampler.c:155-157 */
int nqs_sampler_num_sites(const nqs_sampler_t *s);

#ifdef __cplusplus
}
#endif

#endif /* NQS_SAMPLER_H */
