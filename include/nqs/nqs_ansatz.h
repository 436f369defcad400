/*
 * nqs/nqs_ansatz.h — wavefunction ansatz handle (mean-field, real RBM,
 * complex RBM) exported by libnqs_hip.so.
 *
 * Each entry point replaces the reference function of the same name in
 * include/nqs/nqs_ansatz.h (declarations :34-82, bodies in
 * src/nqs/nqs_ansatz.c).  The parameter vector keeps the reference layout:
 *   mean-field : [theta_0 .. theta_{N-1}]
 *   RBM        : [a (N) | b (M) | W (M x N, W[h*N + i])]
 *   complex RBM: [real block as RBM | imaginary block as RBM]
 * and the reference initialiser stream (xorshift64, nqs_ansatz.c:50-111,
 * 219-248), so a handle created here holds bit-identical parameters.
 *
 * The handle owns a host master copy (what params_raw returns) and a device
 * mirror in HBM; every compute entry point re-uploads the host copy first,
 * so external writes through params_raw are always honoured.
 */
#ifndef NQS_ANSATZ_H
#define NQS_ANSATZ_H

#include <stddef.h>
#include "nqs_config.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nqs_ansatz nqs_ansatz_t;

/* nqs_ansatz.h:34 / nqs_ansatz.c:65-252.  NULL for kinds this build does
 * not provide (ViT, factored ViT, autoregressive, KAN). */
nqs_ansatz_t *nqs_ansatz_create(const nqs_config_t *cfg, int num_sites);

/* nqs_ansatz.h:36 / nqs_ansatz.c:254-258 */
void nqs_ansatz_free(nqs_ansatz_t *a);

/* nqs_ansatz.h:39 / nqs_ansatz.c:260-262 */
long nqs_ansatz_num_params(const nqs_ansatz_t *a);

/* nqs_ansatz.h:43-46 / nqs_ansatz.c:846-873.  Callback-compatible
 * (nqs_log_amp_fn_t).  Evaluated on the GPU.  Writes (0, 0) when the
 * handle is NULL or num_sites does not match, as the reference does. */
void nqs_ansatz_log_amp(const int *spins, int num_sites, void *ansatz_user,
                        double *out_log_abs, double *out_arg);

/* nqs_ansatz.h:57-59 / nqs_ansatz.c:997-1008.  d log|psi| / d theta
 * (real-projected for the complex RBM).  0 on success, -1 on bad args. */
int nqs_ansatz_logpsi_gradient(nqs_ansatz_t *a, const int *spins,
                               int num_sites, double *out_grad);

/* nqs_ansatz.h:65-67 / nqs_ansatz.c:1258-1273.  Holomorphic gradient,
 * Re and Im parts; Im is zero for real kinds. */
int nqs_ansatz_logpsi_gradient_complex(nqs_ansatz_t *a, const int *spins,
                                       int num_sites, double *out_re,
                                       double *out_im);

/* nqs_ansatz.h:72 / nqs_ansatz.c:1010-1014 */
int nqs_ansatz_is_complex(const nqs_ansatz_t *a);

/* nqs_ansatz.h:76-77 / nqs_ansatz.c:1275-1282.  params += step * delta,
 * applied on the device and mirrored back to the host copy. */
int nqs_ansatz_apply_update(nqs_ansatz_t *a, const double *delta, double step);

/* nqs_ansatz.h:82 / nqs_ansatz.c:1284-1286.  Mutable host master copy. */
double *nqs_ansatz_params_raw(nqs_ansatz_t *a);

#ifdef __cplusplus
}
#endif

#endif /* NQS_ANSATZ_H */
