/*
 * nqs/nqs_translation.h — zero-momentum translation projection on Lx x Ly
 * square lattices (1D chains: size_y = 1), the reference's
 * include/nqs/nqs_translation.h interface (src/nqs/nqs_translation.c).
 * psi_sym(s) = (1/sqrt(Lx Ly)) sum_(tx,ty) psi_base(T s), (T s)[x Ly + y]
 * = s[((x + tx) mod Lx) Ly + (y + ty) mod Ly], combined as in
 * nqs_symproj_log_amp with all characters 1.
 */
#ifndef NQS_TRANSLATION_H
#define NQS_TRANSLATION_H

#include "nqs_sampler.h"
#include "nqs_ansatz.h"

#ifdef __cplusplus
extern "C" {
#endif

/* reference nqs_translation.h:28-33 */
typedef struct {
    nqs_log_amp_fn_t base_log_amp;
    void            *base_user;
    int              size_x;
    int              size_y;
} nqs_translation_wrapper_t;

/* reference nqs_translation.c:24-72 */
void nqs_translation_log_amp(const int *spins, int num_sites, void *user,
                             double *out_log_abs, double *out_arg);

/* reference nqs_translation.c:74-145 (real weights) */
int nqs_translation_gradient(void *grad_user, nqs_ansatz_t *ansatz,
                             const int *spins, int num_sites, double *out_grad);

#ifdef __cplusplus
}
#endif

#endif /* NQS_TRANSLATION_H */
