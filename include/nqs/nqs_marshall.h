/*
 * nqs/nqs_marshall.h — Marshall sign rule wrapper, exported by
 * libnqs_hip.so.
 *
 * Replaces include/nqs/nqs_marshall.h (:35-52) / src/nqs/nqs_marshall.c.
 * psi_phys(s) = (-1)^{n_down on sublattice A} psi_base(s), A = {(x+y) even},
 * site index x*size_y + y.  When base_log_amp is nqs_ansatz_log_amp the
 * wrapper is recognised by the device path and the parity is folded into
 * the kernels (no per-configuration host call).
 */
#ifndef NQS_MARSHALL_H
#define NQS_MARSHALL_H

#include "nqs_sampler.h"

#ifdef __cplusplus
extern "C" {
#endif

/* nqs_marshall.h:35-40 — field order is ABI. */
typedef struct {
    nqs_log_amp_fn_t base_log_amp;
    void            *base_user;
    int              size_x;
    int              size_y;
} nqs_marshall_wrapper_t;

/* nqs_marshall.h:45 / nqs_marshall.c:16-27: 0 even, 1 odd. */
int nqs_marshall_parity(const int *spins, int size_x, int size_y);

/* nqs_marshall.h:49-52 / nqs_marshall.c:29-40: base magnitude, base phase
 * plus pi on odd parity. */
void nqs_marshall_log_amp(const int *spins, int num_sites, void *user,
                          double *out_log_abs, double *out_arg);

#ifdef __cplusplus
}
#endif

#endif /* NQS_MARSHALL_H */
