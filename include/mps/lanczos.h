/*
 * mps/lanczos.h — the result record of the reference's Lanczos solvers
 * (include/mps/lanczos.h:34-42), for the exact-diagonalisation entry points
 * of nqs/nqs_lanczos.h.  The matrix-free solvers themselves
 * (lanczos_smallest*, src/mps/lanczos.c) are not exported by
 * libnqs_hip.so: its Lanczos runs on the device behind the nqs_lanczos_*
 * entry points.
 */
#ifndef MPS_LANCZOS_H
#define MPS_LANCZOS_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* H·v callback type of the reference solvers (layout compatibility) */
typedef void (*lanczos_matvec_fn_t)(const double *in, double *out, long dim, void *user_data);

typedef struct {
    double eigenvalue;       /* smallest eigenvalue found                         */
    int    iterations;       /* Krylov steps used                                  */
    int    converged;        /* 1 when the stopping rule was met                   */
    double residual_norm;    /* lean Lanczos: |change of the Ritz value| at the end */
} lanczos_result_t;

#ifdef __cplusplus
}
#endif

#endif /* MPS_LANCZOS_H */
