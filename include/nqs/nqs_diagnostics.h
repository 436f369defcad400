/*
 * nqs/nqs_diagnostics.h — sample diagnostics of libnqs_hip.so, the
 * reference's include/nqs/nqs_diagnostics.h interface (same names,
 * arguments, return codes; src/nqs/nqs_diagnostics.c).  Both draw a fresh
 * batch of cfg->num_samples configurations from the sampler exactly as
 * nqs_sampler_batch does and leave the ansatz unchanged.
 */
#ifndef NQS_DIAGNOSTICS_H
#define NQS_DIAGNOSTICS_H

#include "nqs_config.h"
#include "nqs_ansatz.h"
#include "nqs_sampler.h"

#ifdef __cplusplus
extern "C" {
#endif

/* kagome bond classes by sublattice pair: 0 A-B, 1 A-C, 2 B-C
 * (reference include/nqs/nqs_diagnostics.h:33) */
#define NQS_KAGOME_NUM_BOND_CLASSES 3

/* Tr(S) = sum_k ( <|O_k|^2> - |<O_k>|^2 ) of the quantum geometric tensor
 * over the batch, O_k = d log psi / d theta_k (complex for the complex RBM);
 * a column's value in (-1e-12, 0) counts as 0.  *out_per_param (may be NULL)
 * receives Tr(S) / num_params.  size_x, size_y are unused (signature of the
 * SR steps).  0 on success, -1 on bad arguments.
 * Replaces src/nqs/nqs_diagnostics.c:14-92 (O rows on the device, column
 * moments reduced there). */
int nqs_compute_chi_F(const nqs_config_t *cfg, int size_x, int size_y,
                      nqs_ansatz_t *ansatz, nqs_sampler_t *sampler,
                      double *out_trace_S, double *out_per_param);

/* Per-class mean of r = psi(s_ij) / psi(s) over the opposite-spin kagome
 * bonds (i, j) of every sampled s (up triangles, and down triangles where
 * the cell's neighbours exist or cfg->kagome_pbc is set); out_mean_re[3],
 * out_mean_im[3], out_counts[3] (may be NULL) per class.  -2 when
 * cfg->hamiltonian is not NQS_HAM_KAGOME_HEISENBERG, -1 on bad arguments.
 * Replaces src/nqs/nqs_diagnostics.c:113-222 (every ratio from the local
 * energy pair kernel on the device). */
int nqs_compute_kagome_bond_phase(const nqs_config_t *cfg, int Lx_cells, int Ly_cells,
                                  nqs_ansatz_t *ansatz, nqs_sampler_t *sampler,
                                  double *out_mean_re, double *out_mean_im,
                                  long *out_counts);

#ifdef __cplusplus
}
#endif

#endif /* NQS_DIAGNOSTICS_H */
