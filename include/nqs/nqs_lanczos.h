/*
 * nqs/nqs_lanczos.h — exact diagonalisation of the kagome Heisenberg model
 * on the device (reference interface include/nqs/nqs_lanczos.h, same names,
 * arguments and return codes; src/nqs/nqs_lanczos.c).
 *
 * Basis: the 2^N computational states, bit i set <-> s_i = -1, N =
 * 3 * Lx_cells * Ly_cells (three sublattices per cell, site
 * 3 (cx Ly + cy) + sub); bonds as the reference enumerates them (up
 * triangle of every cell; down triangle with periodic cells or cx, cy >= 1).
 * H = J sum_bonds S_u . S_v.  Vectors live in HBM (8 * 2^N bytes each; four
 * at a time: 34 GB at N = 30).
 *
 * Of the reference's nqs_lanczos.h this library provides the kagome entry
 * points below; the TFIM / square-lattice Heisenberg refiners, the
 * full-reorthogonalisation and k-lowest variants are not exported.
 */
#ifndef NQS_LANCZOS_H
#define NQS_LANCZOS_H

#include "nqs_ansatz.h"
#include "nqs_sampler.h"
#include "../mps/lanczos.h"

#ifdef __cplusplus
extern "C" {
#endif

/* <psi|H|psi> / <psi|psi> of psi(s) = Re psi_ansatz(s) = |psi| cos(arg psi)
 * over all 2^N states (N <= 24), amplitudes evaluated on the device.
 * Replaces src/nqs/nqs_lanczos.c:434-451 (materialisation :115-149). */
int nqs_exact_energy_kagome_heisenberg(nqs_ansatz_t *a, int Lx_cells, int Ly_cells,
                                       double J, int pbc, double *out_energy);

/* The same with amplitudes from a log-amplitude callback (device-evaluated
 * when the callback is nqs_ansatz_log_amp or nqs_marshall_log_amp over an
 * ansatz, else one host call per basis state).  Replaces
 * src/nqs/nqs_lanczos.c:453-473. */
int nqs_exact_energy_kagome_heisenberg_with_cb(nqs_log_amp_fn_t log_amp, void *user,
                                               int Lx_cells, int Ly_cells,
                                               double J, int pbc, double *out_energy);

/* Lowest eigenvalue in the sector of the projector
 * P = (1/G) sum_g characters[g] T(g) (perm [G][N]: T(g) maps basis state s
 * to the state whose bit i is bit perm[g][i] of s) by the memory-lean
 * three-term Lanczos with P applied to every new Krylov vector
 * (src/nqs/nqs_lanczos.c:629-647, src/mps/lanczos.c:824-955): the
 * reference's deterministic xorshift start vector, stop when the smallest
 * Ritz value moves by less than tol after more than 5 steps or when
 * beta < 1e-14.  N <= 30.  0 on success, -1 on bad arguments. */
int nqs_lanczos_e0_kagome_heisenberg_projected_lean(int Lx_cells, int Ly_cells, double J,
                                                    int pbc, const int *perm,
                                                    const double *characters, int G,
                                                    int max_iters, double tol,
                                                    double *out_eigenvalue,
                                                    lanczos_result_t *out_result);

#ifdef __cplusplus
}
#endif

#endif /* NQS_LANCZOS_H */
