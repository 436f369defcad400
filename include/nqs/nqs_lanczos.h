/*
 * nqs/nqs_lanczos.h — exact diagonalisation on the device (reference
 * interface include/nqs/nqs_lanczos.h, same names, arguments and return
 * codes; src/nqs/nqs_lanczos.c, src/mps/lanczos.c).
 *
 * Basis: the 2^N computational states, bit i set <-> s_i = -1.  Models:
 * TFIM -J sz sz - Gamma sx and XXZ Heisenberg on the open Lx x Ly square
 * lattice (site x Ly + y, N = Lx Ly); kagome Heisenberg J S_u . S_v with
 * N = 3 Lx_cells Ly_cells (site 3 (cx Ly + cy) + sub; bonds as the
 * reference enumerates them: up triangle of every cell, down triangle with
 * periodic cells or cx, cy >= 1).  Vectors live in HBM (8 * 2^N bytes
 * each); the full-reorthogonalisation drivers keep their max_iters Krylov
 * vectors there too (max_iters * 8 * 2^N bytes; -1 when that does not fit).
 *
 * Materialised state (every seeded driver and exact energy): Re psi(s) =
 * cos(arg psi) exp(log|psi| - max), normalised, over all 2^N states
 * (N <= 24), amplitudes evaluated on the device when the callback is
 * nqs_ansatz_log_amp, nqs_marshall_log_amp or a symmetry / translation
 * projection over them, else one host call per state.  A warning goes to
 * stderr when Im psi holds more than 1e-6 of the norm.
 *
 * Refiners (smallest eigenvalue): full-reorthogonalisation Lanczos from the
 * materialised state (from the reference's deterministic xorshift vector
 * when N > 24, and always for the TFIM refiner), stop when
 * |beta_{k+1} z_k| < tol after at least two steps; out_eigenvalue is set on
 * success whether or not out_result is given.  k-lowest drivers run
 * max_iters steps (fewer on Krylov breakdown) and return the k smallest
 * Ritz values; converged = (last beta < 1e-10).  Projected variants apply
 * P = (1/G) sum_g characters[g] T(g) (perm [G][N]: T(g) maps basis state s
 * to the state whose bit i is bit perm[g][i] of s) to every new Krylov
 * vector.  N <= 30 for the Lanczos drivers.
 */
#ifndef NQS_LANCZOS_H
#define NQS_LANCZOS_H

#include "nqs_ansatz.h"
#include "nqs_sampler.h"
#include "../mps/lanczos.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Re psi of the ansatz on the Lx x Ly lattice (N = Lx Ly <= 24) as a
 * malloc'd vector of 2^N doubles (caller frees).  Replaces
 * src/nqs/nqs_lanczos.c:104-113. */
int nqs_materialise_state(nqs_ansatz_t *a, int Lx, int Ly, double **out_psi, long *out_dim);

/* The same from a log-amplitude callback (:65-102, :115-149). */
int nqs_materialise_state_with_cb(nqs_log_amp_fn_t log_amp, void *user, int Lx, int Ly,
                                  double **out_psi, long *out_dim);
int nqs_materialise_state_with_cb_N(nqs_log_amp_fn_t log_amp, void *user, int N,
                                    double **out_psi, long *out_dim);

/* <psi|H|psi> / <psi|psi> of the materialised ansatz for the open-lattice
 * TFIM (:193-208) and XXZ Heisenberg (:283-298). */
int nqs_exact_energy_tfim(nqs_ansatz_t *a, int Lx, int Ly, double J, double Gamma,
                          double *out_energy);
int nqs_exact_energy_heisenberg(nqs_ansatz_t *a, int Lx, int Ly, double J, double Jz,
                                double *out_energy);

/* Smallest eigenvalue (and unit eigenvector, 2^N doubles, when
 * out_eigenvector is non-null) of the open-lattice TFIM from the
 * deterministic start vector (:210-222) and of the XXZ Heisenberg model
 * seeded from the ansatz (:300-322). */
int nqs_lanczos_refine_tfim(nqs_ansatz_t *a, int Lx, int Ly, double J, double Gamma,
                            int max_iters, double tol, double *out_eigenvalue,
                            double *out_eigenvector, lanczos_result_t *out_result);
int nqs_lanczos_refine_heisenberg(nqs_ansatz_t *a, int Lx, int Ly, double J, double Jz,
                                  int max_iters, double tol, double *out_eigenvalue,
                                  double *out_eigenvector, lanczos_result_t *out_result);

/* <psi|H|psi> / <psi|psi> of the materialised state on the kagome cluster
 * (N <= 24).  Replaces src/nqs/nqs_lanczos.c:434-451 and :453-473. */
int nqs_exact_energy_kagome_heisenberg(nqs_ansatz_t *a, int Lx_cells, int Ly_cells,
                                       double J, int pbc, double *out_energy);
int nqs_exact_energy_kagome_heisenberg_with_cb(nqs_log_amp_fn_t log_amp, void *user,
                                               int Lx_cells, int Ly_cells,
                                               double J, int pbc, double *out_energy);

/* Kagome refiners seeded from the ansatz / callback (:477-530). */
int nqs_lanczos_refine_kagome_heisenberg(nqs_ansatz_t *a, int Lx_cells, int Ly_cells, double J,
                                         int pbc, int max_iters, double tol,
                                         double *out_eigenvalue, double *out_eigenvector,
                                         lanczos_result_t *out_result);
int nqs_lanczos_refine_kagome_heisenberg_with_cb(nqs_log_amp_fn_t log_amp, void *user,
                                                 int Lx_cells, int Ly_cells, double J, int pbc,
                                                 int max_iters, double tol,
                                                 double *out_eigenvalue, double *out_eigenvector,
                                                 lanczos_result_t *out_result);

/* The k smallest Ritz values, seeded from the ansatz / callback
 * (:532-577). */
int nqs_lanczos_k_lowest_kagome_heisenberg(nqs_ansatz_t *a, int Lx_cells, int Ly_cells, double J,
                                           int pbc, int max_iters, int k,
                                           double *out_eigenvalues, lanczos_result_t *out_result);
int nqs_lanczos_k_lowest_kagome_heisenberg_with_cb(nqs_log_amp_fn_t log_amp, void *user,
                                                   int Lx_cells, int Ly_cells, double J, int pbc,
                                                   int max_iters, int k, double *out_eigenvalues,
                                                   lanczos_result_t *out_result);

/* Sector-projected full-reorthogonalisation drivers (:593-627, :674-704):
 * the seed is not projected; the deterministic start vector (when there is
 * no seed) and every new Krylov vector are. */
int nqs_lanczos_k_lowest_kagome_heisenberg_projected(
    nqs_log_amp_fn_t log_amp, void *user, int Lx_cells, int Ly_cells, double J, int pbc,
    const int *perm, const double *characters, int G, int max_iters, int k,
    double *out_eigenvalues, lanczos_result_t *out_result);
int nqs_lanczos_refine_kagome_heisenberg_projected(
    nqs_log_amp_fn_t log_amp, void *user, int Lx_cells, int Ly_cells, double J, int pbc,
    const int *perm, const double *characters, int G, int max_iters, double tol,
    double *out_eigenvalue, double *out_eigenvector, lanczos_result_t *out_result);

/* Lowest eigenvalue in the sector of P by the memory-lean three-term
 * Lanczos with P applied to every new Krylov vector (:629-647,
 * src/mps/lanczos.c:824-970): the reference's deterministic xorshift start
 * vector, stop when the smallest Ritz value moves by less than tol after
 * more than 5 steps or when beta < 1e-14.  N <= 30.  0 on success, -1 on
 * bad arguments. */
int nqs_lanczos_e0_kagome_heisenberg_projected_lean(int Lx_cells, int Ly_cells, double J,
                                                    int pbc, const int *perm,
                                                    const double *characters, int G,
                                                    int max_iters, double tol,
                                                    double *out_eigenvalue,
                                                    lanczos_result_t *out_result);

/* The same plus its unit eigenvector (2^N doubles) by a replay of the
 * recurrence (:651-672, src/mps/lanczos.c:972-1130); the eigenvalue is
 * re-taken from the final tridiagonal and converged = (K >= 6 and the last
 * Ritz change < tol). */
int nqs_lanczos_e0_kagome_heisenberg_projected_lean_eigvec(
    int Lx_cells, int Ly_cells, double J, int pbc, const int *perm, const double *characters,
    int G, int max_iters, double tol, double *out_eigenvalue, double *out_eigenvector,
    lanczos_result_t *out_result);

#ifdef __cplusplus
}
#endif

#endif /* NQS_LANCZOS_H */
