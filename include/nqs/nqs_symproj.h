/*
 * nqs/nqs_symproj.h — finite-group symmetry projection of an amplitude, the
 * reference's include/nqs/nqs_symproj.h interface (same types, names,
 * arguments and return codes; src/nqs/nqs_symproj.c).
 *
 *     psi_sym(s) = (1/sqrt|G|) sum_g chi(g) psi_base(g.s),   (g.s)_i = s_{perm[g][i]}
 *
 * nqs_symproj_log_amp keeps the reference's real combination: the orbit
 * members enter as chi(g) |psi_base| cos(arg psi_base), so arg psi_sym is 0
 * or pi; the gradients weight the orbit members' base gradients by
 * w_g = chi(g) psi_base(g.s) / sum_g' chi(g') psi_base(g'.s) (real weights
 * from the cosines for nqs_symproj_gradient, complex weights for
 * nqs_symproj_gradient_complex).
 *
 * When the base amplitude is this library's ansatz (nqs_ansatz_log_amp, or
 * nqs_marshall_log_amp over it) all |G| orbit members of a configuration
 * are evaluated in one device batch; samplers, local energies and the SR
 * steps that receive these callbacks evaluate whole batches of
 * configurations and their orbits on the device.
 */
#ifndef NQS_SYMPROJ_H
#define NQS_SYMPROJ_H

#include "nqs_sampler.h"
#include "nqs_ansatz.h"

#ifdef __cplusplus
extern "C" {
#endif

/* reference nqs_symproj.h:40-47 */
typedef struct {
    nqs_log_amp_fn_t base_log_amp;
    void            *base_user;
    int              num_sites;          /* N                                  */
    int              num_group_elements; /* |G|                                */
    const int       *perm;               /* [|G| * N], perm[g * N + i] = pi_g(i) */
    const double    *characters;         /* [|G|]                              */
} nqs_symproj_wrapper_t;

/* log|psi_sym|, arg psi_sym (0 or pi); both 0 on bad arguments, log|psi_sym|
 * = -1e300 when the sum vanishes (reference nqs_symproj.c:30-80). */
void nqs_symproj_log_amp(const int *spins, int num_sites, void *user,
                         double *out_log_abs, double *out_arg);

/* sum_g w_g d log psi_base(g.s) / d theta with the real weights; zero when
 * the weights' denominator vanishes (reference nqs_symproj.c:82-153). */
int nqs_symproj_gradient(void *grad_user, nqs_ansatz_t *ansatz,
                         const int *spins, int num_sites, double *out_grad);

/* holomorphic version with complex weights (reference nqs_symproj.c:155-263) */
int nqs_symproj_gradient_complex(void *grad_user, nqs_ansatz_t *ansatz,
                                 const int *spins, int num_sites,
                                 double *out_grad_re, double *out_grad_im);

/* kagome groups on an Lx x Ly (p2) or L x L torus (p3, p6, p6m), PBC; site
 * 3 (cx Ly + cy) + sub.  The tables are malloc'ed (caller frees); rows in
 * the reference's element order (nqs_symproj.c:269-785). */
int nqs_kagome_translation_perm(int Lx, int Ly, int **out_perm, int *out_num_elements);
int nqs_kagome_p2_perm(int Lx, int Ly, int **out_perm, double **out_characters,
                       int *out_num_elements);
int nqs_kagome_p3_perm(int L, int **out_perm, double **out_characters,
                       int *out_num_elements);
int nqs_kagome_p6_perm(int L, int **out_perm, double **out_characters,
                       int *out_num_elements);
int nqs_kagome_p6m_perm(int L, int **out_perm, double **out_characters,
                        int *out_num_elements);

/* C_6v irreps at Gamma (reference nqs_symproj.h:178-185); the 2D irreps'
 * characters are returned multiplied by d_E = 2 */
typedef enum {
    NQS_SYMPROJ_KAGOME_GAMMA_A1 = 0,
    NQS_SYMPROJ_KAGOME_GAMMA_A2 = 1,
    NQS_SYMPROJ_KAGOME_GAMMA_B1 = 2,
    NQS_SYMPROJ_KAGOME_GAMMA_B2 = 3,
    NQS_SYMPROJ_KAGOME_GAMMA_E1 = 4,
    NQS_SYMPROJ_KAGOME_GAMMA_E2 = 5
} nqs_symproj_kagome_irrep_t;

int nqs_kagome_p6m_perm_irrep(int L, nqs_symproj_kagome_irrep_t irrep_name,
                              int **out_perm, double **out_characters,
                              int *out_num_elements);

/* psi <- (1/G) sum_g chi(g) T(g) psi on a basis-indexed vector of length
 * 2^N (N <= 30), (T(g) psi)(s) = psi(s_g) with bit i of s_g equal to bit
 * perm[g][i] of s; computed on the device (reference nqs_symproj.c:786-851). */
void nqs_kagome_p6m_project_inplace(double *psi, int N, int G, const int *perm,
                                    const double *characters);

#ifdef __cplusplus
}
#endif

#endif /* NQS_SYMPROJ_H */
