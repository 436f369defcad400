/*
 * nqs/nqs_config.h — run configuration for the NQS variational Monte Carlo
 * path, exported by libnqs_hip.so.
 *
 * Drop-in for the reference header include/nqs/nqs_config.h:
 *   enums          include/nqs/nqs_config.h:25-79
 *   nqs_config_t   include/nqs/nqs_config.h:82-120
 *   defaults       include/nqs/nqs_config.h:123-152
 * The struct is the ABI every entry point takes, so member order, types and
 * enum values are identical to the reference; the defaults are the same
 * numbers.  Engine-only knobs (chain count, proposal kind, boundary,
 * precision) live in nqs_hip/nqs_hip.h, never in this struct.
 */
#ifndef NQS_CONFIG_H
#define NQS_CONFIG_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    NQS_ANSATZ_LEGACY_MLP           = 0, /* mean-field: log psi = sum_i theta_i s_i */
    NQS_ANSATZ_VIT                  = 1, /* not provided: create returns NULL      */
    NQS_ANSATZ_FACTORED_VIT         = 2, /* not provided: create returns NULL      */
    NQS_ANSATZ_AUTOREGRESSIVE       = 3, /* not provided: create returns NULL      */
    NQS_ANSATZ_KAN                  = 4, /* not provided: create returns NULL      */
    NQS_ANSATZ_RBM                  = 5, /* real RBM, N + M + M*N parameters       */
    NQS_ANSATZ_COMPLEX_RBM          = 6, /* complex RBM, 2*(N + M + M*N)           */
    NQS_ANSATZ_FACTORED_VIT_COMPLEX = 7  /* not provided: create returns NULL      */
} nqs_ansatz_kind_t;

typedef enum {
    NQS_SYM_NONE        = 0,
    NQS_SYM_TRANSLATION = 1 << 0,
    NQS_SYM_SPIN_FLIP   = 1 << 1,
    NQS_SYM_U1          = 1 << 2,
    NQS_SYM_POINT_GROUP = 1 << 3,
    NQS_SYM_SU2         = 1 << 4
} nqs_symmetry_flags_t;

typedef enum {
    NQS_HAM_TFIM              = 0, /* -J sum zz - Gamma sum x                    */
    NQS_HAM_HEISENBERG        = 1, /* J sum S.S                                  */
    NQS_HAM_J1_J2             = 2, /* J1 nearest + J2 both diagonals             */
    NQS_HAM_KITAEV_HONEYCOMB  = 3, /* not on the device path                     */
    NQS_HAM_XXZ               = 4, /* Jxy (x,y) + Jz zz                          */
    NQS_HAM_KITAEV_HEISENBERG = 5, /* not on the device path                     */
    NQS_HAM_KAGOME_HEISENBERG = 6  /* J sum S.S on kagome, 3 sites per cell      */
} nqs_hamiltonian_kind_t;

typedef struct {
    /* ansatz */
    nqs_ansatz_kind_t    ansatz;
    int                  patch_size;
    int                  depth;
    int                  width;
    int                  heads;
    int                  rbm_hidden_units;  /* M; 0 selects 2*N                 */
    double               rbm_init_scale;
    nqs_symmetry_flags_t symmetries;

    /* Hamiltonian */
    nqs_hamiltonian_kind_t hamiltonian;
    double               j_coupling;
    double               j2_coupling;
    double               j_z_coupling;
    double               transverse_field;
    double               kh_K;
    double               kh_J;
    int                  kagome_pbc;

    /* sampling */
    int                  num_samples;
    int                  num_thermalize;
    int                  num_decorrelate;
    int                  cluster_moves;
    unsigned             rng_seed;

    /* optimisation */
    int                  num_iterations;
    double               learning_rate;
    double               sr_diag_shift;
    int                  sr_cg_max_iters;
    double               sr_cg_tol;
} nqs_config_t;

/* Same default values as the reference (6x6 TFIM, 1024 samples, lr 1e-2,
 * shift 1e-3, CG 100 iterations / tol 1e-6, seed 0xC0FFEE). */
static inline nqs_config_t nqs_config_defaults(void) {
    nqs_config_t c;
    c.ansatz           = NQS_ANSATZ_LEGACY_MLP;
    c.patch_size       = 1;
    c.depth            = 3;
    c.width            = 128;
    c.heads            = 4;
    c.rbm_hidden_units = 0;
    c.rbm_init_scale   = 0.01;
    c.symmetries       = (nqs_symmetry_flags_t)(NQS_SYM_SPIN_FLIP | NQS_SYM_TRANSLATION);
    c.hamiltonian      = NQS_HAM_TFIM;
    c.j_coupling       = 1.0;
    c.j2_coupling      = 0.0;
    c.j_z_coupling     = 1.0;
    c.transverse_field = 1.0;
    c.kh_K             = 1.0;
    c.kh_J             = 0.0;
    c.kagome_pbc       = 1;
    c.num_samples      = 1024;
    c.num_thermalize   = 256;
    c.num_decorrelate  = 4;
    c.cluster_moves    = 0;
    c.rng_seed         = 0xC0FFEEu;
    c.num_iterations   = 200;
    c.learning_rate    = 1e-2;
    c.sr_diag_shift    = 1e-3;
    c.sr_cg_max_iters  = 100;
    c.sr_cg_tol        = 1e-6;
    return c;
}

#ifdef __cplusplus
}
#endif

#endif /* NQS_CONFIG_H */
