/*
 * oracle/nqs_oracle.h — CPU oracle for the NQS variational Monte Carlo path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C FP64 restatement of the
 * reference algorithm (tsotchke/spin_based_neural_network, the src/nqs sources),
 * written from the reference's behaviour, used as the checker by tests/,
 * by __graft_entry__.smoke() and by bench.py's cpu_baseline leg.  Nothing
 * in the product (libnqs_hip.so) links, loads or calls it.
 *
 * Every entry point mirrors a reference function (orc_X <-> nqs_X) with the
 * same arguments and error behaviour, so an oracle run and a product run
 * given the same config and seed walk the same chain.  Extensions defined
 * by include/nqs_hip/nqs_hip.h (chains, exchange moves, Sz=0 start,
 * periodic square lattices) are restated here with the semantics the
 * device path must reproduce.
 *
 * Parity pins (what checks each function; see tests/):
 *   ansatz init / log_amp / gradients : golden vectors from the compiled
 *       reference (tests/golden), finite differences, Re(holomorphic) ==
 *       real-projected identities
 *   sampler (single chain)            : golden sample sequences
 *   sampler extensions                : Sz conservation, exact stationary
 *       distribution by enumeration (N <= 10), chain-0 == reference chain
 *   local energies                    : golden vectors, uniform / linear
 *       ansatz closed forms of the reference tests, ED consistency
 *       (<E_loc> over |psi|^2 == <psi|H|psi> by enumeration)
 *   SR / holomorphic SR / MinSR        : golden parameter updates and traces,
 *       dense (S + eps I)^-1 F, push-through identity MinSR == dense
 *   ED (Lanczos)                       : closed forms (TFIM free fermions),
 *       reference Lanczos anchors, dense diagonalisation for N <= 10
 */
#ifndef NQS_ORACLE_H
#define NQS_ORACLE_H

#include "../include/nqs/nqs_config.h"
#include "../include/nqs_hip/nqs_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef void (*orc_log_amp_fn)(const int *spins, int num_sites, void *user,
                               double *out_re, double *out_im);

/* ---- ansatz (src/nqs/nqs_ansatz.c) ---------------------------------- */
typedef struct orc_ansatz orc_ansatz_t;
orc_ansatz_t *orc_ansatz_create(const nqs_config_t *cfg, int num_sites);
void   orc_ansatz_free(orc_ansatz_t *a);
long   orc_ansatz_num_params(const orc_ansatz_t *a);
int    orc_ansatz_num_hidden(const orc_ansatz_t *a);
void   orc_ansatz_log_amp(const int *spins, int num_sites, void *user,
                          double *out_log_abs, double *out_arg);
int    orc_ansatz_logpsi_gradient(orc_ansatz_t *a, const int *spins,
                                  int num_sites, double *out);
int    orc_ansatz_logpsi_gradient_complex(orc_ansatz_t *a, const int *spins,
                                          int num_sites, double *re, double *im);
int    orc_ansatz_is_complex(const orc_ansatz_t *a);
int    orc_ansatz_apply_update(orc_ansatz_t *a, const double *delta, double step);
double *orc_ansatz_params_raw(orc_ansatz_t *a);

/* ---- Marshall (src/nqs/nqs_marshall.c) ------------------------------ */
typedef struct {
    orc_log_amp_fn base_log_amp;
    void *base_user;
    int size_x, size_y;
} orc_marshall_wrapper_t;
int  orc_marshall_parity(const int *spins, int size_x, int size_y);
void orc_marshall_log_amp(const int *spins, int num_sites, void *user,
                          double *out_log_abs, double *out_arg);

/* ---- sampler (src/nqs/nqs_sampler.c + extensions) ------------------- */
typedef struct orc_sampler orc_sampler_t;
orc_sampler_t *orc_sampler_create(int num_sites, const nqs_config_t *cfg,
                                  orc_log_amp_fn fn, void *user);
orc_sampler_t *orc_sampler_create_ex(int num_sites, const nqs_config_t *cfg,
                                     const nqs_hip_options_t *opt,
                                     orc_log_amp_fn fn, void *user);
void   orc_sampler_free(orc_sampler_t *s);
void   orc_sampler_thermalize(orc_sampler_t *s);
const int *orc_sampler_next(orc_sampler_t *s);
int    orc_sampler_batch(orc_sampler_t *s, int n, int *out);
double orc_sampler_acceptance_ratio(const orc_sampler_t *s);
const int *orc_sampler_current(const orc_sampler_t *s);
const int *orc_sampler_chain_state(const orc_sampler_t *s, int chain);
int    orc_sampler_num_sites(const orc_sampler_t *s);
unsigned long long orc_sampler_rng_state(const orc_sampler_t *s, int chain);
unsigned long long orc_chain_seed(unsigned rng_seed, int chain, int num_chains);

/* ---- local energy (src/nqs/nqs_gradient.c + periodic extension) ----- */
double orc_local_energy(const nqs_config_t *cfg, int size_x, int size_y,
                        const int *spins, orc_log_amp_fn fn, void *user);
void   orc_local_energy_batch(const nqs_config_t *cfg, int size_x, int size_y,
                              const int *spins, int B, orc_log_amp_fn fn,
                              void *user, double *out);
void   orc_local_energy_complex(const nqs_config_t *cfg, int size_x, int size_y,
                                const int *spins, orc_log_amp_fn fn, void *user,
                                double *re, double *im);
void   orc_local_energy_batch_complex(const nqs_config_t *cfg, int size_x,
                                      int size_y, const int *spins, int B,
                                      orc_log_amp_fn fn, void *user,
                                      double *re, double *im);
void   orc_local_energy_batch_ex(const nqs_config_t *cfg, int size_x, int size_y,
                                 int periodic, const int *spins, int B,
                                 orc_log_amp_fn fn, void *user,
                                 double *re, double *im);

/* ---- optimiser (src/nqs/nqs_optimizer.c) ----------------------------- */
typedef struct {
    double mean_energy, variance_energy, update_norm, acceptance_ratio;
    int cg_iterations, converged;
} orc_sr_info_t;
typedef int (*orc_gradient_fn)(void *u, orc_ansatz_t *a, const int *s, int n,
                               double *out);
typedef int (*orc_holo_gradient_fn)(void *u, orc_ansatz_t *a, const int *s,
                                    int n, double *re, double *im);

int orc_sr_step_custom_full(const nqs_config_t *cfg, int Lx, int Ly,
                            orc_ansatz_t *a, orc_sampler_t *s,
                            orc_log_amp_fn fn, void *user,
                            orc_gradient_fn gfn, void *guser, orc_sr_info_t *info);
int orc_sr_run_custom_full(const nqs_config_t *cfg, int Lx, int Ly,
                           orc_ansatz_t *a, orc_sampler_t *s,
                           orc_log_amp_fn fn, void *user,
                           orc_gradient_fn gfn, void *guser, double *trace);
int orc_sr_step_holomorphic_full(const nqs_config_t *cfg, int Lx, int Ly,
                                 orc_ansatz_t *a, orc_sampler_t *s,
                                 orc_log_amp_fn fn, void *user,
                                 orc_holo_gradient_fn gfn, void *guser,
                                 orc_sr_info_t *info);
int orc_sr_run_holomorphic_full(const nqs_config_t *cfg, int Lx, int Ly,
                                orc_ansatz_t *a, orc_sampler_t *s,
                                orc_log_amp_fn fn, void *user,
                                orc_holo_gradient_fn gfn, void *guser,
                                double *trace);
int orc_sr_step_minsr_full(const nqs_config_t *cfg, int Lx, int Ly,
                           orc_ansatz_t *a, orc_sampler_t *s,
                           orc_log_amp_fn fn, void *user,
                           orc_gradient_fn gfn, void *guser, orc_sr_info_t *info);
int orc_sr_run_minsr(const nqs_config_t *cfg, int Lx, int Ly, orc_ansatz_t *a,
                     orc_sampler_t *s, orc_log_amp_fn fn, void *user,
                     double *trace);

int orc_sr_step(const nqs_config_t *cfg, int Lx, int Ly, orc_ansatz_t *a,
                orc_sampler_t *s, orc_sr_info_t *info);
int orc_sr_run(const nqs_config_t *cfg, int Lx, int Ly, orc_ansatz_t *a,
               orc_sampler_t *s, double *trace);
int orc_sr_step_custom(const nqs_config_t *cfg, int Lx, int Ly, orc_ansatz_t *a,
                       orc_sampler_t *s, orc_log_amp_fn fn, void *user,
                       orc_sr_info_t *info);
int orc_sr_run_custom(const nqs_config_t *cfg, int Lx, int Ly, orc_ansatz_t *a,
                      orc_sampler_t *s, orc_log_amp_fn fn, void *user, double *trace);
int orc_sr_step_holomorphic(const nqs_config_t *cfg, int Lx, int Ly, orc_ansatz_t *a,
                            orc_sampler_t *s, orc_log_amp_fn fn, void *user,
                            orc_sr_info_t *info);
int orc_sr_run_holomorphic(const nqs_config_t *cfg, int Lx, int Ly, orc_ansatz_t *a,
                           orc_sampler_t *s, orc_log_amp_fn fn, void *user, double *trace);

/* Excited-state SR (orthogonality penalty) and real-time tVMC
 * (src/nqs/nqs_optimizer.c:562-993). */
int orc_sr_step_excited(const nqs_config_t *cfg, int Lx, int Ly, orc_ansatz_t *a,
                        orc_sampler_t *s, orc_log_amp_fn fn, void *user,
                        orc_log_amp_fn ref_fn, void *ref_user, double mu,
                        orc_sr_info_t *info);
int orc_sr_run_excited(const nqs_config_t *cfg, int Lx, int Ly, orc_ansatz_t *a,
                       orc_sampler_t *s, orc_log_amp_fn fn, void *user,
                       orc_log_amp_fn ref_fn, void *ref_user, double mu, double *trace);
int orc_tvmc_step_real_time(const nqs_config_t *cfg, double dt, int Lx, int Ly,
                            orc_ansatz_t *a, orc_sampler_t *s, orc_log_amp_fn fn,
                            void *user, orc_sr_info_t *info);
int orc_tvmc_step_heun(const nqs_config_t *cfg, double dt, int Lx, int Ly,
                       orc_ansatz_t *a, orc_sampler_t *s, orc_log_amp_fn fn,
                       void *user, orc_sr_info_t *info);

/* Solver building blocks on caller-provided matrices (row-major O[Ns][P]). */
int  orc_cg_solve(const double *O, long Ns, long P, const double *Omean,
                  double eps, const double *F, int max_iters, double tol,
                  double *delta, int *converged);
int  orc_cg_solve_complex(const double *R, const double *I, long Ns, long P,
                          const double *Rm, const double *Im, double eps,
                          const double *F, int max_iters, double tol,
                          double *delta, int *converged);
int  orc_cholesky_lower(double *A, int n);
void orc_cholesky_solve(const double *L, int n, const double *b, double *x);
/* Full MinSR solve: O (centred in place), energies -> delta; optional T out
 * (Ns x Ns, before factorisation, including the shift) and y out. */
int  orc_minsr_solve(double *O, long Ns, long P, const double *E, double eps,
                     double *delta, double *T_out, double *y_out);

/* ---- exact diagonalisation (own Lanczos) ----------------------------- */
/* Ground energy of the Hamiltonian cfg describes on the lattice (same bond
 * sets as the local energy).  n_down < 0: full 2^N space; otherwise the
 * sector with n_down down spins (pair-flip Hamiltonians only). */
double orc_ed_ground_energy(const nqs_config_t *cfg, int size_x, int size_y,
                            int periodic, int n_down, int max_iters, double tol);
/* <psi|H|psi>/<psi|psi> by enumeration of all 2^N configurations
 * (psi from the callback, complex). N <= 20. */
double orc_exact_energy(const nqs_config_t *cfg, int size_x, int size_y,
                        int periodic, orc_log_amp_fn fn, void *user);

/* ---- symmetry projection (src/nqs/nqs_symproj.c, nqs_translation.c) -- */
typedef struct {                 /* layout of nqs_symproj_wrapper_t */
    orc_log_amp_fn base_log_amp;
    void *base_user;
    int num_sites, num_group_elements;
    const int *perm;
    const double *characters;
} orc_symproj_wrapper_t;
typedef struct {                 /* layout of nqs_translation_wrapper_t */
    orc_log_amp_fn base_log_amp;
    void *base_user;
    int size_x, size_y;
} orc_translation_wrapper_t;
void orc_symproj_log_amp(const int *spins, int num_sites, void *user,
                         double *out_log_abs, double *out_arg);
int  orc_symproj_gradient(void *grad_user, orc_ansatz_t *a, const int *spins,
                          int num_sites, double *out_grad);
int  orc_symproj_gradient_complex(void *grad_user, orc_ansatz_t *a, const int *spins,
                                  int num_sites, double *out_re, double *out_im);
void orc_translation_log_amp(const int *spins, int num_sites, void *user,
                             double *out_log_abs, double *out_arg);
int  orc_translation_gradient(void *grad_user, orc_ansatz_t *a, const int *spins,
                              int num_sites, double *out_grad);

/* ---- sample diagnostics (src/nqs/nqs_diagnostics.c) ------------------ */
int orc_compute_chi_F(const nqs_config_t *cfg, int size_x, int size_y,
                      orc_ansatz_t *a, orc_sampler_t *s,
                      double *out_trace_S, double *out_per_param);
int orc_compute_kagome_bond_phase(const nqs_config_t *cfg, int Lx_cells, int Ly_cells,
                                  orc_ansatz_t *a, orc_sampler_t *s,
                                  double *out_mean_re, double *out_mean_im,
                                  long *out_counts);

/* ---- statistics ----------------------------------------------------- */
double orc_tau_int(const double *x, int n);

/* ---- learned proposal network (ar_oracle.c) ----------------------------- */
long orc_ar_num_params(int N);
void orc_ar_init(int N, unsigned long long seed, double scale, float *w);
int  orc_ar_log_prob(int N, const float *w, const int *spins, int B, int constrained,
                     int quantize, double *out);
int  orc_ar_gradient(int N, const float *w, const int *spins, int B, int constrained,
                     int quantize, double *grad, double *mean_nll);
void orc_ar_adam_step(int N, float *w, double *m, double *v, const double *grad, long t,
                      double lr);

#ifdef __cplusplus
}
#endif

#endif /* NQS_ORACLE_H */
