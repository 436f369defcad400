/*
 * nqs/nqs_gradient.h — local energy E_loc(s) = sum_s' H_ss' psi(s')/psi(s),
 * exported by libnqs_hip.so.
 *
 * Replaces include/nqs/nqs_gradient.h (:39-101) / src/nqs/nqs_gradient.c.
 * Lattices and bond order follow the reference: square lattices index
 * x*size_y + y with open boundaries (nqs_gradient.c:50-205); kagome uses
 * size_x x size_y unit cells, site 3*(cx*Ly+cy)+sub, up triangle plus the
 * down triangle anchored at A, periodic when cfg->kagome_pbc
 * (nqs_gradient.c:374-445).  On the device path each connected
 * configuration is priced from the sample's hidden activations by an
 * incremental update instead of a full re-evaluation.
 */
#ifndef NQS_GRADIENT_H
#define NQS_GRADIENT_H

#include "nqs_config.h"
#include "nqs_sampler.h"

#ifdef __cplusplus
extern "C" {
#endif

/* nqs_gradient.h:39-43 / nqs_gradient.c:447-491.  0.0 on bad args. */
double nqs_local_energy(const nqs_config_t *cfg, int size_x, int size_y,
                        const int *spins, nqs_log_amp_fn_t log_amp,
                        void *log_amp_user);

/* nqs_gradient.h:46-52 / nqs_gradient.c:894-909 */
void nqs_local_energy_batch(const nqs_config_t *cfg, int size_x, int size_y,
                            const int *spins_batch, int batch_size,
                            nqs_log_amp_fn_t log_amp, void *log_amp_user,
                            double *out_energies);

/* nqs_gradient.h:60-65 / nqs_gradient.c:807-875 */
void nqs_local_energy_complex(const nqs_config_t *cfg, int size_x, int size_y,
                              const int *spins, nqs_log_amp_fn_t log_amp,
                              void *log_amp_user, double *out_re,
                              double *out_im);

/* nqs_gradient.h:66-71 / nqs_gradient.c:877-892 */
void nqs_local_energy_batch_complex(const nqs_config_t *cfg, int size_x,
                                    int size_y, const int *spins_batch,
                                    int batch_size, nqs_log_amp_fn_t log_amp,
                                    void *log_amp_user, double *out_re,
                                    double *out_im);

/* nqs_gradient.h:74-101: running mean / population variance, host inline. */
typedef struct {
    double sum;
    double sum_sq;
    long   count;
} nqs_energy_accumulator_t;

static inline void nqs_energy_accumulator_init(nqs_energy_accumulator_t *a) {
    a->sum = 0.0;
    a->sum_sq = 0.0;
    a->count = 0;
}

static inline void nqs_energy_accumulator_add(nqs_energy_accumulator_t *a,
                                              double sample) {
    a->sum += sample;
    a->sum_sq += sample * sample;
    a->count += 1;
}

static inline double nqs_energy_accumulator_mean(const nqs_energy_accumulator_t *a) {
    return a->count > 0 ? a->sum / (double)a->count : 0.0;
}

static inline double nqs_energy_accumulator_variance(const nqs_energy_accumulator_t *a) {
    if (a->count < 2) return 0.0;
    double m = nqs_energy_accumulator_mean(a);
    return a->sum_sq / (double)a->count - m * m;
}

#ifdef __cplusplus
}
#endif

#endif /* NQS_GRADIENT_H */
