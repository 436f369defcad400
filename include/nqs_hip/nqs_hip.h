/*
 * nqs_hip/nqs_hip.h — engine extensions of libnqs_hip.so that have no
 * counterpart in the reference API (include/nqs has no chain count, no
 * move kind, no periodic square lattice and no precision switch; see
 * include/nqs/nqs_config.h:82-120).
 *
 * A sampler created with nqs_hip_sampler_create_ex carries these options;
 * every reference entry point that takes that sampler (nqs_sampler_*,
 * nqs_sr_step*, nqs_sr_run*) honours them.  nqs_sampler_create uses the
 * defaults, which reproduce the reference exactly (one chain, single-site
 * flips, uniform random start, open square lattices, FP64), except that the
 * environment variable NQS_HIP_CHAINS=<n> selects n chains.
 *
 * Multi-chain semantics (also restated by the CPU oracle):
 *   chain 0 uses the reference seed (rng_seed, or 0x9E3779B97F4A7C15 when
 *   zero); chain c > 0 uses splitmix64(seed0 ^ (c * 0xD1B54A32D192ED03)),
 *   replaced by 1 if zero.  nqs_sampler_batch(s, n, out) runs rounds; in
 *   each round every chain makes max(num_decorrelate, 1) moves, then the
 *   configuration of chain c is stored as sample r*num_chains + c when that
 *   index is below n.  thermalize makes num_thermalize moves per chain.
 *   The acceptance ratio counts all chains.
 */
#ifndef NQS_HIP_H
#define NQS_HIP_H

#include "../nqs/nqs_config.h"
#include "../nqs/nqs_ansatz.h"
#include "../nqs/nqs_sampler.h"
#include "../nqs/nqs_optimizer.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    NQS_HIP_MOVE_FLIP     = 0, /* reference single-site flip                  */
    NQS_HIP_MOVE_EXCHANGE = 1, /* swap a random up site with a random down    */
    NQS_HIP_MOVE_LEARNED  = 2  /* autoregressive proposal, MH-corrected       */
} nqs_hip_move_t;

typedef enum {
    NQS_HIP_FP64 = 0,          /* all arithmetic FP64 (parity mode)           */
    NQS_HIP_FP32 = 1,          /* amplitudes, E_loc, moves and O in FP32;
                                  reductions, Gram, solve and update FP64     */
    NQS_HIP_MIXED = 2          /* as FP32, with theta = S W^T of the O rows
                                  on BF16 MFMA (spins exact, weights split
                                  in three BF16 parts) and the dense MinSR
                                  Gram (O_c^T O_c or O_c O_c^T) on FP16
                                  MFMA: entries scaled by a power of two and
                                  split into hi + lo halves, products as
                                  hi hi + hi lo + lo hi (about 3 * 2^-22 per
                                  product), FP32 accumulation over at most
                                  2048 terms, FP64 across them; the shifted
                                  solve, reductions and delta stay FP64     */
} nqs_hip_precision_t;

typedef enum {
    NQS_HIP_GRAM_DIRECT   = 0, /* T = O_c O_c^T over the materialised O       */
    NQS_HIP_GRAM_FACTORED = 1, /* RBM / complex RBM: T from (S S^T, Th Th^T)  */
    NQS_HIP_GRAM_AUTO     = 2, /* the cheaper of the sample-space solve
                                  (factored Gram where the ansatz allows it)
                                  and the parameter-space solve
                                  (O_c^T O_c / Ns + eps I) delta = O_c^T eps / Ns,
                                  which gives the same delta (push-through
                                  identity), by operation count              */
    NQS_HIP_GRAM_PARAMETER = 3 /* always the parameter-space solve           */
} nqs_hip_gram_t;

typedef struct {
    int num_chains;      /* >= 1                                               */
    int move;            /* nqs_hip_move_t                                     */
    int sz_zero_start;   /* 1: each chain starts at a random Sz=0 state        */
    int periodic;        /* 1: periodic square lattice for TFIM, Heisenberg,
                            XXZ, J1-J2 (dimensions of length >= 3 wrap)        */
    int precision;       /* nqs_hip_precision_t                                */
    int gram;            /* nqs_hip_gram_t (MinSR)                             */
    /* learned proposal (move == NQS_HIP_MOVE_LEARNED) */
    int ar_hidden;       /* hidden width of the masked MLP (64; 0 = 64)        */
    int ar_train_every;  /* retrain every k batches (default 10)               */
    int ar_train_window; /* samples kept for training (default 16384)          */
    int ar_epochs;       /* passes over the window per retrain (default 2)     */
    int ar_batch;        /* minibatch size (default 2048)                      */
    double ar_lr;        /* Adam learning rate (default 3e-3)                  */
    int ar_local;        /* exchange moves before each learned move (0)        */
    unsigned ar_seed;    /* network initialisation seed (0: chain 0's seed)    */
} nqs_hip_options_t;

/* Defaults: one chain, flips, random start, open boundaries, FP64, direct
 * Gram, learned-proposal defaults as listed above.
 *
 * Learned proposal (move = NQS_HIP_MOVE_LEARNED, device amplitudes only,
 * N <= 64).  A masked autoregressive network q (N -> 64 -> 64 -> N, ReLU,
 * BF16 weights and activations, FP32 accumulation) proposes whole
 * configurations in the chain's own magnetisation sector; each proposal is
 * accepted with min(1, |psi(s')|^2 q(s) / (|psi(s)|^2 q(s'))), so the chains
 * sample |psi|^2 exactly whatever the network has learned.  Thermalisation
 * and nqs_sampler_next use exchange moves.  Every recorded batch
 * (nqs_sampler_batch, the sampling stage of an SR step) appends its samples
 * to a window of ar_train_window configurations; before a batch the network
 * is trained by maximum likelihood on the window (ar_epochs passes, Adam,
 * minibatch ar_batch, rate ar_lr): first as soon as the window holds
 * samples, then every ar_train_every batches.  A batch before the first
 * training uses exchange moves (num_decorrelate per round); afterwards each
 * round is ar_local exchange moves followed by one learned move.  A due
 * retraining starts right after the batch before it, on a second stream,
 * so it runs beside the caller's work between batches (the SR solve) with
 * the same window and the same result; the next use of the network waits
 * for it.  After a sampler's last batch it may thus hold one training more
 * than batches used.  NQS_HIP_AR_OVERLAP=0 trains at the batch start. */
nqs_hip_options_t nqs_hip_options_defaults(void);

/* nqs_sampler_create with options.  NULL on bad args or unsupported
 * combinations (message on stderr). */
nqs_sampler_t *nqs_hip_sampler_create_ex(int num_sites, const nqs_config_t *cfg,
                                         const nqs_hip_options_t *opt,
                                         nqs_log_amp_fn_t log_amp,
                                         void *log_amp_user);

/* Copy of the options a sampler carries (0 on success). */
int nqs_hip_sampler_get_options(const nqs_sampler_t *s, nqs_hip_options_t *out);

/* Per-chain energy series of the last SR step: E_loc of sample r*C + c is
 * element [c][r] of out (num_chains x rounds, row-major).  Returns the
 * number of rounds, or -1.  Used for tau_int / ESS. */
int nqs_hip_sampler_last_energies(const nqs_sampler_t *s, double *out, int capacity);

/* Number of samples per chain in the last SR step (rounds). */
int nqs_hip_sampler_rounds(const nqs_sampler_t *s);

/* Local energies with options (periodic boundaries; FP32 arithmetic). */
int nqs_hip_local_energy_batch_ex(const nqs_config_t *cfg, int size_x, int size_y,
                                  const nqs_hip_options_t *opt,
                                  const int *spins_batch, int batch_size,
                                  nqs_log_amp_fn_t log_amp, void *log_amp_user,
                                  double *out_re, double *out_im);

/* Batched amplitudes: log|psi| and arg psi of B configurations. */
int nqs_hip_log_amp_batch(nqs_ansatz_t *a, const int *spins_batch, int batch_size,
                          double *out_log_abs, double *out_arg);

/* Batched O matrix, row-major [B][P]: mode 0 real-projected (out_re only),
 * mode 1 holomorphic (out_re, out_im). */
int nqs_hip_logpsi_gradient_batch(nqs_ansatz_t *a, const int *spins_batch,
                                  int batch_size, int mode,
                                  double *out_re, double *out_im);

/* Learned-proposal network on its own (the sampler's, or a standalone one
 * for tests and offline training).  Parameters are FP32, layout
 * [W1 (64 x N) | b1 (64) | W2 (64 x 64) | b2 (64) | W3 (N x 64) | b3 (N)],
 * with MADE masks (hidden unit k has degree 1 + floor(k (N-1) / 64)); masked
 * entries stay 0.
 * `constrained` restricts a configuration to the magnetisation of its
 * reference (sampling: ref_spins; evaluation and training: itself). */
typedef struct nqs_hip_ar nqs_hip_ar_t;
nqs_hip_ar_t *nqs_hip_ar_create(int num_sites, unsigned long long seed, double init_scale);
void nqs_hip_ar_free(nqs_hip_ar_t *ar);
long nqs_hip_ar_num_params(const nqs_hip_ar_t *ar);
int nqs_hip_ar_get_params(nqs_hip_ar_t *ar, float *out);
int nqs_hip_ar_set_params(nqs_hip_ar_t *ar, const float *in);
/* log q of B configurations [B][N] (the sampling walk in evaluation mode) */
int nqs_hip_ar_log_prob(nqs_hip_ar_t *ar, const int *spins, int batch, int constrained,
                        double *out_logq);
/* B draws; row b uses xorshift64 state seeds[b] (nonzero) */
int nqs_hip_ar_sample(nqs_hip_ar_t *ar, const int *ref_spins, int batch, int constrained,
                      const unsigned long long *seeds, int *out_spins, double *out_logq);
/* gradient of the mean negative log-likelihood over B configurations */
int nqs_hip_ar_gradient(nqs_hip_ar_t *ar, const int *spins, int batch, int constrained,
                        float *grad, double *mean_nll);
/* Adam training on n configurations; mean NLL of the last epoch */
int nqs_hip_ar_train(nqs_hip_ar_t *ar, const int *spins, int n, int constrained, int epochs,
                     int minibatch, double lr, double *mean_nll);
/* the network of a learned-proposal sampler (NULL otherwise) */
nqs_hip_ar_t *nqs_hip_sampler_ar(nqs_sampler_t *s);
/* learned moves: acceptance ratio, proposals made, trainings, last mean NLL */
int nqs_hip_sampler_learned_stats(const nqs_sampler_t *s, double *acceptance,
                                  long long *proposals, int *trainings, double *last_nll);

/* Batched independent problems.  n (ansatz, sampler) pairs, each as
 * nqs_sr_step(&cfgs[i], size_x[i], size_y[i], ansatz[i], samplers[i], ...)
 * would step it (CG-SR, src/nqs/nqs_optimizer.c:255-262), stepped together:
 * every problem's sampling and local energies, then one launch sequence for
 * all problems' force, CG iterations and updates; from the second step on,
 * the whole step is one replayed hipGraph.  Requirements per problem: a real
 * RBM ansatz (NQS_ANSATZ_RBM) and a device sampler of that same ansatz
 * (nqs_ansatz_log_amp, FP64, flip or exchange moves); each ansatz and
 * sampler appears once.  The batch owns its problems' device state from
 * create to free: parameters are updated on the device, and
 * nqs_hip_multi_sync copies them (and each sampler's current
 * configuration) back to the host.  CG results equal nqs_sr_step's up to
 * the order of floating-point sums (the QGT products run through the RBM
 * factors instead of an explicit O matrix).
 * run: thermalizes each sampler once, then num_iterations steps;
 * out_traces [n][num_iterations] receives the mean energies (may be NULL).
 * tuning: the split-factor targets timed at the first step and their CG
 * product times (microseconds), then (target 0) the persistent launch's time
 * for the same product at the chosen split; returns the chosen target.
 * use_graph(m, 0): later steps are issued launch by launch instead of as
 * the captured graph (same kernels and results; for per-kernel timing).
 * use_persistent(m, on): the CG iterations of all problems in one
 * cooperative launch with grid barriers between phases (on), or one launch
 * per phase and iteration (off); same results bit for bit.  Without this
 * call or NQS_HIP_MULTI_PERSIST=0/1 at creation the first step's autotuning
 * times both and keeps the faster (on the MI355X the launch sequence: the
 * grid barriers cost more than the captured graph's launch gaps).  A device
 * that refuses the cooperative launch gets the launch sequence.
 * last_energies: the local energies of problem i's last batch, sample
 * r*C + c in slot r*C + c (chain c, round r); returns the count. */
typedef struct nqs_hip_multi nqs_hip_multi_t;
nqs_hip_multi_t *nqs_hip_multi_create(int n, const nqs_config_t *cfgs, const int *size_x,
                                      const int *size_y, nqs_ansatz_t *const *ansatz,
                                      nqs_sampler_t *const *samplers);
int nqs_hip_multi_step(nqs_hip_multi_t *m, nqs_sr_step_info_t *out_infos);
int nqs_hip_multi_run(nqs_hip_multi_t *m, int num_iterations, double *out_traces);
int nqs_hip_multi_sync(nqs_hip_multi_t *m);
int nqs_hip_multi_tuning(const nqs_hip_multi_t *m, int *targets, double *matvec_us, int capacity);
int nqs_hip_multi_use_graph(nqs_hip_multi_t *m, int on);
int nqs_hip_multi_use_persistent(nqs_hip_multi_t *m, int on);
/* device time of the last step's Metropolis launches (all problems), ms */
int nqs_hip_multi_last_sample_ms(const nqs_hip_multi_t *m, double *ms);
int nqs_hip_multi_last_energies(const nqs_hip_multi_t *m, int i, double *out, int capacity);
void nqs_hip_multi_free(nqs_hip_multi_t *m);

/* Integrated autocorrelation time of a series (Sokal window, c = 5). */
double nqs_hip_tau_int(const double *x, int n);

/* Stage timings of the last SR step on the sampler's stream, milliseconds:
 * [0] sampling, [1] local energy (with its forward pass), [2] O_k, means and
 * force, [3] solve, [4] update, [5] unused, [6] total.  0 when available.
 * A CG-SR step run as a batch of one problem reports [0] and [6] only. */
int nqs_hip_last_stage_ms(const nqs_sampler_t *s, double out[7]);

/* MinSR building blocks on host arrays (both run on the device).
 * Gram: T = O O^T * scale for row-major O [B][P], full symmetric T [B][B].
 * Cholesky solve: factor the symmetric positive T [n][n] in place (lower
 * triangle holds L, as src/nqs/nqs_optimizer.c:1026-1043 leaves it, upper
 * untouched) and solve T x = b.  On a non-positive pivot it retries once
 * with +1e-6 on the diagonal of the partially factored matrix, like
 * nqs_optimizer.c:1156-1160; returns 1 on success, 0 when both attempts
 * fail (x = 0), -1 on errors. */
int nqs_hip_gram(const double *O, int B, long P, double scale, double *T);
int nqs_hip_cholesky_solve(double *T, int n, const double *b, double *x);

/* nqs_hip_cholesky_solve on device memory, allocating nothing: no
 * counterpart in the reference; for callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid hipMalloc).  It waits
 * for its own work before returning (the failure flag decides the retry), so
 * it is not itself a graph-capturable call.  d_T [n][n] (row-major, lower
 * triangle read; factored in place as nqs_hip_cholesky_solve leaves it),
 * d_b [n] and d_x [n] are device pointers the caller owns (d_b and d_x
 * distinct).  Every work buffer lives
 * in one static device array of 64 MiB that the library holds on each GPU it
 * runs on until the process ends: n <= 2319 fits, larger n is refused with a
 * message (-1).  Calls that overlap on one GPU share that array and race.
 * Returns 1 on success, 0 when both factorisation attempts fail (x = 0),
 * -1 on errors. */
int nqs_hip_cholesky_solve_noalloc(double *d_T, int n, const double *d_b, double *d_x);

/* The MinSR update of nqs_sr_step_minsr_full (src/nqs/nqs_optimizer.c:
 * 1119-1177) for caller-given samples spins [B][N] and local energies
 * e_loc [B], with the ansatz's current parameters (real O for a real RBM,
 * the real-projected O for a complex RBM, as the reference's MinSR uses):
 * eps_i = E_i - mean E, O_c = O - mean O, and
 *   gram = NQS_HIP_GRAM_DIRECT:   (O_c O_c^T / B + shift I) y = eps over
 *                                 the materialised O, delta = O_c^T y / B;
 *   gram = NQS_HIP_GRAM_FACTORED: the same system with the Gram formed from
 *                                 the RBM factors (RBM / complex RBM only);
 *   gram = NQS_HIP_GRAM_PARAMETER: (O_c^T O_c / B + shift I) delta =
 *                                 O_c^T eps / B (same delta by the
 *                                 push-through identity).
 * Failure rule as the reference: one retry with +1e-6 on the diagonal of
 * the partially factored matrix, then delta = 0.  out_delta [P] receives
 * delta (the parameters are not changed).  Sample space only: out_y [B]
 * (may be NULL) receives y, and out_T [B][B] (may be NULL) the shifted
 * Gram before factorisation, its lower triangle (j <= i) defined.
 * Returns 1 when a factorisation succeeded, 0 when both attempts failed,
 * -1 on errors. */
int nqs_hip_minsr_solve(nqs_ansatz_t *a, const int *spins, int B, const double *e_loc,
                        double shift, int gram, double *out_delta, double *out_y,
                        double *out_T);

/* nqs_hip_minsr_solve with a sampler's options: opt->gram selects the Gram
 * as above, opt->precision the arithmetic (NQS_HIP_FP32: O from FP32
 * amplitudes; NQS_HIP_MIXED: additionally the dense Gram on the split
 * 16-bit MFMA).  Same returns. */
int nqs_hip_minsr_solve_ex(nqs_ansatz_t *a, const int *spins, int B, const double *e_loc,
                           double shift, const nqs_hip_options_t *opt, double *out_delta,
                           double *out_y, double *out_T);

/* Live kernel timing with HIP events on the library's stream.
 * nqs_hip_kernel_timing("eloc,gemm") starts timing the listed kernel tags
 * (forward, gen_O, mh, eloc, colsum, gemm, gemm_potrf, potrf_diag, trsm;
 * NULL or "" stops) and resets the counts; nqs_hip_kernel_stats returns the
 * number of timed launches of a tag and their total duration. */
int nqs_hip_kernel_timing(const char *names);
int nqs_hip_kernel_stats(const char *name, long long *count, double *total_ms);

/* HBM probe: streaming copy of `bytes` bytes between two device buffers,
 * `reps` times; average milliseconds per copy (the reference point for
 * hardware byte counters), < 0 on error. */
double nqs_hip_copy_probe(long bytes, int reps);

/* Diagnostics of the device path: how many kernel launches the library has
 * made, and how many host-callback adaptor calls it has made (a callback
 * that is not nqs_ansatz_log_amp or nqs_marshall_log_amp over it runs on
 * the host). */
long long nqs_hip_device_launches(void);
long long nqs_hip_host_callback_calls(void);

/* ABI layout check: writes sizeof(nqs_config_t), the offset of each of its
 * members in declaration order, sizeof(nqs_sr_step_info_t),
 * sizeof(nqs_marshall_wrapper_t) and sizeof(nqs_hip_options_t); returns the
 * count written or -1 when capacity is too small. */
int nqs_hip_config_layout(long *out, int capacity);

/* Library version string. */
const char *nqs_hip_version(void);

#ifdef __cplusplus
}
#endif

#endif /* NQS_HIP_H */
