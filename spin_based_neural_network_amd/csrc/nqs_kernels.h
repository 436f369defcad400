// nqs_kernels.h — launch interface between the host engine (engine.cpp) and
// the gfx950 kernels (k_*.hip) of libnqs_hip.so.
#pragma once

#include "nqs_common.h"

namespace nqs {

// Device view of an ansatz's derived tables in arithmetic type R.  Built by
// launch_prep from the FP64 parameter vector (reference layout,
// src/nqs/nqs_ansatz.c:36-48): a (N), b (M), W (M x N) per real / imaginary
// block.  EP = e^{2W} and EM = e^{-2W} give the ratio of one hidden factor
// after flipping sites with the logistic pair (p, q) of theta
// (logistic_small in nqs_common.h): cosh(theta + d)/cosh(theta) =
// p e^{d} + q e^{-d}, with e^{d} = EM, e^{-d} = EP for the flip of an up site
// (d = -2 s_i W_hi) and the products of two sites' entries for a pair.
// Imaginary arrays are null for real kinds.
template <typename R>
struct ModelView {
  int N = 0, M = 0, cplx = 0;
  const R *a_re = nullptr, *a_im = nullptr;
  const R *b_re = nullptr, *b_im = nullptr;
  const R *WT_re = nullptr, *WT_im = nullptr;      // [N][M]
  const R *EP_re = nullptr, *EP_im = nullptr;      // [M][N]  e^{2W}
  const R *EM_re = nullptr, *EM_im = nullptr;      // [M][N]  e^{-2W}
  const R *EPT_re = nullptr, *EPT_im = nullptr;    // [N][M]
  const R *EMT_re = nullptr, *EMT_im = nullptr;    // [N][M]
};

// Largest |Re W| the ratio tables serve, and how many hidden units the
// kernels multiply between two rescalings of a running product.  A pair
// changes Re theta by up to 4 max|Re W| per unit, so the bounds keep every
// table entry and every product between rescalings inside the type's range
// (FP64: 8 x 80 < 709; FP32: 2 x 32 < 88), and a small logistic member that
// underflows to zero is negligible beside the other term of its ratio
// (2 x 80 < 745 for FP64, 2 x 32 < 87 for FP32).  launch_prep refuses
// larger weights.
template <typename R>
struct RatioRange;
template <>
struct RatioRange<double> { static constexpr double max_w = 20.0; static constexpr int rescale = 8; };
template <>
struct RatioRange<float> { static constexpr double max_w = 8.0; static constexpr int rescale = 2; };

template <typename R>
struct ModelBuffers {
  R* base = nullptr;
  size_t count = 0;
  ModelView<R> view;
  int valid = 0;
  int* range_flag = nullptr;   // device: set by the table kernel when |Re W| > max_w
  int ratio_ok = -1;           // 1: ratio tables usable, 0: refused, -1: not read
};

// Hamiltonian as term lists (built on the host in the reference's bond
// order, see lattice.cpp).  E_loc(s) = sum_d c_d s_a s_b
//   + sum_f c_f psi(s^f)/psi(s) + sum_p [s_a != s_b] c_p psi(s^{ab})/psi(s).
struct HamDev {
  int n_diag = 0, n_flip = 0, n_pair = 0;
  const int* diag_a = nullptr;
  const int* diag_b = nullptr;
  const double* diag_c = nullptr;
  const int* flip_i = nullptr;
  const double* flip_c = nullptr;
  const int* pair_a = nullptr;
  const int* pair_b = nullptr;
  const double* pair_c = nullptr;
  const signed char* sublat = nullptr;  // Marshall: 1 on sublattice A, or null
};

// Metropolis chain state on the device.
struct ChainState {
  int C = 0, N = 0, nw = 0;
  uint64_t* spins = nullptr;      // [C][nw] bitmask, bit=1 <-> s=-1
  uint64_t* rng = nullptr;        // [C]
  unsigned long long* acc = nullptr;  // [C]
  unsigned long long* tot = nullptr;  // [C]
  // log|psi| of each chain's configuration as of its last accepted move (or
  // creation), evaluated with the parameters of that moment: the reference
  // sampler keeps this value across parameter updates
  // (src/nqs/nqs_sampler.c:50-76, 106-113) and the acceptance test uses it.
  double* la = nullptr;               // [C]
};

// ---- model (k_model.hip) --------------------------------------------------
// Tables of the ansatz.  check = 1 waits for them and sets mb.ratio_ok
// (0: a weight beyond RatioRange, the ratio tables must not be used);
// check = 0 (graph capture) leaves the verdict in *mb.range_flag (nonzero:
// refused) for the caller to read later.  Amplitudes (a, b, W^T) are valid
// either way.
template <typename R>
int launch_prep(const double* d_par, int kind, int N, int M, ModelBuffers<R>& mb,
                hipStream_t st, int check = 1);
// message on stderr naming the largest |Re W| (scratch: mb.range_flag)
template <typename R>
int report_ratio_range(const double* d_par, int N, int M, int* scratch, hipStream_t st);
template <typename R>
void free_model(ModelBuffers<R>& mb);
// theta = b + W s per sample; writes t = tanh(theta) [B][M] (re / im),
// when q_re != null the small member of the logistic pair of theta [B][M]
// (re / im; which member it is follows from the sign of Re t: p when
// Re t < 0), and, when la != null, log|psi| and arg psi [B] in FP64.
template <typename R>
int launch_forward(const ModelView<R>& m, const int* spins, int B, R* t_re, R* t_im,
                   double* la, double* ar, hipStream_t st, R* q_re = nullptr,
                   R* q_im = nullptr);
// t (and optionally the small logistic member) only, FP64, theta = S W^T + b
// on FP64 MFMA (d_par: the parameter vector, sd: B*N doubles of scratch)
int launch_forward_t_gemm(const ModelView<double>& m, const double* d_par, const int* spins, int B,
                          double* t_re, double* t_im, double* sd, hipStream_t st,
                          double* q_re = nullptr, double* q_im = nullptr);
// O rows [B][P] from spins and t: mode 0 real-projected, 1 holomorphic.
template <typename R>
int launch_gen_O(const ModelView<R>& m, int kind, const int* spins, int B,
                 const R* t_re, const R* t_im, int mode, double* O_re, double* O_im,
                 long P, hipStream_t st);

// ---- sampler (k_sampler.hip) ---------------------------------------------
struct MHLaunch {
  int move = 0;              // 0 flip, 1 exchange
  int steps = 1;             // moves per round
  int rounds = 1;
  int record = 0;            // write samples
  int n_out = 0;             // samples to write (index r*C + c < n_out)
  int* out = nullptr;        // [n_out][N] int32
};
template <typename R>
int launch_mh(const ModelView<R>& m, ChainState& cs, const MHLaunch& L, hipStream_t st);
// Several FP64 samplers in one launch.  A variant (0..5) fixes the lanes per
// chain and hidden units per lane (by M) and real / complex; every job of a
// launch has the same variant, jobs[j].blk0 is its first block (jobs in
// block order, mh_blocks(variant, C) blocks each), `jobs` is device memory.
struct MHJob {
  ModelView<double> m;
  ChainState cs;
  MHLaunch ml;
  int blk0 = 0;
};
int mh_variant(const ModelView<double>& m);
int mh_blocks(int variant, int C);
int launch_mh_batch(int variant, const MHJob* jobs, int njobs, int nblocks, hipStream_t st);
int launch_chain_init(ChainState& cs, const uint64_t* seeds_dev, int sz_zero, hipStream_t st);
int launch_unpack(const ChainState& cs, int chain, int* out, hipStream_t st);
// all chains' configurations as int32 rows [C][N]
int launch_unpack_all(const ChainState& cs, int* out, hipStream_t st);

// ---- learned proposal (k_ar.hip; network layout in ar_layout.h) -----------
// Device copy of the masked autoregressive network: FP32 master weights and
// Adam moments [P], BF16 MFMA operand copies, FP32 biases, and W1m
// transposed [N][64] (BF16-rounded values) for the walk's per-site update.
struct ArDev {
  int N = 0;
  float *w = nullptr, *m = nullptr, *v = nullptr;
  uint16_t *op1f = nullptr, *op2f = nullptr, *op2b = nullptr, *op3f = nullptr, *op3b = nullptr;
  float *w1t = nullptr, *b1 = nullptr, *b2 = nullptr, *b3 = nullptr;
};
// Rows [0, n_sample) draw a configuration with xorshift state rng[r] and
// the magnetisation of bits[r]; rows [n_sample, rows) evaluate the
// configuration bits[r - n_sample].  lq[r] = log q of the row's result.
struct ArWalk {
  int rows = 0, n_sample = 0, constrained = 1;
  const uint64_t* bits = nullptr;
  uint64_t* rng = nullptr;
  uint64_t* out_bits = nullptr;   // [n_sample], optional
  int* out_spins = nullptr;       // [n_sample][N], optional
  double* lq = nullptr;           // [rows]
};
// Gradient of the summed negative log-likelihood times `scale` over rows
// (start + b) mod n_avail, b < count; partial[blk][P] and loss[blk] per
// 64-row block.
struct ArGrad {
  const uint64_t* rows = nullptr;
  long n_avail = 0, start = 0;
  int count = 0, constrained = 1;
  int rows_per_block = 32;   // set by launch_ar_grad
  float scale = 1.f;
  float* partial = nullptr;
  double* loss = nullptr;
};
// k_ar_grad blocks: 32 rows per block up to kArGradMaxBlocks blocks, then
// more rows per block (a multiple of 32); each block writes one partial
// gradient of P floats, which the Adam kernel sums
constexpr int kArGradMaxBlocks = 32;
inline int ar_grad_rows(long count) {
  const long per = (count + kArGradMaxBlocks - 1) / kArGradMaxBlocks;
  return (int)(per <= 32 ? 32 : (per + 31) / 32 * 32);
}
inline int ar_grad_blocks(long count) {
  const int r = ar_grad_rows(count);
  return (int)((count + r - 1) / r);
}
struct ArAdam {
  const float* partial = nullptr;
  int nblk = 0;
  float lr = 1e-3f, bc1 = 1.f, bc2 = 1.f, beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
};
int launch_ar_walk(const ArDev& v, const ArWalk& w, hipStream_t st);
int launch_ar_grad(const ArDev& v, const ArGrad& g, hipStream_t st);
int launch_ar_adam(const ArDev& v, const ArAdam& a, hipStream_t st);
int launch_ar_pack(const ArDev& v, hipStream_t st);
int launch_ar_sum_grad(const ArDev& v, const float* partial, int nblk, float* out, hipStream_t st);
// Learned Metropolis-Hastings rounds (nw == 1): rounds global moves per
// chain; chain c's state after round r is written as sample r*C + c of out
// when that index is below n_out (out may be null).  lacc[c] counts the
// chain's accepted learned moves.
struct ArRounds {
  int rounds = 0, constrained = 1;
  int w_lds = 0;                 // set by the launcher: W^T staged in LDS
  long n_out = 0;
  int* out = nullptr;
  unsigned long long* lacc = nullptr;
};
template <typename R>
int launch_ar_mh(const ArDev& v, const ModelView<R>& m, ChainState& cs, const ArRounds& ar,
                 hipStream_t st);
// Learned rounds in bulk (independence proposals drawn for all rounds at
// once; see k_ar.hip): prop [rounds*C], spins [rounds*C][N], lq
// [(rounds+1)*C] (the last C: the chains' current states), lp [rounds*C]
// (log|psi| of the proposals, filled between the two launches).
struct ArBulk {
  int rounds = 0, constrained = 1;
  uint64_t* prop = nullptr;
  int* spins = nullptr;
  double* lq = nullptr;
  double* lp = nullptr;
  long n_out = 0;
  int* out = nullptr;
  unsigned long long* lacc = nullptr;
  // training window (optional): samples i in [win_first, n_out) also go to
  // win[(win_pos + i - win_first) % win_cap] as N-bit masks (bit 1: s = -1)
  uint64_t* win = nullptr;
  long win_cap = 1, win_pos = 0, win_first = 0;
};
int launch_ar_propose(const ArDev& v, ChainState& cs, const ArBulk& b, hipStream_t st);
int launch_ar_accept_rounds(ChainState& cs, const ArBulk& b, hipStream_t st);
// int32 rows -> bitmasks (N <= 64): ring positions (pos + b) mod cap
int launch_bits_append(const int* spins, long n, int N, uint64_t* win, long cap, long pos,
                       hipStream_t st);
int launch_spins_to_bits(const int* spins, long n, int N, uint64_t* out, hipStream_t st);

// ---- local energy (k_eloc.hip) -------------------------------------------
// With a workspace of at least eloc_ws_bytes(m, h, B) bytes the term-major
// kernels run (term tables, 64 samples per wave, fixed-order sums);
// without one, the per-sample item-list kernel.
// t and q: tanh(theta) and the small logistic member of every sample's
// hidden inputs, as launch_forward writes them.
template <typename R>
size_t eloc_ws_bytes(const ModelView<R>& m, const HamDev& h, int B);
template <typename R>
int launch_eloc(const ModelView<R>& m, const HamDev& h, const int* spins, int B,
                const R* t_re, const R* t_im, const R* q_re, const R* q_im, double* e_re,
                double* e_im, hipStream_t st, void* ws = nullptr, size_t ws_bytes = 0);
// psi(s')/psi(s) of every (term, sample) pair of h with coefficient and
// sublattice sign applied (zero for inactive pairs), term-major [T][B], in
// the workspace (eloc_ws_bytes): res[0] real parts, res[1] imaginary parts.
template <typename R>
int launch_pair_ratios(const ModelView<R>& m, const HamDev& h, const int* spins, int B,
                       const R* t_re, const R* q_re, const R* q_im, void* ws, size_t ws_bytes,
                       double** res, hipStream_t st);
// E_loc from externally evaluated log-amplitudes of every connected
// configuration (host-callback adaptor): la/ar of sample b at [b][0], of
// its k-th connected configuration at [b][1+k] in term order.
int launch_eloc_from_logamps(const HamDev& h, const int* spins, int B, int N,
                             int n_terms, const double* la, const double* ar,
                             double* e_re, double* e_im, hipStream_t st);

// ---- group orbits of configurations (k_orbit.hip) ------------------------
// out [B*G][N]: row b G + g is g.s_b, (g.s)_i = s_{perm[g][i]}
int launch_orbit_spins(const int* spins, int B, int N, const int* perm, int G, int* out,
                       hipStream_t st);
// real combination of the orbit rows' (lp, ar) [B][G] into log|psi|, arg
int launch_orbit_combine(const double* lp, const double* ar, const double* chi, int B, int G,
                         double* la_out, double* ar_out, hipStream_t st);
// weights w [B][G] (w_re, w_im scratch) and O_sym [B][P] = sum_g w O[bG+g]
int launch_orbit_grad(const double* lp, const double* ar, const double* chi, int B, int G,
                      int cplx, const double* O_re, const double* O_im, long P, double* w_re,
                      double* w_im, double* out_re, double* out_im, hipStream_t st);

// ---- mixed precision: split 16-bit MFMA products (k_mixed.hip) -------------
// C = alpha A B^T with A [M][K], B [N][K] FP32 (K a multiple of 32, rows
// 16-byte aligned), C FP64: entries scaled by a power of two from the
// largest |entry| (device amax / bmax) and split into hi + lo half
// precision, each product as three FP16 MFMA terms, FP32 accumulation;
// lower = 1: entries with column <= row of a square product only.  K sums
// longer than 2048 run as splits of at most 2048 columns into part
// (x3_part_elems(M, N, K) floats), added in FP64.
long x3_pad(long k);   // K padded to what the split product reads (zeros)
size_t x3_part_elems(int M, int N, long K);
int launch_gemm_x3(int M, int N, int K, double alpha, const float* A, long lda, const float* B,
                   long ldb, const float* amax, const float* bmax, double* C, long ldc, int lower,
                   float* part, hipStream_t st);
// FP32 staging of a Gram operand with amax[0] = largest |entry|; amax holds
// x3_max_elems(...) floats (block maxima after the first)
long x3_max_elems(long R, long C, long ld, int transpose);
// out [C][ldk] = FP32 transpose of in [R][C], zero for rows R..ldk-1
// out [P][ldk] = (O - Om)^T in FP32 for an RBM's real-projected O rows
// (s, u, u (x) s; then 0, -v, -v (x) s with v = U2 when given), generated
// from the spins [B][N] and the FP32 factors U1, U2 [B][M] (columns B..ldk
// zero), with the block maxima of |out| and their maximum in amax[0]
// (x3_max_elems(B, P, ldk, 1) floats); N + M (+ M) <= 256 (kOctRows), else
// -1
int launch_oct_factored(const int* spins, int B, int N, int M, const float* U1, const float* U2,
                        const double* Om, long P, float* out, long ldk, float* amax,
                        hipStream_t st);
int launch_to_f32_t(const double* in, long R, long C, float* out, long ldk, float* amax,
                    hipStream_t st);
// out [R][ldc] = FP32 copy of in [R][C], zero for columns C..ldc-1
int launch_to_f32_pad(const double* in, long R, long C, float* out, long ldc, float* amax,
                      hipStream_t st);
// t = tanh(S W^T + b) in FP32 for B configurations (RBM / complex RBM), theta
// on the BF16 MFMA with W split three ways; wsplit: theta_x3_split_elems(M, N)
// bf16 per weight block (twice that for a complex RBM)
int theta_x3_split_elems(int M, int N);
int launch_forward_t_x3(const ModelView<float>& m, const double* d_par, const int* spins, int B,
                        float* t_re, float* t_im, __bf16* wsplit, hipStream_t st);

// ---- sample diagnostics (k_diag.hip) --------------------------------------
// out[0] = sum_k (<|O_k|^2> - |<O_k>|^2) over rows O_re [B][P] (+ i O_im,
// may be null); work: chi_work_elems(P) doubles
size_t chi_work_elems(long P);
int launch_chi_trace(const double* O_re, const double* O_im, long B, long P, double* work,
                     double* out, hipStream_t st);

// ---- exact diagonalisation on the 2^N basis (k_ed.hip) --------------------
// Every reduction writes partials (part: at least ed_partials() doubles) and
// reduces them into a device scalar in a fixed order.
int ed_partials();
// H = sum_bonds [jd s_u s_v + jo flip(u, v) if s_u != s_v] + field sum_i flip(i)
struct EdOp {
  const int* bonds = nullptr;   // device [nb][2] site pairs
  int nb = 0;
  double jd = 0.0, jo = 0.0, field = 0.0;
  int N = 0;                    // sites flipped by the field term
};
int launch_ed_matvec(const EdOp& op, const double* in, double* out, long dim, double* part,
                     double* dot, hipStream_t st);
int launch_ed_axpy2(double* w, const double* v, double a, const double* u, double b, long dim,
                    double* part, double* norm2, hipStream_t st);
// tab: device [G][nchunk][8] OR-masks, chi: device [G]
int launch_ed_project(const double* in, double* out, long dim, const unsigned long long* tab,
                      const double* chi, int G, int nchunk, double* part, double* norm2,
                      hipStream_t st);
int launch_ed_scale(const double* x, double* y, double c, long dim, double* part, double* norm2,
                    hipStream_t st);
int launch_ed_state_spins(long s0, long n, int N, int* out, hipStream_t st);
// Krylov basis V [nv][dim]: h[j] = <V_j, w> (part: ed_mdot_partials(dim, nv)
// doubles); w = (keep ? w : 0) - sum_j h[j] V_j, norm2 = |w|^2 when given
size_t ed_mdot_partials(long dim, int nv);
int launch_ed_mdot(const double* V, long dim, int nv, const double* w, double* part, double* h,
                   hipStream_t st);
int launch_ed_mupdate(const double* V, long dim, int nv, const double* h, double* w, int keep,
                      double* part, double* norm2, hipStream_t st);
// psi = cos(ar) exp(la - max la); scal[0] = max la, scal[1] = sum psi^2,
// scal[2] = sum exp(2 (la - max la))
int launch_ed_psi(const double* la, const double* ar, long dim, double* psi, double* part,
                  double* scal, hipStream_t st);

// ---- linear algebra (k_linalg.hip), FP64 ---------------------------------
// out[k] = alpha * sum_b O[b][k] * (w ? w[b] : 1)  (deterministic order)
int launch_colsum(const double* O, long B, long P, const double* w, double alpha,
                  double* out, double* work, hipStream_t st);
size_t colsum_work_elems(long B, long P);
// u[b] = sum_k O[b][k] v[k]
int launch_rowdot(const double* O, long B, long P, const double* v, double* u,
                  hipStream_t st);
// deterministic sum of x[0..n) times alpha into *out (device scalar)
int launch_sum(const double* x, long n, const double* y, double alpha, double* out,
               hipStream_t st);
// O[b][k] -= mean[k]
int launch_center(double* O, long B, long P, const double* mean, hipStream_t st);
// C = alpha * A B^T + beta C  (row-major; A [M][K] lda, B [N][K] ldb, C ldc);
// lower != 0: only tiles with any element on or below the diagonal.
int launch_gemm_nt(int M, int N, int K, double alpha, const double* A, long lda,
                   const double* B, long ldb, double beta, double* C, long ldc,
                   int lower, hipStream_t st);
// In-place lower Cholesky of the n x n row-major matrix (lower triangle read
// and written, upper untouched).  Returns 0 on success or 1 + the first
// column whose pivot is not positive, -1 on errors.  Linv receives the
// inverses of the 64 x 64 diagonal blocks; it also holds launch_potrs's
// super-block inverses (potrf_linv_elems(n) doubles in all).
size_t potrf_linv_elems(int n);
// trailing: an event after which rows and columns >= potrf_first_width(n)
// of A hold their input (nullptr: all of A does on entry).  The first wide
// panel, columns [0, potrf_first_width(n)), is factored before the wait.
int launch_potrf(double* A, int n, double* Linv, int* fail_dev, hipStream_t st,
                 hipEvent_t trailing = nullptr);
int potrf_first_width(int n);
// the factorisation's second stream (created on first use): work issued on
// it ahead of launch_potrf runs before the lookahead updates
hipStream_t potrf_side_stream();
// After a failure at column jf: bring A to the state the unblocked
// left-looking factorisation leaves (columns < jf hold L, columns >= jf
// their input values, taken from the lower triangle of orig).
int launch_potrf_repair(double* A, int n, int jf, const double* orig, hipStream_t st);
// x = (L L^T)^-1 b with L, Linv from launch_potrf; w is scratch [n].
int launch_potrs(const double* L, const double* Linv, int n, const double* b, double* x,
                 double* w, hipStream_t st);
// Factored MinSR Gram T = (O_c O_c^T)/B + eps I (lower triangle) for O rows
// (s, u, u (x) s); u . u' = U1 U1^T (+ U2 U2^T when U2 is given), Ku columns
// each (Ku = 0 for mean-field).  work: gram_factored_work_elems(B, N, Ku,
// 1 or 2) doubles, Sd: B*(N+2) scratch, r: B scratch (row means),
// g: device scalar.
size_t gram_factored_work_elems(int B, int N, int M, int nblk);
int launch_gram_factored(const int* spins, int B, int N, const double* U1, const double* U2, int Ku,
                         double eps, double* T, double* work, double* Sd, double* r, double* g,
                         hipStream_t st);
// MinSR update of an RBM from its factors: delta = alpha * O^T w over the
// parameter blocks (O rows (s, u, u (x) s); u = U1, then u = -U2 with a
// zero visible part when U2 is given), w = centred solution (null: unit
// weights).  work:
// factored_delta_work_elems(B, N, M, 1 or 2) doubles.
int factored_delta_work_elems(int B, int N, int M, int nblk);
int launch_factored_delta(const int* spins, int B, int N, int M, const double* U1,
                          const double* U2, const double* w, double alpha, double* work,
                          double* delta, hipStream_t st);
// the same sums over FP32 factors (each factor widened exactly, FP64 MFMA)
int launch_factored_delta(const int* spins, int B, int N, int M, const float* U1,
                          const float* U2, const double* w, double alpha, double* work,
                          double* delta, hipStream_t st);
int launch_add_diag(double* A, int n, long ld, double v, hipStream_t st);
// out [C][R] = in [R][C]^T
int launch_transpose(const double* in, long R, long C, double* out, hipStream_t st);
int launch_copy_diag(const double* A, int n, double* d, hipStream_t st);
int launch_mirror_lower(double* A, int n, hipStream_t st);

// streaming copy of n doubles (n a multiple of 4)
int launch_copy_probe(const double* src, double* dst, long n, hipStream_t st);

// elementwise helpers
int launch_axpy(double* y, const double* x, double a, long n, hipStream_t st);
int launch_fill(double* x, double v, long n, hipStream_t st);
// y[i] = x[i] - *s (s a device scalar)
int launch_sub_scalar(double* y, const double* x, long n, const double* s, hipStream_t st);

// ---- CG on the device (k_linalg.hip) ----------------------------------------
struct CGScalars {   // lives in device memory
  double rs, rs0, pAp, alpha, beta, rs_new;
  int iters, converged, breakdown, done;
};
int launch_cg_init(const double* F, long P, double* x, double* r, double* p,
                   CGScalars* sc, hipStream_t st);
int launch_cg_alpha(CGScalars* sc, const double* p, const double* Ap, long P,
                    hipStream_t st);
int launch_cg_update(CGScalars* sc, double* x, double* r, const double* p,
                     const double* Ap, long P, double tol, hipStream_t st);
int launch_cg_direction(CGScalars* sc, double* p, const double* r, long P,
                        hipStream_t st);
// out[k] = a * (x[k] - m[k] * s[0]) + eps * v[k] accumulate helpers for QGT
int launch_qgt_finish(double* out, const double* c1, const double* m1, const double* u1mean,
                      const double* c2, const double* m2, const double* u2mean,
                      double scale, double eps, const double* v, long P, hipStream_t st);

// ---- batched independent CG-SR problems (k_multi.hip) -------------------------
// One real-RBM problem of a batch.  Vectors over the parameters live in a
// padded (M+1) x (N+1) row-major layout: W[h][i] at (h, i), b[h] at (h, N),
// a[i] at (M, i), and (M, N) is always 0.  In that layout the O row of
// sample b is the outer product [t_b, 1] (x) [s_b, 1] (t = tanh theta), so
// O v = rowsum([t, 1] * ([s, 1] V^T)) and O^T w = ([t, 1] * w)^T [s, 1]:
// two small GEMMs through the spins and t instead of two passes over O.
constexpr int kMultiVecBlock = 256;      // threads of the vector kernels
constexpr int kMultiMaxPart = 512;       // block partials per problem
struct MultiProb {
  int B = 0, N = 0, M = 0, nsplit = 1, max_it = 0, C = 0;
  double eps = 0.0, tol = 0.0, lr = 0.0;
  const unsigned long long* acc = nullptr;   // chains' accepted / proposed moves [C]
  const unsigned long long* tot = nullptr;
  const int* spins = nullptr;    // [B][N]
  const double* t = nullptr;     // [B][M]
  const double* E = nullptr;     // [B]
  double* S1 = nullptr;          // [B][N+1]   [s_b, 1]
  double* S1t = nullptr;         // [N+1][B]
  double* T1t = nullptr;         // [M+1][B]   [t_b, 1]^T
  double* u = nullptr;           // [(M+64)/64][B] row weights of the CG product,
                                 // (O p)_b, in partial sums over 64-row tiles
  double* Gp = nullptr;          // [2][nsplit][M+1][N+1] split-K partials
  double *x = nullptr, *r = nullptr, *p = nullptr, *Ap = nullptr, *Om = nullptr,
         *F = nullptr;           // padded vectors
  double* part = nullptr;        // [kMultiMaxPart] block partials
  unsigned* ticket = nullptr;    // arrival counter of the last-block reductions
  CGScalars* sc = nullptr;
  const double* stats = nullptr; // [0] <E> (reference order), read by the force
  double* norm2 = nullptr;       // |delta|^2
  double* par = nullptr;         // [P] parameters, reference layout
  const int* range_flag = nullptr;  // the ansatz tables' range verdict (nonzero: refused)
  // first block of this problem in each batched launch
  // (blk_g: one G product per problem, CG; blk_g2: two, the force)
  int blk_prep = 0, blk_z = 0, blk_g = 0, blk_g2 = 0, blk_vec = 0;
};
struct MultiBatch {
  int n = 0;
  const MultiProb* probs = nullptr;   // device copy
  int nblk_prep = 0, nblk_z = 0, nblk_g = 0, nblk_g2 = 0, nblk_vec = 0;
};
// fills the problems' first-block offsets of every launch and the totals
void multi_layout(MultiProb* probs, int n, MultiBatch& mb);
// [s, 1], [t, 1] operands and the two weight sets (1, E) of the force
int launch_multi_prep(const MultiBatch& mb, hipStream_t st);
// Om, F, CG start (x = 0, r = p = F, rs = |F|^2)
int launch_multi_force(const MultiBatch& mb, hipStream_t st);
// one CG iteration (A p, alpha, x / r update, beta, p) on every problem
// not yet done; `matvec_only` stops after A p and alpha (idempotent at the
// CG start, used to time the split factors)
int launch_multi_cg_iter(const MultiBatch& mb, int matvec_only, hipStream_t st);
// up to `iters` CG iterations of every problem in one cooperative launch
// (same results as as many launch_multi_cg_iter calls; matvec_only: one
// iteration's A p and alpha, as launch_multi_cg_iter's); -1 when the device
// refuses the cooperative launch
int launch_multi_cg_persist(const MultiBatch& mb, int iters, hipStream_t st, int matvec_only = 0);
// par -= lr * x (reference layout), |x|^2
int launch_multi_apply(const MultiBatch& mb, hipStream_t st);
// per problem kMultiOut doubles: <E>, <E^2>, |delta|^2, CG iterations,
// converged, breakdown, pAp, rs_new, rs0, accepted / proposed moves, the
// tables' range flag
constexpr int kMultiOut = 11;
int launch_multi_gather(const MultiBatch& mb, double* out, hipStream_t st);

// ---- excited-state SR and Marshall sign helpers (k_dyn.hip) -----------------
// r = psi_ref / psi per sample from the two log-amplitudes (log-ratio
// clamped to [-10, 10])
int launch_excited_ratio(const double* la_p, const double* ar_p, const double* la_r,
                         const double* ar_r, int B, double* r_re, double* r_im, hipStream_t st);
// E += mu r conj(<r>), rm = {<r_re>, <r_im>} on the device
int launch_excited_augment(double* e_re, double* e_im, const double* r_re, const double* r_im,
                           const double* rm, double mu, int B, hipStream_t st);
// closed-form augmented means of <E_re>, <E_im> (device scalars) -> out[2]
int launch_excited_means(const double* er, const double* ei, const double* rm, double mu,
                         double* out, hipStream_t st);
// arg += pi where the Marshall sublattice holds an odd number of down spins
int launch_marshall_arg(const int* spins, int B, int N, int Lx, int Ly, double* ar,
                        hipStream_t st);

// ---- bookkeeping ------------------------------------------------------------
void count_launch(int n = 1);

// Live per-kernel timing with HIP events on the launch stream, enabled by
// name through nqs_hip_kernel_timing(); a no-op otherwise.
struct KScope {
  int slot = -1;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipStream_t st = nullptr;
  KScope(const char* name, hipStream_t stream);
  ~KScope();
};

}  // namespace nqs
