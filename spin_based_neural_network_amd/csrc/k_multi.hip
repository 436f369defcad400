// k_multi.hip — many independent real-RBM CG-SR problems stepped together.
//
// Reference semantics per problem: nqs_sr_step_custom_full with the ansatz's
// own gradient (src/nqs/nqs_optimizer.c:141-241): F_k = <O_k E> - <O_k> <E>,
// CG on (S + eps I) delta = F with S v = O^T (O v) / Ns - <O> mean(O v)
// (qgt_apply, :44-80) and cg_solve's stopping rules (:84-139), then
// theta -= lr delta.
//
// MI355X layout: problems of different sizes share every launch; a block
// finds its problem from the per-launch block offsets.  The O matrix is never
// formed: in the padded (M+1) x (N+1) parameter layout an O row is
// [t_b, 1] (x) [s_b, 1], so
//   (O v)_b     = sum_h [t_b, 1]_h (V [s_b, 1]^T)_h       (Z = V S1^T, row weights u)
//   (O^T w)_hi  = sum_b w_b [t_b, 1]_h [s_b, 1]_i         (GEMM (T1 w)^T S1)
// on FP64 MFMA, and the last entry of O^T w is sum_b w_b, the mean that the
// QGT needs.  Per CG iteration: Z product and row weights u in one pass
// (Z never leaves the registers), split-K G GEMM with its operand [t_b, 1]
// scaled by u_b as it is staged, finish (A p, p.Ap, alpha), update (x, r,
// r.r, beta), direction.
// Scalar reductions over a problem's blocks use one arrival counter: every
// block stores its partial, the block that arrives last sums the partials in
// block order (deterministic) and resets the counter.
#include <hip/hip_cooperative_groups.h>

#include "nqs_kernels.h"

namespace nqs {

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));
constexpr int MT = 64, MKS = 16;     // GEMM tile and K step

enum { F_PREP = 0, F_Z = 1, F_G = 3, F_G2 = 4, F_VEC = 5 };

template <int FIELD>
__device__ inline int first_blk(const MultiProb& q) {
  return FIELD == F_PREP ? q.blk_prep
       : FIELD == F_Z    ? q.blk_z
       : FIELD == F_G    ? q.blk_g
       : FIELD == F_G2   ? q.blk_g2
                         : q.blk_vec;
}

// problem owning block `blk` of a launch (largest q with first block <= blk)
template <int FIELD>
__device__ inline int find_prob(const MultiBatch& mb, int blk) {
  int lo = 0, hi = mb.n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (first_blk<FIELD>(mb.probs[mid]) <= blk) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// acc = the 64 x 64 tile [row0.., col0..] of sum_{k0 <= k < k1} A[r][k] w_k
// Bm[c][k] with w_k = sum_{t < nw} w[t ldw + k] added in t order (w null:
// 1), 4 waves of 32 x 32, v_mfma_f64_16x16x4f64: lane l
// supplies A[l%16][l/16] and B[l%16][l/16]; result element j of lane l is
// D[l/16 + 4j][l%16] (wave rows wr = 32 (wave / 2), columns wc = 32 (wave % 2)).
// A row times its weight is rounded once as it is staged.
__device__ inline void tile_acc(int M, int N, int k0, int k1, const double* __restrict__ A,
                                long lda, const double* __restrict__ w, int nw, long ldw,
                                const double* __restrict__ Bm, long ldb, int row0, int col0,
                                v4d (&acc)[2][2]) {
  __shared__ double As[MT][MKS + 1];
  __shared__ double Bs[MT][MKS + 1];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int lr = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = (v4d){0.0, 0.0, 0.0, 0.0};
  for (int kb = k0; kb < k1; kb += MKS) {
    // a thread stages K index kb + tid % 16 in all four of its rows: its
    // weight once per step
    const int gk = kb + (tid & 15);
    double wk = 1.0;
    if (w && gk < k1) {
      wk = w[gk];
      for (int t = 1; t < nw; t++) wk += w[t * ldw + gk];
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int e = tid + 256 * q;
      const int r = e >> 4, kk = e & 15;
      const int ga = row0 + r, gb = col0 + r;
      double a = 0.0;
      if (ga < M && gk < k1) {
        a = A[(long)ga * lda + gk];
        if (w) a *= wk;
      }
      As[r][kk] = a;
      Bs[r][kk] = (gb < N && gk < k1) ? Bm[(long)gb * ldb + gk] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < MKS; kk += 4) {
      const double a0 = As[wr + lr][kk + lk];
      const double a1 = As[wr + 16 + lr][kk + lk];
      const double b0 = Bs[wc + lr][kk + lk];
      const double b1 = Bs[wc + 16 + lr][kk + lk];
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
    __syncthreads();
  }
}

// C[row0.., col0..] = the tile of tile_acc
__device__ inline void tile_nt(int M, int N, int k0, int k1, const double* __restrict__ A,
                               long lda, const double* __restrict__ w, int nw, long ldw,
                               const double* __restrict__ Bm, long ldb, double* C, long ldc,
                               int row0, int col0) {
  v4d acc[2][2];
  tile_acc(M, N, k0, k1, A, lda, w, nw, ldw, Bm, ldb, row0, col0, acc);
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int lr = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int gr = row0 + wr + 16 * i + lk + 4 * r;
        const int gc = col0 + wc + 16 * j + lr;
        if (gr < M && gc < N) C[(long)gr * ldc + gc] = acc[i][j][r];
      }
}

// Block partial -> problem scalar.  Every thread passes its value; the
// block that arrives last among `nblk` returns true with the sum of all
// partials (block order) in `total`; the others return false.
__device__ inline bool reduce_last(double v, double* red, double* part, unsigned* ticket,
                                   int blk, int nblk, double& total) {
  __shared__ int last;
  const double s = block_sum(v, red);
  if (threadIdx.x == 0) {
    part[blk] = s;
    __threadfence();
    const unsigned t = atomicAdd(ticket, 1u);
    last = (t == (unsigned)nblk - 1u);
  }
  __syncthreads();
  if (!last) return false;
  __threadfence();
  double w = 0.0;
  for (int i = threadIdx.x; i < nblk; i += blockDim.x) w += __builtin_nontemporal_load(&part[i]);
  // fixed assignment of partials to threads and a fixed tree: deterministic
  total = block_sum(w, red);
  if (threadIdx.x == 0) atomicExch(ticket, 0u);
  return true;
}

__device__ inline int vec_blocks(const MultiProb& q) {
  return ((q.M + 1) * (q.N + 1) + kMultiVecBlock - 1) / kMultiVecBlock;
}
__device__ inline bool cg_done(const MultiProb& q) {
  return __builtin_nontemporal_load(&q.sc->done) != 0;
}

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_mb_prep(MultiBatch mb) {
  const int qi = find_prob<F_PREP>(mb, blockIdx.x);
  const MultiProb& q = mb.probs[qi];
  const int b = (blockIdx.x - q.blk_prep) * 256 + threadIdx.x;
  if (b >= q.B) return;
  const int N = q.N, M = q.M, N1 = N + 1, B = q.B;
  for (int j = 0; j <= N; j++) {
    const double v = j < N ? (double)q.spins[(long)b * N + j] : 1.0;
    q.S1[(long)b * N1 + j] = v;
    q.S1t[(long)j * B + b] = v;
  }
  for (int h = 0; h <= M; h++) q.T1t[(long)h * B + b] = h < M ? q.t[(long)b * M + h] : 1.0;
}

// Row weights of the CG product, u_b = sum_h [t_b, 1]_h (V [s_b, 1]^T)_h, in
// partial sums over 64-row tiles: work item (tile tm, samples 64 tb + [0,
// 64)) forms its 64 x 64 tile of Z = V S1^T on MFMA (K = N+1), weights it by
// [t_b, 1]_h as it leaves the accumulators (Z is not stored) and writes the
// tile's column sums to u[tm][b].  A thread sums its 8 rows of the tile for
// its 2 samples; the 8 threads that share a sample are added in a fixed
// order.  The G product adds the tiles' partials in tile order.
__device__ inline void mb_zu(const MultiBatch& mb, int blk) {
  __shared__ double ured[8][MT];
  const int qi = find_prob<F_Z>(mb, blk);
  const MultiProb& q = mb.probs[qi];
  if (cg_done(q)) return;
  const int M1 = q.M + 1, N1 = q.N + 1;
  const long B = q.B;
  const int tiles_n = (q.B + MT - 1) / MT;
  const int t = blk - q.blk_z, tm = t / tiles_n;
  const int row0 = tm * MT, col0 = (t % tiles_n) * MT;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int lr = lane & 15, lk = lane >> 4;
  v4d acc[2][2];
  tile_acc(M1, q.B, 0, N1, q.p, N1, nullptr, 0, 0, q.S1, N1, row0, col0, acc);
  double part[2] = {0.0, 0.0};
#pragma unroll
  for (int j = 0; j < 2; j++) {
    const long gc = col0 + wc + 16 * j + lr;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int gr = row0 + wr + 16 * i + lk + 4 * r;
        if (gr < M1 && gc < B) part[j] += acc[i][j][r] * q.T1t[(long)gr * B + gc];
      }
  }
  // slot (wave / 2) * 4 + lk of column wc + 16 j + lr: every (slot, column) once
#pragma unroll
  for (int j = 0; j < 2; j++) ured[(wave >> 1) * 4 + lk][wc + 16 * j + lr] = part[j];
  __syncthreads();
  if (tid < MT) {
    double u = 0.0;
#pragma unroll
    for (int k = 0; k < 8; k++) u += ured[k][tid];
    if (col0 + tid < B) q.u[(long)tm * B + col0 + tid] = u;
  }
  __syncthreads();   // ured is read before a later work item of this block writes it
}

// Gp[slot][split] = (T1t w)[:, Kchunk] S1t[:, Kchunk]^T, (M+1) x (N+1): the
// operand [t_b, 1]_h times the sample weight w_b.  With two slots (the force)
// slot 0 takes w = 1 and slot 1 w = E; with one slot (CG) w = u = O p, the
// sum of the row-tile partials of mb_zu.
__device__ inline void mb_gemm_g(const MultiBatch& mb, int blk, int force) {
  const int qi = force ? find_prob<F_G2>(mb, blk) : find_prob<F_G>(mb, blk);
  const MultiProb& q = mb.probs[qi];
  if (!force && cg_done(q)) return;
  const int M1 = q.M + 1, N1 = q.N + 1;
  const int tm = (M1 + MT - 1) / MT, tn = (N1 + MT - 1) / MT;
  int t = blk - (force ? q.blk_g2 : q.blk_g);
  const int per_slot = tm * tn * q.nsplit;
  const int slot = t / per_slot;
  t -= slot * per_slot;
  const int split = t / (tm * tn);
  t -= split * tm * tn;
  const int chunk = ((q.B + q.nsplit - 1) / q.nsplit + MKS - 1) / MKS * MKS;
  const int k0 = split * chunk, k1 = min(q.B, k0 + chunk);
  const double* w = force ? (slot == 0 ? nullptr : q.E) : q.u;
  const int nw = force ? 1 : tm;   // row tiles of mb_zu
  double* C = q.Gp + ((long)slot * q.nsplit + split) * M1 * N1;
  if (k0 >= k1) {
    // empty chunk: its partial is zero
    for (int e = threadIdx.x; e < MT * MT; e += 256) {
      const int r = (t / tn) * MT + e / MT, c = (t % tn) * MT + e % MT;
      if (r < M1 && c < N1) C[(long)r * N1 + c] = 0.0;
    }
    return;
  }
  tile_nt(M1, N1, k0, k1, q.T1t, q.B, w, nw, q.B, q.S1t, q.B, C, N1, (t / tn) * MT,
          (t % tn) * MT);
}

// Om = (O^T 1)/B, F = (O^T E)/B - Om <E>; CG start x = 0, r = p = F,
// rs = rs0 = F.F (cg_solve, nqs_optimizer.c:84-95)
__global__ __launch_bounds__(kMultiVecBlock) void k_mb_force(MultiBatch mb) {
  __shared__ double red[kMultiVecBlock];
  const int qi = find_prob<F_VEC>(mb, blockIdx.x);
  const MultiProb& q = mb.probs[qi];
  const int blk = blockIdx.x - q.blk_vec, nblk = vec_blocks(q);
  const int N1 = q.N + 1, L = (q.M + 1) * N1;
  const int e = blk * kMultiVecBlock + threadIdx.x;
  double ff = 0.0;
  if (e < L) {
    const long plane = (long)q.nsplit * L;
    double g1 = 0.0, gE = 0.0;
    for (int k = 0; k < q.nsplit; k++) {
      g1 += q.Gp[(long)k * L + e];
      gE += q.Gp[plane + (long)k * L + e];
    }
    const double inv = 1.0 / (double)q.B;
    const double om = g1 * inv;
    const double f = (e == L - 1) ? 0.0 : gE * inv - om * q.stats[0];
    q.Om[e] = om;
    q.F[e] = f;
    q.x[e] = 0.0;
    q.r[e] = f;
    q.p[e] = f;
    ff = f * f;
  }
  double tot;
  if (reduce_last(ff, red, q.part, q.ticket, blk, nblk, tot) && threadIdx.x == 0) {
    CGScalars* sc = q.sc;
    sc->rs = tot;
    sc->rs0 = tot;
    sc->beta = 0.0;
    sc->iters = 0;
    sc->converged = 0;
    sc->breakdown = 0;
    sc->done = q.max_it <= 0 ? 1 : 0;
  }
}

// Ap = (O^T u)/B - Om mean(u) + eps p, pAp; alpha = rs / pAp or breakdown
// when |pAp| < 1e-30 (nqs_optimizer.c:103-112)
__device__ inline void mb_finish(const MultiBatch& mb, int gblk) {
  __shared__ double red[kMultiVecBlock];
  const int qi = find_prob<F_VEC>(mb, gblk);
  const MultiProb& q = mb.probs[qi];
  if (cg_done(q)) return;
  const int blk = gblk - q.blk_vec, nblk = vec_blocks(q);
  const int L = (q.M + 1) * (q.N + 1);
  const int e = blk * kMultiVecBlock + threadIdx.x;
  double pa = 0.0;
  if (e < L) {
    double g = 0.0, gu = 0.0;
    for (int k = 0; k < q.nsplit; k++) {
      g += q.Gp[(long)k * L + e];
      gu += q.Gp[(long)k * L + L - 1];
    }
    const double inv = 1.0 / (double)q.B;
    const double ap = (e == L - 1) ? 0.0 : g * inv - q.Om[e] * (gu * inv) + q.eps * q.p[e];
    q.Ap[e] = ap;
    pa = q.p[e] * ap;
  }
  double tot;
  if (reduce_last(pa, red, q.part, q.ticket, blk, nblk, tot) && threadIdx.x == 0) {
    CGScalars* sc = q.sc;
    sc->pAp = tot;
    if (!(fabs(tot) >= 1e-30)) {
      sc->breakdown = 1;
      sc->done = 1;
    } else {
      sc->alpha = sc->rs / tot;
    }
  }
}

// x += alpha p, r -= alpha Ap, rs_new = r.r; converged when
// rs_new <= tol^2 rs0, else beta = rs_new / rs (nqs_optimizer.c:113-136)
__device__ inline void mb_update(const MultiBatch& mb, int gblk) {
  __shared__ double red[kMultiVecBlock];
  const int qi = find_prob<F_VEC>(mb, gblk);
  const MultiProb& q = mb.probs[qi];
  if (cg_done(q)) return;
  const int blk = gblk - q.blk_vec, nblk = vec_blocks(q);
  const int L = (q.M + 1) * (q.N + 1);
  const int e = blk * kMultiVecBlock + threadIdx.x;
  const double a = q.sc->alpha;
  double rr = 0.0;
  if (e < L) {
    q.x[e] += a * q.p[e];
    const double rn = q.r[e] + (-a) * q.Ap[e];
    q.r[e] = rn;
    rr = rn * rn;
  }
  double tot;
  if (reduce_last(rr, red, q.part, q.ticket, blk, nblk, tot) && threadIdx.x == 0) {
    CGScalars* sc = q.sc;
    sc->iters += 1;
    sc->rs_new = tot;
    if (tot <= q.tol * q.tol * sc->rs0) {
      sc->converged = 1;
      sc->done = 1;
    } else {
      sc->beta = tot / sc->rs;
      sc->rs = tot;
      if (sc->iters >= q.max_it) sc->done = 1;
    }
  }
}

// p = r + beta p
__device__ inline void mb_dir(const MultiBatch& mb, int gblk) {
  const int qi = find_prob<F_VEC>(mb, gblk);
  const MultiProb& q = mb.probs[qi];
  if (cg_done(q)) return;
  const int L = (q.M + 1) * (q.N + 1);
  const int e = (gblk - q.blk_vec) * kMultiVecBlock + threadIdx.x;
  if (e < L) q.p[e] = q.r[e] + q.sc->beta * q.p[e];
}

// the CG launches: one block per work item
__global__ __launch_bounds__(256) void k_mb_zu(MultiBatch mb) { mb_zu(mb, blockIdx.x); }
__global__ __launch_bounds__(256) void k_mb_gemm_g(MultiBatch mb, int force) {
  mb_gemm_g(mb, blockIdx.x, force);
}
__global__ __launch_bounds__(kMultiVecBlock) void k_mb_finish(MultiBatch mb) {
  mb_finish(mb, blockIdx.x);
}
__global__ __launch_bounds__(kMultiVecBlock) void k_mb_update(MultiBatch mb) {
  mb_update(mb, blockIdx.x);
}
__global__ __launch_bounds__(kMultiVecBlock) void k_mb_dir(MultiBatch mb) { mb_dir(mb, blockIdx.x); }

// Every CG iteration of every problem in one cooperative launch: the five
// phases of an iteration walk their work items (the blocks of the launch
// sequence above, same device functions, so the same results bit for bit)
// grid-stride, with a grid barrier between phases.  Every block passes every
// barrier: the phase loops only skip items, and the loop ends for all blocks
// together when, after an iteration's last barrier, every problem reads done
// (its flags were written before that barrier).  The launch is cooperative,
// so it either has all its blocks resident or does not start.
// matvec_only: one iteration's A p, p.Ap and alpha only (idempotent at the
// CG start; the autotuner times it against the launch sequence's).
__global__ __launch_bounds__(256) void k_mb_cg_persist(MultiBatch mb, int iters, int matvec_only) {
  static_assert(kMultiVecBlock == 256, "one block size for every phase");
  cooperative_groups::grid_group grid = cooperative_groups::this_grid();
  for (int it = 0; it < iters; it++) {
    for (int b = blockIdx.x; b < mb.nblk_z; b += gridDim.x) mb_zu(mb, b);
    grid.sync();
    for (int b = blockIdx.x; b < mb.nblk_g; b += gridDim.x) mb_gemm_g(mb, b, 0);
    grid.sync();
    for (int b = blockIdx.x; b < mb.nblk_vec; b += gridDim.x) mb_finish(mb, b);
    if (matvec_only) break;   // uniform: a kernel argument
    grid.sync();
    for (int b = blockIdx.x; b < mb.nblk_vec; b += gridDim.x) mb_update(mb, b);
    grid.sync();
    for (int b = blockIdx.x; b < mb.nblk_vec; b += gridDim.x) mb_dir(mb, b);
    grid.sync();
    bool all = true;
    for (int i = 0; i < mb.n; i++) all = all && cg_done(mb.probs[i]);
    if (all) break;
  }
}

// theta -= lr delta in the reference layout [a | b | W], |delta|^2
__global__ __launch_bounds__(kMultiVecBlock) void k_mb_apply(MultiBatch mb) {
  __shared__ double red[kMultiVecBlock];
  const int qi = find_prob<F_VEC>(mb, blockIdx.x);
  const MultiProb& q = mb.probs[qi];
  const int blk = blockIdx.x - q.blk_vec, nblk = vec_blocks(q);
  const int N = q.N, M = q.M, N1 = N + 1, L = (M + 1) * N1;
  const int e = blk * kMultiVecBlock + threadIdx.x;
  double xx = 0.0;
  if (e < L - 1) {
    const int h = e / N1, i = e - h * N1;
    const long k = (h == M) ? i : (i == N ? (long)N + h : (long)N + M + (long)h * N + i);
    const double d = q.x[e];
    q.par[k] += -q.lr * d;
    xx = d * d;
  }
  double tot;
  if (reduce_last(xx, red, q.part, q.ticket, blk, nblk, tot) && threadIdx.x == 0) *q.norm2 = tot;
}

// one thread per problem
__global__ void k_mb_gather(MultiBatch mb, double* out) {
  const int qi = blockIdx.x * blockDim.x + threadIdx.x;
  if (qi >= mb.n) return;
  const MultiProb& q = mb.probs[qi];
  const CGScalars* sc = q.sc;
  unsigned long long a = 0, t = 0;
  for (int c = 0; c < q.C; c++) {
    a += q.acc[c];
    t += q.tot[c];
  }
  double* o = out + (long)qi * kMultiOut;
  o[0] = q.stats[0];
  o[1] = q.stats[1];
  o[2] = *q.norm2;
  o[3] = (double)sc->iters;
  o[4] = (double)sc->converged;
  o[5] = (double)sc->breakdown;
  o[6] = sc->pAp;
  o[7] = sc->rs_new;
  o[8] = sc->rs0;
  o[9] = t ? (double)a / (double)t : 0.0;
  o[10] = q.range_flag ? (double)*q.range_flag : 0.0;
}

}  // namespace

void multi_layout(MultiProb* probs, int n, MultiBatch& mb) {
  int prep = 0, z = 0, g = 0, g2 = 0, vec = 0;
  for (int i = 0; i < n; i++) {
    MultiProb& q = probs[i];
    const int M1 = q.M + 1, N1 = q.N + 1;
    q.blk_prep = prep;
    q.blk_z = z;
    q.blk_g = g;
    q.blk_g2 = g2;
    q.blk_vec = vec;
    prep += (q.B + 255) / 256;
    z += ((M1 + MT - 1) / MT) * ((q.B + MT - 1) / MT);
    const int gt = ((M1 + MT - 1) / MT) * ((N1 + MT - 1) / MT) * q.nsplit;
    g += gt;
    g2 += 2 * gt;
    vec += (M1 * N1 + kMultiVecBlock - 1) / kMultiVecBlock;
  }
  mb.n = n;
  mb.nblk_prep = prep;
  mb.nblk_z = z;
  mb.nblk_g = g;
  mb.nblk_g2 = g2;
  mb.nblk_vec = vec;
}

int launch_multi_prep(const MultiBatch& mb, hipStream_t st) {
  if (mb.n <= 0) return 0;
  KScope ks("multi_prep", st);
  hipLaunchKernelGGL(k_mb_prep, dim3((unsigned)mb.nblk_prep), dim3(256), 0, st, mb);
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_multi_force(const MultiBatch& mb, hipStream_t st) {
  if (mb.n <= 0) return 0;
  {
    KScope ks("multi_gemm_g", st);
    hipLaunchKernelGGL(k_mb_gemm_g, dim3((unsigned)mb.nblk_g2), dim3(256), 0, st, mb, 1);
  }
  hipLaunchKernelGGL(k_mb_force, dim3((unsigned)mb.nblk_vec), dim3(kMultiVecBlock), 0, st, mb);
  count_launch(2);
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_multi_cg_iter(const MultiBatch& mb, int matvec_only, hipStream_t st) {
  if (mb.n <= 0) return 0;
  {
    KScope ks("multi_gemm_z", st);
    hipLaunchKernelGGL(k_mb_zu, dim3((unsigned)mb.nblk_z), dim3(256), 0, st, mb);
  }
  {
    KScope ks("multi_gemm_g", st);
    hipLaunchKernelGGL(k_mb_gemm_g, dim3((unsigned)mb.nblk_g), dim3(256), 0, st, mb, 0);
  }
  {
    KScope ks("multi_vec", st);
    hipLaunchKernelGGL(k_mb_finish, dim3((unsigned)mb.nblk_vec), dim3(kMultiVecBlock), 0, st, mb);
    if (!matvec_only) {
      hipLaunchKernelGGL(k_mb_update, dim3((unsigned)mb.nblk_vec), dim3(kMultiVecBlock), 0, st,
                         mb);
      hipLaunchKernelGGL(k_mb_dir, dim3((unsigned)mb.nblk_vec), dim3(kMultiVecBlock), 0, st, mb);
    }
  }
  count_launch(matvec_only ? 3 : 5);
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

int multi_persist_grid(const MultiBatch& mb) {
  static int cap = -1;
  if (cap < 0) {
    int dev = 0, per_cu = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_mb_cg_persist, 256, 0) !=
            hipSuccess) {
      (void)hipGetLastError();
      cap = 0;
    } else {
      // two blocks per CU: the grid barrier's cost grows with the block
      // count (measured 768 us per CG iteration at 512 blocks, 926 at 768)
      cap = (per_cu < 2 ? per_cu : 2) * prop.multiProcessorCount;
    }
  }
  int need = mb.nblk_z;
  if (mb.nblk_g > need) need = mb.nblk_g;
  if (mb.nblk_vec > need) need = mb.nblk_vec;
  return need < cap ? need : cap;
}

int launch_multi_cg_persist(const MultiBatch& mb, int iters, hipStream_t st, int matvec_only) {
  if (mb.n <= 0 || iters <= 0) return 0;
  const int grid = multi_persist_grid(mb);
  if (grid <= 0) {
    fprintf(stderr, "nqs_hip: no occupancy for the persistent CG launch\n");
    return -1;
  }
  MultiBatch a = mb;
  int it = matvec_only ? 1 : iters, mv = matvec_only ? 1 : 0;
  void* args[] = {&a, &it, &mv};
  KScope ks("multi_cg", st);
  const hipError_t e = hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_mb_cg_persist),
                                                  dim3((unsigned)grid), dim3(256), args, 0, st);
  count_launch();
  if (e != hipSuccess) {
    (void)hipGetLastError();
    fprintf(stderr, "nqs_hip: persistent CG launch refused (%s)\n", hipGetErrorString(e));
    return -1;
  }
  return 0;
}

int launch_multi_apply(const MultiBatch& mb, hipStream_t st) {
  if (mb.n <= 0) return 0;
  hipLaunchKernelGGL(k_mb_apply, dim3((unsigned)mb.nblk_vec), dim3(kMultiVecBlock), 0, st, mb);
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_multi_gather(const MultiBatch& mb, double* out, hipStream_t st) {
  if (mb.n <= 0) return 0;
  hipLaunchKernelGGL(k_mb_gather, dim3((unsigned)((mb.n + 63) / 64)), dim3(64), 0, st, mb, out);
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

}  // namespace nqs
