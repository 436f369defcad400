// k_linalg.hip — FP64 reductions, CG pieces, MFMA GEMM, blocked Cholesky
// and triangular solves for the SR / MinSR updates.
//
// Reference semantics: src/nqs/nqs_optimizer.c — QGT product on uncentred O
// (:44-80), CG (:84-139), holomorphic QGT (:317-356), Cholesky (:1026-1043),
// triangular solves (:1046-1061), MinSR linear algebra (:1101-1177).  Every
// reduction here runs in a fixed order, so results are reproducible run to
// run; they differ from the reference's sequential sums only by rounding.
#include "nqs_kernels.h"

namespace nqs {

typedef double v4d __attribute__((ext_vector_type(4)));

// --------------------------------------------------------------------------
// column sums  out[k] = alpha * sum_b O[b][k] w[b]   (two fixed-order stages)
// --------------------------------------------------------------------------
constexpr int kColChunk = 128;

size_t colsum_work_elems(long B, long P) {
  long nch = (B + kColChunk - 1) / kColChunk;
  return (size_t)nch * (size_t)P;
}

__global__ __launch_bounds__(256) void k_colsum_partial(const double* __restrict__ O, long B,
                                                        long P, const double* __restrict__ w,
                                                        double* __restrict__ work) {
  const long k = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long c = blockIdx.y;
  if (k >= P) return;
  const long b0 = c * kColChunk, b1 = (b0 + kColChunk < B) ? b0 + kColChunk : B;
  double s = 0.0;
  if (w) {
    for (long b = b0; b < b1; b++) s += O[b * P + k] * w[b];
  } else {
    for (long b = b0; b < b1; b++) s += O[b * P + k];
  }
  work[c * P + k] = s;
}

__global__ __launch_bounds__(256) void k_colsum_final(const double* __restrict__ work, long nch,
                                                      long P, double alpha, double* __restrict__ out) {
  const long k = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (k >= P) return;
  double s = 0.0;
  for (long c = 0; c < nch; c++) s += work[c * P + k];
  out[k] = alpha * s;
}

int launch_colsum(const double* O, long B, long P, const double* w, double alpha, double* out,
                  double* work, hipStream_t st) {
  if (P <= 0) return 0;
  const long nch = (B + kColChunk - 1) / kColChunk;
  if (nch > 65535) return -1;
  dim3 g1((unsigned)((P + 255) / 256), (unsigned)nch);
  KScope ks("colsum", st);
  hipLaunchKernelGGL(k_colsum_partial, g1, dim3(256), 0, st, O, B, P, w, work);
  hipLaunchKernelGGL(k_colsum_final, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, st, work,
                     nch, P, alpha, out);
  count_launch(2);
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

// --------------------------------------------------------------------------
// row dots  u[b] = sum_k O[b][k] v[k]  (one wave per row)
// --------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_rowdot(const double* __restrict__ O, long B, long P,
                                                const double* __restrict__ v, double* __restrict__ u) {
  const long row = blockIdx.x * 4L + threadIdx.x / 64;
  const int lane = threadIdx.x & 63;
  if (row >= B) return;
  const double* o = O + row * P;
  double s = 0.0;
  for (long k = lane; k < P; k += 64) s += o[k] * v[k];
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m, 64);
  if (lane == 0) u[row] = s;
}

int launch_rowdot(const double* O, long B, long P, const double* v, double* u, hipStream_t st) {
  if (B <= 0) return 0;
  hipLaunchKernelGGL(k_rowdot, dim3((unsigned)((B + 3) / 4)), dim3(256), 0, st, O, B, P, v, u);
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

// --------------------------------------------------------------------------
// scalar sum / dot into a device scalar (one block, fixed tree)
// --------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_sum(const double* __restrict__ x, long n,
                                              const double* __restrict__ y, double alpha,
                                              double* out) {
  __shared__ double red[1024];
  double s = 0.0;
  if (y) {
    for (long i = threadIdx.x; i < n; i += blockDim.x) s += x[i] * y[i];
  } else {
    for (long i = threadIdx.x; i < n; i += blockDim.x) s += x[i];
  }
  s = block_sum(s, red);
  if (threadIdx.x == 0) *out = alpha * s;
}

int launch_sum(const double* x, long n, const double* y, double alpha, double* out,
               hipStream_t st) {
  hipLaunchKernelGGL(k_sum, dim3(1), dim3(1024), 0, st, x, n, y, alpha, out);
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

__global__ void k_center(double* O, long B, long P, const double* __restrict__ mean) {
  const long total = B * P;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
       e += (long)gridDim.x * blockDim.x)
    O[e] -= mean[e % P];
}

int launch_center(double* O, long B, long P, const double* mean, hipStream_t st) {
  long total = B * P;
  long blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_center, dim3((unsigned)blocks), dim3(256), 0, st, O, B, P, mean);
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

__global__ void k_axpy(double* y, const double* __restrict__ x, double a, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    y[i] += a * x[i];
}
int launch_axpy(double* y, const double* x, double a, long n, hipStream_t st) {
  long blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_axpy, dim3((unsigned)blocks), dim3(256), 0, st, y, x, a, n);
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}
__global__ void k_fill(double* x, double v, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    x[i] = v;
}
int launch_fill(double* x, double v, long n, hipStream_t st) {
  long blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)blocks), dim3(256), 0, st, x, v, n);
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

// --------------------------------------------------------------------------
// QGT finish: out = scale * (c1 - m1 u1 + c2 - m2 u2) + eps * v
// (nqs_optimizer.c:70-78 real with scale 1 and no second term,
//  :346-353 holomorphic with scale 2)
// --------------------------------------------------------------------------
__global__ void k_qgt_finish(double* out, const double* c1, const double* m1, const double* u1,
                             const double* c2, const double* m2, const double* u2, double scale,
                             double eps, const double* v, long P) {
  const long k = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (k >= P) return;
  double a = c1[k] - m1[k] * u1[0];
  if (c2) a = (c1[k] + c2[k]) - m1[k] * u1[0] - m2[k] * u2[0];
  out[k] = scale * a + eps * v[k];
}
int launch_qgt_finish(double* out, const double* c1, const double* m1, const double* u1mean,
                      const double* c2, const double* m2, const double* u2mean, double scale,
                      double eps, const double* v, long P, hipStream_t st) {
  hipLaunchKernelGGL(k_qgt_finish, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, st, out, c1,
                     m1, u1mean, c2, m2, u2mean, scale, eps, v, P);
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

// --------------------------------------------------------------------------
// CG scalar / vector steps (x0 = 0; stop when |r|^2 <= tol^2 |r0|^2, on
// |pAp| < 1e-30, or at the iteration cap — nqs_optimizer.c:84-139)
// --------------------------------------------------------------------------
__global__ void k_cg_init(const double* F, long P, double* x, double* r, double* p) {
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < P; k += (long)gridDim.x * blockDim.x) {
    x[k] = 0.0;
    r[k] = F[k];
    p[k] = F[k];
  }
}
__global__ void k_cg_set0(CGScalars* sc) {
  sc->rs0 = sc->rs;
  sc->iters = 0;
  sc->converged = 0;
  sc->breakdown = 0;
  sc->done = 0;
}
int launch_cg_init(const double* F, long P, double* x, double* r, double* p, CGScalars* sc,
                   hipStream_t st) {
  long blocks = (P + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_cg_init, dim3((unsigned)blocks), dim3(256), 0, st, F, P, x, r, p);
  count_launch();
  if (launch_sum(F, P, F, 1.0, &sc->rs, st)) return -1;
  hipLaunchKernelGGL(k_cg_set0, dim3(1), dim3(1), 0, st, sc);
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

__global__ void k_cg_alpha(CGScalars* sc) {
  if (sc->done) return;
  const double pAp = sc->pAp;
  if (!(fabs(pAp) >= 1e-30)) {
    sc->breakdown = 1;
    sc->done = 1;
    return;
  }
  sc->alpha = sc->rs / pAp;
}
int launch_cg_alpha(CGScalars* sc, const double* p, const double* Ap, long P, hipStream_t st) {
  if (launch_sum(p, P, Ap, 1.0, &sc->pAp, st)) return -1;
  hipLaunchKernelGGL(k_cg_alpha, dim3(1), dim3(1), 0, st, sc);
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

__global__ void k_cg_xr(const CGScalars* sc, double* x, double* r, const double* p,
                        const double* Ap, long P) {
  if (sc->done) return;
  const double a = sc->alpha;
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < P; k += (long)gridDim.x * blockDim.x) {
    x[k] += a * p[k];
    r[k] += -a * Ap[k];
  }
}
__global__ void k_cg_check(CGScalars* sc, double tol) {
  if (sc->done) return;
  sc->iters += 1;
  const double rn = sc->rs_new;
  if (rn <= tol * tol * sc->rs0) {
    sc->converged = 1;
    sc->done = 1;
    return;
  }
  sc->beta = rn / sc->rs;
  sc->rs = rn;
}
int launch_cg_update(CGScalars* sc, double* x, double* r, const double* p, const double* Ap,
                     long P, double tol, hipStream_t st) {
  long blocks = (P + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_cg_xr, dim3((unsigned)blocks), dim3(256), 0, st, sc, x, r, p, Ap, P);
  count_launch();
  if (launch_sum(r, P, r, 1.0, &sc->rs_new, st)) return -1;
  hipLaunchKernelGGL(k_cg_check, dim3(1), dim3(1), 0, st, sc, tol);
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}
__global__ void k_cg_dir(const CGScalars* sc, double* p, const double* r, long P) {
  if (sc->done) return;
  const double b = sc->beta;
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < P; k += (long)gridDim.x * blockDim.x)
    p[k] = r[k] + b * p[k];
}
int launch_cg_direction(CGScalars* sc, double* p, const double* r, long P, hipStream_t st) {
  long blocks = (P + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_cg_dir, dim3((unsigned)blocks), dim3(256), 0, st, sc, p, r, P);
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

// --------------------------------------------------------------------------
// C = alpha A B^T + beta C with v_mfma_f64_16x16x4f64.
// Workgroup tile 64x64 (4 waves, 32x32 each = 2x2 MFMA blocks), K step 16
// staged in LDS.  MFMA lane layout (16x16x4, FP64, checked on gfx950): lane
// l supplies A[l%16][l/16] and B^T[l/16][l%16]; result element r of lane l
// is D[(l/16) + 4r][l%16].  `lower`: the tile grid covers only tiles with
// tn <= tm and elements strictly above the diagonal are never written.
// `skip`: device flag; the kernel does nothing when it is non-zero.
// --------------------------------------------------------------------------
constexpr int GT = 64, GKS = 16;

__device__ inline void lower_tile(int t, int& tm, int& tn) {
  int m = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while (m * (m + 1) / 2 > t) m--;
  while ((m + 1) * (m + 2) / 2 <= t) m++;
  tm = m;
  tn = t - m * (m + 1) / 2;
}

__global__ __launch_bounds__(256) void k_gemm_nt(int M, int N, int K, double alpha,
                                                 const double* A, long lda,
                                                 const double* __restrict__ Bm, long ldb,
                                                 double beta, double* C, long ldc,
                                                 int lower, int tiles_n, const int* skip) {
  if (skip && *skip) return;
  __shared__ double As[GT][GKS + 1];
  __shared__ double Bs[GT][GKS + 1];
  int tm, tn;
  if (lower) lower_tile(blockIdx.x, tm, tn);
  else { tm = blockIdx.x / tiles_n; tn = blockIdx.x % tiles_n; }
  const int row0 = tm * GT, col0 = tn * GT;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  v4d acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = (v4d){0.0, 0.0, 0.0, 0.0};
  const int lr = lane & 15, lk = lane >> 4;
  for (int k0 = 0; k0 < K; k0 += GKS) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int e = tid + 256 * q;
      const int r = e >> 4, kk = e & 15;
      const int gk = k0 + kk;
      const int ga = row0 + r, gb = col0 + r;
      As[r][kk] = (ga < M && gk < K) ? A[(long)ga * lda + gk] : 0.0;
      Bs[r][kk] = (gb < N && gk < K) ? Bm[(long)gb * ldb + gk] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < GKS; kk += 4) {
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
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int gr = row0 + wr + 16 * i + lk + 4 * r;
        const int gc = col0 + wc + 16 * j + lr;
        if (gr < M && gc < N && (!lower || gc <= gr)) {
          double* c = C + (long)gr * ldc + gc;
          *c = (beta == 0.0) ? alpha * acc[i][j][r] : alpha * acc[i][j][r] + beta * (*c);
        }
      }
}

static int gemm_nt_impl(int M, int N, int K, double alpha, const double* A, long lda,
                        const double* B, long ldb, double beta, double* C, long ldc, int lower,
                        const int* skip, hipStream_t st) {
  if (M <= 0 || N <= 0) return 0;
  const int tm = (M + GT - 1) / GT, tn = (N + GT - 1) / GT;
  long tiles = lower ? (long)tm * (tm + 1) / 2 : (long)tm * tn;
  if (lower && M != N) return -1;
  KScope ks(skip ? (lower ? "gemm_potrf" : "gemm_trsm") : "gemm", st);
  hipLaunchKernelGGL(k_gemm_nt, dim3((unsigned)tiles), dim3(256), 0, st, M, N, K, alpha, A, lda,
                     B, ldb, beta, C, ldc, lower, tn, skip);
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_gemm_nt(int M, int N, int K, double alpha, const double* A, long lda, const double* B,
                   long ldb, double beta, double* C, long ldc, int lower, hipStream_t st) {
  return gemm_nt_impl(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, lower, nullptr, st);
}

// --------------------------------------------------------------------------
// Cholesky, right-looking with 64-wide panels.  The panel kernel factors the
// diagonal block in LDS (right-looking, one column per step) and also
// inverts it; the rows below are then A21 <- A21 Linv^T on the MFMA GEMM and
// the trailing matrix is updated on its lower triangle by the same GEMM.
// The inverted diagonal blocks are kept for the triangular solves.
// --------------------------------------------------------------------------
constexpr int NB = 64;

size_t potrf_linv_elems(int n) { return (size_t)((n + NB - 1) / NB) * NB * NB; }

__global__ __launch_bounds__(256) void k_potrf_panel(double* A, int n, int k0, int nb,
                                                     double* Linv, int* fail) {
  if (*fail) return;
  // lower triangle (incl. diagonal) holds L; the strict upper triangle holds
  // the inverse transposed (X[i][c] at L[c][i], c < i); column NB holds the
  // inverse's diagonal
  __shared__ double L[NB][NB + 1];
  __shared__ int bad;
  const int tid = threadIdx.x;
  for (int e = tid; e < nb * nb; e += blockDim.x) {
    const int i = e / nb, j = e % nb;
    if (j <= i) L[i][j] = A[(long)(k0 + i) * n + k0 + j];
  }
  if (tid == 0) bad = -1;
  __syncthreads();
  int jdone = nb;
  for (int j = 0; j < nb; j++) {
    if (tid == 0) {
      const double d = L[j][j];
      if (!(d > 0.0)) bad = j;
      else L[j][j] = sqrt(d);
    }
    __syncthreads();
    if (bad >= 0) { jdone = bad; break; }
    const double inv = 1.0 / L[j][j];
    for (int i = j + 1 + tid; i < nb; i += blockDim.x) L[i][j] *= inv;
    __syncthreads();
    const int m = nb - 1 - j;
    for (int e = tid; e < m * m; e += blockDim.x) {
      const int i = j + 1 + e / m, c = j + 1 + e % m;
      if (c <= i) L[i][c] -= L[i][j] * L[c][j];
    }
    __syncthreads();
  }
  for (int e = tid; e < nb * nb; e += blockDim.x) {
    const int i = e / nb, j = e % nb;
    if (j <= i && j < jdone) A[(long)(k0 + i) * n + k0 + j] = L[i][j];
  }
  if (jdone < nb) {
    if (tid == 0) *fail = k0 + jdone + 1;
    return;
  }
  // invert the diagonal block: thread c solves L x = e_c
  if (tid < nb) {
    const int c = tid;
    L[c][NB] = 1.0 / L[c][c];
    for (int i = c + 1; i < nb; i++) {
      double s = 0.0;
      for (int k = c; k < i; k++) s += L[i][k] * (k == c ? L[c][NB] : L[c][k]);
      L[c][i] = -s / L[i][i];
    }
  }
  __syncthreads();
  double* X = Linv + (size_t)(k0 / NB) * NB * NB;
  for (int e = tid; e < NB * NB; e += blockDim.x) {
    const int i = e / NB, c = e % NB;
    double v = 0.0;
    if (i < nb && c < nb) {
      if (c == i) v = L[i][NB];
      else if (c < i) v = L[c][i];
    }
    X[e] = v;
  }
}

// rows i >= r0: A[i][k0 + j] for j < ncols solved against the diagonal block
// (x L^T = a, forward over j); thread per row.  Used to finish the failing
// panel's factored columns when bringing a failed factorisation to the
// unblocked algorithm's state.
__global__ __launch_bounds__(256) void k_trsm_rows(double* A, int n, int k0, int ncols, int r0) {
  __shared__ double L[NB][NB + 1];
  for (int e = threadIdx.x; e < ncols * ncols; e += blockDim.x) {
    const int i = e / ncols, j = e % ncols;
    if (j <= i) L[i][j] = A[(long)(k0 + i) * n + k0 + j];
  }
  __syncthreads();
  const long i = (long)r0 + blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double* row = A + i * n + k0;
  double x[NB];
#pragma unroll
  for (int j = 0; j < NB; j++) x[j] = (j < ncols) ? row[j] : 0.0;
#pragma unroll
  for (int j = 0; j < NB; j++) {
    if (j < ncols) {
      double s = x[j];
#pragma unroll
      for (int k = 0; k < NB; k++)
        if (k < j) s -= x[k] * L[j][k];
      x[j] = s / L[j][j];
    }
  }
#pragma unroll
  for (int j = 0; j < NB; j++)
    if (j < ncols) row[j] = x[j];
}

// lower-triangle columns >= jf of A restored from the lower triangle of orig
__global__ void k_restore_cols(double* A, const double* orig, int n, int jf) {
  const long total = (long)n * n;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
       e += (long)gridDim.x * blockDim.x) {
    const int i = (int)(e / n), j = (int)(e % n);
    if (j >= jf && j <= i) A[e] = orig[e];
  }
}

__global__ void k_zero_int(int* p) { *p = 0; }

int launch_potrf(double* A, int n, double* Linv, int* fail_dev, hipStream_t st) {
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, st, fail_dev);
  count_launch();
  for (int k0 = 0; k0 < n; k0 += NB) {
    const int nb = (n - k0) < NB ? (n - k0) : NB;
    {
      KScope ks("potrf_diag", st);
      hipLaunchKernelGGL(k_potrf_panel, dim3(1), dim3(256), 0, st, A, n, k0, nb, Linv, fail_dev);
    }
    count_launch();
    const int r0 = k0 + nb;
    if (r0 < n) {
      const int rows = n - r0;
      double* panel = A + (long)r0 * n + k0;
      // A21 <- A21 Linv^T (in place: each workgroup reads its rows before writing)
      {
        KScope ks("trsm", st);
        if (gemm_nt_impl(rows, nb, nb, 1.0, panel, n, Linv + (size_t)(k0 / NB) * NB * NB, NB, 0.0,
                         panel, n, 0, fail_dev, st))
          return -1;
      }
      if (gemm_nt_impl(rows, rows, nb, -1.0, panel, n, panel, n, 1.0, A + (long)r0 * n + r0, n, 1,
                       fail_dev, st))
        return -1;
    }
  }
  NQS_HIP_CHECK(hipGetLastError());
  int fail = 0;
  NQS_HIP_CHECK(hipMemcpyAsync(&fail, fail_dev, sizeof(int), hipMemcpyDeviceToHost, st));
  NQS_HIP_CHECK(hipStreamSynchronize(st));
  return fail;
}

int launch_potrf_repair(double* A, int n, int jf, const double* orig, hipStream_t st) {
  // finish the failing panel's factored columns below the block, then
  // restore the untouched columns from the input
  const int k0 = (jf / NB) * NB;
  const int ncols = jf - k0;
  const int nb = (n - k0) < NB ? (n - k0) : NB;
  const int r0 = k0 + nb;
  if (ncols > 0 && r0 < n) {
    hipLaunchKernelGGL(k_trsm_rows, dim3((n - r0 + 255) / 256), dim3(256), 0, st, A, n, k0, ncols, r0);
    count_launch();
  }
  long blocks = ((long)n * n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_restore_cols, dim3((unsigned)blocks), dim3(256), 0, st, A, orig, n, jf);
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

// --------------------------------------------------------------------------
// triangular solves L L^T x = b with the inverted diagonal blocks: one
// launch per block and direction; every workgroup recomputes the block's
// solution (64x64 product) and workgroup 0 publishes it
// --------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_trsv_fwd(const double* L, const double* Linv, int n,
                                                  int k0, int nb, double* w, double* x) {
  __shared__ double wb[NB], xb[NB];
  const int tid = threadIdx.x;
  if (tid < nb) wb[tid] = w[k0 + tid];
  __syncthreads();
  const double* X = Linv + (size_t)(k0 / NB) * NB * NB;
  if (tid < nb) {
    double s = 0.0;
    for (int c = 0; c <= tid; c++) s += X[tid * NB + c] * wb[c];
    xb[tid] = s;
    if (blockIdx.x == 0) x[k0 + tid] = s;
  }
  __syncthreads();
  const long i = (long)k0 + nb + blockIdx.x * (long)blockDim.x + tid;
  if (i < n) {
    const double* row = L + i * n + k0;
    double s = 0.0;
    for (int c = 0; c < nb; c++) s += row[c] * xb[c];
    w[i] -= s;
  }
}

__global__ __launch_bounds__(256) void k_trsv_bwd(const double* L, const double* Linv, int n,
                                                  int k0, int nb, double* z, double* x) {
  __shared__ double zb[NB], xb[NB];
  const int tid = threadIdx.x;
  if (tid < nb) zb[tid] = z[k0 + tid];
  __syncthreads();
  const double* X = Linv + (size_t)(k0 / NB) * NB * NB;
  if (tid < nb) {
    double s = 0.0;
    for (int i = tid; i < nb; i++) s += X[i * NB + tid] * zb[i];
    xb[tid] = s;
    if (blockIdx.x == 0) x[k0 + tid] = s;
  }
  __syncthreads();
  const long j = blockIdx.x * (long)blockDim.x + tid;
  if (j < k0) {
    double s = 0.0;
    for (int r = 0; r < nb; r++) s += L[(long)(k0 + r) * n + j] * xb[r];
    z[j] -= s;
  }
}

int launch_potrs(const double* L, const double* Linv, int n, const double* b, double* x,
                 double* w, hipStream_t st) {
  NQS_HIP_CHECK(hipMemcpyAsync(w, b, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, st));
  for (int k0 = 0; k0 < n; k0 += NB) {
    const int nb = (n - k0) < NB ? (n - k0) : NB;
    const int rows = n - k0 - nb;
    const int grid = rows > 0 ? (rows + 255) / 256 : 1;
    hipLaunchKernelGGL(k_trsv_fwd, dim3(grid), dim3(256), 0, st, L, Linv, n, k0, nb, w, x);
    count_launch();
  }
  // x now holds z = L^-1 b; back substitution in place on w = z
  NQS_HIP_CHECK(hipMemcpyAsync(w, x, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, st));
  const int last = ((n - 1) / NB) * NB;
  for (int k0 = last; k0 >= 0; k0 -= NB) {
    const int nb = (n - k0) < NB ? (n - k0) : NB;
    const int grid = k0 > 0 ? (k0 + 255) / 256 : 1;
    hipLaunchKernelGGL(k_trsv_bwd, dim3(grid), dim3(256), 0, st, L, Linv, n, k0, nb, w, x);
    count_launch();
  }
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

__global__ void k_sub_scalar(double* y, const double* x, long n, const double* s) {
  const double v = *s;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    y[i] = x[i] - v;
}
int launch_sub_scalar(double* y, const double* x, long n, const double* s, hipStream_t st) {
  long blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_sub_scalar, dim3((unsigned)blocks), dim3(256), 0, st, y, x, n, s);
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

__global__ void k_add_diag(double* A, int n, long ld, double v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) A[(long)i * ld + i] += v;
}
int launch_add_diag(double* A, int n, long ld, double v, hipStream_t st) {
  hipLaunchKernelGGL(k_add_diag, dim3((n + 255) / 256), dim3(256), 0, st, A, n, ld, v);
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}
__global__ void k_copy_diag(const double* A, int n, double* d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = A[(long)i * n + i];
}
int launch_copy_diag(const double* A, int n, double* d, hipStream_t st) {
  hipLaunchKernelGGL(k_copy_diag, dim3((n + 255) / 256), dim3(256), 0, st, A, n, d);
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}
__global__ void k_mirror_lower(double* A, int n) {
  const long total = (long)n * n;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
       e += (long)gridDim.x * blockDim.x) {
    const int i = (int)(e / n), j = (int)(e % n);
    if (j > i) A[e] = A[(long)j * n + i];
  }
}
int launch_mirror_lower(double* A, int n, hipStream_t st) {
  long blocks = ((long)n * n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_mirror_lower, dim3((unsigned)blocks), dim3(256), 0, st, A, n);
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

}  // namespace nqs
