// k_model.hip — ansatz tables, batched forward pass and O_k generation.
//
// Reference semantics: RBM / complex RBM amplitude src/nqs/nqs_ansatz.c:
// 277-355 (log psi = sum_i a_i s_i + sum_h log 2cosh(theta_h),
// theta_h = b_h + sum_i W_hi s_i), gradients :885-995 (real-projected) and
// :1025-1091 (holomorphic); mean-field is the M = 0 case.
#include "nqs_kernels.h"

namespace nqs {

// --------------------------------------------------------------------------
// tables: a, b, W^T, e^{2W} and e^{-2W} in both h-major and site-major order
// --------------------------------------------------------------------------
template <typename R>
__global__ void k_prep(const double* __restrict__ par, int kind, int N, int M,
                       R* a_re, R* a_im, R* b_re, R* b_im, R* WT_re, R* WT_im, R* EP_re,
                       R* EP_im, R* EM_re, R* EM_im, R* EPT_re, R* EPT_im, R* EMT_re,
                       R* EMT_im, int* range_flag) {
  const long Pr = (long)N + M + (long)M * N;
  const int cplx = (kind == 6);
  const long total = (long)M * N;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total + N + M;
       idx += (long)gridDim.x * blockDim.x) {
    if (idx < N) {
      a_re[idx] = (R)par[idx];
      if (cplx) a_im[idx] = (R)par[Pr + idx];
      continue;
    }
    if (idx < (long)N + M) {
      long h = idx - N;
      b_re[h] = (R)par[N + h];
      if (cplx) b_im[h] = (R)par[Pr + N + h];
      continue;
    }
    long e = idx - N - M;       // e = h*N + i
    int h = (int)(e / N), i = (int)(e % N);
    double wr = par[N + M + e];
    double wi = cplx ? par[Pr + N + M + e] : 0.0;
    long et = (long)i * M + h;
    WT_re[et] = (R)wr;
    // beyond the served range the entries are not formed (the launcher
    // reports the weight and refuses the tables)
    if (!(fabs(wr) <= RatioRange<R>::max_w)) {
      *range_flag = 1;
      wr = 0.0;
    }
    // e^{+-2w} for complex w = x + iy: e^{+-2x} (cos 2y +- i sin 2y)
    const double ep = exp(2.0 * wr), em = exp(-2.0 * wr);
    double sy = 0.0, cy = 1.0;
    if (cplx) sincos(2.0 * wi, &sy, &cy);
    const R pr = (R)(ep * cy), mr = (R)(em * cy);
    EP_re[e] = pr;  EM_re[e] = mr;  EPT_re[et] = pr;  EMT_re[et] = mr;
    if (cplx) {
      const R pi = (R)(ep * sy), mi = (R)(-em * sy);
      WT_im[et] = (R)wi;
      EP_im[e] = pi;  EM_im[e] = mi;  EPT_im[et] = pi;  EMT_im[et] = mi;
    }
  }
}

__global__ void k_max_abs(const double* __restrict__ x, long n, double* __restrict__ out) {
  __shared__ double red[256];
  double m = 0.0;
  for (long e = threadIdx.x; e < n; e += blockDim.x) m = fmax(m, fabs(x[e]));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

__global__ void k_flag_reset(int* flag) { *flag = 0; }

// check = 1: wait for the tables and set mb.ratio_ok (0: a weight lies
// beyond the served range); check = 0 (inside a captured graph): the flag is
// left in mb.range_flag for the caller to read after the stream synchronises
// (mb.ratio_ok = -1).  a, b and W^T serve amplitudes either way.
template <typename R>
int launch_prep(const double* d_par, int kind, int N, int M, ModelBuffers<R>& mb,
                hipStream_t st, int check) {
  const int cplx = (kind == 6);
  const size_t MN = (size_t)M * N;
  const size_t need = 2 * (size_t)N + 2 * (size_t)M + 10 * MN + 16;
  if (mb.count < need) {
    if (mb.base) (void)hipFree(mb.base);
    mb.base = nullptr;
    NQS_HIP_CHECK(hipMalloc(&mb.base, need * sizeof(R)));
    mb.count = need;
  }
  if (!mb.range_flag) NQS_HIP_CHECK(hipMalloc(&mb.range_flag, 2 * sizeof(double)));
  R* p = mb.base;
  auto take = [&](size_t n) { R* q = p; p += n; return q; };
  R *a_re = take(N), *a_im = take(N), *b_re = take(M), *b_im = take(M);
  R *WT_re = take(MN), *WT_im = take(MN), *EP_re = take(MN), *EP_im = take(MN);
  R *EM_re = take(MN), *EM_im = take(MN), *EPT_re = take(MN), *EPT_im = take(MN);
  R *EMT_re = take(MN), *EMT_im = take(MN);
  ModelView<R>& v = mb.view;
  v.N = N;
  v.M = M;
  v.cplx = cplx;
  v.a_re = a_re;  v.a_im = cplx ? a_im : nullptr;
  v.b_re = b_re;  v.b_im = cplx ? b_im : nullptr;
  v.WT_re = WT_re;  v.WT_im = cplx ? WT_im : nullptr;
  v.EP_re = EP_re;  v.EP_im = cplx ? EP_im : nullptr;
  v.EM_re = EM_re;  v.EM_im = cplx ? EM_im : nullptr;
  v.EPT_re = EPT_re;  v.EPT_im = cplx ? EPT_im : nullptr;
  v.EMT_re = EMT_re;  v.EMT_im = cplx ? EMT_im : nullptr;
  hipLaunchKernelGGL(k_flag_reset, dim3(1), dim3(1), 0, st, mb.range_flag);
  count_launch();
  long total = (long)MN + N + M;
  int blocks = (int)((total + 255) / 256);
  if (blocks > 4096) blocks = 4096;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_prep<R>, dim3(blocks), dim3(256), 0, st, d_par, kind, N, M,
                     a_re, a_im, b_re, b_im, WT_re, WT_im, EP_re, EP_im, EM_re, EM_im,
                     EPT_re, EPT_im, EMT_re, EMT_im, mb.range_flag);
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  mb.valid = 1;
  mb.ratio_ok = -1;
  if (!check) return 0;
  int flag = 0;
  NQS_HIP_CHECK(hipMemcpyAsync(&flag, mb.range_flag, sizeof(int), hipMemcpyDeviceToHost, st));
  NQS_HIP_CHECK(hipStreamSynchronize(st));
  mb.ratio_ok = flag ? 0 : 1;
  return 0;
}

// the largest |Re W| (real block [N + M, N + M + M N) of the parameters),
// reported on stderr as the reason the ratio tables are refused
template <typename R>
int report_ratio_range(const double* d_par, int N, int M, int* scratch, hipStream_t st) {
  double* dmax = reinterpret_cast<double*>(scratch) + 1;
  hipLaunchKernelGGL(k_max_abs, dim3(1), dim3(256), 0, st, d_par + N + M, (long)M * N, dmax);
  count_launch();
  double wmax = 0.0;
  NQS_HIP_CHECK(hipMemcpyAsync(&wmax, dmax, sizeof(double), hipMemcpyDeviceToHost, st));
  NQS_HIP_CHECK(hipStreamSynchronize(st));
  fprintf(stderr, "nqs_hip: a weight |Re W| = %.6g lies beyond the range of the %s amplitude-ratio "
                  "tables (|Re W| <= %g); moves and local energies are refused\n", wmax,
          sizeof(R) == 8 ? "FP64" : "FP32", RatioRange<R>::max_w);
  return 0;
}

template <typename R>
void free_model(ModelBuffers<R>& mb) {
  if (mb.base) (void)hipFree(mb.base);
  if (mb.range_flag) (void)hipFree(mb.range_flag);
  mb.range_flag = nullptr;
  mb.base = nullptr;
  mb.count = 0;
  mb.valid = 0;
}

// --------------------------------------------------------------------------
// forward pass: one wave per configuration (4 per 256-thread block), lane h
// owns hidden units h, h + 64, ... (or half a wave per configuration, lane
// h of the half units h, h + 32, ...); the sums over hidden units are xor
// butterflies inside the wave, so there is no block barrier.
// --------------------------------------------------------------------------

// STAGED: the transposed weights and hidden biases are first copied into
// LDS (when they fit, kFwdStage bytes), and the workgroup's waves take
// several configurations each, so that the site loop reads LDS instead of
// re-reading the weights through the caches for every configuration.
constexpr size_t kFwdStage = 48 * 1024;

// LPC lanes per configuration (64, or 32 with two configurations per wave
// when the hidden units fill 32-lane passes better); the sums over hidden
// units are butterflies within the LPC lanes.
template <int LPC>
__device__ inline double group_sum(double v) {
#pragma unroll
  for (int m = LPC / 2; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

template <typename R, bool CPLX, bool STAGED, int LPC>
__global__ __launch_bounds__(256) void k_forward(ModelView<R> m, const int* __restrict__ spins,
                                                 int B, R* __restrict__ t_re,
                                                 R* __restrict__ t_im, double* la,
                                                 double* ar, R* __restrict__ q_re,
                                                 R* __restrict__ q_im) {
  const int N = m.N, M = m.M;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  extern __shared__ char fwd_lds[];
  // [N][M] re, [N][M] im, then b_re [M], b_im [M]
  R* const sw = reinterpret_cast<R*>(fwd_lds);
  const R* w_re = m.WT_re;
  const R* w_im = m.WT_im;
  const R* bb_re = m.b_re;
  const R* bb_im = m.b_im;
  if (STAGED) {
    const int nw = N * M;
    for (int e = threadIdx.x; e < nw; e += 256) {
      sw[e] = m.WT_re[e];
      if (CPLX) sw[nw + e] = m.WT_im[e];
    }
    R* const sb = sw + (CPLX ? 2 : 1) * nw;
    for (int e = threadIdx.x; e < M; e += 256) {
      sb[e] = m.b_re[e];
      if (CPLX) sb[M + e] = m.b_im[e];
    }
    __syncthreads();
    w_re = sw;
    w_im = sw + nw;
    bb_re = sb;
    bb_im = sb + M;
  }
  constexpr int CPW = 64 / LPC;
  const int sub = lane / LPC, ll = lane % LPC;
  for (int b = (blockIdx.x * 4 + wave) * CPW + sub; b < B; b += gridDim.x * 4 * CPW) {
    const int* sp = spins + (size_t)b * N;
    double lsum = 0.0, asum = 0.0;
    for (int h = ll; h < M; h += LPC) {
      // theta_h = b_h + sum_i W_hi s_i, accumulated in site order as in
      // nqs_ansatz.c:287-290 (s = +-1 makes every product exact)
      R xr = bb_re[h];
      R xi = CPLX ? bb_im[h] : (R)0;
      for (int i = 0; i < N; i++) {
        const R s = (R)sp[i];
        xr += w_re[(size_t)i * M + h] * s;
        if (CPLX) xi += w_im[(size_t)i * M + h] * s;
      }
      if (t_re) {
        if (CPLX) {
          R tr, ti;
          ctanh<R>(xr, xi, tr, ti);
          t_re[(size_t)b * M + h] = tr;
          t_im[(size_t)b * M + h] = ti;
        } else {
          t_re[(size_t)b * M + h] = tanh(xr);
        }
      }
      if (q_re) {
        R sr, si;
        int is_p;
        logistic_small<R>(xr, xi, sr, si, is_p);
        q_re[(size_t)b * M + h] = sr;
        if (CPLX) q_im[(size_t)b * M + h] = si;
      }
      if (la) {
        if (CPLX && ar) {
          double l1, a1;
          logcosh2_c((double)xr, (double)xi, l1, a1);
          lsum += l1;
          asum += a1;
        } else if (CPLX) {
          // log|2 cosh z| only: |x| + log|(1 + e) cos y + i sgn(x) (1 - e) sin y|,
          // e = exp(-2|x|) (both parts are bounded by 2, no scaling needed)
          const double ax = fabs((double)xr), e = exp(-2.0 * ax);
          double sy, cy;
          sincos((double)xi, &sy, &cy);
          const double re = (1.0 + e) * cy, im = (1.0 - e) * sy;
          const double m2 = re * re + im * im;
          lsum += ax + (m2 > 0.0 ? 0.5 * log(m2) : -1e300);
        } else {
          double ax = fabs((double)xr);
          lsum += ax + log1p(exp(-2.0 * ax));
        }
      }
    }
    if (la) {
      // visible term sum_i a_i s_i (nqs_ansatz.c:284-285, 333-336)
      for (int i = ll; i < N; i += LPC) {
        lsum += (double)m.a_re[i] * (double)sp[i];
        if (CPLX) asum += (double)m.a_im[i] * (double)sp[i];
      }
      const double L = group_sum<LPC>(lsum);
      const double A = CPLX ? group_sum<LPC>(asum) : 0.0;
      if (ll == 0) {
        la[b] = L;
        if (ar) ar[b] = A;
      }
    }
  }
}

template <typename R>
int launch_forward(const ModelView<R>& m, const int* spins, int B, R* t_re, R* t_im,
                   double* la, double* ar, hipStream_t st, R* q_re, R* q_im) {
  if (B <= 0) return 0;
  if (m.N > kMaxSites) {
    fprintf(stderr, "nqs_hip: amplitudes support N <= %d sites (got %d)\n", kMaxSites, m.N);
    return -1;
  }
  int grid = (B + 3) / 4 < 65535 ? (B + 3) / 4 : 65535;
  KScope ks("forward", st);
  const size_t stage = sizeof(R) * ((size_t)(m.cplx ? 2 : 1) * ((size_t)m.N * m.M + m.M));
  // FP32: 32 lanes per configuration where that takes fewer lane passes per
  // configuration (M = 96: three full passes for two configurations instead
  // of a full and a half-empty one for one).  FP64 keeps one configuration
  // per wave and with it its order of the sums.
  const bool half = sizeof(R) == 4 && (m.M + 31) / 32 < 2 * ((m.M + 63) / 64);
  if (stage <= kFwdStage) {
    // eight configurations per wave at most, at least 2048 workgroups
    const int g2 = (B + 31) / 32 > 2048 ? (B + 31) / 32 : 2048;
    if (g2 < grid) grid = g2;
    if (m.cplx && half)
      hipLaunchKernelGGL((k_forward<R, true, true, 32>), dim3(grid), dim3(256), stage, st, m,
                         spins, B, t_re, t_im, la, ar, q_re, q_im);
    else if (m.cplx)
      hipLaunchKernelGGL((k_forward<R, true, true, 64>), dim3(grid), dim3(256), stage, st, m,
                         spins, B, t_re, t_im, la, ar, q_re, q_im);
    else if (half)
      hipLaunchKernelGGL((k_forward<R, false, true, 32>), dim3(grid), dim3(256), stage, st, m,
                         spins, B, t_re, t_im, la, ar, q_re, q_im);
    else
      hipLaunchKernelGGL((k_forward<R, false, true, 64>), dim3(grid), dim3(256), stage, st, m,
                         spins, B, t_re, t_im, la, ar, q_re, q_im);
  } else if (m.cplx) {
    hipLaunchKernelGGL((k_forward<R, true, false, 64>), dim3(grid), dim3(256), 0, st, m, spins, B,
                       t_re, t_im, la, ar, q_re, q_im);
  } else {
    hipLaunchKernelGGL((k_forward<R, false, false, 64>), dim3(grid), dim3(256), 0, st, m, spins,
                       B, t_re, t_im, la, ar, q_re, q_im);
  }
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

// --------------------------------------------------------------------------
// t = tanh(theta) of many configurations with theta = S W^T + b formed on
// FP64 MFMA (launch_gemm_nt): the spins as doubles (B x N) times the
// row-major W (M x N) of the parameter vector, into t itself; then
// t <- tanh(t + b) in place (complex: both parts from W_re and W_im).
// --------------------------------------------------------------------------
__global__ void k_spins_f64(const int* __restrict__ s, long n, double* __restrict__ out) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
    out[e] = (double)s[e];
}

template <bool CPLX>
__global__ void k_theta_finish(double* __restrict__ t_re, double* __restrict__ t_im,
                               const double* __restrict__ b_re, const double* __restrict__ b_im,
                               long B, int M, double* __restrict__ q_re,
                               double* __restrict__ q_im) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < B * M;
       e += (long)gridDim.x * blockDim.x) {
    const int h = (int)(e % M);
    const double xr = t_re[e] + b_re[h], xi = CPLX ? t_im[e] + b_im[h] : 0.0;
    if (q_re) {
      double sr, si;
      int is_p;
      logistic_small<double>(xr, xi, sr, si, is_p);
      q_re[e] = sr;
      if (CPLX) q_im[e] = si;
    }
    if (CPLX) {
      double tr, ti;
      ctanh<double>(xr, xi, tr, ti);
      t_re[e] = tr;
      t_im[e] = ti;
    } else {
      t_re[e] = tanh(xr);
    }
  }
}

int launch_forward_t_gemm(const ModelView<double>& m, const double* d_par, const int* spins, int B,
                          double* t_re, double* t_im, double* sd, hipStream_t st, double* q_re,
                          double* q_im) {
  if (B <= 0) return 0;
  const int N = m.N, M = m.M;
  const long Pr = (long)N + M + (long)M * N;
  KScope ks("forward", st);
  long blocks = ((long)B * N + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_spins_f64, dim3((unsigned)blocks), dim3(256), 0, st, spins, (long)B * N, sd);
  count_launch();
  if (launch_gemm_nt(B, M, N, 1.0, sd, N, d_par + N + M, N, 0.0, t_re, M, 0, st)) return -1;
  if (m.cplx &&
      launch_gemm_nt(B, M, N, 1.0, sd, N, d_par + Pr + N + M, N, 0.0, t_im, M, 0, st))
    return -1;
  blocks = ((long)B * M + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (m.cplx)
    hipLaunchKernelGGL(k_theta_finish<true>, dim3((unsigned)blocks), dim3(256), 0, st, t_re, t_im,
                       m.b_re, m.b_im, (long)B, M, q_re, q_im);
  else
    hipLaunchKernelGGL(k_theta_finish<false>, dim3((unsigned)blocks), dim3(256), 0, st, t_re, t_im,
                       m.b_re, m.b_im, (long)B, M, q_re, q_im);
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

// --------------------------------------------------------------------------
// O_k rows.  Real block layout [a (N) | b (M) | W (M x N)], imaginary block
// (complex RBM) the same at offset Pr.
//   real RBM / mean-field : [s | t | t s]
//   complex, mode 0 (d log|psi|):  re [s | tR | tR s]  [0 | -tI | -tI s]
//   complex, mode 1 (holomorphic): re as mode 0, im [0 | tI | tI s] [s | tR | tR s]
// --------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(256) void k_gen_O(int N, int M, int cplx, int mode,
                                               const int* __restrict__ spins, int B,
                                               const R* __restrict__ t_re,
                                               const R* __restrict__ t_im,
                                               double* __restrict__ O_re,
                                               double* __restrict__ O_im, long P) {
  const long Pr = (long)N + M + (long)M * N;
  for (int b = blockIdx.y; b < B; b += gridDim.y) {
    const int* sp = spins + (size_t)b * N;
    const R* tr = t_re + (size_t)b * M;
    const R* ti = cplx ? t_im + (size_t)b * M : nullptr;
    double* orow = O_re + (size_t)b * P;
    double* irow = (mode == 1) ? O_im + (size_t)b * P : nullptr;
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < P;
         k += (long)gridDim.x * blockDim.x) {
      const int blk = (k >= Pr) ? 1 : 0;
      const long kk = k - blk * Pr;
      double vr, vi = 0.0;
      if (kk < N) {
        double s = (double)sp[kk];
        if (!cplx) { vr = s; }
        else if (blk == 0) { vr = s; vi = 0.0; }
        else { vr = 0.0; vi = s; }
      } else {
        long e = kk - N;
        double s = 1.0;
        int h;
        if (e < M) {
          h = (int)e;
        } else {
          e -= M;
          h = (int)(e / N);
          s = (double)sp[e % N];
        }
        double a = (double)tr[h];
        if (!cplx) {
          vr = a * s;
        } else {
          double c = (double)ti[h];
          if (blk == 0) { vr = a * s; vi = c * s; }
          else { vr = -c * s; vi = a * s; }
        }
      }
      orow[k] = vr;
      if (irow) irow[k] = vi;
    }
  }
}

template <typename R>
int launch_gen_O(const ModelView<R>& m, int kind, const int* spins, int B,
                 const R* t_re, const R* t_im, int mode, double* O_re, double* O_im,
                 long P, hipStream_t st) {
  (void)kind;
  if (B <= 0) return 0;
  int gx = (int)((P + 255) / 256);
  if (gx > 1024) gx = 1024;
  int gy = B < 65535 ? B : 65535;
  KScope ks("gen_O", st);
  hipLaunchKernelGGL(k_gen_O<R>, dim3(gx, gy), dim3(256), 0, st, m.N, m.M, m.cplx, mode,
                     spins, B, t_re, t_im, O_re, O_im, P);
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

template int launch_prep<double>(const double*, int, int, int, ModelBuffers<double>&, hipStream_t, int);
template int launch_prep<float>(const double*, int, int, int, ModelBuffers<float>&, hipStream_t, int);
template int report_ratio_range<double>(const double*, int, int, int*, hipStream_t);
template int report_ratio_range<float>(const double*, int, int, int*, hipStream_t);
template void free_model<double>(ModelBuffers<double>&);
template void free_model<float>(ModelBuffers<float>&);
template int launch_forward<double>(const ModelView<double>&, const int*, int, double*, double*, double*, double*, hipStream_t, double*, double*);
template int launch_forward<float>(const ModelView<float>&, const int*, int, float*, float*, double*, double*, hipStream_t, float*, float*);
template int launch_gen_O<double>(const ModelView<double>&, int, const int*, int, const double*, const double*, int, double*, double*, long, hipStream_t);
template int launch_gen_O<float>(const ModelView<float>&, int, const int*, int, const float*, const float*, int, double*, double*, long, hipStream_t);

}  // namespace nqs
