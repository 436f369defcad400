// k_mixed.hip — dense products of the mixed-precision mode on 16-bit MFMA
// with FP32 accumulation (the MinSR Gram matrices O_c^T O_c and O_c O_c^T,
// the batched forward theta = S W^T).
//
// Gram ("x3" products): the FP32 entries are scaled by a power of two sc
// chosen from the largest |entry| (largest scaled entry in [2^13, 2^14),
// inside the half-precision range) and split as hi = fp16(sc x),
// lo = fp16(sc x - hi), both rounded to nearest even, so
// |sc x - hi - lo| <= 2^-22 |sc x| (plus the half subnormal spacing 2^-24,
// below 2^-37 of the largest entry).  A product is taken as
// hi_a hi_b + hi_a lo_b + lo_a hi_b; the dropped lo_a lo_b and the two
// representation errors bound its error by about 3 * 2^-22 |a b|.  Each term
// is one v_mfma_f32_16x16x32_f16 (operand lane l: row l % 16,
// k = 8 (l / 16) + [0, 8); result lane l, register r: row 4 (l / 16) + r,
// column l % 16), accumulated in FP32; the result is divided by sc^2 (exact)
// and written as FP64.
//
// Tiles: 128 x 128 outputs per workgroup of four waves (2 x 2, 64 x 64 each,
// 4 x 4 MFMA blocks of 16 x 16), K in chunks of 32.  Long K sums are split
// into pieces of about 2048 (second grid dimension), each accumulated in
// FP32 and written to a partial buffer, and the pieces added in FP64 by
// k_x3_reduce: the FP32 accumulation error stays that of a 2048-term sum.
// Global FP32 loads of the next chunk are in flight while the current
// chunk's MFMAs issue; the split into hi / lo happens once per element when
// it is staged in LDS (double-buffered; rows of 32 halves, 64 bytes, whose
// four 16-byte groups are permuted by an XOR with (row / 2) % 4: the eight
// lanes of an operand read (rows r..r+7, one group) and the sixteen of a
// staging store (two rows, eight 8-byte pieces each) then fall in distinct
// banks, and the 64 KB of a workgroup's buffers leave room for two
// workgroups per CU).  With
// lower = 1 only tiles on or below the diagonal run and only entries with
// column <= row are written.
//
// Forward theta = S W^T on the bfloat16 MFMA: spins +-1 are exact in
// bfloat16 and every weight is held as hi + mid + lo bfloat16 (each the
// rounding of what the previous ones left, |w - hi - mid - lo| <= 2^-24 |w|),
// so the three products per term are exact and theta carries only the FP32
// accumulation error over N terms; t = tanh(theta + b) in FP32.
#include "nqs_kernels.h"
#include "nqs_common.h"

namespace nqs {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int XT = 128;       // output tile edge
constexpr int XK = 32;        // K per chunk
constexpr int XLD = XK;       // LDS row stride (halves)

// lane l's value of v in every lane (l wave-uniform)
__device__ inline double x3_readlane(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// half offset of the 8-half group g of row r in the swizzled LDS rows
__device__ inline int x3_grp(int r, int g) { return 8 * (g ^ ((r >> 1) & 3)); }

__device__ inline void x3_lower_tile(int t, int& tm, int& tn) {
  int m = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while (m * (m + 1) / 2 > t) m--;
  while ((m + 1) * (m + 2) / 2 <= t) m++;
  tm = m;
  tn = t - m * (m + 1) / 2;
}

__device__ inline void x3_split(const float4 v, float sc, f16x4& hi, f16x4& lo) {
  const float x0 = v.x * sc, x1 = v.y * sc, x2 = v.z * sc, x3 = v.w * sc;
  hi[0] = (_Float16)x0;
  hi[1] = (_Float16)x1;
  hi[2] = (_Float16)x2;
  hi[3] = (_Float16)x3;
  lo[0] = (_Float16)(x0 - (float)hi[0]);
  lo[1] = (_Float16)(x1 - (float)hi[1]);
  lo[2] = (_Float16)(x2 - (float)hi[2]);
  lo[3] = (_Float16)(x3 - (float)hi[3]);
}

// power-of-two scale putting the largest |entry| m in [2^13, 2^14)
__device__ inline float x3_scale(float m) {
  if (!(m > 0.f) || !isfinite(m)) return 1.f;
  int e;
  (void)frexpf(m, &e);   // m = f 2^e, f in [0.5, 1)
  return ldexpf(1.f, 14 - e);
}

// C = alpha A B^T: A [M][K] (lda), B [N][K] (ldb) FP32, C [M][N] FP64 (ldc);
// K a multiple of 32, lda and ldb multiples of 4
__global__ __launch_bounds__(256) void k_gemm_x3(int M, int N, int K, double alpha,
                                                 const float* __restrict__ A, long lda,
                                                 const float* __restrict__ Bm, long ldb,
                                                 const float* __restrict__ amax,
                                                 const float* __restrict__ bmax,
                                                 double* __restrict__ C, long ldc, int lower,
                                                 int tiles_n, float* __restrict__ part) {
  __shared__ __attribute__((aligned(16))) _Float16 sAh[2][XT][XLD];
  __shared__ __attribute__((aligned(16))) _Float16 sAl[2][XT][XLD];
  __shared__ __attribute__((aligned(16))) _Float16 sBh[2][XT][XLD];
  __shared__ __attribute__((aligned(16))) _Float16 sBl[2][XT][XLD];
  const float sa = x3_scale(*amax), sb = x3_scale(*bmax);
  int tm, tn;
  if (lower) x3_lower_tile(blockIdx.x, tm, tn);
  else {
    tm = blockIdx.x / tiles_n;
    tn = blockIdx.x % tiles_n;
  }
  const int row0 = tm * XT, col0 = tn * XT;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int wr = (wave >> 1) * 64, wc = (wave & 1) * 64;
  const int l16 = lane & 15, lg = lane >> 4;
  // split K: workgroup row y sums K columns [y K, (y + 1) K)
  const long kbase = (long)blockIdx.y * K;
  // loader: float4 number e = 256 q + tid of the chunk, row e / 8, columns
  // 4 (e % 8) + [0, 4); rows past M / N repeat the last row (never written)
  const float* pa[4];
  const float* pb[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int e = 256 * q + tid, r = e >> 3, c4 = 4 * (e & 7);
    pa[q] = A + (long)min(row0 + r, M - 1) * lda + kbase + c4;
    pb[q] = Bm + (long)min(col0 + r, N - 1) * ldb + kbase + c4;
  }
  float4 ra[4], rb[4];
  auto gload = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      ra[q] = *(const float4*)(pa[q] + k0);
      rb[q] = *(const float4*)(pb[q] + k0);
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int e = 256 * q + tid, r = e >> 3, c4 = 4 * (e & 7);
      const int pc = x3_grp(r, c4 >> 3) + (c4 & 7);
      f16x4 hi, lo;
      x3_split(ra[q], sa, hi, lo);
      *(f16x4*)&sAh[buf][r][pc] = hi;
      *(f16x4*)&sAl[buf][r][pc] = lo;
      x3_split(rb[q], sb, hi, lo);
      *(f16x4*)&sBh[buf][r][pc] = hi;
      *(f16x4*)&sBl[buf][r][pc] = lo;
    }
  };
  f32x4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
  const int nch = K / XK;
  if (nch > 0) {
    gload(0);
    sstore(0);
  }
  __syncthreads();
  for (int c = 0; c < nch; c++) {
    const int buf = c & 1;
    if (c + 1 < nch) gload((c + 1) * XK);
    f16x8 ah[4], al[4], bh[4], bl[4];
    // rows 16 i + l16 of the wave's blocks: (row / 2) % 4 = (l16 / 2) % 4
    const int pg = x3_grp(l16, lg);
#pragma unroll
    for (int i = 0; i < 4; i++) {
      ah[i] = *(const f16x8*)&sAh[buf][wr + 16 * i + l16][pg];
      al[i] = *(const f16x8*)&sAl[buf][wr + 16 * i + l16][pg];
      bh[i] = *(const f16x8*)&sBh[buf][wc + 16 * i + l16][pg];
      bl[i] = *(const f16x8*)&sBl[buf][wc + 16 * i + l16][pg];
    }
    // the small terms first, sixteen independent accumulators between
    // dependent MFMAs
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 4; j++)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[i], bh[j], acc[i][j], 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 4; j++)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[i], bl[j], acc[i][j], 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 4; j++)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[i], bh[j], acc[i][j], 0, 0, 0);
    // the other buffer's last reads were before the previous barrier
    if (c + 1 < nch) sstore(buf ^ 1);
    __syncthreads();
  }
  // undo the power-of-two scales (exact)
  const float inv = (1.f / sa) * (1.f / sb);
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int gr = row0 + wr + 16 * i + 4 * lg + r;
        const int gc = col0 + wc + 16 * j + l16;
        if (gr < M && gc < N && (!lower || gc <= gr)) {
          if (part) part[((long)blockIdx.y * M + gr) * N + gc] = acc[i][j][r] * inv;
          else C[(long)gr * ldc + gc] = alpha * (double)(acc[i][j][r] * inv);
        }
      }
}

// C = alpha sum_s part[s] in FP64 over the split-K partial products.
// Workgroups (p, y) take rows p and M - 1 - p (lower: p + 1 and M - p
// elements, M + 1 together, so every pair has the same work) as one run of
// coalesced columns, interleaved over the gridDim.y workgroups of the pair.
__global__ __launch_bounds__(256) void k_x3_reduce(const float* __restrict__ part, int S, int M,
                                                   int N, double alpha, double* __restrict__ C,
                                                   long ldc, int lower) {
  const long n = (long)M * N;
  const int p = blockIdx.x, q = M - 1 - p;
  const int len_p = lower ? p + 1 : N;
  const int len = len_p + (q > p ? (lower ? q + 1 : N) : 0);
  for (int t = blockIdx.y * 256 + threadIdx.x; t < len; t += 256 * gridDim.y) {
    const bool first = t < len_p;
    const long i = first ? p : q;
    const long e = i * N + (first ? t : t - len_p);
    // up to eight slices' loads in flight together; the sum runs over s in
    // order
    double acc = 0.0;
    for (int s0 = 0; s0 < S; s0 += 8) {
      float v[8];
#pragma unroll
      for (int u = 0; u < 8; u++) v[u] = s0 + u < S ? part[(long)(s0 + u) * n + e] : 0.f;
#pragma unroll
      for (int u = 0; u < 8; u++)
        if (s0 + u < S) acc += (double)v[u];
    }
    C[i * ldc + (e - i * N)] = alpha * acc;
  }
}

// out [C][ldk] FP32: out[c][r] = in[r][c] for r < R, 0 for R <= r < ldk
// (64 x 64 tiles through LDS); bmax[block] = largest |out| of the block
__global__ __launch_bounds__(256) void k_to_f32_t(const double* __restrict__ in, long R, long C,
                                                  float* __restrict__ out, long ldk,
                                                  float* __restrict__ bmax) {
  __shared__ float t[64][65];
  __shared__ float red[256];
  const long r0 = (long)blockIdx.y * 64, c0 = (long)blockIdx.x * 64;
  float m = 0.f;
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int r = e >> 6, c = e & 63;
    const float v = (r0 + r < R && c0 + c < C) ? (float)in[(r0 + r) * C + c0 + c] : 0.f;
    t[r][c] = v;
    m = fmaxf(m, fabsf(v));
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) bmax[(long)blockIdx.y * gridDim.x + blockIdx.x] = red[0];
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int c = e >> 6, r = e & 63;
    if (c0 + c < C && r0 + r < ldk) out[(c0 + c) * ldk + r0 + r] = t[r][c];
  }
}

// out [R][ldc] FP32: out[r][c] = in[r][c] for c < C, 0 for C <= c < ldc;
// bmax[block] = largest |out| of the block
__global__ __launch_bounds__(256) void k_to_f32_pad(const double* __restrict__ in, long R, long C,
                                                    float* __restrict__ out, long ldc,
                                                    float* __restrict__ bmax) {
  __shared__ float red[256];
  const long n = R * ldc;
  float m = 0.f;
  for (long e = blockIdx.x * 256L + threadIdx.x; e < n; e += (long)gridDim.x * 256) {
    const long r = e / ldc, c = e - r * ldc;
    const float v = c < C ? (float)in[r * C + c] : 0.f;
    out[e] = v;
    m = fmaxf(m, fabsf(v));
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) bmax[blockIdx.x] = red[0];
}

// out[0] = largest of n block maxima (one workgroup)
__global__ __launch_bounds__(256) void k_absmax_finish(const float* __restrict__ bmax, long n,
                                                       float* __restrict__ out) {
  __shared__ float red[256];
  float m = 0.f;
  for (long i = threadIdx.x; i < n; i += 256) m = fmaxf(m, bmax[i]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

// K padded for the split-K product: S splits of a multiple of 32 columns
long x3_pad(long k) {
  long kp = (k + XK - 1) / XK * XK;
  if (kp > 2048) {
    const long S = (kp + 2047) / 2048;
    const long ks = ((kp + S - 1) / S + XK - 1) / XK * XK;
    kp = S * ks;
  }
  return kp;
}

// Wsplit [3][M][Kp] bf16: hi, mid, lo of the FP64 weights W [M][N] (row
// h contiguous), zero for columns N..Kp-1
__global__ void k_w_split3(const double* __restrict__ W, int M, int N, int Kp,
                           __bf16* __restrict__ out) {
  const long n = (long)M * Kp;
  for (long e = blockIdx.x * 256L + threadIdx.x; e < n; e += (long)gridDim.x * 256) {
    const long h = e / Kp;
    const int k = (int)(e - h * Kp);
    const double w = k < N ? W[h * N + k] : 0.0;
    const __bf16 hi = (__bf16)(float)w;
    const double r1 = w - (double)(float)hi;
    const __bf16 mid = (__bf16)(float)r1;
    const double r2 = r1 - (double)(float)mid;
    out[e] = hi;
    out[n + e] = mid;
    out[2 * n + e] = (__bf16)(float)r2;
  }
}

// theta [B][M] FP32 = S W^T (+ the imaginary part from a second split when
// CPLX): workgroup of 4 waves, 64 configurations x 64 hidden units, wave w
// the configurations 16 w + [0, 16) as four 16 x 16 MFMA blocks; operands
// straight from global memory (spin rows and the small split weights stay
// in cache)
template <bool CPLX>
__global__ __launch_bounds__(256) void k_theta_x3(const int* __restrict__ spins, int B, int N,
                                                  int M, int Kp, const __bf16* __restrict__ wr,
                                                  const __bf16* __restrict__ wi,
                                                  float* __restrict__ th_re,
                                                  float* __restrict__ th_im) {
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int l16 = lane & 15, lg = lane >> 4;
  const int row0 = blockIdx.x * 64 + 16 * wave, col0 = blockIdx.y * 64;
  const int cfg = min(row0 + l16, B - 1);
  const long n = (long)M * Kp;
  f32x4 ar[4], ai[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    ar[j] = (f32x4){0.f, 0.f, 0.f, 0.f};
    ai[j] = (f32x4){0.f, 0.f, 0.f, 0.f};
  }
  for (int k0 = 0; k0 < Kp; k0 += XK) {
    bf16x8 a;
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int k = k0 + 8 * lg + q;
      a[q] = (__bf16)(k < N ? (float)spins[(long)cfg * N + k] : 0.f);
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const long h = min(col0 + 16 * j + l16, M - 1);
      const long off = h * Kp + k0 + 8 * lg;
#pragma unroll
      for (int sp = 2; sp >= 0; sp--) {
        ar[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, *(const bf16x8*)(wr + sp * n + off),
                                                        ar[j], 0, 0, 0);
        if (CPLX)
          ai[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, *(const bf16x8*)(wi + sp * n + off),
                                                          ai[j], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int j = 0; j < 4; j++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int b = row0 + 4 * lg + r, h = col0 + 16 * j + l16;
      if (b < B && h < M) {
        th_re[(long)b * M + h] = ar[j][r];
        if (CPLX) th_im[(long)b * M + h] = ai[j][r];
      }
    }
}

// t <- tanh(theta + b) in FP32, in place (complex: both parts)
template <bool CPLX>
__global__ void k_theta_finish_f32(float* __restrict__ t_re, float* __restrict__ t_im,
                                   const float* __restrict__ b_re, const float* __restrict__ b_im,
                                   long B, int M) {
  for (long e = blockIdx.x * 256L + threadIdx.x; e < B * M; e += (long)gridDim.x * 256) {
    const int h = (int)(e % M);
    const float xr = t_re[e] + b_re[h];
    if (CPLX) {
      float tr, ti;
      ctanh<float>(xr, t_im[e] + b_im[h], tr, ti);
      t_re[e] = tr;
      t_im[e] = ti;
    } else {
      t_re[e] = tanhf(xr);
    }
  }
}

int theta_x3_split_elems(int M, int N) { return 3 * M * (int)x3_pad(N); }

int launch_forward_t_x3(const ModelView<float>& m, const double* d_par, const int* spins, int B,
                        float* t_re, float* t_im, __bf16* wsplit, hipStream_t st) {
  if (B <= 0) return 0;
  const int N = m.N, M = m.M, Kp = (int)x3_pad(N);
  const long Pr = (long)N + M + (long)M * N, n = (long)M * Kp;
  KScope ks("forward", st);
  long g = (n + 255) / 256;
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL(k_w_split3, dim3((unsigned)g), dim3(256), 0, st, d_par + N + M, M, N, Kp,
                     wsplit);
  if (m.cplx)
    hipLaunchKernelGGL(k_w_split3, dim3((unsigned)g), dim3(256), 0, st, d_par + Pr + N + M, M, N,
                       Kp, wsplit + 3 * n);
  const dim3 grid((unsigned)((B + 63) / 64), (unsigned)((M + 63) / 64));
  if (m.cplx)
    hipLaunchKernelGGL(k_theta_x3<true>, grid, dim3(256), 0, st, spins, B, N, M, Kp, wsplit,
                       wsplit + 3 * n, t_re, t_im);
  else
    hipLaunchKernelGGL(k_theta_x3<false>, grid, dim3(256), 0, st, spins, B, N, M, Kp, wsplit,
                       nullptr, t_re, nullptr);
  long fb = ((long)B * M + 255) / 256;
  if (fb > 65536) fb = 65536;
  if (m.cplx)
    hipLaunchKernelGGL(k_theta_finish_f32<true>, dim3((unsigned)fb), dim3(256), 0, st, t_re, t_im,
                       m.b_re, m.b_im, (long)B, M);
  else
    hipLaunchKernelGGL(k_theta_finish_f32<false>, dim3((unsigned)fb), dim3(256), 0, st, t_re,
                       t_im, m.b_re, m.b_im, (long)B, M);
  for (int i = 0; i < (m.cplx ? 4 : 3); i++) count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}



// K columns per split of the K sum: the FP32 accumulation of each split is
// short (at most 3 * 2048 / 32 = 192 roundings), the splits add in FP64
constexpr int kX3SplitK = 2048;

int x3_splits(long K) { return K > kX3SplitK ? (int)((K + kX3SplitK - 1) / kX3SplitK) : 1; }

size_t x3_part_elems(int M, int N, long K) {
  const int S = x3_splits(K);
  return S > 1 ? (size_t)S * M * N : 0;
}

int launch_gemm_x3(int M, int N, int K, double alpha, const float* A, long lda, const float* B,
                   long ldb, const float* amax, const float* bmax, double* C, long ldc, int lower,
                   float* part, hipStream_t st) {
  if (M <= 0 || N <= 0) return 0;
  if (K % XK != 0 || lda % 4 != 0 || ldb % 4 != 0 || lda < K || ldb < K || ((size_t)A & 15) ||
      ((size_t)B & 15) || (lower && M != N)) {
    fprintf(stderr, "nqs_hip: gemm_x3 needs K %% 32 == 0, 16-byte aligned rows and a square "
                    "lower product\n");
    return -1;
  }
  // S splits of Ks columns each (Ks a multiple of 32; the last split's
  // columns past K are zero padding the caller provides up to S Ks <= lda)
  int S = x3_splits(K);
  long Ks = (K + S - 1) / S;
  Ks = (Ks + XK - 1) / XK * XK;
  S = (int)((K + Ks - 1) / Ks);
  if (S > 1 && (!part || S * Ks > lda || S * Ks > ldb)) {
    fprintf(stderr, "nqs_hip: gemm_x3 split-K needs a partial buffer and K padding\n");
    return -1;
  }
  const int tm = (M + XT - 1) / XT, tn = (N + XT - 1) / XT;
  const long tiles = lower ? (long)tm * (tm + 1) / 2 : (long)tm * tn;
  KScope ks("gram_x3", st);
  hipLaunchKernelGGL(k_gemm_x3, dim3((unsigned)tiles, (unsigned)S), dim3(256), 0, st, M, N,
                     (int)(S > 1 ? Ks : K), alpha, A, lda, B, ldb, amax, bmax, C, ldc, lower,
                     tn, S > 1 ? part : nullptr);
  count_launch();
  if (S > 1) {
    hipLaunchKernelGGL(k_x3_reduce, dim3((unsigned)((M + 1) / 2), 8), dim3(256), 0, st, part, S,
                       M, N, alpha, C, ldc, lower);
    count_launch();
  }
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

long x3_max_elems(long R, long C, long ld, int transpose) {
  if (transpose) return ((C + 63) / 64) * ((ld + 63) / 64) + 1;
  long g = (R * ld + 255) / 256;
  return (g > 16384 ? 16384 : g) + 1;
}

// Block (x, y): samples 64 x + [0, 64) and the hidden units [hlo, hhi) of
// slice y of gridDim.y: their bias and weight rows (slice 0 also the
// visible-bias rows).  The block stages only what those rows read, the
// spins and its slice's factors, in LDS sample-minor (row stride 65 floats:
// conflict-free stores and reads; a few KB, so that many blocks share a CU
// and their stores overlap), then wave w takes rows w, w + 4, ... (the
// parameter decode is wave-uniform) and its 64 lanes the samples, so each
// row segment is one 256-byte store.  An entry is formed in FP64 (an FP32
// factor times a spin, exact), centred there and rounded once; samples
// past B give zeros.
constexpr int kOctRows = 256;   // N + M (+ M) at most

__global__ __launch_bounds__(256) void k_oct_factored(const int* __restrict__ spins, int B, int N,
                                                      int M, const float* __restrict__ U1,
                                                      const float* __restrict__ U2,
                                                      const double* __restrict__ Om, long P,
                                                      float* __restrict__ out, long ldk,
                                                      float* __restrict__ bmax) {
  // rows [0, N): spins; rows N + z ns + j: factor z of hidden unit hlo + j
  extern __shared__ float stg[];
  __shared__ float red[256];
  const long b0 = (long)blockIdx.x * 64;
  const int nz = U2 ? 2 : 1;
  const int nvalid = (int)min(64L, max(0L, (long)B - b0));
  const int hlo = (int)((long)blockIdx.y * M / gridDim.y);
  const int hhi = (int)((long)(blockIdx.y + 1) * M / gridDim.y);
  const int ns = hhi - hlo;
  for (int e = threadIdx.x; e < 64 * N; e += 256) {
    const int bl = e / N, i = e - bl * N;
    stg[i * 65 + bl] = bl < nvalid ? (float)spins[(b0 + bl) * N + i] : 0.f;
  }
  for (int z = 0; z < nz; z++) {
    const float* U = z ? U2 : U1;
    for (int e = threadIdx.x; e < 64 * ns; e += 256) {
      const int bl = e / ns, j = e - bl * ns;
      stg[(N + z * ns + j) * 65 + bl] = bl < nvalid ? U[(b0 + bl) * M + hlo + j] : 0.f;
    }
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const bool live = lane < nvalid, inpad = b0 + lane < ldk;
  const long Pr = (long)N + M + (long)M * N;
  float mx = 0.f;
  auto put = [&](long k, double o, double om) {
    const float v = live ? (float)(o - om) : 0.f;
    if (inpad) out[k * ldk + b0 + lane] = v;
    mx = fmaxf(mx, fabsf(v));
  };
  const int nvis = blockIdx.y ? 0 : N;
  for (int z = 0; z < nz; z++) {
    const double sg = z ? -1.0 : 1.0;
    const long kz = z * Pr;
    const float* su = stg + (N + z * ns) * 65;
    // visible biases (slice 0; zero in the second block), then the
    // slice's hidden biases
    for (int r = wave; r < nvis + ns; r += 4) {
      if (r < nvis) {
        put(kz + r, z ? 0.0 : (double)stg[r * 65 + lane], Om[kz + r]);
      } else {
        const int j = r - nvis;
        put(kz + N + hlo + j, sg * (double)su[j * 65 + lane], Om[kz + N + hlo + j]);
      }
    }
    // weights, row h*N + i of the block; with N <= 64 the means of the
    // next unit's rows are loaded (lane i: row i) while this unit's are
    // stored, and taken by readlane
    const long kw = kz + N + M;
    const bool pre = N <= 64;
    double omv = (pre && hlo < hhi && lane < N) ? Om[kw + (long)hlo * N + lane] : 0.0;
    for (int h = hlo; h < hhi; h++) {
      const double omn = (pre && h + 1 < hhi && lane < N) ? Om[kw + (long)(h + 1) * N + lane] : 0.0;
      const double u = sg * (double)su[(h - hlo) * 65 + lane];
      const long kh = kw + (long)h * N;
      for (int i = wave; i < N; i += 4)
        put(kh + i, u * (double)stg[i * 65 + lane], pre ? x3_readlane(omv, i) : Om[kh + i]);
      omv = omn;
    }
  }
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) bmax[(long)blockIdx.y * gridDim.x + blockIdx.x] = red[0];
}

int launch_oct_factored(const int* spins, int B, int N, int M, const float* U1, const float* U2,
                        const double* Om, long P, float* out, long ldk, float* amax,
                        hipStream_t st) {
  if (B <= 0 || P <= 0) return 0;
  const long Pr = (long)N + M + (long)M * N;
  if (P != (U2 ? 2 * Pr : Pr) || ldk < B || N + (U2 ? 2 : 1) * M > kOctRows) return -1;
  // up to 8 hidden-unit slices per 64 samples (more blocks in flight), and
  // at most (P + 63) / 64 of them: one block maximum each, within the
  // x3_max_elems(B, P, ldk, 1) = ((P + 63) / 64) ((ldk + 63) / 64) + 1 floats
  int gy = M < 8 ? (M > 0 ? M : 1) : 8;
  if (gy > (P + 63) / 64) gy = (int)((P + 63) / 64);
  const dim3 grid((unsigned)((ldk + 63) / 64), (unsigned)gy);
  const long nb = (long)grid.x * grid.y;
  // staged rows: the spins and the largest slice's factors
  const int ns_max = (M + gy - 1) / gy;
  const size_t lds = sizeof(float) * 65 * (size_t)(N + (U2 ? 2 : 1) * ns_max);
  if (lds > 65536)   // up to 66.5 KB (256 rows): above the default dynamic limit
    NQS_HIP_CHECK(hipFuncSetAttribute((const void*)k_oct_factored,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  KScope ks("gen_O", st);
  hipLaunchKernelGGL(k_oct_factored, grid, dim3(256), lds, st, spins, B, N, M, U1, U2, Om, P, out,
                     ldk, amax + 1);
  hipLaunchKernelGGL(k_absmax_finish, dim3(1), dim3(256), 0, st, amax + 1, nb, amax);
  count_launch();
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_to_f32_t(const double* in, long R, long C, float* out, long ldk, float* amax,
                    hipStream_t st) {
  if (R <= 0 || C <= 0) return 0;
  if (ldk < R || (ldk + 63) / 64 > 65535) return -1;
  const dim3 grid((unsigned)((C + 63) / 64), (unsigned)((ldk + 63) / 64));
  const long nb = (long)grid.x * grid.y;
  hipLaunchKernelGGL(k_to_f32_t, grid, dim3(256), 0, st, in, R, C, out, ldk, amax + 1);
  hipLaunchKernelGGL(k_absmax_finish, dim3(1), dim3(256), 0, st, amax + 1, nb, amax);
  count_launch();
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_to_f32_pad(const double* in, long R, long C, float* out, long ldc, float* amax,
                      hipStream_t st) {
  if (R <= 0 || C <= 0) return 0;
  if (ldc < C) return -1;
  long g = (R * ldc + 255) / 256;
  if (g > 16384) g = 16384;
  hipLaunchKernelGGL(k_to_f32_pad, dim3((unsigned)g), dim3(256), 0, st, in, R, C, out, ldc,
                     amax + 1);
  hipLaunchKernelGGL(k_absmax_finish, dim3(1), dim3(256), 0, st, amax + 1, g, amax);
  count_launch();
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

}  // namespace nqs
