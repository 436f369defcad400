// k_ar.hip — learned proposal: masked autoregressive network (layout in
// ar_layout.h) on gfx950 BF16 MFMA with FP32 accumulation.
//
// k_ar_walk   one wave per 16 configurations walks the sites in order:
//             at site j it forms a1 = bf16(relu(z1)), z2 = b2 + W2m a1 and
//             the logits of j's 16-output tile on MFMA, then either draws
//             s_j (sampling rows) or reads it (evaluation rows), adds
//             log p(s_j | s_<j) in FP64 and updates z1 += W1m[:, j] s_j.
//             Sampling and evaluation run the same instructions on the same
//             operand values, so log q of a configuration is one fixed
//             function of it and the weights, and the Metropolis-Hastings
//             ratio q(s)/q(s') below is exact.  With `constrained`, a site
//             is forced once the remaining sites must all be up or all down,
//             which keeps every proposal in the chain's magnetisation sector.
// k_ar_accept accept / reject of one global move per chain:
//             d = 2(log|psi(s')| - log|psi(s)|) + log q(s) - log q(s');
//             accept when d >= 0 or log u < d (u from the chain's xorshift).
// k_ar_grad   forward + backward of the mean negative log-likelihood over a
//             minibatch; 64 rows per single-wave block, weight gradients as
//             MFMA products over the rows, partial sums per block.
// k_ar_adam   sums the partials in block order (deterministic), Adam on the
//             FP32 master weights, and writes the BF16 operand copies.
#include "ar_layout.h"
#include "nqs_kernels.h"

namespace nqs {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

__device__ inline bf16x8 ld_op(const uint16_t* base, long off) {
  const u16x8 v = *(const u16x8*)(base + off);
  return __builtin_bit_cast(bf16x8, v);
}
__device__ inline f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ inline float relu(float v) { return v > 0.f ? v : 0.f; }
// chunk c operand from per-tile values z[t][r] (feature 16t + 4q + r)
__device__ inline bf16x8 relu_op(const f32x4* z, int c) {
  u16x8 u;
#pragma unroll
  for (int e = 0; e < 8; e++) u[e] = ar::f2bf(relu(z[2 * c + (e >> 2)][e & 3]));
  return __builtin_bit_cast(bf16x8, u);
}
template <int T>
__device__ inline bf16x8 plain_op(const f32x4* z, int c) {
  u16x8 u;
#pragma unroll
  for (int e = 0; e < 8; e++) {
    const int t = 2 * c + (e >> 2);
    u[e] = t < T ? ar::f2bf(z[t][e & 3]) : (unsigned short)0;
  }
  return __builtin_bit_cast(bf16x8, u);
}
__device__ inline double log_sigmoid(double x) {
  return x >= 0.0 ? -log1p(exp(-x)) : x - log1p(exp(x));
}
__device__ inline f32x4 zero4() { return f32x4{0.f, 0.f, 0.f, 0.f}; }
__device__ inline uint64_t site_mask(int N) { return N >= 64 ? ~0ULL : ((1ULL << N) - 1ULL); }

// ---------------------------------------------------------------------------
template <int NT>
__global__ __launch_bounds__(64) void k_ar_walk(ArDev v, ArWalk w) {
  __shared__ float s_w1t[ar::kMaxN * ar::H];
  const int N = v.N, lane = threadIdx.x, rl = lane & 15, q = lane >> 4;
  for (int e = lane; e < N * ar::H; e += 64) s_w1t[e] = v.w1t[e];
  const int row = blockIdx.x * 16 + rl;
  const bool live = row < w.rows;
  const bool samp = row < w.n_sample;

  bf16x8 A2[4][2], A3[NT][2];
#pragma unroll
  for (int mt = 0; mt < 4; mt++)
#pragma unroll
    for (int c = 0; c < 2; c++) A2[mt][c] = ld_op(v.op2f, ((long)(mt * 2 + c) * 64 + lane) * 8);
#pragma unroll
  for (int mt = 0; mt < NT; mt++)
#pragma unroll
    for (int c = 0; c < 2; c++) A3[mt][c] = ld_op(v.op3f, ((long)(mt * 2 + c) * 64 + lane) * 8);
  f32x4 B2[4], B3[NT], z1[4];
#pragma unroll
  for (int t = 0; t < 4; t++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      B2[t][r] = v.b2[16 * t + 4 * q + r];
      z1[t][r] = v.b1[16 * t + 4 * q + r];
    }
#pragma unroll
  for (int t = 0; t < NT; t++)
#pragma unroll
    for (int r = 0; r < 4; r++) B3[t][r] = v.b3[16 * t + 4 * q + r];

  uint64_t bits = 0;
  if (live) bits = w.bits[samp ? row : row - w.n_sample];
  const int target = __popcll(bits & site_mask(N));
  uint64_t rng = (live && samp) ? w.rng[row] : 0ULL;
  uint64_t ob = 0;
  int downs = 0;
  double lq = 0.0;
  __syncthreads();

  for (int j = 0; j < N; j++) {
    const bf16x8 a0 = relu_op(z1, 0), a1 = relu_op(z1, 1);
    f32x4 z2[4];
#pragma unroll
    for (int mt = 0; mt < 4; mt++) z2[mt] = mfma(A2[mt][1], a1, mfma(A2[mt][0], a0, B2[mt]));
    const bf16x8 c0 = relu_op(z2, 0), c1 = relu_op(z2, 1);
    const int mj = j >> 4;
    f32x4 lg = zero4();
#pragma unroll
    for (int t = 0; t < NT; t++)
      if (t == mj) lg = mfma(A3[t][1], c1, mfma(A3[t][0], c0, B3[t]));
    const int rj = j & 3, qj = (j >> 2) & 3;
    const float mine = rj == 0 ? lg[0] : rj == 1 ? lg[1] : rj == 2 ? lg[2] : lg[3];
    const float l = __shfl(mine, rl + 16 * qj, 64);

    const int rem_d = target - downs, rem = N - j;
    const bool forced = w.constrained && (rem_d <= 0 || rem_d >= rem);
    const double u = samp ? xs_uniform(rng) : 0.0;
    const double ld = log_sigmoid((double)l);   // log p(down)
    int bit;
    if (forced) bit = rem_d > 0 ? 1 : 0;
    else if (samp) bit = u < exp(ld) ? 1 : 0;
    else bit = (int)((bits >> j) & 1ULL);
    if (!forced) lq += bit ? ld : ld - (double)l;
    downs += bit;
    ob |= (uint64_t)bit << j;
    const float s = bit ? -1.f : 1.f;
    const float* wr = s_w1t + j * ar::H + 4 * q;
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
      for (int r = 0; r < 4; r++) z1[t][r] = fmaf(wr[16 * t + r], s, z1[t][r]);
  }
  if (!live) return;
  if (q == 0) w.lq[row] = lq;
  if (samp) {
    if (q == 0) w.rng[row] = rng;
    if (q == 0 && w.out_bits) w.out_bits[row] = ob;
    if (w.out_spins)
      for (int i = q; i < N; i += 4) w.out_spins[(size_t)row * N + i] = ((ob >> i) & 1ULL) ? -1 : 1;
  }
}

// ---------------------------------------------------------------------------
__global__ void k_ar_accept(ChainState cs, const uint64_t* __restrict__ prop,
                            const double* __restrict__ la_prop, const double* __restrict__ lq,
                            unsigned long long* __restrict__ lacc) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cs.C) return;
  uint64_t rng = cs.rng[c];
  const double d = 2.0 * (la_prop[c] - cs.la[c]) + (lq[cs.C + c] - lq[c]);
  const bool acc = d >= 0.0 ? true : (log(xs_uniform(rng)) < d);
  cs.rng[c] = rng;
  cs.tot[c] += 1;
  if (acc) {
    cs.spins[c] = prop[c];
    cs.la[c] = la_prop[c];
    cs.acc[c] += 1;
    lacc[c] += 1;
  }
}

// rows first + c (c < C, first + c < n_out) of the sample buffer
__global__ void k_unpack_rows(ChainState cs, long first, long n_out, int* out) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= (long)cs.C * cs.N) return;
  const int c = (int)(e / cs.N), i = (int)(e % cs.N);
  if (first + c >= n_out) return;
  const uint64_t wd = cs.spins[(size_t)c * cs.nw + (i >> 6)];
  out[(first + c) * cs.N + i] = ((wd >> (i & 63)) & 1ULL) ? -1 : 1;
}

__global__ void k_bits_append(const int* __restrict__ spins, long n, int N, uint64_t* win,
                              long cap, long pos) {
  const long b = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (b >= n) return;
  uint64_t x = 0;
  const int* s = spins + b * N;
  for (int i = 0; i < N; i++) x |= (uint64_t)(s[i] < 0) << i;
  win[(pos + b) % cap] = x;
}

__global__ void k_spins_to_bits(const int* __restrict__ spins, long n, int N, uint64_t* out) {
  const long b = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (b >= n) return;
  uint64_t x = 0;
  const int* s = spins + b * N;
  for (int i = 0; i < N; i++) x |= (uint64_t)(s[i] < 0) << i;
  out[b] = x;
}

// ---------------------------------------------------------------------------
constexpr long kMaxP = 2L * ar::H * ar::kMaxN + (long)ar::H * ar::H + 2 * ar::H + ar::kMaxN;

__device__ inline bf16x8 ld_rows(const uint16_t (*t)[32], int f, int q) {
  const u16x8 v = *(const u16x8*)&t[f][8 * q];
  return __builtin_bit_cast(bf16x8, v);
}

template <int NT>
__global__ __launch_bounds__(64) void k_ar_grad(ArDev v, ArGrad g) {
  constexpr int C3 = (NT + 1) / 2;
  __shared__ float acc[kMaxP];
  __shared__ __attribute__((aligned(16))) uint16_t sX[64][32];
  __shared__ __attribute__((aligned(16))) uint16_t sA1[64][32];
  __shared__ __attribute__((aligned(16))) uint16_t sA2[64][32];
  __shared__ __attribute__((aligned(16))) uint16_t sG1[64][32];
  __shared__ __attribute__((aligned(16))) uint16_t sG2[64][32];
  __shared__ __attribute__((aligned(16))) uint16_t sG3[64][32];
  const int N = v.N, lane = threadIdx.x, rl = lane & 15, q = lane >> 4;
  const ar::Dims d = ar::dims(N);
  const int c1 = d.c1, ns = (N + 15) / 16;
  for (long p = lane; p < d.P; p += 64) acc[p] = 0.f;

  bf16x8 A2[4][2], A3[NT][2];
  f32x4 B1[4], B2[4], B3[NT];
#pragma unroll
  for (int mt = 0; mt < 4; mt++)
#pragma unroll
    for (int c = 0; c < 2; c++) A2[mt][c] = ld_op(v.op2f, ((long)(mt * 2 + c) * 64 + lane) * 8);
#pragma unroll
  for (int mt = 0; mt < NT; mt++)
#pragma unroll
    for (int c = 0; c < 2; c++) A3[mt][c] = ld_op(v.op3f, ((long)(mt * 2 + c) * 64 + lane) * 8);
#pragma unroll
  for (int t = 0; t < 4; t++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      B1[t][r] = v.b1[16 * t + 4 * q + r];
      B2[t][r] = v.b2[16 * t + 4 * q + r];
    }
#pragma unroll
  for (int t = 0; t < NT; t++)
#pragma unroll
    for (int r = 0; r < 4; r++) B3[t][r] = v.b3[16 * t + 4 * q + r];
  const uint64_t msk = site_mask(N);
  double loss = 0.0;
  __syncthreads();

  for (int h = 0; h < 2; h++) {
#pragma unroll 1
    for (int tile = 0; tile < 2; tile++) {
      const int r32 = tile * 16 + rl;
      const long mrow = (long)blockIdx.x * 64 + h * 32 + r32;
      const bool valid = mrow < g.count;
      const uint64_t bits = valid ? g.rows[(g.start + mrow) % g.n_avail] : 0ULL;
      // inputs, K slot 32c + 8q + e
      bf16x8 X[2];
#pragma unroll
      for (int c = 0; c < 2; c++) {
        u16x8 u;
#pragma unroll
        for (int e = 0; e < 8; e++) {
          const int i = 32 * c + 8 * q + e;
          u[e] = i < N ? (((bits >> i) & 1ULL) ? (unsigned short)0xBF80 : (unsigned short)0x3F80)
                       : (unsigned short)0;
          sX[i][r32] = u[e];
        }
        X[c] = __builtin_bit_cast(bf16x8, u);
      }
      f32x4 z1[4], z2[4], lg[NT];
#pragma unroll
      for (int mt = 0; mt < 4; mt++) {
        f32x4 a = B1[mt];
        for (int c = 0; c < c1; c++)
          a = mfma(ld_op(v.op1f, ((long)(mt * c1 + c) * 64 + lane) * 8), X[c], a);
        z1[mt] = a;
      }
      const bf16x8 a0 = relu_op(z1, 0), a1 = relu_op(z1, 1);
#pragma unroll
      for (int mt = 0; mt < 4; mt++) z2[mt] = mfma(A2[mt][1], a1, mfma(A2[mt][0], a0, B2[mt]));
      const bf16x8 b0 = relu_op(z2, 0), b1 = relu_op(z2, 1);
#pragma unroll
      for (int mt = 0; mt < NT; mt++) lg[mt] = mfma(A3[mt][1], b1, mfma(A3[mt][0], b0, B3[mt]));
      // dL/dl_j = (sigmoid(l_j) - x_j) / count on free sites
      const int D = __popcll(bits & msk);
      f32x4 g3[NT];
#pragma unroll
      for (int mt = 0; mt < NT; mt++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int o = 16 * mt + 4 * q + r;
          float gv = 0.f;
          if (valid && o < N) {
            const int before = __popcll(bits & ((1ULL << o) - 1ULL));
            const int rem_d = D - before, rem = N - o;
            const bool forced = g.constrained && (rem_d <= 0 || rem_d >= rem);
            if (!forced) {
              const int x = (int)((bits >> o) & 1ULL);
              const double l = (double)lg[mt][r];
              const double ls = log_sigmoid(l);
              loss -= x ? ls : ls - l;
              gv = (float)((exp(ls) - (double)x) * (double)g.scale);
            }
          }
          g3[mt][r] = gv;
        }
      bf16x8 G3[C3];
#pragma unroll
      for (int c = 0; c < C3; c++) G3[c] = plain_op<NT>(g3, c);
      f32x4 g2[4], g1[4];
#pragma unroll
      for (int mt = 0; mt < 4; mt++) {
        f32x4 a = zero4();
#pragma unroll
        for (int c = 0; c < C3; c++)
          a = mfma(ld_op(v.op3b, ((long)(mt * C3 + c) * 64 + lane) * 8), G3[c], a);
#pragma unroll
        for (int r = 0; r < 4; r++) g2[mt][r] = z2[mt][r] > 0.f ? a[r] : 0.f;
      }
      const bf16x8 G20 = plain_op<4>(g2, 0), G21 = plain_op<4>(g2, 1);
#pragma unroll
      for (int mt = 0; mt < 4; mt++) {
        f32x4 a = mfma(ld_op(v.op2b, ((long)(mt * 2 + 1) * 64 + lane) * 8), G21,
                       mfma(ld_op(v.op2b, ((long)(mt * 2) * 64 + lane) * 8), G20, zero4()));
#pragma unroll
        for (int r = 0; r < 4; r++) g1[mt][r] = z1[mt][r] > 0.f ? a[r] : 0.f;
      }
      // stage [feature][row] BF16 tiles for the weight gradients
#pragma unroll
      for (int t = 0; t < 4; t++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int f = 16 * t + 4 * q + r;
          sA1[f][r32] = ar::f2bf(relu(z1[t][r]));
          sA2[f][r32] = ar::f2bf(relu(z2[t][r]));
          sG1[f][r32] = ar::f2bf(g1[t][r]);
          sG2[f][r32] = ar::f2bf(g2[t][r]);
        }
#pragma unroll
      for (int t = 0; t < 4; t++)
#pragma unroll
        for (int r = 0; r < 4; r++) sG3[16 * t + 4 * q + r][r32] = t < NT ? ar::f2bf(g3[t < NT ? t : 0][r]) : (uint16_t)0;
    }
    __syncthreads();
    // weight gradients: sums over the 32 staged rows (K = rows 8q + e)
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
      const bf16x8 ga = ld_rows(sG2, 16 * mt + rl, q);
#pragma unroll
      for (int nt = 0; nt < 4; nt++) {
        const f32x4 dd = mfma(ga, ld_rows(sA1, 16 * nt + rl, q), zero4());
#pragma unroll
        for (int r = 0; r < 4; r++) acc[d.oW2 + (16 * mt + 4 * q + r) * ar::H + 16 * nt + rl] += dd[r];
      }
      const bf16x8 g1a = ld_rows(sG1, 16 * mt + rl, q);
      for (int nt = 0; nt < ns; nt++) {
        const f32x4 dd = mfma(g1a, ld_rows(sX, 16 * nt + rl, q), zero4());
        const int i = 16 * nt + rl;
        if (i < N)
#pragma unroll
          for (int r = 0; r < 4; r++) acc[d.oW1 + (long)(16 * mt + 4 * q + r) * N + i] += dd[r];
      }
    }
#pragma unroll
    for (int mt = 0; mt < NT; mt++) {
      const bf16x8 g3a = ld_rows(sG3, 16 * mt + rl, q);
#pragma unroll
      for (int nt = 0; nt < 4; nt++) {
        const f32x4 dd = mfma(g3a, ld_rows(sA2, 16 * nt + rl, q), zero4());
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int j = 16 * mt + 4 * q + r;
          if (j < N) acc[d.oW3 + (long)j * ar::H + 16 * nt + rl] += dd[r];
        }
      }
    }
    {
      float s1 = 0.f, s2 = 0.f, s3 = 0.f;
      for (int r = 0; r < 32; r++) {
        s1 += ar::bf2f(sG1[lane][r]);
        s2 += ar::bf2f(sG2[lane][r]);
        s3 += ar::bf2f(sG3[lane][r]);
      }
      acc[d.ob1 + lane] += s1;
      acc[d.ob2 + lane] += s2;
      if (lane < N) acc[d.ob3 + lane] += s3;
    }
    __syncthreads();
  }
  for (long p = lane; p < d.P; p += 64)
    g.partial[(long)blockIdx.x * d.P + p] = ar::free_param(d, p) ? acc[p] : 0.f;
  loss = group_allreduce_sum<64>(loss);
  if (lane == 0) g.loss[blockIdx.x] = loss;
}

// ---------------------------------------------------------------------------
__device__ inline void ar_store(const ArDev& v, const ar::Dims& d, long p, float w) {
  const int N = d.N;
  const uint16_t hb = ar::f2bf(w);
  if (p < d.ob1) {
    const int k = (int)(p / N), i = (int)(p % N);
    v.w1t[(long)i * ar::H + k] = ar::bf2f(hb);
    v.op1f[ar::pos_op1f(d, k, i)] = hb;
  } else if (p < d.oW2) {
    v.b1[p - d.ob1] = w;
  } else if (p < d.ob2) {
    const long x = p - d.oW2;
    const int kp = (int)(x / ar::H), k = (int)(x % ar::H);
    v.op2f[ar::pos_op2f(kp, k)] = hb;
    v.op2b[ar::pos_op2b(kp, k)] = hb;
  } else if (p < d.oW3) {
    v.b2[p - d.ob2] = w;
  } else if (p < d.ob3) {
    const long x = p - d.oW3;
    const int j = (int)(x / ar::H), kp = (int)(x % ar::H);
    v.op3f[ar::pos_op3f(j, kp)] = hb;
    v.op3b[ar::pos_op3b(d, j, kp)] = hb;
  } else {
    v.b3[p - d.ob3] = w;
  }
}

__global__ void k_ar_pack(ArDev v) {
  const ar::Dims d = ar::dims(v.N);
  const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (p >= d.P) return;
  const bool fr = ar::free_param(d, p);
  if (!fr) v.w[p] = 0.f;
  ar_store(v, d, p, fr ? v.w[p] : 0.f);
}

__global__ void k_ar_adam(ArDev v, ArAdam a) {
  const ar::Dims d = ar::dims(v.N);
  const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (p >= d.P) return;
  if (!ar::free_param(d, p)) return;
  float gs = 0.f;
  for (int b = 0; b < a.nblk; b++) gs += a.partial[(long)b * d.P + p];
  const float m = a.beta1 * v.m[p] + (1.f - a.beta1) * gs;
  const float s = a.beta2 * v.v[p] + (1.f - a.beta2) * gs * gs;
  v.m[p] = m;
  v.v[p] = s;
  const float w = v.w[p] - a.lr * (m / a.bc1) / (sqrtf(s / a.bc2) + a.eps);
  v.w[p] = w;
  ar_store(v, d, p, w);
}

__global__ void k_ar_sum_grad(ArDev v, const float* __restrict__ partial, int nblk, float* out) {
  const ar::Dims d = ar::dims(v.N);
  const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (p >= d.P) return;
  float gs = 0.f;
  for (int b = 0; b < nblk; b++) gs += partial[(long)b * d.P + p];
  out[p] = gs;
}

}  // namespace

// ---------------------------------------------------------------------------
int launch_ar_walk(const ArDev& v, const ArWalk& w, hipStream_t st) {
  if (w.rows <= 0) return 0;
  if (v.N < 1 || v.N > ar::kMaxN) return -1;
  const int nt = (v.N + 15) / 16;
  const dim3 grid((w.rows + 15) / 16), blk(64);
  KScope ks("ar_walk", st);
  switch (nt) {
    case 1: hipLaunchKernelGGL(k_ar_walk<1>, grid, blk, 0, st, v, w); break;
    case 2: hipLaunchKernelGGL(k_ar_walk<2>, grid, blk, 0, st, v, w); break;
    case 3: hipLaunchKernelGGL(k_ar_walk<3>, grid, blk, 0, st, v, w); break;
    default: hipLaunchKernelGGL(k_ar_walk<4>, grid, blk, 0, st, v, w); break;
  }
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_ar_accept(ChainState& cs, const uint64_t* prop, const double* la_prop, const double* lq,
                     unsigned long long* lacc, hipStream_t st) {
  if (cs.nw != 1) return -1;
  hipLaunchKernelGGL(k_ar_accept, dim3((cs.C + 255) / 256), dim3(256), 0, st, cs, prop, la_prop, lq,
                     lacc);
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_unpack_rows(const ChainState& cs, long first, long n_out, int* out, hipStream_t st) {
  const long total = (long)cs.C * cs.N;
  hipLaunchKernelGGL(k_unpack_rows, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, cs,
                     first, n_out, out);
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_bits_append(const int* spins, long n, int N, uint64_t* win, long cap, long pos,
                       hipStream_t st) {
  if (n <= 0) return 0;
  if (N > 64) return -1;
  hipLaunchKernelGGL(k_bits_append, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, spins, n,
                     N, win, cap, pos);
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_spins_to_bits(const int* spins, long n, int N, uint64_t* out, hipStream_t st) {
  if (n <= 0) return 0;
  if (N > 64) return -1;
  hipLaunchKernelGGL(k_spins_to_bits, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, spins,
                     n, N, out);
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_ar_grad(const ArDev& v, const ArGrad& g, hipStream_t st) {
  if (g.count <= 0 || g.n_avail <= 0) return 0;
  if (v.N < 1 || v.N > ar::kMaxN) return -1;
  const int nt = (v.N + 15) / 16;
  const dim3 grid((unsigned)((g.count + 63) / 64)), blk(64);
  KScope ks("ar_grad", st);
  switch (nt) {
    case 1: hipLaunchKernelGGL(k_ar_grad<1>, grid, blk, 0, st, v, g); break;
    case 2: hipLaunchKernelGGL(k_ar_grad<2>, grid, blk, 0, st, v, g); break;
    case 3: hipLaunchKernelGGL(k_ar_grad<3>, grid, blk, 0, st, v, g); break;
    default: hipLaunchKernelGGL(k_ar_grad<4>, grid, blk, 0, st, v, g); break;
  }
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_ar_adam(const ArDev& v, const ArAdam& a, hipStream_t st) {
  const long P = ar::dims(v.N).P;
  KScope ks("ar_adam", st);
  hipLaunchKernelGGL(k_ar_adam, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, st, v, a);
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_ar_pack(const ArDev& v, hipStream_t st) {
  const long P = ar::dims(v.N).P;
  hipLaunchKernelGGL(k_ar_pack, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, st, v);
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_ar_sum_grad(const ArDev& v, const float* partial, int nblk, float* out, hipStream_t st) {
  const long P = ar::dims(v.N).P;
  hipLaunchKernelGGL(k_ar_sum_grad, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, st, v, partial,
                     nblk, out);
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

}  // namespace nqs
