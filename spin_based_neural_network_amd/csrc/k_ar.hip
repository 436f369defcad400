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
// k_ar_mh     learned Metropolis-Hastings rounds of 16 chains per wave:
//             d = 2(log|psi(s')| - log|psi(s)|) + log q(s) - log q(s');
//             accept when d >= 0 or log u < d (u from the chain's xorshift).
// k_ar_propose / k_ar_accept_rounds  the batched form of k_ar_mh: the
//             proposal does not depend on the chain state, so every round's
//             proposals are drawn in one launch, evaluated by one batched
//             amplitude pass, and accepted round by round per chain.
// k_ar_grad   forward + backward of the mean negative log-likelihood over a
//             minibatch; ar_grad_rows() rows per single-wave block, dL/dl in
//             FP32, weight gradients as MFMA products over the rows, partial
//             sums per block.
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
// BF16 round-to-nearest-even (v_cvt_pk_bf16_f32; the same rounding as
// ar::f2bf for finite values)
__device__ inline uint16_t to_bf(float v) { return __builtin_bit_cast(uint16_t, (__bf16)v); }
// chunk c operand from per-tile values z[t][r] (feature 16t + 4q + r)
__device__ inline bf16x8 relu_op(const f32x4* z, int c) {
  bf16x8 r;
#pragma unroll
  for (int e = 0; e < 8; e++) r[e] = (__bf16)relu(z[2 * c + (e >> 2)][e & 3]);
  return r;
}
template <int T>
__device__ inline bf16x8 plain_op(const f32x4* z, int c) {
  bf16x8 r;
#pragma unroll
  for (int e = 0; e < 8; e++) {
    const int t = 2 * c + (e >> 2);
    r[e] = (__bf16)(t < T ? z[t][e & 3] : 0.f);
  }
  return r;
}
// log sigmoid(x) and sigmoid(x) in FP32 from one exponential of -|x|
__device__ inline float log_sigmoid_f(float x, float& sig) {
  const float e = __expf(-fabsf(x));
  const float r = __frcp_rn(1.f + e);
  sig = x >= 0.f ? r : e * r;
  return fminf(x, 0.f) - __logf(1.f + e);
}
__device__ inline f32x4 zero4() { return f32x4{0.f, 0.f, 0.f, 0.f}; }
__device__ inline uint64_t site_mask(int N) { return N >= 64 ? ~0ULL : ((1ULL << N) - 1ULL); }

// ---------------------------------------------------------------------------
// The walk over one 16-row tile, shared by every kernel that samples or
// evaluates q (so that all of them compute the same function).
template <int NT>
struct WalkRegs {
  bf16x8 A2[4][2], A3[NT][2];
  f32x4 B2[4], B3[NT], Z1[4];
};

template <int NT>
__device__ inline void walk_load(const ArDev& v, float* s_w1t, WalkRegs<NT>& W) {
  const int N = v.N, lane = threadIdx.x, q = lane >> 4;
  for (int e = lane; e < N * ar::H; e += 64) s_w1t[e] = v.w1t[e];
#pragma unroll
  for (int mt = 0; mt < 4; mt++)
#pragma unroll
    for (int c = 0; c < 2; c++) W.A2[mt][c] = ld_op(v.op2f, ((long)(mt * 2 + c) * 64 + lane) * 8);
#pragma unroll
  for (int mt = 0; mt < NT; mt++)
#pragma unroll
    for (int c = 0; c < 2; c++) W.A3[mt][c] = ld_op(v.op3f, ((long)(mt * 2 + c) * 64 + lane) * 8);
#pragma unroll
  for (int t = 0; t < 4; t++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      W.B2[t][r] = v.b2[16 * t + 4 * q + r];
      W.Z1[t][r] = v.b1[16 * t + 4 * q + r];
    }
#pragma unroll
  for (int t = 0; t < NT; t++)
#pragma unroll
    for (int r = 0; r < 4; r++) W.B3[t][r] = v.b3[16 * t + 4 * q + r];
}

// Lane (row rl = lane%16, group q = lane/16).  samp: draw each site with
// the row's xorshift state `rng` (one draw per site); otherwise read the
// sites from `bits`.  `target`: down spins of the row's sector.  Returns the
// configuration in `ob` and log q in `lq`, identical in the row's 4 lanes.
// The conditional is p = sigmoid(l) in FP32 with l clamped to [-16, 16]
// (so 0 < p < 1 in FP32); site j is down when u < p, which happens with
// probability p, and up with probability 1 - p (FP64), so q is the product
// of exactly the probabilities the draws use.  The product is kept in FP64
// as a mantissa and a power of two (exact frexp renormalisation), and its
// logarithm is taken once at the end.
template <int NT>
__device__ inline void walk_tile(const WalkRegs<NT>& W, const float* s_w1t, int N, bool samp,
                                 int constrained, uint64_t bits, int target, uint64_t& rng,
                                 uint64_t& ob, double& lq) {
  const int lane = threadIdx.x, rl = lane & 15, q = lane >> 4;
  f32x4 z1[4];
#pragma unroll
  for (int t = 0; t < 4; t++) z1[t] = W.Z1[t];
  ob = 0;
  lq = 0.0;
  double pr = 1.0;   // product of the conditionals as pr * 2^pe
  int pe = 0;
  int downs = 0;
  for (int j = 0; j < N; j++) {
    const bf16x8 a0 = relu_op(z1, 0), a1 = relu_op(z1, 1);
    // all eight products issue back to back (four independent chains) so
    // their latencies overlap; masked weights contribute exact zeros
    // the two K chunks of each product are independent MFMAs added on the
    // VALU, so a layer costs one MFMA latency
    f32x4 z2[4], z2b[4];
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
      z2[mt] = mfma(W.A2[mt][0], a0, W.B2[mt]);
      z2b[mt] = mfma(W.A2[mt][1], a1, zero4());
    }
#pragma unroll
    for (int mt = 0; mt < 4; mt++) z2[mt] += z2b[mt];
    const bf16x8 c0 = relu_op(z2, 0), c1 = relu_op(z2, 1);
    const int mj = j >> 4;
    f32x4 lg = zero4();
#pragma unroll
    for (int t = 0; t < NT; t++)
      if (t == mj) lg = mfma(W.A3[t][0], c0, W.B3[t]) + mfma(W.A3[t][1], c1, zero4());
    const int rj = j & 3, qj = (j >> 2) & 3;
    const float mine = rj == 0 ? lg[0] : rj == 1 ? lg[1] : rj == 2 ? lg[2] : lg[3];
    const float l = __shfl(mine, rl + 16 * qj, 64);

    const int rem_d = target - downs, rem = N - j;
    const bool forced = constrained && (rem_d <= 0 || rem_d >= rem);
    // the 53-bit uniform u = k 2^-53 of the row's stream; p is an FP32
    // value >= 2^-24, so u < p is the exact integer test k < p 2^53
    const uint64_t k53 = samp ? (xs_next(rng) >> 11) : 0ULL;
    const float lc = fminf(fmaxf(l, -16.f), 16.f);
    const float pf = __builtin_amdgcn_rcpf(1.f + __expf(-lc));   // p(down), v_rcp_f32
    const double p = (double)pf;
    // p 2^53 as an integer from the bits of pf (normal, 2^-24 <= pf < 1)
    const uint32_t fb = __float_as_uint(pf);
    const uint64_t thr = (uint64_t)((fb & 0x7FFFFFu) | 0x800000u) << (((fb >> 23) & 0xFF) - 127 - 23 + 53);
    const int drawn = samp ? (k53 < thr ? 1 : 0) : (int)((bits >> j) & 1ULL);
    const int bit = forced ? (rem_d > 0 ? 1 : 0) : drawn;
    int e2;
    pr = frexp(pr * (forced ? 1.0 : (bit ? p : 1.0 - p)), &e2);
    pe += e2;
    downs += bit;
    ob |= (uint64_t)bit << j;
    const float sg = bit ? -1.f : 1.f;
    const float* wr = s_w1t + j * ar::H + 4 * q;
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
      for (int r = 0; r < 4; r++) z1[t][r] = fmaf(wr[16 * t + r], sg, z1[t][r]);
  }
  lq = log(pr) + (double)pe * 0.69314718055994530942;
}

template <int NT>
__global__ __launch_bounds__(64) void k_ar_walk(ArDev v, ArWalk w) {
  __shared__ float s_w1t[ar::kMaxN * ar::H];
  WalkRegs<NT> W;
  walk_load<NT>(v, s_w1t, W);
  const int N = v.N, lane = threadIdx.x, rl = lane & 15, q = lane >> 4;
  const int row = blockIdx.x * 16 + rl;
  const bool live = row < w.rows;
  const bool samp = row < w.n_sample;
  uint64_t bits = 0;
  if (live) bits = w.bits[samp ? row : row - w.n_sample];
  const int target = __popcll(bits & site_mask(N));
  uint64_t rng = (live && samp) ? w.rng[row] : 0ULL;
  __syncthreads();
  uint64_t ob;
  double lq;
  walk_tile<NT>(W, s_w1t, N, samp, w.constrained, bits, target, rng, ob, lq);
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
// Learned Metropolis-Hastings rounds, CPW chains per wave (rows 0..CPW-1 of
// the walk tile), all rounds in one launch.  The chains' log q is evaluated
// once at entry (the network or the states may have changed since the last
// launch) and afterwards taken from the accepted proposal.  Per round: walk
// a proposal s' in the chain's sector; log|psi(s')| from scratch, every lane
// taking hidden units h = sub, sub + 64/CPW, ... of chain lane % CPW (W^T
// from LDS when it fits), summed over the chain's lanes by a butterfly;
//   d = 2(log|psi(s')| - cached log|psi(s)|) + log q(s) - log q(s');
// accept when d >= 0 or log u < d; record the chain as sample r*C + c.
template <int NT, typename R, bool CPLX, int CPW>
__global__ __launch_bounds__(64) void k_ar_mh(ArDev v, ModelView<R> m, ChainState cs, ArRounds ar) {
  extern __shared__ __attribute__((aligned(16))) unsigned char dyn_lds[];
  __shared__ float s_w1t[ar::kMaxN * ar::H];
  constexpr int LPC = 64 / CPW;
  WalkRegs<NT> W;
  walk_load<NT>(v, s_w1t, W);
  const int N = v.N, M = m.M, lane = threadIdx.x, rl = lane & 15, q = lane >> 4;
  const int sub = lane / CPW, own = lane % CPW;
  const R* Wr = m.WT_re;
  const R* Wi = m.WT_im;
  if (ar.w_lds) {
    R* sr = (R*)dyn_lds;
    R* si = sr + (size_t)N * M;
    for (int e = lane; e < N * M; e += 64) {
      sr[e] = m.WT_re[e];
      if (CPLX) si[e] = m.WT_im[e];
    }
    Wr = sr;
    Wi = si;
  }
  const int c = blockIdx.x * CPW + rl;
  const bool live = rl < CPW && c < cs.C;
  uint64_t bits = live ? cs.spins[c] : 0ULL;
  uint64_t rng = live ? cs.rng[c] : 0ULL;
  double la = live ? cs.la[c] : 0.0;
  const int target = __popcll(bits & site_mask(N));
  __syncthreads();
  uint64_t ob, none = 0;
  double lq_cur;
  walk_tile<NT>(W, s_w1t, N, false, ar.constrained, bits, target, none, ob, lq_cur);
  unsigned long long nacc = 0;
  for (int r = 0; r < ar.rounds; r++) {
    uint64_t prop;
    double lq_p;
    walk_tile<NT>(W, s_w1t, N, true, ar.constrained, bits, target, rng, prop, lq_p);
    const uint64_t pp = __shfl((unsigned long long)prop, own, 64);
    // log|2 cosh(x + iy)| = |x| + log((1 - e)^2 + 4 e cos^2 y) / 2 with
    // e = exp(-2|x|) (both terms non-negative: no cancellation); the factors
    // of a lane's hidden units are multiplied (each lies in (0, 4]) and one
    // logarithm is taken per lane, or earlier if the product gets small
    double lp = 0.0, prod = 1.0;
    for (int h = sub; h < M; h += LPC) {
      R xr = m.b_re[h], xi = CPLX ? m.b_im[h] : (R)0;
#pragma unroll 8
      for (int i = 0; i < N; i++) {
        const R sg = ((pp >> i) & 1ULL) ? (R)-1 : (R)1;
        xr += Wr[(size_t)i * M + h] * sg;
        if (CPLX) xi += Wi[(size_t)i * M + h] * sg;
      }
      const double ax = fabs((double)xr), e = exp(-2.0 * ax);
      lp += ax;
      if (CPLX) {
        const double cy = cos((double)xi);
        prod *= (1.0 - e) * (1.0 - e) + 4.0 * e * cy * cy;
      } else {
        prod *= 1.0 + e;
      }
      if (prod < 1e-200) {
        lp += (CPLX ? 0.5 : 1.0) * log(prod);
        prod = 1.0;
      }
    }
    lp += (CPLX ? 0.5 : 1.0) * log(prod);
    for (int i = sub; i < N; i += LPC) lp += (double)m.a_re[i] * (((pp >> i) & 1ULL) ? -1.0 : 1.0);
#pragma unroll
    for (int msk = CPW; msk < 64; msk <<= 1) lp += __shfl_xor(lp, msk, 64);
    // row lanes (rl < CPW) hold their own chain's sum: lane % CPW == rl
    const double d = 2.0 * (lp - la) + (lq_cur - lq_p);
    const bool acc = d >= 0.0 ? true : (log(xs_uniform(rng)) < d);
    if (acc) {
      bits = prop;
      la = lp;
      lq_cur = lq_p;
      nacc++;
    }
    if (live && ar.out) {
      const long b = (long)r * cs.C + c;
      if (b < ar.n_out)
        for (int i = q; i < N; i += 4) ar.out[b * N + i] = ((bits >> i) & 1ULL) ? -1 : 1;
    }
  }
  if (live && q == 0) {
    cs.spins[c] = bits;
    cs.rng[c] = rng;
    cs.la[c] = la;
    cs.acc[c] += nacc;
    cs.tot[c] += (unsigned long long)ar.rounds;
    ar.lacc[c] += nacc;
  }
}

// ---------------------------------------------------------------------------
// Learned rounds in bulk.  The proposal distribution q does not depend on the
// chain's state, so the proposals of all R rounds of a batch are drawn at once
// (k_ar_propose: R*C sampling rows plus C rows evaluating the chains' current
// states), their amplitudes are evaluated at once, and only the accept /
// reject decisions run round by round per chain (k_ar_accept_rounds).  This
// is the same Markov chain as drawing each proposal when its round comes.
// Row r*C + c draws with the xorshift state splitmix64(rng_c + (r+1) G) (1 if
// zero), G = 0x9E3779B97F4A7C15, rng_c the chain's state at the batch start.
template <int NT>
__global__ __launch_bounds__(64) void k_ar_propose(ArDev v, ChainState cs, ArBulk b) {
  __shared__ float s_w1t[ar::kMaxN * ar::H];
  WalkRegs<NT> W;
  walk_load<NT>(v, s_w1t, W);
  const int N = v.N, lane = threadIdx.x, rl = lane & 15, q = lane >> 4;
  const long C = cs.C, nsamp = (long)b.rounds * C;
  const long row = (long)blockIdx.x * 16 + rl;
  const bool live = row < nsamp + C;
  const bool samp = row < nsamp;
  const int c = live ? (int)(row % C) : 0;
  const uint64_t bits = live ? cs.spins[c] : 0ULL;
  const int target = __popcll(bits & site_mask(N));
  uint64_t rng = 0ULL;
  if (live && samp) {
    rng = splitmix64(cs.rng[c] + (uint64_t)(row / C + 1) * 0x9E3779B97F4A7C15ULL);
    if (!rng) rng = 1ULL;
  }
  __syncthreads();
  uint64_t ob;
  double lq;
  walk_tile<NT>(W, s_w1t, N, samp, b.constrained, bits, target, rng, ob, lq);
  if (!live) return;
  if (q == 0) b.lq[row] = lq;
  if (samp) {
    if (q == 0) b.prop[row] = ob;
    for (int i = q; i < N; i += 4) b.spins[row * N + i] = ((ob >> i) & 1ULL) ? -1 : 1;
  }
}

// One thread per chain: rounds of accept / reject against the pre-drawn
// proposals (log|psi| in b.lp, log q in b.lq; the chains' own log q at
// b.lq[R*C + c]); d = 2(log|psi(s')| - cached log|psi(s)|) + log q(s) -
// log q(s'), accept when d >= 0 or log u < d (u from the chain's xorshift).
// The chain's state after round r is sample r*C + c of b.out when that index
// is below b.n_out.  The proposals of up to kAccPre rounds are loaded before
// the first decision (the chain of decisions then waits on no load), and
// the launch uses 64-thread workgroups so that the samples' stores spread
// over C / 64 CUs.
constexpr int kAccPre = 8;
__global__ __launch_bounds__(64) void k_ar_accept_rounds(ChainState cs, ArBulk b) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cs.C) return;
  const long C = cs.C;
  const int N = cs.N;
  uint64_t bits = cs.spins[c], rng = cs.rng[c];
  double la = cs.la[c], lqc = b.lq[(long)b.rounds * C + c];
  unsigned long long nacc = 0;
  for (int r0 = 0; r0 < b.rounds; r0 += kAccPre) {
    double plp[kAccPre], plq[kAccPre];
    uint64_t pb[kAccPre];
#pragma unroll
    for (int u = 0; u < kAccPre; u++) {
      const long i = (long)(r0 + u) * C + c;
      const bool in = r0 + u < b.rounds;
      plp[u] = in ? b.lp[i] : 0.0;
      plq[u] = in ? b.lq[i] : 0.0;
      pb[u] = in ? b.prop[i] : 0ULL;
    }
#pragma unroll
    for (int u = 0; u < kAccPre; u++) {
      if (r0 + u >= b.rounds) break;
      const long i = (long)(r0 + u) * C + c;
      const double d = 2.0 * (plp[u] - la) + (lqc - plq[u]);
      // one draw per round whatever d is, so that the chain's stream (which
      // seeds the next batch's proposals) advances every batch
      const double uu = xs_uniform(rng);
      const bool acc = d >= 0.0 ? true : (log(uu) < d);
      if (acc) {
        bits = pb[u];
        la = plp[u];
        lqc = plq[u];
        nacc++;
      }
      if (b.out && i < b.n_out) {
        for (int k = 0; k < N; k++) b.out[i * N + k] = ((bits >> k) & 1ULL) ? -1 : 1;
        if (b.win && i >= b.win_first)
          b.win[(b.win_pos + i - b.win_first) % b.win_cap] = bits & site_mask(N);
      }
    }
  }
  cs.spins[c] = bits;
  cs.rng[c] = rng;
  cs.la[c] = la;
  cs.acc[c] += nacc;
  cs.tot[c] += (unsigned long long)b.rounds;
  b.lacc[c] += nacc;
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
  for (int p = lane; p < (int)d.P; p += 64) acc[p] = 0.f;

  // every operand of the forward and backward products stays in registers
  // for the whole block (one wave per block, so no other wave hides loads)
  bf16x8 A1[4][2], A2[4][2], A3[NT][2], A2b[4][2], A3b[4][C3];
  f32x4 B1[4], B2[4], B3[NT];
#pragma unroll
  for (int mt = 0; mt < 4; mt++)
#pragma unroll
    for (int c = 0; c < 2; c++) {
      A1[mt][c] = c < c1 ? ld_op(v.op1f, ((long)(mt * c1 + c) * 64 + lane) * 8) : bf16x8{};
      A2[mt][c] = ld_op(v.op2f, ((long)(mt * 2 + c) * 64 + lane) * 8);
      A2b[mt][c] = ld_op(v.op2b, ((long)(mt * 2 + c) * 64 + lane) * 8);
    }
#pragma unroll
  for (int mt = 0; mt < 4; mt++)
#pragma unroll
    for (int c = 0; c < C3; c++) A3b[mt][c] = ld_op(v.op3b, ((long)(mt * C3 + c) * 64 + lane) * 8);
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

  for (int h = 0; h < g.rows_per_block / 32; h++) {
    uint64_t rbits[2];
    bool rvalid[2];
#pragma unroll
    for (int tile = 0; tile < 2; tile++) {
      const long mrow = (long)blockIdx.x * g.rows_per_block + h * 32 + tile * 16 + rl;
      rvalid[tile] = mrow < g.count;
      rbits[tile] = rvalid[tile] ? g.rows[(g.start + mrow) % g.n_avail] : 0ULL;
    }
#pragma unroll 1
    for (int tile = 0; tile < 2; tile++) {
      const int r32 = tile * 16 + rl;
      const bool valid = rvalid[tile];
      const uint64_t bits = rbits[tile];
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
        f32x4 a = mfma(A1[mt][0], X[0], B1[mt]);
        if (c1 > 1) a = mfma(A1[mt][1], X[1], a);
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
              const float l = lg[mt][r];
              float sig;
              const float ls = log_sigmoid_f(l, sig);
              loss -= (double)(x ? ls : ls - l);
              gv = (sig - (float)x) * g.scale;
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
        for (int c = 0; c < C3; c++) a = mfma(A3b[mt][c], G3[c], a);
#pragma unroll
        for (int r = 0; r < 4; r++) g2[mt][r] = z2[mt][r] > 0.f ? a[r] : 0.f;
      }
      const bf16x8 G20 = plain_op<4>(g2, 0), G21 = plain_op<4>(g2, 1);
#pragma unroll
      for (int mt = 0; mt < 4; mt++) {
        f32x4 a = mfma(A2b[mt][1], G21, mfma(A2b[mt][0], G20, zero4()));
#pragma unroll
        for (int r = 0; r < 4; r++) g1[mt][r] = z1[mt][r] > 0.f ? a[r] : 0.f;
      }
      // stage [feature][row] BF16 tiles for the weight gradients
#pragma unroll
      for (int t = 0; t < 4; t++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int f = 16 * t + 4 * q + r;
          sA1[f][r32] = to_bf(relu(z1[t][r]));
          sA2[f][r32] = to_bf(relu(z2[t][r]));
          sG1[f][r32] = to_bf(g1[t][r]);
          sG2[f][r32] = to_bf(g2[t][r]);
        }
#pragma unroll
      for (int t = 0; t < 4; t++)
#pragma unroll
        for (int r = 0; r < 4; r++) sG3[16 * t + 4 * q + r][r32] = t < NT ? to_bf(g3[t < NT ? t : 0][r]) : (uint16_t)0;
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
  // masked entries are written as computed; the consumers of the partials
  // (k_ar_adam, k_ar_sum_grad) skip or zero them
  for (int p = lane; p < (int)d.P; p += 64) g.partial[(long)blockIdx.x * d.P + p] = acc[p];
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
  out[p] = ar::free_param(d, p) ? gs : 0.f;
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

template <typename R, int CPW>
static void ar_mh_go(int nt, const ArDev& v, const ModelView<R>& m, ChainState& cs,
                     const ArRounds& ar, size_t lds, hipStream_t st) {
  const dim3 grid((cs.C + CPW - 1) / CPW), blk(64);
#define NQS_AR_MH(NTV)                                                                         \
  if (m.cplx) hipLaunchKernelGGL((k_ar_mh<NTV, R, true, CPW>), grid, blk, lds, st, v, m, cs, ar); \
  else hipLaunchKernelGGL((k_ar_mh<NTV, R, false, CPW>), grid, blk, lds, st, v, m, cs, ar);
  switch (nt) {
    case 1: NQS_AR_MH(1) break;
    case 2: NQS_AR_MH(2) break;
    case 3: NQS_AR_MH(3) break;
    default: NQS_AR_MH(4) break;
  }
#undef NQS_AR_MH
}

// chains per wave: enough waves for the 256 CUs (16 when C >= 4096, 8 when
// C >= 2048, else 4); W^T in LDS when it takes at most 48 KB
template <typename R>
int launch_ar_mh(const ArDev& v, const ModelView<R>& m, ChainState& cs, const ArRounds& ar_in,
                 hipStream_t st) {
  if (cs.C <= 0 || ar_in.rounds <= 0) return 0;
  if (cs.nw != 1 || v.N != cs.N || m.N != cs.N) return -1;
  ArRounds ar = ar_in;
  const size_t wbytes = (size_t)m.N * m.M * sizeof(R) * (m.cplx ? 2 : 1);
  ar.w_lds = wbytes <= 48 * 1024 ? 1 : 0;
  const size_t lds = ar.w_lds ? wbytes : 0;
  const int nt = (v.N + 15) / 16;
  KScope ks("ar_mh", st);
  if (cs.C >= 4096) ar_mh_go<R, 16>(nt, v, m, cs, ar, lds, st);
  else if (cs.C >= 2048) ar_mh_go<R, 8>(nt, v, m, cs, ar, lds, st);
  else ar_mh_go<R, 4>(nt, v, m, cs, ar, lds, st);
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}
template int launch_ar_mh<double>(const ArDev&, const ModelView<double>&, ChainState&, const ArRounds&,
                                  hipStream_t);
template int launch_ar_mh<float>(const ArDev&, const ModelView<float>&, ChainState&, const ArRounds&,
                                 hipStream_t);

int launch_ar_propose(const ArDev& v, ChainState& cs, const ArBulk& b, hipStream_t st) {
  if (cs.C <= 0 || b.rounds <= 0) return 0;
  if (cs.nw != 1 || v.N != cs.N) return -1;
  const long rows = ((long)b.rounds + 1) * cs.C;
  const int nt = (v.N + 15) / 16;
  const dim3 grid((unsigned)((rows + 15) / 16)), blk(64);
  KScope ks("ar_walk", st);
  switch (nt) {
    case 1: hipLaunchKernelGGL(k_ar_propose<1>, grid, blk, 0, st, v, cs, b); break;
    case 2: hipLaunchKernelGGL(k_ar_propose<2>, grid, blk, 0, st, v, cs, b); break;
    case 3: hipLaunchKernelGGL(k_ar_propose<3>, grid, blk, 0, st, v, cs, b); break;
    default: hipLaunchKernelGGL(k_ar_propose<4>, grid, blk, 0, st, v, cs, b); break;
  }
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_ar_accept_rounds(ChainState& cs, const ArBulk& b, hipStream_t st) {
  if (cs.C <= 0 || b.rounds <= 0) return 0;
  KScope ks("ar_accept", st);
  hipLaunchKernelGGL(k_ar_accept_rounds, dim3((cs.C + 63) / 64), dim3(64), 0, st, cs, b);
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

int launch_ar_grad(const ArDev& v, const ArGrad& g0, hipStream_t st) {
  if (g0.count <= 0 || g0.n_avail <= 0) return 0;
  if (v.N < 1 || v.N > ar::kMaxN) return -1;
  ArGrad g = g0;
  g.rows_per_block = ar_grad_rows(g.count);
  const int nt = (v.N + 15) / 16;
  const dim3 grid((unsigned)ar_grad_blocks(g.count)), blk(64);
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
