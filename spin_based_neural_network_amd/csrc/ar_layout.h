// ar_layout.h — the masked autoregressive network of the learned proposal:
// parameter layout, MADE degrees and masks, BF16 rounding, and the order in
// which its weights are stored as gfx950 MFMA operands.  Shared by the host
// engine and the kernels (k_ar.hip).
//
// Network (N sites, H = 64 hidden units per layer, ReLU):
//   z1 = b1 + W1m s          (s_i = +-1, W1m = W1 * mask1)        [H]
//   z2 = b2 + W2m a1,  a1 = bf16(relu(z1))                        [H]
//   l  = b3 + W3m a2,  a2 = bf16(relu(z2))                        [N]
//   p(s_j = -1 | s_<j) = sigmoid(l_j)
// Hidden unit k has degree m(k) = 1 + floor(k (N-1) / H), non-decreasing in k,
// so 16-unit tiles and 32-unit chunks cover contiguous degree ranges; site i
// has degree i+1.
// mask1[k][i] = m(k) >= i+1, mask2[k'][k] = m(k') >= m(k),
// mask3[j][k'] = j+1 > m(k'), so l_j depends on s_0..s_{j-1} only.
// Parameters (FP32 master), P = 2HN + H^2 + 2H + N:
//   [W1 (H x N) | b1 (H) | W2 (H x H) | b2 (H) | W3 (N x H) | b3 (N)].
//
// MFMA operand order (v_mfma_f32_16x16x32_bf16, wave64): lane l feeds row
// (A) or column (B) l%16 and eight K entries e = 0..7 from lane group
// q = l/16.  Every product below sums over a hidden or output index f whose
// K slot is (chunk c, group q, entry e) with
//   f = pi(c, q, e) = 32c + 16(e/4) + 4q + e%4,
// which is exactly where the MFMA result D[4q + r][l%16] of tile t puts
// f = 16t + 4q + r.  A layer's output therefore feeds the next product from
// the same registers, with no transpose.  Site (input) K slots use the plain
// order 32c + 8q + e.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define NQS_AR_HD __host__ __device__
#else
#define NQS_AR_HD
#endif

namespace nqs {
namespace ar {

constexpr int H = 64;
constexpr int kMaxN = 64;

struct Dims {
  int N = 0, nt = 0, c1 = 0, c3 = 0;
  long P = 0, oW1 = 0, ob1 = 0, oW2 = 0, ob2 = 0, oW3 = 0, ob3 = 0;
  // operand array sizes in 16-bit elements
  long n_op1f = 0, n_op2 = 0, n_op3f = 0, n_op3b = 0;
};

NQS_AR_HD inline Dims dims(int N) {
  Dims d;
  d.N = N;
  d.nt = (N + 15) / 16;
  d.c1 = (N + 31) / 32;
  d.c3 = (d.nt + 1) / 2;
  d.oW1 = 0;
  d.ob1 = (long)H * N;
  d.oW2 = d.ob1 + H;
  d.ob2 = d.oW2 + (long)H * H;
  d.oW3 = d.ob2 + H;
  d.ob3 = d.oW3 + (long)N * H;
  d.P = d.ob3 + N;
  d.n_op1f = 4L * d.c1 * 64 * 8;
  d.n_op2 = 4L * 2 * 64 * 8;
  d.n_op3f = (long)d.nt * 2 * 64 * 8;
  d.n_op3b = 4L * d.c3 * 64 * 8;
  return d;
}

NQS_AR_HD inline int degree(int k, int N) { return N > 1 ? 1 + (k * (N - 1)) / H : 1; }
NQS_AR_HD inline bool mask1(int k, int i, int N) { return degree(k, N) >= i + 1; }
NQS_AR_HD inline bool mask2(int kp, int k, int N) { return degree(kp, N) >= degree(k, N); }
NQS_AR_HD inline bool mask3(int j, int kp, int N) { return j + 1 > degree(kp, N); }

// whether parameter p is a free (unmasked) parameter
// (P < 2^15, so the index arithmetic is 32-bit)
NQS_AR_HD inline bool free_param(const Dims& d, long p) {
  const int N = d.N, x = (int)p;
  if (x < (int)d.ob1) return mask1(x / N, x % N, N);
  if (x < (int)d.oW2) return true;
  if (x < (int)d.ob2) {
    const int q = x - (int)d.oW2;
    return mask2(q / H, q % H, N);
  }
  if (x < (int)d.oW3) return true;
  if (x < (int)d.ob3) {
    const int q = x - (int)d.oW3;
    return mask3(q / H, q % H, N);
  }
  return true;
}

// K slot of feature f: chunk c, lane group q, entry e
NQS_AR_HD inline void slot_of(int f, int& c, int& q, int& e) {
  c = f >> 5;
  q = (f >> 2) & 3;
  e = ((f >> 4) & 1) * 4 + (f & 3);
}
NQS_AR_HD inline int feature_of(int c, int q, int e) { return 32 * c + 16 * (e >> 2) + 4 * q + (e & 3); }

// BF16 round-to-nearest-even of an FP32 value, and back.
NQS_AR_HD inline uint16_t f2bf(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
NQS_AR_HD inline float bf2f(uint16_t h) {
  return __builtin_bit_cast(float, (uint32_t)h << 16);
}

// Operand positions (16-bit element index) of one weight.
// op1f[mt][c][lane][e] = W1m[16mt + lane%16][32c + 8(lane/16) + e]
NQS_AR_HD inline long pos_op1f(const Dims& d, int k, int i) {
  const int mt = k >> 4, c = i >> 5, q = (i >> 3) & 3, e = i & 7;
  return (((long)mt * d.c1 + c) * 64 + (k & 15) + 16 * q) * 8 + e;
}
// op2f[mt][c][lane][e] = W2m[16mt + lane%16][pi(c, lane/16, e)]
NQS_AR_HD inline long pos_op2f(int kp, int k) {
  int c, q, e;
  slot_of(k, c, q, e);
  return (((long)(kp >> 4) * 2 + c) * 64 + (kp & 15) + 16 * q) * 8 + e;
}
// op2b[mt][c][lane][e] = W2m[pi(c, lane/16, e)][16mt + lane%16]
NQS_AR_HD inline long pos_op2b(int kp, int k) {
  int c, q, e;
  slot_of(kp, c, q, e);
  return (((long)(k >> 4) * 2 + c) * 64 + (k & 15) + 16 * q) * 8 + e;
}
// op3f[mt][c][lane][e] = W3m[16mt + lane%16][pi(c, lane/16, e)]
NQS_AR_HD inline long pos_op3f(int j, int kp) {
  int c, q, e;
  slot_of(kp, c, q, e);
  return (((long)(j >> 4) * 2 + c) * 64 + (j & 15) + 16 * q) * 8 + e;
}
// op3b[mt][c][lane][e] = W3m[pi(c, lane/16, e)][16mt + lane%16]
NQS_AR_HD inline long pos_op3b(const Dims& d, int j, int kp) {
  int c, q, e;
  slot_of(j, c, q, e);
  return (((long)(kp >> 4) * d.c3 + c) * 64 + (kp & 15) + 16 * q) * 8 + e;
}

}  // namespace ar
}  // namespace nqs
