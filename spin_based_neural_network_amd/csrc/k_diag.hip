// k_diag.hip — reductions of the sample diagnostics.
//
// Tr(S) of the quantum geometric tensor (src/nqs/nqs_diagnostics.c:14-92):
//   Tr(S) = sum_k ( <|O_k|^2> - |<O_k>|^2 ),  |O_k|^2 = Re(O_k)^2 + Im(O_k)^2,
// over the rows O [B][P] of a batch; a column's value in (-1e-12, 0) counts
// as 0 (the reference's clamp of rounding below zero).
//
// k_chi_partial: workgroup (x, y) owns 256 columns and the y-th slice of
// the samples; each thread sums its column over the slice in sample order
// (coalesced rows), partials [S][3][P].  k_chi_final: one workgroup; the
// partials of a column are added in slice order, the per-thread totals in a
// fixed tree.  The result does not depend on timing.
#include "nqs_kernels.h"

namespace nqs {

constexpr int kChiSlices = 64;

__global__ __launch_bounds__(256) void k_chi_partial(const double* __restrict__ O_re,
                                                     const double* __restrict__ O_im, long B,
                                                     long P, int S, double* __restrict__ part) {
  const long k = blockIdx.x * 256L + threadIdx.x;
  if (k >= P) return;
  const long per = (B + S - 1) / S;
  const long b0 = blockIdx.y * per, b1 = b0 + per < B ? b0 + per : B;
  double sr = 0.0, si = 0.0, sq = 0.0;
  for (long b = b0; b < b1; b++) {
    const double r = O_re[b * P + k];
    const double i = O_im ? O_im[b * P + k] : 0.0;
    sr += r;
    si += i;
    sq += r * r + i * i;
  }
  double* p = part + (size_t)blockIdx.y * 3 * P;
  p[k] = sr;
  p[P + k] = si;
  p[2 * P + k] = sq;
}

__global__ __launch_bounds__(256) void k_chi_final(const double* __restrict__ part, long B, long P,
                                                   int S, double* __restrict__ out) {
  __shared__ double red[256];
  const int tid = threadIdx.x;
  const double inv = 1.0 / (double)B;
  double acc = 0.0;
  for (long k = tid; k < P; k += 256) {
    double sr = 0.0, si = 0.0, sq = 0.0;
    for (int y = 0; y < S; y++) {
      const double* p = part + (size_t)y * 3 * P;
      sr += p[k];
      si += p[P + k];
      sq += p[2 * P + k];
    }
    const double mr = sr * inv, mi = si * inv;
    double v = sq * inv - (mr * mr + mi * mi);
    if (v < 0.0 && v > -1e-12) v = 0.0;
    acc += v;
  }
  red[tid] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) red[tid] += red[tid + w];
    __syncthreads();
  }
  if (tid == 0) out[0] = red[0];
}

size_t chi_work_elems(long P) { return (size_t)kChiSlices * 3 * (size_t)P; }

int launch_chi_trace(const double* O_re, const double* O_im, long B, long P, double* work,
                     double* out, hipStream_t st) {
  if (B <= 0 || P <= 0) return -1;
  const int S = B < kChiSlices ? (int)B : kChiSlices;
  hipLaunchKernelGGL(k_chi_partial, dim3((unsigned)((P + 255) / 256), S), dim3(256), 0, st, O_re,
                     O_im, B, P, S, work);
  count_launch();
  hipLaunchKernelGGL(k_chi_final, dim3(1), dim3(256), 0, st, work, B, P, S, out);
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

}  // namespace nqs
