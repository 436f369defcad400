// k_dyn.hip — per-sample helpers of the excited-state SR step and of the
// Marshall sign on device log-amplitudes.
//
// Excited-state SR (src/nqs/nqs_optimizer.c:562-716): r(s) = psi_ref(s) /
// psi(s) with the log-ratio clamped to [-10, 10]; the local energies are
// augmented by mu r(s) conj(<r>), and the force uses the augmented means in
// the closed form the reference writes, <E_re> + mu |<r>|^2 and
// <E_im> + mu (<r_im><r_re> - <r_re><r_im>).
#include "nqs_kernels.h"

namespace nqs {

__global__ void k_excited_ratio(const double* __restrict__ la_p, const double* __restrict__ ar_p,
                                const double* __restrict__ la_r, const double* __restrict__ ar_r,
                                int B, double* __restrict__ r_re, double* __restrict__ r_im) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  double dl = la_r[b] - la_p[b];
  dl = dl > 10.0 ? 10.0 : (dl < -10.0 ? -10.0 : dl);
  const double m = exp(dl), th = ar_r[b] - ar_p[b];
  double s, c;
  sincos(th, &s, &c);
  r_re[b] = m * c;
  r_im[b] = m * s;
}

// E += mu r conj(<r>) per sample; rm = {<r_re>, <r_im>} (device scalars)
__global__ void k_excited_augment(double* __restrict__ e_re, double* __restrict__ e_im,
                                  const double* __restrict__ r_re, const double* __restrict__ r_im,
                                  const double* __restrict__ rm, double mu, int B) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const double Rr = rm[0], Ri = rm[1];
  e_re[b] += mu * (r_re[b] * Rr + r_im[b] * Ri);
  e_im[b] += mu * (r_im[b] * Rr - r_re[b] * Ri);
}

// out[0] = <E_re> + mu (Rr^2 + Ri^2), out[1] = <E_im> + mu (Ri Rr - Rr Ri)
__global__ void k_excited_means(const double* __restrict__ er, const double* __restrict__ ei,
                                const double* __restrict__ rm, double mu,
                                double* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double Rr = rm[0], Ri = rm[1];
  out[0] = er[0] + mu * (Rr * Rr + Ri * Ri);
  out[1] = ei[0] + mu * (Ri * Rr - Rr * Ri);
}

// arg psi += pi for configurations with an odd number of down spins on the
// sublattice (x + y) even, site x * Ly + y (src/nqs/nqs_marshall.c:16-40)
__global__ void k_marshall_arg(const int* __restrict__ spins, int B, int N, int Lx, int Ly,
                               double* __restrict__ ar) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const int* s = spins + (size_t)b * N;
  int n = 0;
  for (int x = 0; x < Lx; x++)
    for (int y = 0; y < Ly; y++)
      if (((x + y) & 1) == 0 && x * Ly + y < N && s[x * Ly + y] < 0) n++;
  if (n & 1) ar[b] += 3.14159265358979323846;
}

int launch_excited_ratio(const double* la_p, const double* ar_p, const double* la_r,
                         const double* ar_r, int B, double* r_re, double* r_im, hipStream_t st) {
  if (B <= 0) return 0;
  hipLaunchKernelGGL(k_excited_ratio, dim3((B + 255) / 256), dim3(256), 0, st, la_p, ar_p, la_r,
                     ar_r, B, r_re, r_im);
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_excited_augment(double* e_re, double* e_im, const double* r_re, const double* r_im,
                           const double* rm, double mu, int B, hipStream_t st) {
  if (B <= 0) return 0;
  hipLaunchKernelGGL(k_excited_augment, dim3((B + 255) / 256), dim3(256), 0, st, e_re, e_im, r_re,
                     r_im, rm, mu, B);
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_excited_means(const double* er, const double* ei, const double* rm, double mu,
                         double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_excited_means, dim3(1), dim3(64), 0, st, er, ei, rm, mu, out);
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_marshall_arg(const int* spins, int B, int N, int Lx, int Ly, double* ar,
                        hipStream_t st) {
  if (B <= 0) return 0;
  hipLaunchKernelGGL(k_marshall_arg, dim3((B + 255) / 256), dim3(256), 0, st, spins, B, N, Lx, Ly,
                     ar);
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

}  // namespace nqs
