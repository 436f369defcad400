// k_orbit.hip — group orbits of configurations for the symmetry-projected
// amplitudes (src/nqs/nqs_symproj.c:30-263, src/nqs/nqs_translation.c).
//
// A batch of B configurations becomes B x G orbit rows (g.s)_i =
// s_{perm[g][i]}, evaluated by the ordinary batched forward pass; the rows'
// log-amplitudes are then combined per configuration:
//   real combination (nqs_symproj_log_amp):   m = max_g lp_g,
//     S = sum_g chi_g cos(arg_g) e^{lp_g - m},
//     log|psi| = log|S| + m - log(G)/2, arg = pi if S < 0 else 0
//     (S = 0: log|psi| = -1e300, arg = 0);
//   gradient weights: w_g = c_g / sum_g' c_g' with c_g = chi_g cos(arg_g)
//     e^{lp_g - m} (real) or c_g = chi_g e^{lp_g - m + i arg_g} (complex),
//     zero weights when the denominator vanishes;
// sums over g run in g order, one thread per configuration.
#include "nqs_kernels.h"

namespace nqs {

__global__ void k_orbit_spins(const int* __restrict__ spins, int B, int N,
                              const int* __restrict__ perm, int G, int* __restrict__ out) {
  const long total = (long)B * G * N;
  for (long e = blockIdx.x * 256L + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    const long row = e / N;
    const int i = (int)(e - row * N);
    const long b = row / G;
    const int g = (int)(row - b * G);
    out[e] = spins[b * N + perm[(size_t)g * N + i]];
  }
}

__global__ void k_orbit_combine(const double* __restrict__ lp, const double* __restrict__ ar,
                                const double* __restrict__ chi, int B, int G,
                                double* __restrict__ la_out, double* __restrict__ ar_out) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const double* l = lp + (size_t)b * G;
  const double* a = ar + (size_t)b * G;
  double m = -1e300;
  for (int g = 0; g < G; g++) m = l[g] > m ? l[g] : m;
  double S = 0.0;
  for (int g = 0; g < G; g++) S += chi[g] * cos(a[g]) * exp(l[g] - m);
  const double aS = fabs(S);
  if (aS > 0.0) {
    la_out[b] = log(aS) + m - 0.5 * log((double)G);
    if (ar_out) ar_out[b] = S < 0.0 ? 3.14159265358979323846 : 0.0;
  } else {
    la_out[b] = -1e300;
    if (ar_out) ar_out[b] = 0.0;
  }
}

// w [B][G] (re, im) from the orbit log-amplitudes
__global__ void k_orbit_weights(const double* __restrict__ lp, const double* __restrict__ ar,
                                const double* __restrict__ chi, int B, int G, int cplx,
                                double* __restrict__ w_re, double* __restrict__ w_im) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const double* l = lp + (size_t)b * G;
  const double* a = ar + (size_t)b * G;
  double* wr = w_re + (size_t)b * G;
  double* wi = w_im + (size_t)b * G;
  double m = -1e300;
  for (int g = 0; g < G; g++) m = l[g] > m ? l[g] : m;
  double dr = 0.0, di = 0.0;
  for (int g = 0; g < G; g++) {
    const double mag = chi[g] * exp(l[g] - m);
    dr += cplx ? mag * cos(a[g]) : chi[g] * cos(a[g]) * exp(l[g] - m);
    if (cplx) di += mag * sin(a[g]);
  }
  const double d2 = dr * dr + di * di;
  for (int g = 0; g < G; g++) {
    if (d2 == 0.0) {
      wr[g] = 0.0;
      wi[g] = 0.0;
      continue;
    }
    if (cplx) {
      const double mag = chi[g] * exp(l[g] - m);
      const double cr = mag * cos(a[g]), ci = mag * sin(a[g]);
      wr[g] = (cr * dr + ci * di) / d2;
      wi[g] = (ci * dr - cr * di) / d2;
    } else {
      wr[g] = chi[g] * cos(a[g]) * exp(l[g] - m) / dr;
      wi[g] = 0.0;
    }
  }
}

// O_sym[b][k] = sum_g w_bg O[b G + g][k] (complex when O_im / w_im given)
__global__ __launch_bounds__(256) void k_orbit_grad(const double* __restrict__ O_re,
                                                    const double* __restrict__ O_im,
                                                    const double* __restrict__ w_re,
                                                    const double* __restrict__ w_im, int B, int G,
                                                    long P, double* __restrict__ out_re,
                                                    double* __restrict__ out_im) {
  for (int b = blockIdx.y; b < B; b += gridDim.y) {
    for (long k = blockIdx.x * 256L + threadIdx.x; k < P; k += (long)gridDim.x * 256) {
      double sr = 0.0, si = 0.0;
      for (int g = 0; g < G; g++) {
        const double wr = w_re[(size_t)b * G + g];
        const double wi = w_im ? w_im[(size_t)b * G + g] : 0.0;
        if (wr == 0.0 && wi == 0.0) continue;
        const size_t row = ((size_t)b * G + g) * P + k;
        const double gr = O_re[row], gi = O_im ? O_im[row] : 0.0;
        sr += wr * gr - wi * gi;
        si += wr * gi + wi * gr;
      }
      out_re[(size_t)b * P + k] = sr;
      if (out_im) out_im[(size_t)b * P + k] = si;
    }
  }
}

int launch_orbit_spins(const int* spins, int B, int N, const int* perm, int G, int* out,
                       hipStream_t st) {
  const long total = (long)B * G * N;
  long g = (total + 255) / 256;
  if (g > 16384) g = 16384;
  if (total <= 0) return 0;
  hipLaunchKernelGGL(k_orbit_spins, dim3((unsigned)g), dim3(256), 0, st, spins, B, N, perm, G, out);
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_orbit_combine(const double* lp, const double* ar, const double* chi, int B, int G,
                         double* la_out, double* ar_out, hipStream_t st) {
  if (B <= 0) return 0;
  hipLaunchKernelGGL(k_orbit_combine, dim3((B + 127) / 128), dim3(128), 0, st, lp, ar, chi, B, G,
                     la_out, ar_out);
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_orbit_grad(const double* lp, const double* ar, const double* chi, int B, int G,
                      int cplx, const double* O_re, const double* O_im, long P, double* w_re,
                      double* w_im, double* out_re, double* out_im, hipStream_t st) {
  if (B <= 0) return 0;
  hipLaunchKernelGGL(k_orbit_weights, dim3((B + 127) / 128), dim3(128), 0, st, lp, ar, chi, B, G,
                     cplx, w_re, w_im);
  int gx = (int)((P + 255) / 256);
  if (gx > 512) gx = 512;
  const int gy = B < 65535 ? B : 65535;
  hipLaunchKernelGGL(k_orbit_grad, dim3(gx, gy), dim3(256), 0, st, O_re, O_im, w_re,
                     cplx ? w_im : nullptr, B, G, P, out_re, out_im);
  count_launch();
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

}  // namespace nqs
