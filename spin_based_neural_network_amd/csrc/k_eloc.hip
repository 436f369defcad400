// k_eloc.hip — local energies E_loc(s) = sum_s' H_ss' psi(s')/psi(s).
//
// Reference semantics: src/nqs/nqs_gradient.c (TFIM :50-82, XXZ /
// Heisenberg :87-139, J1-J2 :146-205, kagome :374-445, real path = real
// part of the complex path :495-875).  The reference re-evaluates
// log psi(s') from scratch for every connected configuration (:24-45);
// here each connected configuration is priced from the sample's
// tanh(theta) and the tables C = cosh(2W), S = sinh(2W):
//   psi(s')/psi(s) = exp(sum_flipped -2 a_i s_i) * prod_h [cosh d_h + t_h sinh d_h]
// which needs no transcendental per hidden unit.
//
// Work decomposition: a workgroup takes SB consecutive samples, compacts the
// active (sample, term) pairs (all flips; pairs with s_a != s_b) into a
// sample-major list in LDS, and gives one lane per active pair.  The hidden
// dimension is walked in tiles of HT units whose table rows (HT x N) and
// t values (SB x HT) are staged in LDS; each lane keeps a running product
// renormalised once per tile.  Per-sample sums run in a fixed order.
#include "nqs_kernels.h"

namespace nqs {

__device__ inline int block_exclusive_scan(int v, int* sh, int& total) {
  // Hillis-Steele over blockDim.x (256) entries
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = 1; off < (int)blockDim.x; off <<= 1) {
    int x = (t >= off) ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += x;
    __syncthreads();
  }
  int incl = sh[t];
  total = sh[blockDim.x - 1];
  __syncthreads();
  return incl - v;
}

template <typename R, bool CPLX, int IPT>
__global__ __launch_bounds__(256) void k_eloc(ModelView<R> m, HamDev H,
                                              const int* __restrict__ spins, int B,
                                              const R* __restrict__ t_re,
                                              const R* __restrict__ t_im,
                                              double* __restrict__ e_re,
                                              double* __restrict__ e_im, int SB, int HT) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int scan_sh[256];
  __shared__ int n_items_sh;
  const int N = m.N, M = m.M;
  const int T = H.n_flip + H.n_pair;
  const int tid = threadIdx.x;
  // carve dynamic LDS (all offsets 16-byte aligned)
  unsigned char* p = smem;
  auto carve = [&](size_t bytes) { unsigned char* q = p; p += (bytes + 15) & ~(size_t)15; return q; };
  signed char* sp = (signed char*)carve((size_t)SB * N);
  int* items = (int*)carve(sizeof(int) * (size_t)SB * T);
  int* starts = (int*)carve(sizeof(int) * (SB + 1));
  R* cr = (R*)carve(sizeof(R) * (size_t)HT * N);
  R* sr = (R*)carve(sizeof(R) * (size_t)HT * N);
  R* ci = CPLX ? (R*)carve(sizeof(R) * (size_t)HT * N) : nullptr;
  R* si = CPLX ? (R*)carve(sizeof(R) * (size_t)HT * N) : nullptr;
  R* ttr = (R*)carve(sizeof(R) * (size_t)SB * HT);
  R* tti = CPLX ? (R*)carve(sizeof(R) * (size_t)SB * HT) : nullptr;
  double* res_r = (double*)carve(sizeof(double) * (size_t)SB * T);
  double* res_i = (double*)carve(sizeof(double) * (size_t)SB * T);

  for (int b0 = blockIdx.x * SB; b0 < B; b0 += gridDim.x * SB) {
    const int nsb = (B - b0) < SB ? (B - b0) : SB;
    for (int e = tid; e < nsb * N; e += blockDim.x) sp[e] = (signed char)spins[(size_t)b0 * N + e];
    if (tid < SB + 1) starts[tid] = 0;
    __syncthreads();
    // ---- compaction of active (sample, term) pairs, sample-major order ----
    int base = 0;
    const int ncand = nsb * T;
    for (int c0 = 0; c0 < ncand; c0 += blockDim.x) {
      const int c = c0 + tid;
      int flag = 0, sb = 0, t = 0;
      if (c < ncand) {
        sb = c / T;
        t = c - sb * T;
        if (t < H.n_flip) flag = 1;
        else {
          const int pidx = t - H.n_flip;
          flag = sp[sb * N + H.pair_a[pidx]] != sp[sb * N + H.pair_b[pidx]];
        }
      }
      int total;
      const int off = block_exclusive_scan(flag, scan_sh, total);
      if (flag) {
        items[base + off] = (sb << 16) | t;
        atomicAdd(&starts[sb + 1], 1);
      }
      base += total;
    }
    if (tid == 0) {
      n_items_sh = base;
      for (int s = 0; s < nsb; s++) starts[s + 1] += starts[s];
    }
    __syncthreads();
    const int n_items = n_items_sh;

    // ---- per-item state ----
    R Lm[IPT], Pr[IPT], Pi[IPT], sgn[IPT];
    int isb[IPT], ia[IPT], ib[IPT];
#pragma unroll
    for (int q = 0; q < IPT; q++) {
      const int idx = tid + q * (int)blockDim.x;
      Lm[q] = 0; Pr[q] = 1; Pi[q] = 0; sgn[q] = 1; isb[q] = -1; ia[q] = 0; ib[q] = -1;
      if (idx < n_items) {
        const int it = items[idx];
        const int sb = it >> 16, t = it & 0xFFFF;
        isb[q] = sb;
        if (t < H.n_flip) { ia[q] = H.flip_i[t]; ib[q] = -1; }
        else { ia[q] = H.pair_a[t - H.n_flip]; ib[q] = H.pair_b[t - H.n_flip]; }
        sgn[q] = (R)sp[sb * N + ia[q]];
      }
    }
    // ---- hidden tiles ----
    for (int h0 = 0; h0 < M; h0 += HT) {
      const int hn = (M - h0) < HT ? (M - h0) : HT;
      const size_t toff = (size_t)h0 * N;
      for (int e = tid; e < hn * N; e += blockDim.x) {
        cr[e] = m.C_re[toff + e];
        sr[e] = m.S_re[toff + e];
        if (CPLX) { ci[e] = m.C_im[toff + e]; si[e] = m.S_im[toff + e]; }
      }
      for (int e = tid; e < nsb * hn; e += blockDim.x) {
        const int s = e / hn, hh = e - s * hn;
        const size_t g = (size_t)(b0 + s) * M + h0 + hh;
        ttr[s * HT + hh] = t_re[g];
        if (CPLX) tti[s * HT + hh] = t_im[g];
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < IPT; q++) {
        if (isb[q] < 0) continue;
        const int a = ia[q], bb = ib[q];
        const R s_a = sgn[q];
        const R* tt = ttr + isb[q] * HT;
        const R* ttI = CPLX ? tti + isb[q] * HT : nullptr;
        R pr = Pr[q], pim = Pi[q];
        for (int hh = 0; hh < hn; hh++) {
          const int ea = hh * N + a;
          R chr, chi = 0, shr, shi = 0;
          if (bb < 0) {
            chr = cr[ea];
            shr = -s_a * sr[ea];
            if (CPLX) { chi = ci[ea]; shi = -s_a * si[ea]; }
          } else {
            const int eb = hh * N + bb;
            if (CPLX) {
              const R car = cr[ea], cai = ci[ea], sar = sr[ea], sai = si[ea];
              const R cbr = cr[eb], cbi = ci[eb], sbr = sr[eb], sbi = si[eb];
              chr = (car * cbr - cai * cbi) - (sar * sbr - sai * sbi);
              chi = (car * cbi + cai * cbr) - (sar * sbi + sai * sbr);
              const R xr = (sar * cbr - sai * cbi) - (car * sbr - cai * sbi);
              const R xi = (sar * cbi + sai * cbr) - (car * sbi + cai * sbr);
              shr = -s_a * xr;
              shi = -s_a * xi;
            } else {
              const R car = cr[ea], sar = sr[ea], cbr = cr[eb], sbr = sr[eb];
              chr = car * cbr - sar * sbr;
              shr = -s_a * (sar * cbr - car * sbr);
            }
          }
          if (CPLX) {
            const R t_r = tt[hh], t_i = ttI[hh];
            const R wr = chr + (t_r * shr - t_i * shi);
            const R wi = chi + (t_r * shi + t_i * shr);
            const R nr = pr * wr - pim * wi;
            pim = pr * wi + pim * wr;
            pr = nr;
          } else {
            pr *= chr + tt[hh] * shr;
          }
        }
        // renormalise once per tile
        if (CPLX) {
          const R mag = sqrt(pr * pr + pim * pim);
          Lm[q] += log(mag);
          pr /= mag;
          pim /= mag;
        } else {
          Lm[q] += log(pr);
          pr = 1;
        }
        Pr[q] = pr;
        Pi[q] = pim;
      }
      __syncthreads();
    }
    // ---- item contributions ----
#pragma unroll
    for (int q = 0; q < IPT; q++) {
      const int idx = tid + q * (int)blockDim.x;
      if (isb[q] < 0) continue;
      const int it = items[idx];
      const int t = it & 0xFFFF;
      const int a = ia[q], bb = ib[q];
      const double s_a = (double)sgn[q];
      // visible term exp(sum_flipped -2 a_i s_i); s_b = -s_a for pairs
      double vr = -2.0 * (double)m.a_re[a] * s_a;
      double vi = CPLX ? -2.0 * (double)m.a_im[a] * s_a : 0.0;
      if (bb >= 0) {
        vr += 2.0 * (double)m.a_re[bb] * s_a;
        if (CPLX) vi += 2.0 * (double)m.a_im[bb] * s_a;
      }
      double mag = exp((double)Lm[q] + vr);
      double ph_r = (double)Pr[q], ph_i = CPLX ? (double)Pi[q] : 0.0;
      if (CPLX) {
        const double c = cos(vi), s = sin(vi);
        const double nr = ph_r * c - ph_i * s;
        ph_i = ph_r * s + ph_i * c;
        ph_r = nr;
      }
      double sign = 1.0;
      if (H.sublat) {
        int par = H.sublat[a] + (bb >= 0 ? H.sublat[bb] : 0);
        if (par & 1) sign = -1.0;
      }
      const double coef = (t < H.n_flip) ? H.flip_c[t] : H.pair_c[t - H.n_flip];
      res_r[idx] = coef * sign * mag * ph_r;
      res_i[idx] = coef * sign * mag * ph_i;
    }
    __syncthreads();
    // ---- per-sample sums in term order ----
    if (tid < nsb) {
      const signed char* s = sp + tid * N;
      double diag = 0.0;
      for (int d = 0; d < H.n_diag; d++)
        diag += H.diag_c[d] * (double)(s[H.diag_a[d]] * s[H.diag_b[d]]);
      double orr = 0.0, oi = 0.0;
      for (int k = starts[tid]; k < starts[tid + 1]; k++) {
        orr += res_r[k];
        oi += res_i[k];
      }
      e_re[b0 + tid] = diag + orr;
      if (e_im) e_im[b0 + tid] = oi;
    }
    __syncthreads();
  }
}

template <typename R, bool CPLX, int IPT>
static int eloc_go(const ModelView<R>& m, const HamDev& h, const int* spins, int B,
                   const R* t_re, const R* t_im, double* e_re, double* e_im, hipStream_t st) {
  const int N = m.N, T = h.n_flip + h.n_pair;
  int SB = (256 * IPT) / (T > 0 ? T : 1);
  if (SB < 1) SB = 1;
  if (SB > 64) SB = 64;
  if (SB > B) SB = B;
  const int ntab = CPLX ? 4 : 2, nt = CPLX ? 2 : 1;
  auto al = [](size_t x) { return (x + 15) & ~(size_t)15; };
  const size_t fixed = al((size_t)SB * N) + al(4 * (size_t)SB * T) + al(4 * (size_t)(SB + 1)) +
                       2 * al(8 * (size_t)SB * T);
  const size_t budget = 62 * 1024;
  int HT = 64;
  while (HT > 1 && fixed + ntab * al(sizeof(R) * (size_t)HT * N) + nt * al(sizeof(R) * (size_t)SB * HT) > budget)
    HT--;
  const size_t bytes = fixed + ntab * al(sizeof(R) * (size_t)HT * N) + nt * al(sizeof(R) * (size_t)SB * HT);
  if (bytes > budget) {
    fprintf(stderr, "nqs_hip: local-energy tile does not fit in LDS (N=%d, terms=%d)\n", N, T);
    return -1;
  }
  int grid = (B + SB - 1) / SB;
  if (grid > 65535) grid = 65535;
  KScope ks("eloc", st);
  hipLaunchKernelGGL((k_eloc<R, CPLX, IPT>), dim3(grid), dim3(256), bytes, st, m, h, spins, B,
                     t_re, t_im, e_re, e_im, SB, HT);
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

template <typename R>
int launch_eloc(const ModelView<R>& m, const HamDev& h, const int* spins, int B,
                const R* t_re, const R* t_im, double* e_re, double* e_im, hipStream_t st) {
  if (B <= 0) return 0;
  const int T = h.n_flip + h.n_pair;
  if (m.N > kMaxSites || T >= 65536) return -1;
  const bool c = m.cplx != 0;
  if (T <= 256) {
    return c ? eloc_go<R, true, 1>(m, h, spins, B, t_re, t_im, e_re, e_im, st)
             : eloc_go<R, false, 1>(m, h, spins, B, t_re, t_im, e_re, e_im, st);
  } else if (T <= 512) {
    return c ? eloc_go<R, true, 2>(m, h, spins, B, t_re, t_im, e_re, e_im, st)
             : eloc_go<R, false, 2>(m, h, spins, B, t_re, t_im, e_re, e_im, st);
  } else if (T <= 1024) {
    return c ? eloc_go<R, true, 4>(m, h, spins, B, t_re, t_im, e_re, e_im, st)
             : eloc_go<R, false, 4>(m, h, spins, B, t_re, t_im, e_re, e_im, st);
  }
  fprintf(stderr, "nqs_hip: more than 1024 off-diagonal terms per sample is not supported\n");
  return -1;
}

// E_loc from host-evaluated log-amplitudes (callback adaptor); one thread per
// sample, terms in order.
__global__ void k_eloc_from_logamps(HamDev H, const int* __restrict__ spins, int B, int N,
                                    int n_terms, const double* __restrict__ la,
                                    const double* __restrict__ ar, double* e_re,
                                    double* e_im) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const int* s = spins + (size_t)b * N;
  const double* l = la + (size_t)b * (n_terms + 1);
  const double* a = ar + (size_t)b * (n_terms + 1);
  double diag = 0.0;
  for (int d = 0; d < H.n_diag; d++) diag += H.diag_c[d] * (double)(s[H.diag_a[d]] * s[H.diag_b[d]]);
  double orr = 0.0, oi = 0.0;
  for (int t = 0; t < n_terms; t++) {
    double c;
    if (t < H.n_flip) c = H.flip_c[t];
    else {
      const int p = t - H.n_flip;
      if (s[H.pair_a[p]] == s[H.pair_b[p]]) continue;
      c = H.pair_c[p];
    }
    const double mag = exp(l[1 + t] - l[0]), d = a[1 + t] - a[0];
    orr += c * mag * cos(d);
    oi += c * mag * sin(d);
  }
  e_re[b] = diag + orr;
  if (e_im) e_im[b] = oi;
}

int launch_eloc_from_logamps(const HamDev& h, const int* spins, int B, int N, int n_terms,
                               const double* la, const double* ar, double* e_re, double* e_im,
                               hipStream_t st) {
  if (B <= 0) return 0;
  hipLaunchKernelGGL(k_eloc_from_logamps, dim3((B + 127) / 128), dim3(128), 0, st, h, spins, B,
                     N, n_terms, la, ar, e_re, e_im);
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

template int launch_eloc<double>(const ModelView<double>&, const HamDev&, const int*, int, const double*, const double*, double*, double*, hipStream_t);
template int launch_eloc<float>(const ModelView<float>&, const HamDev&, const int*, int, const float*, const float*, double*, double*, hipStream_t);

}  // namespace nqs
