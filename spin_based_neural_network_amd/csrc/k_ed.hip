// k_ed.hip — exact diagonalisation of the kagome Heisenberg model on the
// full 2^N basis (N <= 30), for the lean Lanczos of engine.hip.
//
// Basis state s: bit i set <-> s_i = -1 (src/nqs/nqs_lanczos.c:16-20).
// H = J sum_bonds S_u . S_v:  diagonal (J/4) s_u s_v per bond, and for an
// antiparallel bond (J/2) |s with u, v flipped> (nqs_lanczos.c:380-432).
// One thread per basis state; a wave's 64 consecutive states read, for one
// bond mask m, the 64 states s ^ m, which lie in one aligned 512-byte block
// (or permute inside it when m touches the low 6 bits), so every gather is
// a coalesced block read.  The bond list (at most 6 per cell, 60 at N = 30)
// is staged in LDS.
//
// Sector projection (nqs_symproj.c:786-851): out[s] = (1/G) sum_g chi_g
// in[s_g], where bit i of s_g is bit perm_g[i] of s; s_g is assembled from
// per-(g, 3-bit chunk) tables of OR-masks staged in LDS.
//
// Reductions (dot products, norms) write one partial per workgroup in a
// fixed grid; k_ed_finish adds them in workgroup order, so every scalar of
// the recurrence is reproducible run to run.
#include "nqs_kernels.h"

namespace nqs {

constexpr int kEdBlocks = 4096;      // fixed grid of the vector kernels
constexpr int kEdMaxBonds = 64;

__device__ inline double block_sum_256(double v, double* red) {
  const int tid = threadIdx.x;
  red[tid] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) red[tid] += red[tid + w];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

// out = H in; part[blockIdx] = sum_s in[s] out[s] over the block's states
__global__ __launch_bounds__(256) void k_ed_matvec_kagome(const double* __restrict__ in,
                                                          double* __restrict__ out, long dim,
                                                          const int* __restrict__ bonds, int nb,
                                                          double J, double* __restrict__ part) {
  __shared__ unsigned long long msk[kEdMaxBonds];
  __shared__ int bu[kEdMaxBonds], bv[kEdMaxBonds];
  __shared__ double red[256];
  for (int b = threadIdx.x; b < nb; b += 256) {
    bu[b] = bonds[2 * b];
    bv[b] = bonds[2 * b + 1];
    msk[b] = (1ULL << bu[b]) | (1ULL << bv[b]);
  }
  __syncthreads();
  const double jd = 0.25 * J, jo = 0.5 * J;
  double dot = 0.0;
  for (long s = blockIdx.x * 256L + threadIdx.x; s < dim; s += (long)gridDim.x * 256) {
    const unsigned long long us = (unsigned long long)s;
    const double x = in[s];
    double diag = 0.0, off = 0.0;
    for (int b = 0; b < nb; b++) {
      const int anti = (int)(((us >> bu[b]) ^ (us >> bv[b])) & 1ULL);
      diag += anti ? -jd : jd;
      if (anti) off += jo * in[(long)(us ^ msk[b])];
    }
    const double y = diag * x + off;
    out[s] = y;
    dot += x * y;
  }
  const double t = block_sum_256(dot, red);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// w -= a v + b u (u may be null); part[blockIdx] = sum w^2 when part != null
__global__ __launch_bounds__(256) void k_ed_axpy2(double* __restrict__ w,
                                                  const double* __restrict__ v, double a,
                                                  const double* __restrict__ u, double b, long dim,
                                                  double* __restrict__ part) {
  __shared__ double red[256];
  double acc = 0.0;
  for (long s = blockIdx.x * 256L + threadIdx.x; s < dim; s += (long)gridDim.x * 256) {
    double x = w[s] - a * v[s];
    if (u) x -= b * u[s];
    w[s] = x;
    acc += x * x;
  }
  if (part) {
    const double t = block_sum_256(acc, red);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
  }
}

// out[s] = (1/G) sum_g chi_g in[s_g]; part[blockIdx] = sum out^2
// tab: [G][nchunk][8] OR-masks (bits of s_g contributed by chunk c of s)
__global__ __launch_bounds__(256) void k_ed_project(const double* __restrict__ in,
                                                    double* __restrict__ out, long dim,
                                                    const unsigned long long* __restrict__ tab,
                                                    const double* __restrict__ chi, int G,
                                                    int nchunk, double* __restrict__ part) {
  extern __shared__ unsigned long long stab[];
  __shared__ double red[256];
  const int nt = G * nchunk * 8;
  for (int e = threadIdx.x; e < nt; e += 256) stab[e] = tab[e];
  __syncthreads();
  const double invG = 1.0 / G;
  double acc2 = 0.0;
  for (long s = blockIdx.x * 256L + threadIdx.x; s < dim; s += (long)gridDim.x * 256) {
    const unsigned long long us = (unsigned long long)s;
    double acc = 0.0;
    for (int g = 0; g < G; g++) {
      const unsigned long long* t = stab + (size_t)g * nchunk * 8;
      unsigned long long sg = 0;
      for (int c = 0; c < nchunk; c++) sg |= t[c * 8 + (int)((us >> (3 * c)) & 7ULL)];
      acc += chi[g] * in[(long)sg];
    }
    const double y = acc * invG;
    out[s] = y;
    acc2 += y * y;
  }
  const double t = block_sum_256(acc2, red);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// y = x * c; part[blockIdx] = sum y^2 when part != null
__global__ __launch_bounds__(256) void k_ed_scale(const double* __restrict__ x,
                                                  double* __restrict__ y, double c, long dim,
                                                  double* __restrict__ part) {
  __shared__ double red[256];
  double acc = 0.0;
  for (long s = blockIdx.x * 256L + threadIdx.x; s < dim; s += (long)gridDim.x * 256) {
    const double v = x[s] * c;
    y[s] = v;
    acc += v * v;
  }
  if (part) {
    const double t = block_sum_256(acc, red);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
  }
}

// out[0] = sum of n partials in order (one workgroup)
__global__ __launch_bounds__(256) void k_ed_finish(const double* __restrict__ part, int n,
                                                   double* __restrict__ out) {
  __shared__ double red[256];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) acc += part[i];
  const double t = block_sum_256(acc, red);
  if (threadIdx.x == 0) out[0] = t;
}

// spins of basis states s0 .. s0 + n - 1 as int32 rows (bit i set -> -1)
__global__ void k_ed_state_spins(long s0, long n, int N, int* __restrict__ out) {
  for (long e = blockIdx.x * 256L + threadIdx.x; e < n * N; e += (long)gridDim.x * 256) {
    const long r = e / N;
    const int i = (int)(e - r * N);
    out[e] = (((unsigned long long)(s0 + r) >> i) & 1ULL) ? -1 : 1;
  }
}

// part[blockIdx] = max over the block's entries
__global__ __launch_bounds__(256) void k_ed_max(const double* __restrict__ x, long dim,
                                                double* __restrict__ part) {
  __shared__ double red[256];
  double m = -INFINITY;
  for (long s = blockIdx.x * 256L + threadIdx.x; s < dim; s += (long)gridDim.x * 256)
    m = fmax(m, x[s]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void k_ed_max_finish(const double* __restrict__ part, int n, double* __restrict__ out) {
  if (threadIdx.x != 0) return;
  double m = -INFINITY;
  for (int i = 0; i < n; i++) m = fmax(m, part[i]);
  out[0] = m;
}

// psi[s] = cos(ar[s]) exp(la[s] - lmax[0]) (real part of the amplitude
// relative to the largest); part: sum psi^2
__global__ __launch_bounds__(256) void k_ed_psi(const double* __restrict__ la,
                                                const double* __restrict__ ar,
                                                const double* __restrict__ lmax, long dim,
                                                double* __restrict__ psi,
                                                double* __restrict__ part) {
  __shared__ double red[256];
  const double m = lmax[0];
  double acc = 0.0;
  for (long s = blockIdx.x * 256L + threadIdx.x; s < dim; s += (long)gridDim.x * 256) {
    const double v = cos(ar ? ar[s] : 0.0) * exp(la[s] - m);
    psi[s] = v;
    acc += v * v;
  }
  const double t = block_sum_256(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

static int ed_grid(long dim) {
  const long g = (dim + 255) / 256;
  return (int)(g < kEdBlocks ? g : kEdBlocks);
}

int ed_partials() { return kEdBlocks; }

int launch_ed_matvec_kagome(const double* in, double* out, long dim, const int* bonds, int nb,
                            double J, double* part, double* dot, hipStream_t st) {
  if (nb > kEdMaxBonds) {
    fprintf(stderr, "nqs_hip: kagome ED supports at most %d bonds\n", kEdMaxBonds);
    return -1;
  }
  const int g = ed_grid(dim);
  KScope ks("ed_matvec", st);
  hipLaunchKernelGGL(k_ed_matvec_kagome, dim3(g), dim3(256), 0, st, in, out, dim, bonds, nb, J,
                     part);
  hipLaunchKernelGGL(k_ed_finish, dim3(1), dim3(256), 0, st, part, g, dot);
  count_launch();
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_ed_axpy2(double* w, const double* v, double a, const double* u, double b, long dim,
                    double* part, double* norm2, hipStream_t st) {
  const int g = ed_grid(dim);
  hipLaunchKernelGGL(k_ed_axpy2, dim3(g), dim3(256), 0, st, w, v, a, u, b, dim,
                     norm2 ? part : nullptr);
  count_launch();
  if (norm2) {
    hipLaunchKernelGGL(k_ed_finish, dim3(1), dim3(256), 0, st, part, g, norm2);
    count_launch();
  }
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_ed_project(const double* in, double* out, long dim, const unsigned long long* tab,
                      const double* chi, int G, int nchunk, double* part, double* norm2,
                      hipStream_t st) {
  const size_t lds = sizeof(unsigned long long) * (size_t)G * nchunk * 8;
  if (lds > 120 * 1024) {
    fprintf(stderr, "nqs_hip: projector tables of %d group elements do not fit in LDS\n", G);
    return -1;
  }
  const int g = ed_grid(dim);
  KScope ks("ed_project", st);
  hipLaunchKernelGGL(k_ed_project, dim3(g), dim3(256), lds, st, in, out, dim, tab, chi, G, nchunk,
                     part);
  hipLaunchKernelGGL(k_ed_finish, dim3(1), dim3(256), 0, st, part, g, norm2);
  count_launch();
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_ed_scale(const double* x, double* y, double c, long dim, double* part, double* norm2,
                    hipStream_t st) {
  const int g = ed_grid(dim);
  hipLaunchKernelGGL(k_ed_scale, dim3(g), dim3(256), 0, st, x, y, c, dim, norm2 ? part : nullptr);
  count_launch();
  if (norm2) {
    hipLaunchKernelGGL(k_ed_finish, dim3(1), dim3(256), 0, st, part, g, norm2);
    count_launch();
  }
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_ed_state_spins(long s0, long n, int N, int* out, hipStream_t st) {
  long g = (n * N + 255) / 256;
  if (g > 16384) g = 16384;
  hipLaunchKernelGGL(k_ed_state_spins, dim3((unsigned)g), dim3(256), 0, st, s0, n, N, out);
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_ed_psi(const double* la, const double* ar, long dim, double* psi, double* part,
                  double* scal, hipStream_t st) {
  const int g = ed_grid(dim);
  hipLaunchKernelGGL(k_ed_max, dim3(g), dim3(256), 0, st, la, dim, part);
  hipLaunchKernelGGL(k_ed_max_finish, dim3(1), dim3(64), 0, st, part, g, scal);
  hipLaunchKernelGGL(k_ed_psi, dim3(g), dim3(256), 0, st, la, ar, scal, dim, psi, part);
  hipLaunchKernelGGL(k_ed_finish, dim3(1), dim3(256), 0, st, part, g, scal + 1);
  for (int i = 0; i < 4; i++) count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

}  // namespace nqs
