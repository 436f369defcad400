// k_ed.hip — exact diagonalisation on the full 2^N basis (N <= 30): the
// matrix-vector products of src/nqs/nqs_lanczos.c and the vector kernels of
// the Lanczos drivers in engine.hip (src/mps/lanczos.c).
//
// Basis state s: bit i set <-> s_i = -1 (src/nqs/nqs_lanczos.c:16-20).
// H = sum_bonds [jd s_u s_v + jo (|s with u, v flipped> if s_u != s_v)]
//     + field sum_i |s with i flipped>:
//   Heisenberg XXZ (nqs_lanczos.c:228-281, kagome :343-432): jd = Jz / 4,
//   jo = J / 2, field = 0;  TFIM (:150-191): jd = -J, jo = 0, field = -Gamma.
// One thread per basis state; a wave's 64 consecutive states read, for one
// flip mask m, the 64 states s ^ m, which lie in one aligned 512-byte block
// (or permute inside it when m touches the low 6 bits), so every gather is
// a coalesced block read.  The bond list (at most 64 bonds) is staged in LDS.
//
// Sector projection (nqs_symproj.c:786-851): out[s] = (1/G) sum_g chi_g
// in[s_g], where bit i of s_g is bit perm_g[i] of s; s_g is assembled from
// per-(g, 3-bit chunk) tables of OR-masks staged in LDS.
//
// Full reorthogonalisation against a Krylov basis V [nv][dim] in HBM:
// k_ed_mdot forms h_j = <V_j, w> for every j (eight basis vectors per sweep
// over w), k_ed_mupdate applies w -= sum_j h_j V_j reading each V_j once.
//
// Reductions (dot products, norms) write one partial per workgroup in a
// fixed grid; the finish kernels add them in a fixed order, so every scalar
// of the recurrence is reproducible run to run.
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
__global__ __launch_bounds__(256) void k_ed_matvec(const double* __restrict__ in,
                                                   double* __restrict__ out, long dim,
                                                   const int* __restrict__ bonds, int nb,
                                                   double jd, double jo, double field, int N,
                                                   double* __restrict__ part) {
  __shared__ unsigned long long msk[kEdMaxBonds];
  __shared__ int bu[kEdMaxBonds], bv[kEdMaxBonds];
  __shared__ double red[256];
  for (int b = threadIdx.x; b < nb; b += 256) {
    bu[b] = bonds[2 * b];
    bv[b] = bonds[2 * b + 1];
    msk[b] = (1ULL << bu[b]) | (1ULL << bv[b]);
  }
  __syncthreads();
  double dot = 0.0;
  for (long s = blockIdx.x * 256L + threadIdx.x; s < dim; s += (long)gridDim.x * 256) {
    const unsigned long long us = (unsigned long long)s;
    const double x = in[s];
    double diag = 0.0, off = 0.0;
    for (int b = 0; b < nb; b++) {
      const int anti = (int)(((us >> bu[b]) ^ (us >> bv[b])) & 1ULL);
      diag += anti ? -jd : jd;
      if (anti && jo != 0.0) off += jo * in[(long)(us ^ msk[b])];
    }
    if (field != 0.0)
      for (int i = 0; i < N; i++) off += field * in[(long)(us ^ (1ULL << i))];
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
// relative to the largest); part: sum psi^2, part2 (when given): sum of
// exp(2 (la[s] - lmax[0])), the squared norm of the complex amplitude
__global__ __launch_bounds__(256) void k_ed_psi(const double* __restrict__ la,
                                                const double* __restrict__ ar,
                                                const double* __restrict__ lmax, long dim,
                                                double* __restrict__ psi,
                                                double* __restrict__ part,
                                                double* __restrict__ part2) {
  __shared__ double red[256];
  const double m = lmax[0];
  double acc = 0.0, tot = 0.0;
  for (long s = blockIdx.x * 256L + threadIdx.x; s < dim; s += (long)gridDim.x * 256) {
    const double a = exp(la[s] - m);
    const double v = cos(ar ? ar[s] : 0.0) * a;
    psi[s] = v;
    acc += v * v;
    tot += a * a;
  }
  const double t = block_sum_256(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
  if (part2) {
    const double t2 = block_sum_256(tot, red);
    if (threadIdx.x == 0) part2[blockIdx.x] = t2;
  }
}

// part[j * gridDim.x + blockIdx] = sum over the block's states of
// V[j][s] w[s], j < nv; eight basis vectors per sweep over w
__global__ __launch_bounds__(256) void k_ed_mdot(const double* __restrict__ V, long dim, int nv,
                                                 const double* __restrict__ w,
                                                 double* __restrict__ part) {
  __shared__ double red[8][256];
  const int tid = threadIdx.x;
  for (int j0 = 0; j0 < nv; j0 += 8) {
    const int nj = nv - j0 < 8 ? nv - j0 : 8;
    double acc[8];
#pragma unroll
    for (int q = 0; q < 8; q++) acc[q] = 0.0;
    for (long s = blockIdx.x * 256L + tid; s < dim; s += (long)gridDim.x * 256) {
      const double x = w[s];
#pragma unroll
      for (int q = 0; q < 8; q++)
        if (q < nj) acc[q] += V[(size_t)(j0 + q) * dim + s] * x;
    }
#pragma unroll
    for (int q = 0; q < 8; q++) red[q][tid] = acc[q];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
      if (tid < h) {
#pragma unroll
        for (int q = 0; q < 8; q++) red[q][tid] += red[q][tid + h];
      }
      __syncthreads();
    }
    if (tid < nj) part[(size_t)(j0 + tid) * gridDim.x + blockIdx.x] = red[tid][0];
    __syncthreads();
  }
}

// out[j] = sum of the n partials of row j (one workgroup per row)
__global__ __launch_bounds__(256) void k_ed_finish_rows(const double* __restrict__ part, int n,
                                                        double* __restrict__ out) {
  __shared__ double red[256];
  const double* p = part + (size_t)blockIdx.x * n;
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) acc += p[i];
  const double t = block_sum_256(acc, red);
  if (threadIdx.x == 0) out[blockIdx.x] = t;
}

// w[s] = (keep ? w[s] : 0) - sum_{j < nv} h[j] V[j][s];
// part[blockIdx] = sum w^2 when part != null
__global__ __launch_bounds__(256) void k_ed_mupdate(const double* __restrict__ V, long dim, int nv,
                                                    const double* __restrict__ h,
                                                    double* __restrict__ w, int keep,
                                                    double* __restrict__ part) {
  __shared__ double red[256];
  double acc = 0.0;
  for (long s = blockIdx.x * 256L + threadIdx.x; s < dim; s += (long)gridDim.x * 256) {
    double x = keep ? w[s] : 0.0;
    for (int j = 0; j < nv; j++) x -= h[j] * V[(size_t)j * dim + s];
    w[s] = x;
    acc += x * x;
  }
  if (part) {
    const double t = block_sum_256(acc, red);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
  }
}

static int ed_grid(long dim) {
  const long g = (dim + 255) / 256;
  return (int)(g < kEdBlocks ? g : kEdBlocks);
}

int ed_partials() { return 2 * kEdBlocks; }

int launch_ed_matvec(const EdOp& op, const double* in, double* out, long dim, double* part,
                     double* dot, hipStream_t st) {
  if (op.nb > kEdMaxBonds) {
    fprintf(stderr, "nqs_hip: exact diagonalisation supports at most %d bonds\n", kEdMaxBonds);
    return -1;
  }
  const int g = ed_grid(dim);
  KScope ks("ed_matvec", st);
  hipLaunchKernelGGL(k_ed_matvec, dim3(g), dim3(256), 0, st, in, out, dim, op.bonds, op.nb, op.jd,
                     op.jo, op.field, op.N, part);
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
  hipLaunchKernelGGL(k_ed_psi, dim3(g), dim3(256), 0, st, la, ar, scal, dim, psi, part,
                     part + kEdBlocks);
  hipLaunchKernelGGL(k_ed_finish, dim3(1), dim3(256), 0, st, part, g, scal + 1);
  hipLaunchKernelGGL(k_ed_finish, dim3(1), dim3(256), 0, st, part + kEdBlocks, g, scal + 2);
  for (int i = 0; i < 5; i++) count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

size_t ed_mdot_partials(long dim, int nv) { return (size_t)ed_grid(dim) * (nv > 0 ? nv : 1); }

int launch_ed_mdot(const double* V, long dim, int nv, const double* w, double* part, double* h,
                   hipStream_t st) {
  if (nv <= 0) return 0;
  const int g = ed_grid(dim);
  KScope ks("ed_reorth", st);
  hipLaunchKernelGGL(k_ed_mdot, dim3(g), dim3(256), 0, st, V, dim, nv, w, part);
  hipLaunchKernelGGL(k_ed_finish_rows, dim3(nv), dim3(256), 0, st, part, g, h);
  count_launch();
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_ed_mupdate(const double* V, long dim, int nv, const double* h, double* w, int keep,
                      double* part, double* norm2, hipStream_t st) {
  const int g = ed_grid(dim);
  KScope ks("ed_reorth", st);
  hipLaunchKernelGGL(k_ed_mupdate, dim3(g), dim3(256), 0, st, V, dim, nv, h, w, keep,
                     norm2 ? part : nullptr);
  count_launch();
  if (norm2) {
    hipLaunchKernelGGL(k_ed_finish, dim3(1), dim3(256), 0, st, part, g, norm2);
    count_launch();
  }
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

}  // namespace nqs
