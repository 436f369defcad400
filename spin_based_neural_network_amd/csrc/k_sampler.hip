// k_sampler.hip — Metropolis chains on the device.
//
// Reference semantics (src/nqs/nqs_sampler.c:31-149): xorshift64 stream per
// chain, site = x mod N, accept when d = 2(log|psi'| - log|psi|) >= 0, else
// draw u and accept when log u < d.  Exchange moves (an up and a down site,
// chosen as the k-th up / l-th down site with k = x mod n_up, l = x mod
// n_down) are this library's extension; both are restated in the oracle.
//
// Layout: L lanes per chain (L | 64), each lane owns hidden units
// h = lane + L*j (j < HPL) and keeps the logistic pair of theta_h
// (p = 1/(1 + e^{-2 theta}), q = 1 - p: the member of smaller modulus and a
// bit saying which, nqs_common.h) in registers; the spin configuration is a
// bitmask in registers and every lane of the chain advances an identical
// copy of the chain's xorshift state, so no broadcast is needed.  A
// proposal's |psi'/psi| is the product over h of
//   cosh(theta_h + d_h)/cosh(theta_h) = p_h e^{d_h} + q_h e^{-d_h}
// with e^{+-d} read from the site-major e^{2W} / e^{-2W} tables (two
// positive terms for a real RBM: no cancellation at any |theta|): one log
// per lane and a log2(L)-step butterfly per proposal; on acceptance the
// pair becomes (p e^{d} / w, q e^{-d} / w), again without cancellation.
#include <type_traits>

#include "nqs_kernels.h"

namespace nqs {

__device__ inline int bit_of(const uint64_t* sp, int i) { return (int)((sp[i >> 6] >> (i & 63)) & 1ULL); }

// a wave-uniform 64-bit value into scalar registers
__device__ inline uint64_t uniform_u64(uint64_t v) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// index of the k-th (0-based) site whose bit equals `want` among N sites
__device__ inline int select_site(const uint64_t* sp, int nw, int N, int want, int k) {
  for (int w = 0; w < nw; w++) {
    int valid = N - 64 * w;
    uint64_t mask = valid >= 64 ? ~0ULL : ((1ULL << valid) - 1ULL);
    uint64_t z = (want ? sp[w] : ~sp[w]) & mask;
    int cnt = __popcll(z);
    if (k < cnt) {
      // position of the (k+1)-th set bit: binary search over halves
      int pos = 0;
#pragma unroll
      for (int width = 32; width >= 1; width >>= 1) {
        const uint64_t low = z & ((1ULL << width) - 1ULL);
        const int c = __popcll(low);
        if (k >= c) {
          k -= c;
          z >>= width;
          pos += width;
        }
      }
      return 64 * w + pos;
    }
    k -= cnt;
  }
  return -1;
}

// chain c of one sampler, lane `lane` of its L lanes
template <typename R, bool CPLX, int L, int HPL>
__device__ inline void mh_chain(const ModelView<R>& m, const ChainState& cs, const MHLaunch& ml,
                                const int c, const int lane) {
  const int N = m.N, M = m.M, nw = cs.nw;
  uint64_t sp[4];
#pragma unroll
  for (int w = 0; w < 4; w++) sp[w] = (w < nw) ? cs.spins[(size_t)c * nw + w] : 0ULL;
  uint64_t rng = cs.rng[c];
  if constexpr (L == 64) {
    // one chain per wave: the configuration and the stream are wave-uniform
    // (scalar registers); site choice and bookkeeping run on the scalar unit
#pragma unroll
    for (int w = 0; w < 4; w++) sp[w] = uniform_u64(sp[w]);
    rng = uniform_u64(rng);
  }

  // theta_h = b_h + sum_i W_hi s_i in site order; the site loop is outer so
  // that the loads of all of a lane's units for one site are in flight
  // together (rows clamped into the table, products s * W exact)
  R thr[HPL], thi[HPL], lr[HPL], li[HPL];
  unsigned isp = 0;   // bit j: lr/li of unit j hold p (else q)
#pragma unroll
  for (int j = 0; j < HPL; j++) {
    const int h = lane + L * j;
    thr[j] = h < M ? m.b_re[h] : (R)0;
    thi[j] = (CPLX && h < M) ? m.b_im[h] : (R)0;
  }
  if (M > 0) {
#pragma unroll 2
    for (int i = 0; i < N; i++) {
      const R s = bit_of(sp, i) ? (R)-1 : (R)1;
      const size_t row = (size_t)i * M;
#pragma unroll
      for (int j = 0; j < HPL; j++) {
        const int h = lane + L * j;
        const int hc = h < M ? h : 0;
        const R wr = m.WT_re[row + hc];
        const R wi = CPLX ? m.WT_im[row + hc] : (R)0;
        if (h < M) {
          thr[j] += wr * s;
          if (CPLX) thi[j] += wi * s;
        }
      }
    }
  }
#pragma unroll
  for (int j = 0; j < HPL; j++) {
    int f;
    logistic_small<R>(thr[j], CPLX ? thi[j] : (R)0, lr[j], li[j], f);
    if (!CPLX) li[j] = 0;
    isp |= (unsigned)f << j;
  }

  // log|psi(s)| with the current parameters, and the chain's cached value
  double la_part = 0.0;
#pragma unroll
  for (int j = 0; j < HPL; j++) {
    const int h = lane + L * j;
    if (h < M) {
      const double x = (double)thr[j];
      if (CPLX) {
        const double ax = fabs(x), e = exp(-2.0 * ax), y = (double)thi[j];
        const double re = (1.0 + e) * cos(y), im = (1.0 - e) * sin(y);
        const double mag = sqrt(re * re + im * im);
        la_part += ax + (mag > 0.0 ? log(mag) : -1e300);
      } else {
        const double ax = fabs(x);
        la_part += ax + log1p(exp(-2.0 * ax));
      }
    }
  }
  for (int i = lane; i < N; i += L) la_part += (double)m.a_re[i] * (bit_of(sp, i) ? -1.0 : 1.0);
  const double la_now = group_allreduce_sum<L>(la_part);
  // off = log|psi_now(s)| - cached; d below is 2 (off + log|psi(s')/psi(s)|)
  double off = la_now - cs.la[c];
  double la_cur = la_now;

  // exchange moves keep the numbers of up and down sites: both counts and
  // the reciprocals of the draw moduli are fixed for the launch
  int nd = 0;
  for (int w = 0; w < nw; w++) {
    int valid = N - 64 * w;
    uint64_t mask = valid >= 64 ? ~0ULL : ((1ULL << valid) - 1ULL);
    nd += __popcll(sp[w] & mask);
  }
  const int nu = N - nd;
  const double rN = 1.0 / N, rnu = nu > 0 ? 1.0 / nu : 0.0, rnd = nd > 0 ? 1.0 / nd : 0.0;

  unsigned long long nacc = 0, ntot = 0;
  for (int r = 0; r < ml.rounds; r++) {
    for (int stp = 0; stp < ml.steps; stp++) {
      int i, j = -1;
      // draws reduced on the vector unit, then back to scalar registers
      auto uni = [](int v) { return L == 64 ? __builtin_amdgcn_readfirstlane(v) : v; };
      if (ml.move == 0) {
        i = uni(mod_recip(xs_next(rng), (uint32_t)N, rN));
      } else {
        if (nu == 0 || nd == 0) { ntot++; continue; }
        const int k = uni(mod_recip(xs_next(rng), (uint32_t)nu, rnu));
        const int l = uni(mod_recip(xs_next(rng), (uint32_t)nd, rnd));
        i = select_site(sp, nw, N, 0, k);   // an up site   (s_i = +1)
        j = select_site(sp, nw, N, 1, l);   // a down site  (s_j = -1)
      }
      const R si = bit_of(sp, i) ? (R)-1 : (R)1;
      // log of the uniform draw that a downhill proposal uses, computed before
      // the proposal's ratio so that its latency overlaps the table loads;
      // the stream advances only if the draw is used
      uint64_t rng_u = rng;
      const double lu = log(xs_uniform(rng_u));
      // hidden factors w = p e^{d} + q e^{-d} of this lane's units, and each
      // unit's next logistic pair (p e^{d} / w, q e^{-d} / w) should the
      // proposal be accepted (computed with the factors, off the decision's
      // path); the running product is rescaled by a power of two every RS
      // units
      constexpr int RS = RatioRange<R>::rescale;
      R pr = 1, pi = 0;
      int pex = 0;
      R nlr[HPL], nli[HPL];
      unsigned nisp = 0;
      // table rows of the flipped sites for the lane's units, in groups of
      // QB units whose loads are issued together (rows clamped into the
      // table; the pair form reads the second site's rows as well)
      auto factors = [&](auto pair_tag) {
        constexpr bool PAIR = decltype(pair_tag)::value;
        constexpr int QB = HPL >= 16 ? 8 : 4;
        const bool up = si > (R)0;
#pragma unroll
        for (int q0 = 0; q0 < HPL; q0 += QB) {
          R pir[QB], mir[QB], pii[QB], mii[QB], pjr[QB], mjr[QB], pji[QB], mji[QB];
#pragma unroll
          for (int u = 0; u < QB; u++) {
            const int q = q0 + u;
            if (q < HPL) {
              const int h = lane + L * q;
              const size_t ei = (size_t)i * M + (h < M ? h : 0);
              pir[u] = m.EPT_re[ei];
              mir[u] = m.EMT_re[ei];
              if (CPLX) { pii[u] = m.EPT_im[ei]; mii[u] = m.EMT_im[ei]; }
              if (PAIR) {
                const size_t ej = (size_t)j * M + (h < M ? h : 0);
                pjr[u] = m.EPT_re[ej];
                mjr[u] = m.EMT_re[ej];
                if (CPLX) { pji[u] = m.EPT_im[ej]; mji[u] = m.EMT_im[ej]; }
              }
            }
          }
#pragma unroll
          for (int u = 0; u < QB; u++) {
            const int q = q0 + u;
            if (q >= HPL) continue;
            const int h = lane + L * q;
            if (h >= M) continue;
            // e^{d} (xr, xi), e^{-d} (yr, yi): d = -2 s_i W_i (+ 2 s_i W_j)
            R xr = up ? mir[u] : pir[u], xi = CPLX ? (up ? mii[u] : pii[u]) : (R)0;
            R yr = up ? pir[u] : mir[u], yi = CPLX ? (up ? pii[u] : mii[u]) : (R)0;
            if (PAIR) {
              const R fr = up ? pjr[u] : mjr[u], gr = up ? mjr[u] : pjr[u];
              if (CPLX) {
                const R fi = up ? pji[u] : mji[u], gi = up ? mji[u] : pji[u];
                const R nxr = xr * fr - xi * fi;
                xi = xr * fi + xi * fr;
                xr = nxr;
                const R nyr = yr * gr - yi * gi;
                yi = yr * gi + yi * gr;
                yr = nyr;
              } else {
                xr *= fr;
                yr *= gr;
              }
            }
            R ppr, ppi, qqr, qqi;
            logistic_unpack<R>(lr[q], CPLX ? li[q] : (R)0, (isp >> q) & 1u, ppr, ppi, qqr, qqi);
            if (CPLX) {
              const R ar = ppr * xr - ppi * xi, ai = ppr * xi + ppi * xr;   // p e^{d}
              const R br = qqr * yr - qqi * yi, bi = qqr * yi + qqi * yr;   // q e^{-d}
              const R wr = ar + br, wi = ai + bi;
              const R nr = pr * wr - pi * wi;
              pi = pr * wi + pi * wr;
              pr = nr;
              const R inv = (R)1 / (wr * wr + wi * wi);
              const R npr = (ar * wr + ai * wi) * inv, npi = (ai * wr - ar * wi) * inv;
              const R nqr = (br * wr + bi * wi) * inv, nqi = (bi * wr - br * wi) * inv;
              const bool small_p = npr * npr + npi * npi < nqr * nqr + nqi * nqi;
              nlr[q] = small_p ? npr : nqr;
              nli[q] = small_p ? npi : nqi;
              nisp |= small_p ? (1u << q) : 0u;
            } else {
              const R a = ppr * xr, b = qqr * yr;    // both positive
              const R w = a + b;
              pr *= w;
              const R inv = (R)1 / w;
              const R np = a * inv, nq = b * inv;
              const bool small_p = np < nq;
              nlr[q] = small_p ? np : nq;
              nisp |= small_p ? (1u << q) : 0u;
            }
            if ((q % RS) == RS - 1) {
              int e2;
              if (CPLX) {
                (void)frexp(fmax(fabs(pr), fabs(pi)), &e2);
                pr = ldexp(pr, -e2);
                pi = ldexp(pi, -e2);
              } else {
                pr = frexp(pr, &e2);
              }
              pex += e2;
            }
          }
        }
      };
      if (M > 0) {
        if (j >= 0) factors(std::true_type{});
        else factors(std::false_type{});
      }
      // log of this lane's product (an exact power of two folded in; a zero
      // factor gives -inf, never accepted)
      R lm = CPLX ? (R)0.5 * log(pr * pr + pi * pi) : log(pr);
      lm += (R)pex * (R)0.69314718055994530942;
      if (M == 0) lm = 0;
      const R lsum = group_allreduce_sum<L>(lm);
      // visible part: Re a_i (s_i' - s_i) = -2 a_i s_i per flipped site
      R dvis = (R)-2 * m.a_re[i] * si;
      if (j >= 0) dvis += (R)2 * m.a_re[j] * si;   // s_j = -s_i
      const double dl = (double)lsum + (double)dvis;
      const double d = 2.0 * (off + dl);
      // every lane of the chain holds the same sums, hence the same decision
      // (wave-uniform with one chain per wave: the stream stays scalar)
      bool down = d < 0.0;
      if constexpr (L == 64) down = __builtin_amdgcn_readfirstlane((int)down) != 0;
      bool acc;
      if (!down) acc = true;
      else {
        acc = lu < d;
        rng = rng_u;
      }
      if constexpr (L == 64) acc = __builtin_amdgcn_readfirstlane((int)acc) != 0;
      ntot++;
      if (acc) {
        nacc++;
        la_cur += dl;
        off = 0.0;
        sp[i >> 6] ^= 1ULL << (i & 63);
        if (j >= 0) sp[j >> 6] ^= 1ULL << (j & 63);
        // the accepted proposal's logistic pairs; their denominator w is not
        // zero (a zero factor gives d = -inf, never accepted)
#pragma unroll
        for (int q = 0; q < HPL; q++) {
          const int h = lane + L * q;
          if (h < M) {
            lr[q] = nlr[q];
            if (CPLX) li[q] = nli[q];
          }
        }
        isp = nisp;
      }
    }
    if (ml.record) {
      const long b = (long)r * cs.C + c;
      if (b < ml.n_out) {
        int* o = ml.out + (size_t)b * N;
        for (int q = lane; q < N; q += L) o[q] = bit_of(sp, q) ? -1 : 1;
      }
    }
  }
  if (lane == 0) {
    for (int w = 0; w < nw; w++) cs.spins[(size_t)c * nw + w] = sp[w];
    cs.rng[c] = rng;
    cs.la[c] = la_cur - off;
    cs.acc[c] += nacc;
    cs.tot[c] += ntot;
  }
}

template <typename R, bool CPLX, int L, int HPL>
__global__ __launch_bounds__(256) void k_mh(ModelView<R> m, ChainState cs, MHLaunch ml) {
  const int lane = threadIdx.x & (L - 1);
  const int c = (int)((blockIdx.x * (long)blockDim.x + threadIdx.x) / L);
  if (c >= cs.C) return;
  mh_chain<R, CPLX, L, HPL>(m, cs, ml, c, lane);
}

// several samplers of one variant in one launch: job j owns the blocks
// [blk0_j, blk0_{j+1})
// a block-uniform struct read word by word into scalar registers (so the
// job's pointers do not occupy vector registers in the chain loop)
template <typename T>
__device__ inline T uniform_load(const T* p) {
  static_assert(sizeof(T) % 4 == 0, "word-sized struct");
  T out;
  int* o = reinterpret_cast<int*>(&out);
  const int* q = reinterpret_cast<const int*>(p);
#pragma unroll
  for (int k = 0; k < (int)(sizeof(T) / 4); k++) o[k] = __builtin_amdgcn_readfirstlane(q[k]);
  return out;
}

template <bool CPLX, int L, int HPL>
__global__ __launch_bounds__(256) void k_mh_batch(const MHJob* __restrict__ jobs, int njobs) {
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (__builtin_amdgcn_readfirstlane(jobs[mid].blk0) <= (int)blockIdx.x) lo = mid;
    else hi = mid - 1;
  }
  const MHJob jb = uniform_load(jobs + lo);
  const int lane = threadIdx.x & (L - 1);
  const int c = (int)(((blockIdx.x - jb.blk0) * 256L + threadIdx.x) / L);
  if (c >= jb.cs.C) return;
  mh_chain<double, CPLX, L, HPL>(jb.m, jb.cs, jb.ml, c, lane);
}

template <typename R, bool CPLX, int L, int HPL>
static void mh_go(const ModelView<R>& m, ChainState& cs, const MHLaunch& ml, hipStream_t st) {
  const long threads = (long)cs.C * L;
  const int blocks = (int)((threads + 255) / 256);
  hipLaunchKernelGGL((k_mh<R, CPLX, L, HPL>), dim3(blocks), dim3(256), 0, st, m, cs, ml);
}

template <typename R>
int launch_mh(const ModelView<R>& m, ChainState& cs, const MHLaunch& ml, hipStream_t st) {
  if (cs.C <= 0 || ml.rounds <= 0) return 0;
  if (m.N > kMaxSites || m.M > kMaxHidden) {
    fprintf(stderr, "nqs_hip: sampler supports N <= %d, M <= %d\n", kMaxSites, kMaxHidden);
    return -1;
  }
  const bool c = m.cplx != 0;
  KScope ks("mh", st);
  if (m.M <= 64) {
    if (c) mh_go<R, true, 16, 4>(m, cs, ml, st); else mh_go<R, false, 16, 4>(m, cs, ml, st);
  } else if (m.M <= 128) {
    // one chain per wave with as few idle unit slots per lane as the
    // register budget allows (the chain's latency grows with the units a
    // lane owns)
    if (c) mh_go<R, true, 64, 2>(m, cs, ml, st); else mh_go<R, false, 64, 2>(m, cs, ml, st);
  } else if (m.M <= 192) {
    if (c) mh_go<R, true, 64, 3>(m, cs, ml, st); else mh_go<R, false, 64, 3>(m, cs, ml, st);
  } else if (m.M <= 256) {
    if (c) mh_go<R, true, 64, 4>(m, cs, ml, st); else mh_go<R, false, 64, 4>(m, cs, ml, st);
  } else if (m.M <= 832) {
    // 13 units per lane: fewer idle slots than 16 for M up to 832 (M = 800:
    // 4% instead of 22%)
    if (c) mh_go<R, true, 64, 13>(m, cs, ml, st); else mh_go<R, false, 64, 13>(m, cs, ml, st);
  } else {
    if (c) mh_go<R, true, 64, 16>(m, cs, ml, st); else mh_go<R, false, 64, 16>(m, cs, ml, st);
  }
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

int mh_variant(const ModelView<double>& m) {
  const int c = m.cplx != 0;
  return (m.M <= 64 ? 0 : (m.M <= 256 ? 2 : 4)) + c;
}

int mh_blocks(int variant, int C) {
  const int L = variant < 2 ? 16 : 64;
  return (int)(((long)C * L + 255) / 256);
}

int launch_mh_batch(int variant, const MHJob* jobs, int njobs, int nblocks, hipStream_t st) {
  if (njobs <= 0 || nblocks <= 0) return 0;
  KScope ks("mh", st);
  const dim3 g((unsigned)nblocks), b(256);
  switch (variant) {
    case 0: hipLaunchKernelGGL((k_mh_batch<false, 16, 4>), g, b, 0, st, jobs, njobs); break;
    case 1: hipLaunchKernelGGL((k_mh_batch<true, 16, 4>), g, b, 0, st, jobs, njobs); break;
    case 2: hipLaunchKernelGGL((k_mh_batch<false, 64, 4>), g, b, 0, st, jobs, njobs); break;
    case 3: hipLaunchKernelGGL((k_mh_batch<true, 64, 4>), g, b, 0, st, jobs, njobs); break;
    case 4: hipLaunchKernelGGL((k_mh_batch<false, 64, 16>), g, b, 0, st, jobs, njobs); break;
    default: hipLaunchKernelGGL((k_mh_batch<true, 64, 16>), g, b, 0, st, jobs, njobs); break;
  }
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

// One thread per chain: initial configuration from the chain's seed.
// Random start: s_i = +1 when the draw is odd (nqs_sampler.c:69-72);
// Sz = 0 start: N/2 up then down, Fisher-Yates from the top index.
__global__ void k_chain_init(ChainState cs, const uint64_t* __restrict__ seeds, int sz0) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cs.C) return;
  uint64_t rng = seeds[c];
  uint64_t sp[4] = {0, 0, 0, 0};
  const int N = cs.N;
  if (!sz0) {
    for (int i = 0; i < N; i++) {
      const bool up = (xs_next(rng) & 1ULL) != 0;
      if (!up) sp[i >> 6] |= 1ULL << (i & 63);
    }
  } else {
    // bit pattern of the ordered start: sites >= N/2 are down
    for (int i = N / 2; i < N; i++) sp[i >> 6] |= 1ULL << (i & 63);
    for (int i = N - 1; i > 0; i--) {
      const int j = mod_small(xs_next(rng), (uint32_t)(i + 1));
      const int bi = (int)((sp[i >> 6] >> (i & 63)) & 1ULL);
      const int bj = (int)((sp[j >> 6] >> (j & 63)) & 1ULL);
      if (bi != bj) {
        sp[i >> 6] ^= 1ULL << (i & 63);
        sp[j >> 6] ^= 1ULL << (j & 63);
      }
    }
  }
  for (int w = 0; w < cs.nw; w++) cs.spins[(size_t)c * cs.nw + w] = sp[w];
  cs.rng[c] = rng;
  cs.acc[c] = 0;
  cs.tot[c] = 0;
}

int launch_chain_init(ChainState& cs, const uint64_t* seeds_dev, int sz0, hipStream_t st) {
  hipLaunchKernelGGL(k_chain_init, dim3((cs.C + 127) / 128), dim3(128), 0, st, cs, seeds_dev, sz0);
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

__global__ void k_unpack(ChainState cs, int chain, int* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cs.N) return;
  const uint64_t w = cs.spins[(size_t)chain * cs.nw + (i >> 6)];
  out[i] = ((w >> (i & 63)) & 1ULL) ? -1 : 1;
}

__global__ void k_unpack_all(ChainState cs, int* out) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= (long)cs.C * cs.N) return;
  const int c = (int)(e / cs.N), i = (int)(e % cs.N);
  const uint64_t w = cs.spins[(size_t)c * cs.nw + (i >> 6)];
  out[e] = ((w >> (i & 63)) & 1ULL) ? -1 : 1;
}

int launch_unpack_all(const ChainState& cs, int* out, hipStream_t st) {
  const long total = (long)cs.C * cs.N;
  hipLaunchKernelGGL(k_unpack_all, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, cs, out);
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_unpack(const ChainState& cs, int chain, int* out, hipStream_t st) {
  hipLaunchKernelGGL(k_unpack, dim3((cs.N + 255) / 256), dim3(256), 0, st, cs, chain, out);
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

template int launch_mh<double>(const ModelView<double>&, ChainState&, const MHLaunch&, hipStream_t);
template int launch_mh<float>(const ModelView<float>&, ChainState&, const MHLaunch&, hipStream_t);

}  // namespace nqs
