// k_eloc.hip — local energies E_loc(s) = sum_s' H_ss' psi(s')/psi(s).
//
// Reference semantics: src/nqs/nqs_gradient.c (TFIM :50-82, XXZ /
// Heisenberg :87-139, J1-J2 :146-205, kagome :374-445, real path = real
// part of the complex path :495-875).  The reference re-evaluates
// log psi(s') from scratch for every connected configuration (:24-45);
// here each connected configuration is priced from the sample's logistic
// pair (p, q) of every hidden input theta (logistic_small, nqs_common.h) and
// the tables EP = e^{2W}, EM = e^{-2W}:
//   psi(s')/psi(s) = exp(sum_flipped -2 a_i s_i) * prod_h [p_h e^{d_h} + q_h e^{-d_h}]
// with d_h the change of theta_h.  For a real RBM both terms of every factor
// are positive, so no factor loses digits however large |theta| is (the
// tanh form cosh d + tanh(theta) sinh d cancels once tanh rounds to +-1), and
// no transcendental is evaluated per hidden unit.  Running products are
// rescaled by powers of two every RatioRange<R>::rescale units.
//
// Work decomposition: a workgroup takes SB consecutive samples, compacts the
// active (sample, term) pairs (all flips; pairs with s_a != s_b) into a
// sample-major list in LDS, and gives one lane per active pair.  The hidden
// dimension is walked in tiles of HT units whose table rows (HT x N) and
// logistic members (SB x HT) are staged in LDS.  Per-sample sums run in a
// fixed order.
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

// p and q of one hidden unit from the staged sign of Re t and small member
template <typename R, bool CPLX>
__device__ inline void pq_of(R t_sign, R sr, R si, R& pr, R& pi, R& qr, R& qi) {
  logistic_unpack<R>(sr, CPLX ? si : (R)0, t_sign < (R)0 ? 1 : 0, pr, pi, qr, qi);
}

// running product rescaled by a power of two (exact), exponent accumulated
template <typename R, bool CPLX>
__device__ inline void rescale2(R& pr, R& pi, int& ex) {
  int e2;
  if (CPLX) {
    (void)frexp(fmax(fabs(pr), fabs(pi)), &e2);
    pr = ldexp(pr, -e2);
    pi = ldexp(pi, -e2);
  } else {
    pr = frexp(pr, &e2);
  }
  ex += e2;
}

template <typename R, bool CPLX, int IPT>
__global__ __launch_bounds__(256) void k_eloc(ModelView<R> m, HamDev H,
                                              const int* __restrict__ spins, int B,
                                              const R* __restrict__ t_re,
                                              const R* __restrict__ q_re,
                                              const R* __restrict__ q_im,
                                              double* __restrict__ e_re,
                                              double* __restrict__ e_im, int SB, int HT) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int scan_sh[256];
  __shared__ int n_items_sh;
  constexpr int RS = RatioRange<R>::rescale;
  const int N = m.N, M = m.M;
  const int T = H.n_flip + H.n_pair;
  const int tid = threadIdx.x;
  // carve dynamic LDS (all offsets 16-byte aligned)
  unsigned char* p = smem;
  auto carve = [&](size_t bytes) { unsigned char* q = p; p += (bytes + 15) & ~(size_t)15; return q; };
  signed char* sp = (signed char*)carve((size_t)SB * N);
  int* items = (int*)carve(sizeof(int) * (size_t)SB * T);
  int* starts = (int*)carve(sizeof(int) * (SB + 1));
  R* epr = (R*)carve(sizeof(R) * (size_t)HT * N);
  R* emr = (R*)carve(sizeof(R) * (size_t)HT * N);
  R* epi = CPLX ? (R*)carve(sizeof(R) * (size_t)HT * N) : nullptr;
  R* emi = CPLX ? (R*)carve(sizeof(R) * (size_t)HT * N) : nullptr;
  R* tsg = (R*)carve(sizeof(R) * (size_t)SB * HT);
  R* lqr = (R*)carve(sizeof(R) * (size_t)SB * HT);
  R* lqi = CPLX ? (R*)carve(sizeof(R) * (size_t)SB * HT) : nullptr;
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
    // every wave's counts must be in starts before thread 0 turns them into
    // offsets
    __syncthreads();
    if (tid == 0) {
      n_items_sh = base;
      for (int s = 0; s < nsb; s++) starts[s + 1] += starts[s];
    }
    __syncthreads();
    const int n_items = n_items_sh;

    // ---- per-item state ----
    R Pr[IPT], Pi[IPT], sgn[IPT];
    int Ex[IPT], isb[IPT], ia[IPT], ib[IPT];
#pragma unroll
    for (int q = 0; q < IPT; q++) {
      const int idx = tid + q * (int)blockDim.x;
      Pr[q] = 1; Pi[q] = 0; Ex[q] = 0; sgn[q] = 1; isb[q] = -1; ia[q] = 0; ib[q] = -1;
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
      // table rows in batches of 8 per thread: every load of a batch is in
      // flight at once (indices clamped into the tile, stores guarded)
      const int tn = hn * N;
      for (int e0 = 0; e0 < tn; e0 += 8 * 256) {
        R vpr[8], vmr[8], vpi[8], vmi[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const int e = min(e0 + tid + 256 * u, tn - 1);
          vpr[u] = m.EP_re[toff + e];
          vmr[u] = m.EM_re[toff + e];
          if (CPLX) { vpi[u] = m.EP_im[toff + e]; vmi[u] = m.EM_im[toff + e]; }
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const int e = e0 + tid + 256 * u;
          if (e < tn) {
            epr[e] = vpr[u];
            emr[e] = vmr[u];
            if (CPLX) { epi[e] = vpi[u]; emi[e] = vmi[u]; }
          }
        }
      }
      for (int e = tid; e < nsb * hn; e += blockDim.x) {
        const int s = e / hn, hh = e - s * hn;
        const size_t g = (size_t)(b0 + s) * M + h0 + hh;
        tsg[s * HT + hh] = t_re[g];
        lqr[s * HT + hh] = q_re[g];
        if (CPLX) lqi[s * HT + hh] = q_im[g];
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < IPT; q++) {
        if (isb[q] < 0) continue;
        const int a = ia[q], bb = ib[q];
        const bool up = sgn[q] > (R)0;
        const R* ts = tsg + isb[q] * HT;
        const R* lr = lqr + isb[q] * HT;
        const R* li = CPLX ? lqi + isb[q] * HT : nullptr;
        R pr = Pr[q], pim = Pi[q];
        int ex = Ex[q];
        for (int hh = 0; hh < hn; hh++) {
          const int ea = hh * N + a;
          // e^{d} (xr, xi) and e^{-d} (yr, yi) of this unit for the term
          R xr, xi = 0, yr, yi = 0;
          {
            const R par_ = epr[ea], mar = emr[ea];
            const R pai = CPLX ? epi[ea] : (R)0, mai = CPLX ? emi[ea] : (R)0;
            // site a alone: s_a = +1 gives e^{d} = EM_a, e^{-d} = EP_a
            R dr = up ? mar : par_, di = up ? mai : pai;
            R er = up ? par_ : mar, ei = up ? pai : mai;
            if (bb >= 0) {
              // second site with s_b = -s_a: e^{d} *= (s_a = +1 ? EP_b : EM_b)
              const int eb = hh * N + bb;
              const R pbr = epr[eb], mbr = emr[eb];
              const R pbi = CPLX ? epi[eb] : (R)0, mbi = CPLX ? emi[eb] : (R)0;
              const R fr = up ? pbr : mbr, fi = up ? pbi : mbi;
              const R gr = up ? mbr : pbr, gi = up ? mbi : pbi;
              if (CPLX) {
                const R nr = dr * fr - di * fi;
                di = dr * fi + di * fr;
                dr = nr;
                const R mr2 = er * gr - ei * gi;
                ei = er * gi + ei * gr;
                er = mr2;
              } else {
                dr *= fr;
                er *= gr;
              }
            }
            xr = dr; xi = di; yr = er; yi = ei;
          }
          R ppr, ppi, qqr, qqi;
          pq_of<R, CPLX>(ts[hh], lr[hh], CPLX ? li[hh] : (R)0, ppr, ppi, qqr, qqi);
          if (CPLX) {
            const R wr = (ppr * xr - ppi * xi) + (qqr * yr - qqi * yi);
            const R wi = (ppr * xi + ppi * xr) + (qqr * yi + qqi * yr);
            const R nr = pr * wr - pim * wi;
            pim = pr * wi + pim * wr;
            pr = nr;
          } else {
            pr *= ppr * xr + qqr * yr;
          }
          if ((hh % RS) == RS - 1) rescale2<R, CPLX>(pr, pim, ex);
        }
        rescale2<R, CPLX>(pr, pim, ex);
        Pr[q] = pr;
        Pi[q] = pim;
        Ex[q] = ex;
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
      double ph_r = (double)Pr[q], ph_i = CPLX ? (double)Pi[q] : 0.0;
      const double mg = CPLX ? sqrt(ph_r * ph_r + ph_i * ph_i) : fabs(ph_r);
      // a zero factor (psi(s') = 0) contributes nothing
      const double lm = mg > 0.0 ? log(mg) + (double)Ex[q] * 0.69314718055994530942 : -INFINITY;
      const double mag = exp(lm + vr);
      ph_r = mg > 0.0 ? ph_r / mg : 0.0;
      ph_i = mg > 0.0 ? ph_i / mg : 0.0;
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

// ---------------------------------------------------------------------------
// Term-major local energies for many samples.  The off-diagonal terms
// (flips, then pairs, in the Hamiltonian's order) get hidden-unit rows of
// their own once per call, e^{d} and e^{-d} of the term for s_a = +1:
//   flip of a:    Fp = EM_a,             Fm = EP_a
//   pair (a, b):  Fp = EM_a EP_b,        Fm = EP_a EM_b
// (complex products for complex RBMs); a sample with s_a = -1 swaps the two,
// so the factor of hidden unit h in psi(s')/psi(s) is
//   s_a = +1: p_h Fp + q_h Fm,   s_a = -1: q_h Fp + p_h Fm.
// A wave holds 64 samples in its lanes and 8 terms in registers; per tile of
// hidden units, the samples' logistic members and the workgroup's 32 term
// rows are staged in LDS by coalesced loads (term rows are read at
// wave-uniform addresses), and the running products are rescaled by a power
// of two every RatioRange<R>::rescale units.  Each (term, sample)
// contribution goes to res[t][b] (zero for inactive pairs); a second kernel
// adds the diagonal part and the contributions in term order, so E_loc sums
// in the same order as the per-sample item lists.
// ---------------------------------------------------------------------------
template <typename R, bool CPLX>
__global__ void k_term_tables(ModelView<R> m, HamDev H, int Tp, R* __restrict__ Fpr,
                              R* __restrict__ Fpi, R* __restrict__ Fmr, R* __restrict__ Fmi) {
  const int T = H.n_flip + H.n_pair, M = m.M;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < (long)Tp * M;
       e += (long)gridDim.x * blockDim.x) {
    const int t = (int)(e / M), h = (int)(e - (long)t * M);
    if (t >= T) {
      // padding rows: factor p + q = 1
      Fpr[e] = (R)1;
      Fmr[e] = (R)1;
      if (CPLX) { Fpi[e] = (R)0; Fmi[e] = (R)0; }
      continue;
    }
    const int a = t < H.n_flip ? H.flip_i[t] : H.pair_a[t - H.n_flip];
    const int b = t < H.n_flip ? -1 : H.pair_b[t - H.n_flip];
    const size_t ea = (size_t)a * M + h;
    const R mar = m.EMT_re[ea], par_ = m.EPT_re[ea];
    const R mai = CPLX ? m.EMT_im[ea] : (R)0, pai = CPLX ? m.EPT_im[ea] : (R)0;
    if (b < 0) {
      Fpr[e] = mar;
      Fmr[e] = par_;
      if (CPLX) { Fpi[e] = mai; Fmi[e] = pai; }
      continue;
    }
    const size_t eb = (size_t)b * M + h;
    const R pbr = m.EPT_re[eb], mbr = m.EMT_re[eb];
    if (CPLX) {
      const R pbi = m.EPT_im[eb], mbi = m.EMT_im[eb];
      Fpr[e] = mar * pbr - mai * pbi;
      Fpi[e] = mar * pbi + mai * pbr;
      Fmr[e] = par_ * mbr - pai * mbi;
      Fmi[e] = par_ * mbi + pai * mbr;
    } else {
      Fpr[e] = mar * pbr;
      Fmr[e] = par_ * mbr;
    }
  }
}

constexpr int kEtTerms = 8;    // terms per wave (4 waves per workgroup)

// hidden units per tile
template <typename R, bool CPLX>
struct EtTile { static constexpr int H = (sizeof(R) == 8 && !CPLX) ? 32 : 16; };

template <typename R, bool CPLX>
__global__ __launch_bounds__(256) void k_eloc_terms(ModelView<R> m, HamDev H,
                                                    const int* __restrict__ spins, int B,
                                                    const R* __restrict__ t_re,
                                                    const R* __restrict__ q_re,
                                                    const R* __restrict__ q_im,
                                                    const R* __restrict__ Fpr,
                                                    const R* __restrict__ Fpi,
                                                    const R* __restrict__ Fmr,
                                                    const R* __restrict__ Fmi,
                                                    double* __restrict__ res_r,
                                                    double* __restrict__ res_i) {
  constexpr int HT = EtTile<R, CPLX>::H;
  constexpr int RS = RatioRange<R>::rescale;
  static_assert(HT % RS == 0, "tiles hold whole rescaling groups");
  constexpr int NTB = 4 * kEtTerms;                 // terms per workgroup
  __shared__ R tl_s[64][HT + 1];                     // Re t (its sign: which member)
  __shared__ R tl_r[64][HT + 1];                     // small logistic member
  __shared__ R tl_i[CPLX ? 64 : 1][HT + 1];
  __shared__ R tp_r[NTB][HT], tm_r[NTB][HT];
  __shared__ R tp_i[CPLX ? NTB : 1][HT], tm_i[CPLX ? NTB : 1][HT];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int T = H.n_flip + H.n_pair, M = m.M, N = m.N;
  const int b0 = blockIdx.x * 64;
  const int bl = min(b0 + lane, B - 1);
  const int tb0 = blockIdx.y * NTB;                 // workgroup's first term
  const int t0 = tb0 + wave * kEtTerms;             // wave's first term
  const int nt = max(0, min(kEtTerms, T - t0));
  unsigned active = 0, upmask = 0;
#pragma unroll
  for (int j = 0; j < kEtTerms; j++) {
    if (j < nt) {
      const int t = t0 + j;
      const int a = t < H.n_flip ? H.flip_i[t] : H.pair_a[t - H.n_flip];
      const int sa = spins[(size_t)bl * N + a];
      upmask |= sa > 0 ? (1u << j) : 0u;
      const bool act = t < H.n_flip || spins[(size_t)bl * N + H.pair_b[t - H.n_flip]] != sa;
      active |= act ? (1u << j) : 0u;
    }
  }
  R pr[kEtTerms], pim[kEtTerms];
  int ex[kEtTerms];
#pragma unroll
  for (int j = 0; j < kEtTerms; j++) { pr[j] = (R)1; pim[j] = (R)0; ex[j] = 0; }
  for (int h0 = 0; h0 < M; h0 += HT) {
    const int hn = min(HT, M - h0);
    __syncthreads();
    {
      // logistic members of the 64 samples and the workgroup's term rows for
      // this tile; all loads of a thread in flight together (indices clamped
      // into the tile)
      constexpr int NQT = 64 * HT / 256, NQC = NTB * HT / 256;
      R vts[NQT], vtr[NQT], vti[NQT], vpr[NQC], vpi[NQC], vmr[NQC], vmi[NQC];
#pragma unroll
      for (int u = 0; u < NQT; u++) {
        const int e = tid + 256 * u, sm = e / HT, hh = e - sm * HT;
        const size_t g = (size_t)min(b0 + sm, B - 1) * M + h0 + min(hh, hn - 1);
        vts[u] = t_re[g];
        vtr[u] = q_re[g];
        if (CPLX) vti[u] = q_im[g];
      }
#pragma unroll
      for (int u = 0; u < NQC; u++) {
        const int e = tid + 256 * u, r = e / HT, hh = e - r * HT;
        const size_t g = (size_t)(tb0 + r) * M + h0 + min(hh, hn - 1);
        vpr[u] = Fpr[g];
        vmr[u] = Fmr[g];
        if (CPLX) { vpi[u] = Fpi[g]; vmi[u] = Fmi[g]; }
      }
#pragma unroll
      for (int u = 0; u < NQT; u++) {
        const int e = tid + 256 * u, sm = e / HT, hh = e - sm * HT;
        tl_s[sm][hh] = vts[u];
        tl_r[sm][hh] = vtr[u];
        if (CPLX) tl_i[sm][hh] = vti[u];
      }
#pragma unroll
      for (int u = 0; u < NQC; u++) {
        const int e = tid + 256 * u, r = e / HT, hh = e - r * HT;
        tp_r[r][hh] = vpr[u];
        tm_r[r][hh] = vmr[u];
        if (CPLX) { tp_i[r][hh] = vpi[u]; tm_i[r][hh] = vmi[u]; }
      }
    }
    __syncthreads();
    if (nt == 0) continue;
    const int rw = wave * kEtTerms;
    for (int g0 = 0; g0 < hn; g0 += RS) {
#pragma unroll
      for (int u = 0; u < RS; u++) {
        const int hh = g0 + u;
        if (hh >= hn) break;
        R ppr, ppi, qqr, qqi;
        pq_of<R, CPLX>(tl_s[lane][hh], tl_r[lane][hh], CPLX ? tl_i[lane][hh] : (R)0, ppr, ppi,
                       qqr, qqi);
        // term rows are padded to whole workgroups: no test on j
#pragma unroll
        for (int j = 0; j < kEtTerms; j++) {
          const bool up = (upmask >> j) & 1u;
          const R ar = up ? ppr : qqr, ai = up ? ppi : qqi;
          const R br = up ? qqr : ppr, bi = up ? qqi : ppi;
          if (CPLX) {
            const R fpr = tp_r[rw + j][hh], fpi = tp_i[rw + j][hh];
            const R fmr = tm_r[rw + j][hh], fmi = tm_i[rw + j][hh];
            const R wr = (ar * fpr - ai * fpi) + (br * fmr - bi * fmi);
            const R wi = (ar * fpi + ai * fpr) + (br * fmi + bi * fmr);
            const R nr = pr[j] * wr - pim[j] * wi;
            pim[j] = pr[j] * wi + pim[j] * wr;
            pr[j] = nr;
          } else {
            pr[j] *= ar * tp_r[rw + j][hh] + br * tm_r[rw + j][hh];
          }
        }
      }
      // rescale by a power of two (exact)
#pragma unroll
      for (int j = 0; j < kEtTerms; j++) rescale2<R, CPLX>(pr[j], pim[j], ex[j]);
    }
  }
  if (b0 + lane >= B) return;
#pragma unroll
  for (int j = 0; j < kEtTerms; j++) {
    if (j >= nt) continue;
    const int t = t0 + j;
    const size_t o = (size_t)t * B + b0 + lane;
    if (!(active & (1u << j))) {
      res_r[o] = 0.0;
      res_i[o] = 0.0;
      continue;
    }
    const int a = t < H.n_flip ? H.flip_i[t] : H.pair_a[t - H.n_flip];
    const int bb = t < H.n_flip ? -1 : H.pair_b[t - H.n_flip];
    const double s_a = ((upmask >> j) & 1u) ? 1.0 : -1.0;
    double vr = -2.0 * (double)m.a_re[a] * s_a;
    double vi = CPLX ? -2.0 * (double)m.a_im[a] * s_a : 0.0;
    if (bb >= 0) {
      vr += 2.0 * (double)m.a_re[bb] * s_a;
      if (CPLX) vi += 2.0 * (double)m.a_im[bb] * s_a;
    }
    const double xr = (double)pr[j], xi = CPLX ? (double)pim[j] : 0.0;
    const double mg = CPLX ? sqrt(xr * xr + xi * xi) : fabs(xr);
    // a zero factor (psi(s') = 0) contributes nothing
    const double lm = mg > 0.0 ? log(mg) + (double)ex[j] * 0.69314718055994530942 : -INFINITY;
    double ph_r = mg > 0.0 ? xr / mg : 0.0, ph_i = mg > 0.0 ? xi / mg : 0.0;
    if (CPLX) {
      const double c = cos(vi), sn = sin(vi);
      const double nr = ph_r * c - ph_i * sn;
      ph_i = ph_r * sn + ph_i * c;
      ph_r = nr;
    }
    const double mag = exp(lm + vr);
    double sign = 1.0;
    if (H.sublat) {
      const int par = H.sublat[a] + (bb >= 0 ? H.sublat[bb] : 0);
      if (par & 1) sign = -1.0;
    }
    const double coef = t < H.n_flip ? H.flip_c[t] : H.pair_c[t - H.n_flip];
    res_r[o] = coef * sign * mag * ph_r;
    res_i[o] = coef * sign * mag * ph_i;
  }
}

// E_loc(b) = diagonal part + sum over terms in order of res[t][b]
__global__ __launch_bounds__(256) void k_eloc_terms_sum(HamDev H, const int* __restrict__ spins,
                                                        int B, int N,
                                                        const double* __restrict__ res_r,
                                                        const double* __restrict__ res_i,
                                                        double* __restrict__ e_re,
                                                        double* __restrict__ e_im) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const int* s = spins + (size_t)b * N;
  double diag = 0.0;
  for (int d = 0; d < H.n_diag; d++)
    diag += H.diag_c[d] * (double)(s[H.diag_a[d]] * s[H.diag_b[d]]);
  const int T = H.n_flip + H.n_pair;
  double orr = 0.0, oi = 0.0;
  for (int t = 0; t < T; t++) {
    orr += res_r[(size_t)t * B + b];
    oi += res_i[(size_t)t * B + b];
  }
  e_re[b] = diag + orr;
  if (e_im) e_im[b] = oi;
}

// term rows padded to whole workgroups (4 waves x 16 terms)
static inline int padded_terms(int T) { return (T + 4 * kEtTerms - 1) / (4 * kEtTerms) * (4 * kEtTerms); }

template <typename R>
size_t eloc_ws_bytes(const ModelView<R>& m, const HamDev& h, int B) {
  const size_t T = (size_t)(h.n_flip + h.n_pair);
  const size_t tab = (size_t)padded_terms((int)T) * (size_t)m.M * sizeof(R) * (m.cplx ? 4 : 2);
  return ((tab + 255) & ~(size_t)255) + 2 * T * (size_t)B * sizeof(double) + 256;
}

// With e_re == nullptr only the per-(term, sample) contributions are formed
// (res_r / res_i in the workspace, term-major); their addresses go to
// *res_out.
template <typename R, bool CPLX>
static int eloc_terms_go(const ModelView<R>& m, const HamDev& h, const int* spins, int B,
                         const R* t_re, const R* q_re, const R* q_im, double* e_re, double* e_im,
                         void* ws, hipStream_t st, double** res_out = nullptr) {
  const int T = h.n_flip + h.n_pair, M = m.M, Tp = padded_terms(T);
  const size_t TM = (size_t)Tp * M;
  unsigned char* p = (unsigned char*)ws;
  R* Fpr = (R*)p;
  R* Fmr = Fpr + TM;
  R* Fpi = CPLX ? Fmr + TM : nullptr;
  R* Fmi = CPLX ? Fpi + TM : nullptr;
  const size_t tab = TM * sizeof(R) * (CPLX ? 4 : 2);
  double* res_r = (double*)(p + ((tab + 255) & ~(size_t)255));
  double* res_i = res_r + (size_t)T * B;
  KScope ks("eloc", st);
  if (T > 0) {
    if (TM > 0) {
      long blocks = ((long)TM + 255) / 256;
      if (blocks > 4096) blocks = 4096;
      hipLaunchKernelGGL((k_term_tables<R, CPLX>), dim3((unsigned)blocks), dim3(256), 0, st, m, h,
                         Tp, Fpr, Fpi, Fmr, Fmi);
      count_launch();
    }
    const dim3 grid((B + 63) / 64, Tp / (4 * kEtTerms));
    hipLaunchKernelGGL((k_eloc_terms<R, CPLX>), grid, dim3(256), 0, st, m, h, spins, B, t_re, q_re,
                       q_im, Fpr, Fpi, Fmr, Fmi, res_r, res_i);
    count_launch();
  }
  if (res_out) {
    res_out[0] = res_r;
    res_out[1] = res_i;
  }
  if (e_re) {
    hipLaunchKernelGGL(k_eloc_terms_sum, dim3((B + 255) / 256), dim3(256), 0, st, h, spins, B,
                       m.N, res_r, res_i, e_re, e_im);
    count_launch();
  }
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

template <typename R>
int launch_pair_ratios(const ModelView<R>& m, const HamDev& h, const int* spins, int B,
                       const R* t_re, const R* q_re, const R* q_im, void* ws, size_t ws_bytes,
                       double** res, hipStream_t st) {
  if (B <= 0) return 0;
  if (m.N > kMaxSites || h.n_flip + h.n_pair >= 65536 || !ws || ws_bytes < eloc_ws_bytes(m, h, B) ||
      (m.M > 0 && !q_re)) {
    fprintf(stderr, "nqs_hip: amplitude ratios need N <= %d, < 65536 terms, the logistic members "
                    "and an E_loc workspace\n", kMaxSites);
    return -1;
  }
  return m.cplx ? eloc_terms_go<R, true>(m, h, spins, B, t_re, q_re, q_im, nullptr, nullptr, ws, st,
                                         res)
                : eloc_terms_go<R, false>(m, h, spins, B, t_re, q_re, q_im, nullptr, nullptr, ws,
                                          st, res);
}

template <typename R, bool CPLX, int IPT>
static int eloc_go(const ModelView<R>& m, const HamDev& h, const int* spins, int B,
                   const R* t_re, const R* q_re, const R* q_im, double* e_re, double* e_im,
                   hipStream_t st) {
  const int N = m.N, T = h.n_flip + h.n_pair;
  int SB = (256 * IPT) / (T > 0 ? T : 1);
  if (SB < 1) SB = 1;
  if (SB > 64) SB = 64;
  if (SB > B) SB = B;
  const int ntab = CPLX ? 4 : 2, nt = CPLX ? 3 : 2;
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
                     t_re, q_re, q_im, e_re, e_im, SB, HT);
  count_launch();
  NQS_HIP_CHECK(hipGetLastError());
  return 0;
}

template <typename R>
int launch_eloc(const ModelView<R>& m, const HamDev& h, const int* spins, int B,
                const R* t_re, const R* t_im, const R* q_re, const R* q_im, double* e_re,
                double* e_im, hipStream_t st, void* ws, size_t ws_bytes) {
  (void)t_im;   // only the sign of Re t is read (which logistic member q holds)
  if (B <= 0) return 0;
  const int T = h.n_flip + h.n_pair;
  if (m.N > kMaxSites || T >= 65536) {
    fprintf(stderr, "nqs_hip: local energies support N <= %d sites and < 65536 off-diagonal "
                    "terms (got N = %d, %d terms)\n", kMaxSites, m.N, T);
    return -1;
  }
  if (m.M > 0 && !q_re) {
    fprintf(stderr, "nqs_hip: local energies need the logistic members of the forward pass\n");
    return -1;
  }
  const bool c = m.cplx != 0;
  if (ws && ws_bytes >= eloc_ws_bytes(m, h, B))
    return c ? eloc_terms_go<R, true>(m, h, spins, B, t_re, q_re, q_im, e_re, e_im, ws, st)
             : eloc_terms_go<R, false>(m, h, spins, B, t_re, q_re, q_im, e_re, e_im, ws, st);
  if (T <= 256) {
    return c ? eloc_go<R, true, 1>(m, h, spins, B, t_re, q_re, q_im, e_re, e_im, st)
             : eloc_go<R, false, 1>(m, h, spins, B, t_re, q_re, q_im, e_re, e_im, st);
  } else if (T <= 512) {
    return c ? eloc_go<R, true, 2>(m, h, spins, B, t_re, q_re, q_im, e_re, e_im, st)
             : eloc_go<R, false, 2>(m, h, spins, B, t_re, q_re, q_im, e_re, e_im, st);
  } else if (T <= 1024) {
    return c ? eloc_go<R, true, 4>(m, h, spins, B, t_re, q_re, q_im, e_re, e_im, st)
             : eloc_go<R, false, 4>(m, h, spins, B, t_re, q_re, q_im, e_re, e_im, st);
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

template int launch_eloc<double>(const ModelView<double>&, const HamDev&, const int*, int, const double*, const double*, const double*, const double*, double*, double*, hipStream_t, void*, size_t);
template int launch_eloc<float>(const ModelView<float>&, const HamDev&, const int*, int, const float*, const float*, const float*, const float*, double*, double*, hipStream_t, void*, size_t);
template int launch_pair_ratios<double>(const ModelView<double>&, const HamDev&, const int*, int, const double*, const double*, const double*, void*, size_t, double**, hipStream_t);
template size_t eloc_ws_bytes<double>(const ModelView<double>&, const HamDev&, int);
template size_t eloc_ws_bytes<float>(const ModelView<float>&, const HamDev&, int);

}  // namespace nqs
