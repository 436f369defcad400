// engine.hip — host engine and the exported C ABI of libnqs_hip.so.
//
// Every reference entry point of include/nqs (ansatz, sampler, local
// energy, SR / holomorphic SR / MinSR, Marshall) is implemented here over
// the gfx950 kernels of k_*.hip.  One HIP stream (default priority) carries
// all work; calls are serialised by a process-wide mutex.
//
// Callback recognition (the reference passes amplitudes as host function
// pointers, include/nqs/nqs_sampler.h:35-37): nqs_ansatz_log_amp over an
// ansatz handle, and nqs_marshall_log_amp wrapping it, are evaluated on the
// device.  Any other callback is a caller's host function; it is called on
// the host for exactly the configurations the reference would evaluate
// (sampler moves, connected configurations) and everything else stays on
// the device.  nqs_hip_host_callback_calls() counts those calls.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/nqs/nqs_ansatz.h"
#include "../../include/nqs/nqs_diagnostics.h"
#include "../../include/nqs/nqs_lanczos.h"
#include "../../include/nqs/nqs_symproj.h"
#include "../../include/nqs/nqs_translation.h"
#include "../../include/nqs/nqs_gradient.h"
#include "../../include/nqs/nqs_marshall.h"
#include "../../include/nqs/nqs_optimizer.h"
#include "../../include/nqs/nqs_sampler.h"
#include "../../include/nqs_hip/nqs_hip.h"
#include "ar_layout.h"
#include "lattice.h"
#include "nqs_kernels.h"

#define NQS_EXPORT extern "C" __attribute__((visibility("default")))

using nqs::ModelView;

namespace {

std::recursive_mutex g_mu;
hipStream_t g_stream = nullptr;
std::atomic<long long> g_launches{0};
std::atomic<long long> g_host_calls{0};

#define CK(expr)                                                                \
  do {                                                                          \
    hipError_t _e = (expr);                                                     \
    if (_e != hipSuccess) {                                                     \
      fprintf(stderr, "nqs_hip: %s failed (%s:%d): %s\n", #expr, __FILE__,      \
              __LINE__, hipGetErrorString(_e));                                 \
      return -1;                                                                \
    }                                                                           \
  } while (0)
#define CKR(expr)            \
  do {                       \
    if ((expr) != 0) return -1; \
  } while (0)

int stream_ready() {
  if (g_stream) return 0;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    fprintf(stderr, "nqs_hip: no HIP device available; this library runs on the GPU only\n");
    return -1;
  }
  CK(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
  return 0;
}

struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
  int ensure(size_t b) {
    if (b <= bytes && p) return 0;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    if (b == 0) b = 16;
    if (hipMalloc(&p, b) != hipSuccess) {
      fprintf(stderr, "nqs_hip: hipMalloc(%zu) failed\n", b);
      (void)hipGetLastError();   // clear the error so later launch checks see their own
      p = nullptr;
      return -1;
    }
    bytes = b;
    return 0;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <typename T>
  T* as() const { return (T*)p; }
};

// ---------------------------------------------------------------------------
// reference initialiser stream, src/nqs/nqs_ansatz.c:65-252 (one thread)
// ---------------------------------------------------------------------------
__device__ double d_gauss(uint64_t& s) {
  double u1 = nqs::xs_uniform(s);
  double u2 = nqs::xs_uniform(s);
  if (u1 < 1e-15) u1 = 1e-15;
  return sqrt(-2.0 * log(u1)) * cos(2.0 * 3.14159265358979323846 * u2);
}

__global__ void k_init_params(double* p, int kind, int N, int M, double scale, uint64_t seed) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint64_t s = seed;
  if (kind == NQS_ANSATZ_LEGACY_MLP) {
    for (int i = 0; i < N; i++) p[i] = 0.01 * (nqs::xs_uniform(s) - 0.5);
    return;
  }
  const long MN = (long)M * N, Pr = (long)N + M + MN;
  const double ws = scale / sqrt((double)N);
  for (int i = 0; i < N; i++) p[i] = scale * (nqs::xs_uniform(s) - 0.5);
  for (int h = 0; h < M; h++) p[N + h] = scale * (nqs::xs_uniform(s) - 0.5);
  for (long k = 0; k < MN; k++) p[N + M + k] = ws * d_gauss(s);
  if (kind == NQS_ANSATZ_COMPLEX_RBM) {
    const double si = 0.1 * scale, wsi = si / sqrt((double)N);
    double* q = p + Pr;
    for (int i = 0; i < N; i++) q[i] = si * (nqs::xs_uniform(s) - 0.5);
    for (int h = 0; h < M; h++) q[N + h] = si * (nqs::xs_uniform(s) - 0.5);
    for (long k = 0; k < MN; k++) q[N + M + k] = wsi * d_gauss(s);
  }
}

}  // namespace

void nqs::count_launch(int n) { g_launches += n; }

namespace {
struct KStat {
  std::string name;
  long long count = 0;
  double total_ms = 0.0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
};
std::mutex g_kt_mu;
std::vector<KStat> g_kstats;
std::atomic<int> g_kt_on{0};

void kt_resolve() {
  for (KStat& k : g_kstats) {
    for (auto& pr : k.pending) {
      float ms = 0.f;
      if (hipEventSynchronize(pr.second) == hipSuccess &&
          hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess) {
        k.total_ms += ms;
        k.count++;
      }
      (void)hipEventDestroy(pr.first);
      (void)hipEventDestroy(pr.second);
    }
    k.pending.clear();
  }
}
}  // namespace

nqs::KScope::KScope(const char* name, hipStream_t stream) {
  if (!g_kt_on.load()) return;
  std::lock_guard<std::mutex> lk(g_kt_mu);
  for (size_t i = 0; i < g_kstats.size(); i++)
    if (g_kstats[i].name == name) slot = (int)i;
  if (slot < 0) return;
  st = stream;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    slot = -1;
    return;
  }
  (void)hipEventRecord(e0, st);
}

nqs::KScope::~KScope() {
  if (slot < 0) return;
  (void)hipEventRecord(e1, st);
  std::lock_guard<std::mutex> lk(g_kt_mu);
  g_kstats[slot].pending.push_back({e0, e1});
}

// names: comma-separated kernel tags to time (NULL or "" disables);
// resets the statistics.
NQS_EXPORT int nqs_hip_kernel_timing(const char* names) {
  std::lock_guard<std::mutex> lk(g_kt_mu);
  kt_resolve();
  g_kstats.clear();
  g_kt_on = 0;
  if (!names || !*names) return 0;
  std::string all(names);
  size_t pos = 0;
  while (pos <= all.size()) {
    size_t q = all.find(',', pos);
    if (q == std::string::npos) q = all.size();
    if (q > pos) {
      KStat k;
      k.name = all.substr(pos, q - pos);
      g_kstats.push_back(k);
    }
    pos = q + 1;
  }
  g_kt_on = 1;
  return (int)g_kstats.size();
}

// launches and total milliseconds of one timed kernel tag (0 on success)
NQS_EXPORT int nqs_hip_kernel_stats(const char* name, long long* count, double* total_ms) {
  std::lock_guard<std::mutex> lk(g_kt_mu);
  kt_resolve();
  for (KStat& k : g_kstats)
    if (name && k.name == name) {
      if (count) *count = k.count;
      if (total_ms) *total_ms = k.total_ms;
      return 0;
    }
  return -1;
}

// ===========================================================================
// ansatz
// ===========================================================================
struct nqs_ansatz {
  int kind = 0, N = 0, M = 0;
  long Pr = 0, P = 0;
  std::vector<double> h;   // host master copy (params_raw)
  DBuf d_par;
  nqs::ModelBuffers<double> m64;
  nqs::ModelBuffers<float> m32;
  int dev_fresh = 0;       // device copy equals host copy (inside a run)
  long long id = 0;        // unique per ansatz created in this process
};

namespace {

int ansatz_upload(nqs_ansatz* a) {
  if (a->dev_fresh) return 0;
  CK(hipMemcpyAsync(a->d_par.p, a->h.data(), sizeof(double) * a->P, hipMemcpyHostToDevice,
                    g_stream));
  a->m64.valid = 0;
  a->m32.valid = 0;
  return 0;
}

int ansatz_download(nqs_ansatz* a) {
  CK(hipMemcpyAsync(a->h.data(), a->d_par.p, sizeof(double) * a->P, hipMemcpyDeviceToHost,
                    g_stream));
  CK(hipStreamSynchronize(g_stream));
  return 0;
}

template <typename R>
nqs::ModelBuffers<R>& model_bufs(nqs_ansatz* a);
template <>
nqs::ModelBuffers<double>& model_bufs<double>(nqs_ansatz* a) { return a->m64; }
template <>
nqs::ModelBuffers<float>& model_bufs<float>(nqs_ansatz* a) { return a->m32; }

// check = 0: tables prepared for amplitudes only, without waiting for the
// ratio tables' range verdict (ratio_view reads it when a caller needs them)
template <typename R>
int model_view(nqs_ansatz* a, ModelView<R>& v, int check = 1) {
  nqs::ModelBuffers<R>& mb = model_bufs<R>(a);
  if (!mb.valid)
    CKR(nqs::launch_prep<R>(a->d_par.as<double>(), a->kind, a->N, a->M, mb, g_stream, check));
  v = mb.view;
  return 0;
}

// the view for kernels that price moves with the ratio tables (sampler,
// local energies): refused, with the reason on stderr, when a weight lies
// beyond the tables' range
template <typename R>
int ratio_view(nqs_ansatz* a, ModelView<R>& v) {
  CKR(model_view<R>(a, v));
  nqs::ModelBuffers<R>& mb = model_bufs<R>(a);
  if (mb.ratio_ok < 0) {
    // prepared without the verdict (amplitude-only use): read it now
    int flag = 0;
    CK(hipMemcpyAsync(&flag, mb.range_flag, sizeof(int), hipMemcpyDeviceToHost, g_stream));
    CK(hipStreamSynchronize(g_stream));
    mb.ratio_ok = flag ? 0 : 1;
  }
  if (mb.ratio_ok == 0) {
    (void)nqs::report_ratio_range<R>(a->d_par.as<double>(), a->N, a->M, mb.range_flag, g_stream);
    return -1;
  }
  return 0;
}

bool is_ansatz_cb(nqs_log_amp_fn_t fn, void* user) {
  return fn == nqs_ansatz_log_amp && user != nullptr;
}

// Device-evaluable amplitude source behind a callback: the ansatz, and a
// Marshall sublattice when wrapped.
struct AmpSource {
  nqs_ansatz* a = nullptr;
  int marshall = 0, mlx = 0, mly = 0;
};

bool recognise(nqs_log_amp_fn_t fn, void* user, AmpSource& src) {
  src = AmpSource();
  if (is_ansatz_cb(fn, user)) {
    src.a = (nqs_ansatz*)user;
    return true;
  }
  if (fn == nqs_marshall_log_amp && user) {
    const nqs_marshall_wrapper_t* w = (const nqs_marshall_wrapper_t*)user;
    if (is_ansatz_cb(w->base_log_amp, w->base_user)) {
      src.a = (nqs_ansatz*)w->base_user;
      src.marshall = 1;
      src.mlx = w->size_x;
      src.mly = w->size_y;
      return true;
    }
  }
  return false;
}

// ---------------------------------------------------------------------------
// group-orbit amplitude sources (nqs_symproj_log_amp, nqs_translation_log_amp
// over a device amplitude source): the (perm, characters) tables on the
// device, cached by address, size and content
// ---------------------------------------------------------------------------
struct OrbitTable {
  const int* hp = nullptr;
  const double* hc = nullptr;
  int G = 0, N = 0;
  uint64_t hash = 0;
  DBuf perm, chi;
};

uint64_t orbit_hash(const int* perm, const double* chi, int G, int N) {
  uint64_t h = 1469598103934665603ULL;
  auto mix = [&](const void* p, size_t n) {
    const unsigned char* c = (const unsigned char*)p;
    for (size_t i = 0; i < n; i++) h = (h ^ c[i]) * 1099511628211ULL;
  };
  mix(perm, sizeof(int) * (size_t)G * N);
  mix(chi, sizeof(double) * (size_t)G);
  return h;
}

OrbitTable* orbit_table(const int* perm, const double* chi, int G, int N) {
  static OrbitTable cache[4];
  static int next = 0;
  const uint64_t h = orbit_hash(perm, chi, G, N);
  for (OrbitTable& t : cache)
    if (t.hp == perm && t.hc == chi && t.G == G && t.N == N && t.hash == h) return &t;
  for (int g = 0; g < G; g++)
    for (int i = 0; i < N; i++)
      if (perm[(size_t)g * N + i] < 0 || perm[(size_t)g * N + i] >= N) {
        fprintf(stderr, "nqs_hip: symmetry table entry perm[%d][%d] = %d is not a site\n", g, i,
                perm[(size_t)g * N + i]);
        return nullptr;
      }
  OrbitTable& t = cache[next];
  next = (next + 1) % 4;
  if (t.perm.ensure(sizeof(int) * (size_t)G * N) || t.chi.ensure(sizeof(double) * G)) return nullptr;
  if (hipMemcpyAsync(t.perm.p, perm, sizeof(int) * (size_t)G * N, hipMemcpyHostToDevice, g_stream) ||
      hipMemcpyAsync(t.chi.p, chi, sizeof(double) * G, hipMemcpyHostToDevice, g_stream) ||
      hipStreamSynchronize(g_stream))
    return nullptr;
  t.hp = perm;
  t.hc = chi;
  t.G = G;
  t.N = N;
  t.hash = h;
  return &t;
}

// nqs_translation's table (nqs_translation.c:36-48): element (tx, ty),
// tx-major, takes site x Ly + y from ((x + tx) mod Lx) Ly + (y + ty) mod Ly
struct TranslationTable {
  int Lx = 0, Ly = 0;
  std::vector<int> perm;
  std::vector<double> chi;
};

const TranslationTable& translation_table(int Lx, int Ly) {
  static TranslationTable t;
  if (t.Lx != Lx || t.Ly != Ly) {
    t.Lx = Lx;
    t.Ly = Ly;
    const int N = Lx * Ly;
    t.perm.assign((size_t)N * N, 0);
    t.chi.assign(N, 1.0);
    int g = 0;
    for (int tx = 0; tx < Lx; tx++)
      for (int ty = 0; ty < Ly; ty++, g++)
        for (int x = 0; x < Lx; x++)
          for (int y = 0; y < Ly; y++)
            t.perm[(size_t)g * N + x * Ly + y] = ((x + tx) % Lx) * Ly + (y + ty) % Ly;
  }
  return t;
}

struct OrbitSource {
  AmpSource base;
  OrbitTable* tab = nullptr;
};

bool recognise_orbit(nqs_log_amp_fn_t fn, void* user, OrbitSource& o) {
  o = OrbitSource();
  if (!user) return false;
  nqs_log_amp_fn_t bfn = nullptr;
  void* buser = nullptr;
  const int* perm = nullptr;
  const double* chi = nullptr;
  int G = 0, N = 0;
  if (fn == nqs_symproj_log_amp) {
    const nqs_symproj_wrapper_t* w = (const nqs_symproj_wrapper_t*)user;
    bfn = w->base_log_amp;
    buser = w->base_user;
    perm = w->perm;
    chi = w->characters;
    G = w->num_group_elements;
    N = w->num_sites;
  } else if (fn == nqs_translation_log_amp) {
    const nqs_translation_wrapper_t* w = (const nqs_translation_wrapper_t*)user;
    if (w->size_x <= 0 || w->size_y <= 0) return false;
    const TranslationTable& t = translation_table(w->size_x, w->size_y);
    bfn = w->base_log_amp;
    buser = w->base_user;
    perm = t.perm.data();
    chi = t.chi.data();
    N = G = w->size_x * w->size_y;
  } else {
    return false;
  }
  if (!bfn || !perm || !chi || G <= 0 || N <= 0) return false;
  if (!recognise(bfn, buser, o.base) || o.base.a->N != N) return false;
  o.tab = orbit_table(perm, chi, G, N);
  return o.tab != nullptr;
}

// ---------------------------------------------------------------------------
// workspace shared by all calls (single stream, serialised)
// ---------------------------------------------------------------------------
struct Workspace {
  DBuf spins, t_re, t_im, q_re, q_im, la, ar, e_re, e_im, O, Oi, colwork, u1, u2, vecs, T, diag,
      y, scal, cg, fail, ham, sublat, hostla, hostar, linv, T2, Sd, elw, fwd, xr, heun, bond_ham,
      bond_sub, orb_rows, orb_lp, orb_ar, orb_O, orb_Oi, orb_w, orb_cfg, o32, x3part, x3max;
};

// FP32 arithmetic for amplitudes, local energies, moves and O (the FP32 and
// mixed modes)
inline bool low_prec(int p) { return p == NQS_HIP_FP32 || p == NQS_HIP_MIXED; }
Workspace g_ws;

// Hamiltonian on the device, rebuilt when the request changes.
struct HamCache {
  nqs_config_t cfg;
  int Lx = -1, Ly = -1, periodic = -1, marshall = 0, mlx = 0, mly = 0;
  nqs::HamHost host;
  nqs::HamDev dev;
  int valid = 0;
};
HamCache g_ham;

bool same_ham(const nqs_config_t* c, int Lx, int Ly, int per, const AmpSource& src) {
  if (!g_ham.valid) return false;
  const nqs_config_t& o = g_ham.cfg;
  return o.hamiltonian == c->hamiltonian && o.j_coupling == c->j_coupling &&
         o.j2_coupling == c->j2_coupling && o.j_z_coupling == c->j_z_coupling &&
         o.transverse_field == c->transverse_field && o.kagome_pbc == c->kagome_pbc &&
         g_ham.Lx == Lx && g_ham.Ly == Ly && g_ham.periodic == per &&
         g_ham.marshall == src.marshall && g_ham.mlx == src.mlx && g_ham.mly == src.mly;
}

int ham_upload(const nqs::HamHost& H, const AmpSource& src, DBuf& store, DBuf& sub_store,
               nqs::HamDev& d);

int ham_device(const nqs_config_t* cfg, int Lx, int Ly, int per, const AmpSource& src,
               nqs::HamDev& out) {
  if (same_ham(cfg, Lx, Ly, per, src)) {
    out = g_ham.dev;
    return 0;
  }
  g_ham.valid = 0;
  if (nqs::build_hamiltonian(cfg, Lx, Ly, per, g_ham.host) != 0) {
    fprintf(stderr, "nqs_hip: Hamiltonian kind %d is not provided by this library\n",
            (int)cfg->hamiltonian);
    return -1;
  }
  nqs::HamDev d;
  CKR(ham_upload(g_ham.host, src, g_ws.ham, g_ws.sublat, d));
  g_ham.cfg = *cfg;
  g_ham.Lx = Lx;
  g_ham.Ly = Ly;
  g_ham.periodic = per;
  g_ham.marshall = src.marshall;
  g_ham.mlx = src.mlx;
  g_ham.mly = src.mly;
  g_ham.dev = d;
  g_ham.valid = 1;
  out = d;
  return 0;
}

// term lists of H (and the Marshall sublattice of src) into device buffers
int ham_upload(const nqs::HamHost& H, const AmpSource& src, DBuf& store, DBuf& sub_store,
               nqs::HamDev& d) {
  const size_t nd = H.diag_a.size(), nf = H.flip_i.size(), np = H.pair_a.size();
  const size_t ints = 2 * nd + nf + 2 * np + 4, dbls = nd + nf + np + 4;
  CKR(store.ensure(ints * sizeof(int) + dbls * sizeof(double) + 64));
  char* base = store.as<char>();
  double* dc = (double*)base;
  double* fc = dc + nd;
  double* pc = fc + nf;
  int* ip = (int*)(pc + np + 1);
  int* da = ip;
  int* db = da + nd;
  int* fi = db + nd;
  int* pa = fi + nf;
  int* pb = pa + np;
  std::vector<char> stage(ints * sizeof(int) + dbls * sizeof(double) + 64);
  double* hdc = (double*)stage.data();
  std::copy(H.diag_c.begin(), H.diag_c.end(), hdc);
  std::copy(H.flip_c.begin(), H.flip_c.end(), hdc + nd);
  std::copy(H.pair_c.begin(), H.pair_c.end(), hdc + nd + nf);
  int* hip_ = (int*)(hdc + nd + nf + np + 1);
  std::copy(H.diag_a.begin(), H.diag_a.end(), hip_);
  std::copy(H.diag_b.begin(), H.diag_b.end(), hip_ + nd);
  std::copy(H.flip_i.begin(), H.flip_i.end(), hip_ + 2 * nd);
  std::copy(H.pair_a.begin(), H.pair_a.end(), hip_ + 2 * nd + nf);
  std::copy(H.pair_b.begin(), H.pair_b.end(), hip_ + 2 * nd + nf + np);
  CK(hipMemcpyAsync(base, stage.data(), stage.size(), hipMemcpyHostToDevice, g_stream));
  d = nqs::HamDev();
  d.n_diag = (int)nd;
  d.n_flip = (int)nf;
  d.n_pair = (int)np;
  d.diag_a = da; d.diag_b = db; d.diag_c = dc;
  d.flip_i = fi; d.flip_c = fc;
  d.pair_a = pa; d.pair_b = pb; d.pair_c = pc;
  d.sublat = nullptr;
  if (src.marshall) {
    const int N = H.n_sites;
    std::vector<signed char> sub(N, 0);
    for (int x = 0; x < src.mlx; x++)
      for (int y = 0; y < src.mly; y++) {
        const int i = x * src.mly + y;
        if (i < N && ((x + y) & 1) == 0) sub[i] = 1;
      }
    CKR(sub_store.ensure(N + 16));
    CK(hipMemcpyAsync(sub_store.p, sub.data(), N, hipMemcpyHostToDevice, g_stream));
    d.sublat = sub_store.as<signed char>();
  }
  CK(hipStreamSynchronize(g_stream));
  return 0;
}

// forward pass of B device configurations into the workspace t buffers
// (and, with want_q, the small logistic members the local energies use)
template <typename R>
int forward_t(nqs_ansatz* a, const int* d_spins, int B, double* la, double* ar,
              bool want_q = false, bool x3 = false) {
  ModelView<R> v;
  CKR(model_view<R>(a, v));
  const size_t tb = sizeof(R) * (size_t)B * (size_t)std::max(a->M, 1);
  CKR(g_ws.t_re.ensure(tb));
  CKR(g_ws.t_im.ensure(tb));
  R *q_re = nullptr, *q_im = nullptr;
  if (want_q) {
    CKR(g_ws.q_re.ensure(tb));
    CKR(g_ws.q_im.ensure(tb));
    q_re = g_ws.q_re.as<R>();
    q_im = g_ws.q_im.as<R>();
  }
  // many FP64 RBM configurations, t only: theta on FP64 MFMA
  if (std::is_same<R, double>::value && !la && !ar && B >= 512 && a->M >= 1 &&
      (a->kind == NQS_ANSATZ_RBM || a->kind == NQS_ANSATZ_COMPLEX_RBM)) {
    CKR(g_ws.fwd.ensure(sizeof(double) * (size_t)B * a->N));
    ModelView<double> vd;
    CKR(model_view<double>(a, vd));
    return nqs::launch_forward_t_gemm(vd, a->d_par.as<double>(), d_spins, B,
                                      g_ws.t_re.as<double>(), g_ws.t_im.as<double>(),
                                      g_ws.fwd.as<double>(), g_stream, (double*)q_re,
                                      (double*)q_im);
  }
  // the mixed mode's FP32 t of an RBM: theta on the split-BF16 MFMA
  if (std::is_same<R, float>::value && x3 && !la && !ar && !want_q && a->M >= 1 &&
      (a->kind == NQS_ANSATZ_RBM || a->kind == NQS_ANSATZ_COMPLEX_RBM)) {
    CKR(g_ws.fwd.ensure(sizeof(__bf16) * 2 * (size_t)nqs::theta_x3_split_elems(a->M, a->N)));
    ModelView<float> vf;
    CKR(model_view<float>(a, vf));
    return nqs::launch_forward_t_x3(vf, a->d_par.as<double>(), d_spins, B,
                                    g_ws.t_re.as<float>(), g_ws.t_im.as<float>(),
                                    g_ws.fwd.as<__bf16>(), g_stream);
  }
  return nqs::launch_forward<R>(v, d_spins, B, g_ws.t_re.as<R>(), g_ws.t_im.as<R>(), la, ar,
                                g_stream, q_re, q_im);
}

// E_loc of B device configurations (device amplitude source)
template <typename R>
int eloc_device_T(const nqs_config_t* cfg, int Lx, int Ly, int per, const AmpSource& src,
                  const int* d_spins, int B, double* d_re, double* d_im) {
  nqs::HamDev H;
  CKR(ham_device(cfg, Lx, Ly, per, src, H));
  if (g_ham.host.n_sites != src.a->N) {
    fprintf(stderr, "nqs_hip: lattice has %d sites but the ansatz %d\n", g_ham.host.n_sites,
            src.a->N);
    return -1;
  }
  ModelView<R> v;
  CKR(ratio_view<R>(src.a, v));
  CKR(forward_t<R>(src.a, d_spins, B, nullptr, nullptr, true));
  const size_t wb = nqs::eloc_ws_bytes<R>(v, H, B);
  CKR(g_ws.elw.ensure(wb));
  return nqs::launch_eloc<R>(v, H, d_spins, B, g_ws.t_re.as<R>(), g_ws.t_im.as<R>(),
                             g_ws.q_re.as<R>(), g_ws.q_im.as<R>(), d_re, d_im, g_stream,
                             g_ws.elw.p, wb);
}

int eloc_device(const nqs_config_t* cfg, int Lx, int Ly, int per, int prec, const AmpSource& src,
                const int* d_spins, int B, double* d_re, double* d_im) {
  if (low_prec(prec))
    return eloc_device_T<float>(cfg, Lx, Ly, per, src, d_spins, B, d_re, d_im);
  return eloc_device_T<double>(cfg, Lx, Ly, per, src, d_spins, B, d_re, d_im);
}

// log|psi| and arg psi of B device configurations under a group-orbit
// source: the orbit rows of a slice of configurations, one batched forward
// of the base ansatz, the real combination per configuration
int orbit_log_amp(const OrbitSource& o, const int* d_spins, long B, double* d_la, double* d_ar) {
  nqs_ansatz* a = o.base.a;
  const int N = a->N, G = o.tab->G;
  CKR(ansatz_upload(a));
  ModelView<double> v;
  CKR(model_view<double>(a, v));
  const long slice = std::max(1L, (1L << 20) / G);
  const long rows = std::min(B, slice) * G;
  CKR(g_ws.orb_rows.ensure(sizeof(int) * (size_t)rows * N));
  CKR(g_ws.orb_lp.ensure(sizeof(double) * (size_t)rows));
  CKR(g_ws.orb_ar.ensure(sizeof(double) * (size_t)rows));
  int* r = g_ws.orb_rows.as<int>();
  double* lp = g_ws.orb_lp.as<double>();
  double* ar = g_ws.orb_ar.as<double>();
  for (long b0 = 0; b0 < B; b0 += slice) {
    const int n = (int)std::min(slice, B - b0);
    CKR(nqs::launch_orbit_spins(d_spins + (size_t)b0 * N, n, N, o.tab->perm.as<int>(), G, r,
                                g_stream));
    CKR(nqs::launch_forward<double>(v, r, n * G, nullptr, nullptr, lp, ar, g_stream));
    if (o.base.marshall)
      CKR(nqs::launch_marshall_arg(r, n * G, N, o.base.mlx, o.base.mly, ar, g_stream));
    CKR(nqs::launch_orbit_combine(lp, ar, o.tab->chi.as<double>(), n, G, d_la + b0,
                                  d_ar ? d_ar + b0 : nullptr, g_stream));
  }
  return 0;
}

// O rows [B][P] (and the imaginary rows with cplx) of the projected
// amplitude: weights from the orbit source's base amplitudes, base
// gradients of ansatz a at the orbit rows (real-projected, or holomorphic
// with cplx), summed per configuration
int orbit_gradients(const OrbitSource& o, nqs_ansatz* a, const int* d_spins, long B, bool cplx,
                    double* O, double* Oi) {
  nqs_ansatz* ab = o.base.a;
  const int N = a->N, G = o.tab->G;
  const long P = a->P;
  CKR(ansatz_upload(ab));
  CKR(ansatz_upload(a));
  ModelView<double> vb, va;
  CKR(model_view<double>(ab, vb));
  CKR(model_view<double>(a, va));
  long slice = std::max(1L, (1L << 26) / ((long)G * P));
  slice = std::min(slice, std::max(1L, (1L << 20) / G));
  const long rows = std::min(B, slice) * G;
  CKR(g_ws.orb_rows.ensure(sizeof(int) * (size_t)rows * N));
  CKR(g_ws.orb_lp.ensure(sizeof(double) * (size_t)rows));
  CKR(g_ws.orb_ar.ensure(sizeof(double) * (size_t)rows));
  CKR(g_ws.orb_O.ensure(sizeof(double) * (size_t)rows * P));
  if (cplx) CKR(g_ws.orb_Oi.ensure(sizeof(double) * (size_t)rows * P));
  CKR(g_ws.orb_w.ensure(sizeof(double) * 2 * (size_t)rows));
  int* r = g_ws.orb_rows.as<int>();
  double* lp = g_ws.orb_lp.as<double>();
  double* ar = g_ws.orb_ar.as<double>();
  double* w = g_ws.orb_w.as<double>();
  for (long b0 = 0; b0 < B; b0 += slice) {
    const int n = (int)std::min(slice, B - b0);
    CKR(nqs::launch_orbit_spins(d_spins + (size_t)b0 * N, n, N, o.tab->perm.as<int>(), G, r,
                                g_stream));
    CKR(nqs::launch_forward<double>(vb, r, n * G, nullptr, nullptr, lp, ar, g_stream));
    if (o.base.marshall)
      CKR(nqs::launch_marshall_arg(r, n * G, N, o.base.mlx, o.base.mly, ar, g_stream));
    CKR(forward_t<double>(a, r, n * G, nullptr, nullptr));
    CKR(nqs::launch_gen_O<double>(va, a->kind, r, n * G, g_ws.t_re.as<double>(),
                                  g_ws.t_im.as<double>(), cplx ? 1 : 0, g_ws.orb_O.as<double>(),
                                  cplx ? g_ws.orb_Oi.as<double>() : nullptr, P, g_stream));
    CKR(nqs::launch_orbit_grad(lp, ar, o.tab->chi.as<double>(), n, G, cplx ? 1 : 0,
                               g_ws.orb_O.as<double>(), cplx ? g_ws.orb_Oi.as<double>() : nullptr,
                               P, w, w + rows, O + (size_t)b0 * P,
                               cplx ? Oi + (size_t)b0 * P : nullptr, g_stream));
  }
  return 0;
}

// the gradient callbacks of the projection wrappers over a device source
bool orbit_gradient_cb(bool holo, nqs_gradient_fn_t gfn, nqs_holomorphic_gradient_fn_t hfn,
                       void* guser, OrbitSource& o) {
  if (holo) return hfn == nqs_symproj_gradient_complex && recognise_orbit(nqs_symproj_log_amp, guser, o);
  if (gfn == nqs_symproj_gradient) return recognise_orbit(nqs_symproj_log_amp, guser, o);
  if (gfn == nqs_translation_gradient) return recognise_orbit(nqs_translation_log_amp, guser, o);
  return false;
}

// E_loc with a host callback: the callback is evaluated on the sample and on
// each connected configuration; the sums run on the device.
int eloc_callback(const nqs_config_t* cfg, int Lx, int Ly, int per, const int* h_spins, int B,
                  nqs_log_amp_fn_t fn, void* user, double* d_re, double* d_im,
                  const int* d_spins) {
  AmpSource none;
  nqs::HamDev H;
  CKR(ham_device(cfg, Lx, Ly, per, none, H));
  const nqs::HamHost& HH = g_ham.host;
  const int N = HH.n_sites;
  const int T = H.n_flip + H.n_pair;
  OrbitSource orb;
  if (recognise_orbit(fn, user, orb) && orb.base.a->N == N) {
    // a group-orbit source over a device ansatz: every sample and each of
    // its connected configurations (the sample itself for inactive pairs)
    // as rows [B][T + 1][N], their amplitudes in device batches
    const size_t R = (size_t)B * (T + 1);
    std::vector<int> rows(R * N);
    for (int b = 0; b < B; b++) {
      const int* sp = h_spins + (size_t)b * N;
      for (int t = 0; t <= T; t++) {
        int* o = rows.data() + ((size_t)b * (T + 1) + t) * N;
        memcpy(o, sp, sizeof(int) * N);
        if (t == 0) continue;
        const int k = t - 1;
        if (k < H.n_flip) o[HH.flip_i[k]] = -o[HH.flip_i[k]];
        else if (sp[HH.pair_a[k - H.n_flip]] != sp[HH.pair_b[k - H.n_flip]]) {
          o[HH.pair_a[k - H.n_flip]] = -o[HH.pair_a[k - H.n_flip]];
          o[HH.pair_b[k - H.n_flip]] = -o[HH.pair_b[k - H.n_flip]];
        }
      }
    }
    CKR(g_ws.orb_cfg.ensure(sizeof(int) * rows.size()));
    CK(hipMemcpyAsync(g_ws.orb_cfg.p, rows.data(), sizeof(int) * rows.size(), hipMemcpyHostToDevice,
                      g_stream));
    CKR(g_ws.hostla.ensure(sizeof(double) * R));
    CKR(g_ws.hostar.ensure(sizeof(double) * R));
    CKR(orbit_log_amp(orb, g_ws.orb_cfg.as<int>(), (long)R, g_ws.hostla.as<double>(),
                      g_ws.hostar.as<double>()));
    CKR(nqs::launch_eloc_from_logamps(H, d_spins, B, N, T, g_ws.hostla.as<double>(),
                                      g_ws.hostar.as<double>(), d_re, d_im, g_stream));
    CK(hipStreamSynchronize(g_stream));
    return 0;
  }
  std::vector<double> la((size_t)B * (T + 1), 0.0), ar((size_t)B * (T + 1), 0.0);
  std::vector<int> scr(N);
  for (int b = 0; b < B; b++) {
    const int* s = h_spins + (size_t)b * N;
    double* l = la.data() + (size_t)b * (T + 1);
    double* r = ar.data() + (size_t)b * (T + 1);
    fn(s, N, user, &l[0], &r[0]);
    g_host_calls++;
    memcpy(scr.data(), s, sizeof(int) * N);
    for (int t = 0; t < T; t++) {
      int i, j = -1;
      if (t < H.n_flip) i = HH.flip_i[t];
      else {
        i = HH.pair_a[t - H.n_flip];
        j = HH.pair_b[t - H.n_flip];
        if (s[i] == s[j]) continue;
      }
      scr[i] = -scr[i];
      if (j >= 0) scr[j] = -scr[j];
      fn(scr.data(), N, user, &l[1 + t], &r[1 + t]);
      g_host_calls++;
      scr[i] = -scr[i];
      if (j >= 0) scr[j] = -scr[j];
    }
  }
  CKR(g_ws.hostla.ensure(sizeof(double) * la.size()));
  CKR(g_ws.hostar.ensure(sizeof(double) * ar.size()));
  CK(hipMemcpyAsync(g_ws.hostla.p, la.data(), sizeof(double) * la.size(), hipMemcpyHostToDevice, g_stream));
  CK(hipMemcpyAsync(g_ws.hostar.p, ar.data(), sizeof(double) * ar.size(), hipMemcpyHostToDevice, g_stream));
  CKR(nqs::launch_eloc_from_logamps(H, d_spins, B, N, T, g_ws.hostla.as<double>(),
                                    g_ws.hostar.as<double>(), d_re, d_im, g_stream));
  CK(hipStreamSynchronize(g_stream));
  return 0;
}

int upload_spins(const int* h, long count) {
  CKR(g_ws.spins.ensure(sizeof(int) * (size_t)count));
  CK(hipMemcpyAsync(g_ws.spins.p, h, sizeof(int) * (size_t)count, hipMemcpyHostToDevice, g_stream));
  return 0;
}

}  // namespace

NQS_EXPORT nqs_ansatz_t* nqs_ansatz_create(const nqs_config_t* cfg, int num_sites) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!cfg || num_sites <= 0) return nullptr;
  const int kind = cfg->ansatz;
  if (kind != NQS_ANSATZ_LEGACY_MLP && kind != NQS_ANSATZ_RBM && kind != NQS_ANSATZ_COMPLEX_RBM)
    return nullptr;
  if (num_sites > nqs::kMaxSites) {
    // every device kernel keeps a configuration as at most 4 64-bit words;
    // refusing here keeps log_amp / local_energy from returning zeros later
    fprintf(stderr, "nqs_hip: nqs_ansatz_create: this library supports N <= %d sites (got %d)\n",
            nqs::kMaxSites, num_sites);
    return nullptr;
  }
  if (stream_ready()) return nullptr;
  nqs_ansatz* a = new nqs_ansatz();
  static long long next_id = 0;
  a->id = ++next_id;
  a->kind = kind;
  a->N = num_sites;
  if (kind == NQS_ANSATZ_LEGACY_MLP) {
    a->M = 0;
    a->Pr = a->P = num_sites;
  } else {
    a->M = cfg->rbm_hidden_units > 0 ? cfg->rbm_hidden_units : 2 * num_sites;
    a->Pr = (long)num_sites + a->M + (long)a->M * num_sites;
    a->P = (kind == NQS_ANSATZ_COMPLEX_RBM) ? 2 * a->Pr : a->Pr;
  }
  a->h.assign(a->P, 0.0);
  if (a->d_par.ensure(sizeof(double) * a->P)) { delete a; return nullptr; }
  const double scale = cfg->rbm_init_scale > 0.0 ? cfg->rbm_init_scale : 0.01;
  const uint64_t seed = cfg->rng_seed ? (uint64_t)cfg->rng_seed : 0xDEADBEEFCAFEBABEULL;
  hipLaunchKernelGGL(k_init_params, dim3(1), dim3(64), 0, g_stream, a->d_par.as<double>(), kind,
                     a->N, a->M, scale, seed);
  nqs::count_launch();
  if (hipGetLastError() != hipSuccess || ansatz_download(a)) {
    a->d_par.release();
    delete a;
    return nullptr;
  }
  return a;
}

NQS_EXPORT void nqs_ansatz_free(nqs_ansatz_t* a) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!a) return;
  (void)hipStreamSynchronize(g_stream);
  a->d_par.release();
  nqs::free_model(a->m64);
  nqs::free_model(a->m32);
  delete a;
}

NQS_EXPORT long nqs_ansatz_num_params(const nqs_ansatz_t* a) { return a ? a->P : 0; }

NQS_EXPORT int nqs_ansatz_is_complex(const nqs_ansatz_t* a) {
  return (a && a->kind == NQS_ANSATZ_COMPLEX_RBM) ? 1 : 0;
}

NQS_EXPORT double* nqs_ansatz_params_raw(nqs_ansatz_t* a) { return a ? a->h.data() : nullptr; }

namespace {
int log_amp_batch(nqs_ansatz* a, const int* h_spins, int B, double* la, double* ar) {
  CKR(ansatz_upload(a));
  CKR(upload_spins(h_spins, (long)B * a->N));
  CKR(g_ws.la.ensure(sizeof(double) * B));
  CKR(g_ws.ar.ensure(sizeof(double) * B));
  CKR(forward_t<double>(a, g_ws.spins.as<int>(), B, g_ws.la.as<double>(), g_ws.ar.as<double>()));
  CK(hipMemcpyAsync(la, g_ws.la.p, sizeof(double) * B, hipMemcpyDeviceToHost, g_stream));
  CK(hipMemcpyAsync(ar, g_ws.ar.p, sizeof(double) * B, hipMemcpyDeviceToHost, g_stream));
  CK(hipStreamSynchronize(g_stream));
  return 0;
}

int gradient_batch(nqs_ansatz* a, const int* h_spins, int B, int mode, double* re, double* im) {
  CKR(ansatz_upload(a));
  CKR(upload_spins(h_spins, (long)B * a->N));
  CKR(forward_t<double>(a, g_ws.spins.as<int>(), B, nullptr, nullptr));
  const size_t ob = sizeof(double) * (size_t)B * a->P;
  CKR(g_ws.O.ensure(ob));
  CKR(g_ws.Oi.ensure(ob));
  ModelView<double> v;
  CKR(model_view<double>(a, v));
  CKR(nqs::launch_gen_O<double>(v, a->kind, g_ws.spins.as<int>(), B, g_ws.t_re.as<double>(),
                                g_ws.t_im.as<double>(), mode, g_ws.O.as<double>(),
                                g_ws.Oi.as<double>(), a->P, g_stream));
  CK(hipMemcpyAsync(re, g_ws.O.p, ob, hipMemcpyDeviceToHost, g_stream));
  if (mode == 1) CK(hipMemcpyAsync(im, g_ws.Oi.p, ob, hipMemcpyDeviceToHost, g_stream));
  CK(hipStreamSynchronize(g_stream));
  return 0;
}
}  // namespace

NQS_EXPORT void nqs_ansatz_log_amp(const int* spins, int num_sites, void* user, double* out_log_abs,
                                   double* out_arg) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  nqs_ansatz* a = (nqs_ansatz*)user;
  double la = 0.0, ar = 0.0;
  if (a && spins && num_sites == a->N) {
    if (log_amp_batch(a, spins, 1, &la, &ar) != 0) {
      la = 0.0;
      ar = 0.0;
    }
  }
  if (out_log_abs) *out_log_abs = la;
  if (out_arg) *out_arg = ar;
}

NQS_EXPORT int nqs_ansatz_logpsi_gradient(nqs_ansatz_t* a, const int* spins, int num_sites,
                                          double* out_grad) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!a || !spins || !out_grad || num_sites != a->N) return -1;
  return gradient_batch(a, spins, 1, 0, out_grad, nullptr);
}

NQS_EXPORT int nqs_ansatz_logpsi_gradient_complex(nqs_ansatz_t* a, const int* spins, int num_sites,
                                                  double* out_re, double* out_im) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!a || !spins || !out_re || !out_im || num_sites != a->N) return -1;
  return gradient_batch(a, spins, 1, 1, out_re, out_im);
}

NQS_EXPORT int nqs_ansatz_apply_update(nqs_ansatz_t* a, const double* delta, double step) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!a || !delta) return -1;
  if (ansatz_upload(a)) return -1;
  if (g_ws.vecs.ensure(sizeof(double) * a->P)) return -1;
  if (hipMemcpyAsync(g_ws.vecs.p, delta, sizeof(double) * a->P, hipMemcpyHostToDevice, g_stream) !=
      hipSuccess)
    return -1;
  if (nqs::launch_axpy(a->d_par.as<double>(), g_ws.vecs.as<double>(), step, a->P, g_stream))
    return -1;
  a->m64.valid = 0;
  a->m32.valid = 0;
  return ansatz_download(a);
}

NQS_EXPORT int nqs_hip_log_amp_batch(nqs_ansatz_t* a, const int* spins, int B, double* la,
                                     double* ar) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!a || !spins || !la || !ar || B <= 0) return -1;
  return log_amp_batch(a, spins, B, la, ar);
}

NQS_EXPORT int nqs_hip_logpsi_gradient_batch(nqs_ansatz_t* a, const int* spins, int B, int mode,
                                             double* re, double* im) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!a || !spins || !re || B <= 0 || (mode == 1 && !im)) return -1;
  return gradient_batch(a, spins, B, mode, re, im);
}

// ===========================================================================
// Marshall
// ===========================================================================
NQS_EXPORT int nqs_marshall_parity(const int* spins, int size_x, int size_y) {
  // nqs_marshall.c:16-27: parity of the down-spin count on (x+y) even sites
  int n = 0;
  for (int x = 0; x < size_x; x++)
    for (int y = 0; y < size_y; y++)
      if (((x + y) & 1) == 0 && spins[x * size_y + y] < 0) n++;
  return n & 1;
}

NQS_EXPORT void nqs_marshall_log_amp(const int* spins, int num_sites, void* user, double* out_log_abs,
                                     double* out_arg) {
  nqs_marshall_wrapper_t* w = (nqs_marshall_wrapper_t*)user;
  double la = 0.0, ar = 0.0;
  w->base_log_amp(spins, num_sites, w->base_user, &la, &ar);
  const int par = nqs_marshall_parity(spins, w->size_x, w->size_y);
  if (out_log_abs) *out_log_abs = la;
  if (out_arg) *out_arg = ar + (par ? 3.14159265358979323846 : 0.0);
}

// ===========================================================================
// sampler
// ===========================================================================
// ---------------------------------------------------------------------------
// learned proposal network (kernels k_ar.hip, layout ar_layout.h)
// ---------------------------------------------------------------------------
struct nqs_hip_ar {
  int N = 0;
  nqs::ar::Dims d;
  DBuf w, m, v, op1f, op2f, op2b, op3f, op3b, w1t, b1, b2, b3;
  DBuf partial, loss, bits, lq, rng, spins, grad;   // scratch
  nqs::ArDev dev;
  long t = 0;                                        // Adam steps taken
  uint64_t shuffle = 0x243F6A8885A308D3ULL;          // epoch offsets
  long loss_n = 0, loss_avail = 0;                   // last training's loss partials
  // a training launched on the side stream after a sampler's batch (its
  // end event; pending until the library's stream has waited for it)
  hipEvent_t ev_trained = nullptr;
  int pending = 0;
  ~nqs_hip_ar() {
    if (ev_trained) {
      (void)hipEventSynchronize(ev_trained);
      (void)hipEventDestroy(ev_trained);
    }
    DBuf* all[] = {&w, &m, &v, &op1f, &op2f, &op2b, &op3f, &op3b, &w1t, &b1, &b2, &b3,
                   &partial, &loss, &bits, &lq, &rng, &spins, &grad};
    for (DBuf* b : all) b->release();
  }
};

struct nqs_hip_multi;
void multi_delete(nqs_hip_multi* m);

struct nqs_sampler {
  int N = 0;
  nqs_config_t cfg;
  nqs_hip_options_t opt;
  nqs_log_amp_fn_t fn = nullptr;
  void* user = nullptr;
  int device = 0;
  AmpSource src;
  // device chains
  nqs::ChainState cs;
  DBuf d_spins, d_rng, d_acc, d_tot, d_la, d_cur, d_samples;
  // host-callback chains
  std::vector<int> hc, hp;
  std::vector<double> hla;
  std::vector<uint64_t> hrng;
  long long hacc = 0, htot = 0;
  std::vector<int> h_cur;  // chain 0 configuration returned to callers
  int cur_dirty = 0;       // h_cur is behind the device chains
  // last SR step bookkeeping
  std::vector<double> last_E;
  int last_rounds = 0;
  double stage_ms[7] = {0, 0, 0, 0, 0, 0, 0};
  int have_stage = 0;
  // learned proposal: network, training window (bit rows, ring), proposal
  // buffers and counters
  nqs_hip_ar* ar = nullptr;
  DBuf win, lacc, prop_bits, prop_spins, lq, la_prop;
  long win_pos = 0, win_n = 0;
  bool win_written = false;   // the last learned batch wrote its window entries
  int since_train = 0, trained = 0, n_train = 0;
  long long ar_props = 0;
  // CG-SR steps of one real RBM run as a batch of one problem (one graph
  // per step); rebuilt when the ansatz, the step's parameters or the sample
  // buffer change
  nqs_hip_multi* solo = nullptr;
  long long solo_ansatz = 0;
  nqs_config_t solo_cfg;
  int solo_Lx = 0, solo_Ly = 0;
  void* solo_samples = nullptr;
  ~nqs_sampler() {
    multi_delete(solo);
    // a training launched ahead reads the window: it ends first
    if (ar && ar->pending && ar->ev_trained) (void)hipEventSynchronize(ar->ev_trained);
    DBuf* all[] = {&d_spins, &d_rng, &d_acc, &d_tot, &d_la, &d_cur, &d_samples,
                   &win, &lacc, &prop_bits, &prop_spins, &lq, &la_prop};
    for (DBuf* b : all) b->release();
    delete ar;
  }
};

namespace {

uint64_t chain_seed(unsigned rng_seed, int c) {
  const uint64_t s0 = rng_seed ? (uint64_t)rng_seed : 0x9E3779B97F4A7C15ULL;
  if (c == 0) return s0;
  uint64_t s = nqs::splitmix64(s0 ^ ((uint64_t)c * 0xD1B54A32D192ED03ULL));
  return s ? s : 1ULL;
}

// ---- learned proposal: host side ------------------------------------------
// Initial weights: one xorshift64 stream over the parameter vector in layout
// order (a draw per entry); W1 entries scale (2u - 1) / sqrt(N), W2 and W3
// entries scale (2u - 1) / sqrt(64), biases 0, masked entries 0.
void ar_init_host(int N, uint64_t seed, double scale, std::vector<float>& w) {
  const nqs::ar::Dims d = nqs::ar::dims(N);
  w.assign(d.P, 0.f);
  uint64_t st = nqs::splitmix64(seed ^ 0xA17E5EEDULL);
  if (!st) st = 1;
  for (long p = 0; p < d.P; p++) {
    const double u = nqs::xs_uniform(st);
    double x = 0.0;
    if (p < d.ob1) x = scale * (2.0 * u - 1.0) / sqrt((double)N);
    else if ((p >= d.oW2 && p < d.ob2) || (p >= d.oW3 && p < d.ob3))
      x = scale * (2.0 * u - 1.0) / 8.0;
    w[p] = nqs::ar::free_param(d, p) ? (float)x : 0.f;
  }
}

int ar_set(nqs_hip_ar* a, const float* host_w) {
  CK(hipMemcpyAsync(a->w.p, host_w, sizeof(float) * a->d.P, hipMemcpyHostToDevice, g_stream));
  return nqs::launch_ar_pack(a->dev, g_stream);
}

nqs_hip_ar* ar_new(int N, uint64_t seed, double scale) {
  if (N < 2 || N > nqs::ar::kMaxN) {
    fprintf(stderr, "nqs_hip: the learned proposal supports 2 <= N <= %d sites\n", nqs::ar::kMaxN);
    return nullptr;
  }
  if (stream_ready()) return nullptr;
  nqs_hip_ar* a = new nqs_hip_ar();
  a->N = N;
  a->d = nqs::ar::dims(N);
  const nqs::ar::Dims& d = a->d;
  struct Z { DBuf* b; size_t bytes; } z[] = {
      {&a->w, 4 * (size_t)d.P},       {&a->m, 4 * (size_t)d.P},      {&a->v, 4 * (size_t)d.P},
      {&a->op1f, 2 * (size_t)d.n_op1f}, {&a->op2f, 2 * (size_t)d.n_op2}, {&a->op2b, 2 * (size_t)d.n_op2},
      {&a->op3f, 2 * (size_t)d.n_op3f}, {&a->op3b, 2 * (size_t)d.n_op3b},
      {&a->w1t, 4 * (size_t)N * nqs::ar::H}, {&a->b1, 4 * (size_t)nqs::ar::H},
      {&a->b2, 4 * (size_t)nqs::ar::H}, {&a->b3, 4 * (size_t)16 * d.nt}};
  for (const Z& e : z) {
    if (e.b->ensure(e.bytes) || hipMemsetAsync(e.b->p, 0, e.bytes, g_stream) != hipSuccess) {
      delete a;
      return nullptr;
    }
  }
  nqs::ArDev& v = a->dev;
  v.N = N;
  v.w = a->w.as<float>();
  v.m = a->m.as<float>();
  v.v = a->v.as<float>();
  v.op1f = a->op1f.as<uint16_t>();
  v.op2f = a->op2f.as<uint16_t>();
  v.op2b = a->op2b.as<uint16_t>();
  v.op3f = a->op3f.as<uint16_t>();
  v.op3b = a->op3b.as<uint16_t>();
  v.w1t = a->w1t.as<float>();
  v.b1 = a->b1.as<float>();
  v.b2 = a->b2.as<float>();
  v.b3 = a->b3.as<float>();
  a->shuffle ^= seed;
  if (!a->shuffle) a->shuffle = 1;
  std::vector<float> hw;
  ar_init_host(N, seed, scale, hw);
  if (ar_set(a, hw.data()) || hipStreamSynchronize(g_stream) != hipSuccess) {
    delete a;
    return nullptr;
  }
  return a;
}

// mean loss of the last epoch of the last training (read back on demand)
int ar_last_loss(nqs_hip_ar* a, double* mean_loss) {
  if (a->loss_n <= 0) {
    *mean_loss = 0.0;
    return 0;
  }
  std::vector<double> l((size_t)a->loss_n);
  CK(hipMemcpyAsync(l.data(), a->loss.p, sizeof(double) * l.size(), hipMemcpyDeviceToHost, g_stream));
  CK(hipStreamSynchronize(g_stream));
  double tot = 0.0;
  for (double x : l) tot += x;
  *mean_loss = tot / (double)a->loss_avail;
  return 0;
}

// the library's stream waits for a training still running on the side
// stream (every use of the network's device state comes after this)
int ar_join(nqs_hip_ar* a) {
  if (a && a->pending) {
    CK(hipStreamWaitEvent(g_stream, a->ev_trained, 0));
    a->pending = 0;
  }
  return 0;
}

// Adam epochs over n_avail device bit rows, on stream strm; mean loss of the
// last epoch into *mean_loss when it is given (a host synchronisation).
int ar_train(nqs_hip_ar* a, const uint64_t* rows, long n_avail, int constrained, int epochs,
             int batch, double lr, double* mean_loss, hipStream_t strm = nullptr) {
  if (!strm) strm = g_stream;
  CKR(ar_join(a));
  if (n_avail <= 0 || epochs <= 0) return 0;
  if (batch < 1) batch = 1;
  const int nb = nqs::ar_grad_blocks(batch);
  const long steps = (n_avail + batch - 1) / batch;
  CKR(a->partial.ensure(sizeof(float) * (size_t)nb * a->d.P));
  CKR(a->loss.ensure(sizeof(double) * (size_t)nb * steps));
  for (int e = 0; e < epochs; e++) {
    if (e == epochs - 1) CKR(nqs::launch_fill(a->loss.as<double>(), 0.0, (long)nb * steps, strm));
    const long off = (long)(nqs::xs_next(a->shuffle) % (uint64_t)n_avail);
    for (long st = 0; st < steps; st++) {
      const int count = (int)std::min<long>(batch, n_avail - st * batch);
      nqs::ArGrad g;
      g.rows = rows;
      g.n_avail = n_avail;
      g.start = (off + st * batch) % n_avail;
      g.count = count;
      g.constrained = constrained;
      g.scale = 1.0f / (float)count;
      g.partial = a->partial.as<float>();
      g.loss = a->loss.as<double>() + st * nb;
      CKR(nqs::launch_ar_grad(a->dev, g, strm));
      a->t++;
      nqs::ArAdam ad;
      ad.partial = a->partial.as<float>();
      ad.nblk = nqs::ar_grad_blocks(count);
      ad.lr = (float)lr;
      ad.bc1 = (float)(1.0 - pow(0.9, (double)a->t));
      ad.bc2 = (float)(1.0 - pow(0.999, (double)a->t));
      CKR(nqs::launch_ar_adam(a->dev, ad, strm));
    }
  }
  a->loss_n = (long)nb * steps;
  a->loss_avail = n_avail;
  if (mean_loss) {
    if (strm != g_stream) return -1;
    CKR(ar_last_loss(a, mean_loss));
  }
  return 0;
}

// the side stream of ahead-of-time trainings (default priority)
hipStream_t g_train_stream = nullptr;

// NQS_HIP_AR_OVERLAP=0: a due retraining runs at the start of the batch
// that needs it instead of beside the work that follows the previous batch
bool ar_overlap_on() {
  const char* e = getenv("NQS_HIP_AR_OVERLAP");
  return !(e && e[0] == '0');
}

// retrain before a batch when due: first once the window holds samples,
// then every ar_train_every batches
int ar_maybe_train(nqs_sampler* s) {
  const int every = std::max(1, s->opt.ar_train_every);
  const bool due = s->trained ? (s->since_train >= every) : (s->win_n > 0);
  if (!due) return 0;
  CKR(ar_train(s->ar, s->win.as<uint64_t>(), s->win_n, 1, s->opt.ar_epochs, s->opt.ar_batch,
               s->opt.ar_lr, nullptr));
  s->trained = 1;
  s->since_train = 0;
  s->n_train++;
  return 0;
}

int ar_window_append(nqs_sampler* s, int n_out, const int* d_out) {
  const long cap = std::max(1, s->opt.ar_train_window);
  const long k = std::min<long>(n_out, cap);
  if (!s->win_written)
    CKR(nqs::launch_bits_append(d_out + (size_t)(n_out - k) * s->N, k, s->N,
                                s->win.as<uint64_t>(), cap, s->win_pos, g_stream));
  s->win_written = false;
  s->win_pos = (s->win_pos + k) % cap;
  s->win_n = std::min(cap, s->win_n + k);
  s->since_train++;
  return 0;
}

// The retraining that the next batch would run at its start (same window,
// same network, same Adam steps: the same result) starts now on the side
// stream, beside whatever the caller does before that batch (the SR solve);
// the next use of the network waits for it (ar_join).
int ar_train_ahead(nqs_sampler* s) {
  const int every = std::max(1, s->opt.ar_train_every);
  const bool due = s->trained ? (s->since_train >= every) : (s->win_n > 0);
  if (!due) return 0;
  nqs_hip_ar* a = s->ar;
  if (!g_train_stream) CK(hipStreamCreateWithFlags(&g_train_stream, hipStreamNonBlocking));
  if (!a->ev_trained) CK(hipEventCreateWithFlags(&a->ev_trained, hipEventDisableTiming));
  CKR(ar_join(a));
  // the window as this batch left it
  CK(hipEventRecord(a->ev_trained, g_stream));
  CK(hipStreamWaitEvent(g_train_stream, a->ev_trained, 0));
  CKR(ar_train(a, s->win.as<uint64_t>(), s->win_n, 1, s->opt.ar_epochs, s->opt.ar_batch,
               s->opt.ar_lr, nullptr, g_train_stream));
  CK(hipEventRecord(a->ev_trained, g_train_stream));
  a->pending = 1;
  s->trained = 1;
  s->since_train = 0;
  s->n_train++;
  return 0;
}

int sampler_sync_cur(nqs_sampler* s) {
  s->cur_dirty = 0;
  if (!s->device) {
    memcpy(s->h_cur.data(), s->hc.data(), sizeof(int) * s->N);
    return 0;
  }
  CKR(nqs::launch_unpack(s->cs, 0, s->d_cur.as<int>(), g_stream));
  CK(hipMemcpyAsync(s->h_cur.data(), s->d_cur.p, sizeof(int) * s->N, hipMemcpyDeviceToHost,
                    g_stream));
  CK(hipStreamSynchronize(g_stream));
  return 0;
}

// one move of host chain c (nqs_sampler.c:86-117 and the exchange extension)
void host_step(nqs_sampler* s, int c) {
  const int N = s->N;
  int* x = s->hc.data() + (size_t)c * N;
  uint64_t& st = s->hrng[c];
  memcpy(s->hp.data(), x, sizeof(int) * N);
  if (s->opt.move == NQS_HIP_MOVE_EXCHANGE) {
    int nu = 0;
    for (int i = 0; i < N; i++) nu += (x[i] > 0);
    const int nd = N - nu;
    if (nu == 0 || nd == 0) { s->htot++; return; }
    const int k = (int)(nqs::xs_next(st) % (uint64_t)nu);
    const int l = (int)(nqs::xs_next(st) % (uint64_t)nd);
    int iu = -1, id = -1;
    for (int i = 0, cu = 0, cd = 0; i < N; i++) {
      if (x[i] > 0) { if (cu++ == k) iu = i; }
      else { if (cd++ == l) id = i; }
    }
    s->hp[iu] = -s->hp[iu];
    s->hp[id] = -s->hp[id];
  } else {
    const int site = (int)(nqs::xs_next(st) % (uint64_t)N);
    s->hp[site] = -s->hp[site];
  }
  double pla = 0.0, par = 0.0;
  s->fn(s->hp.data(), N, s->user, &pla, &par);
  g_host_calls++;
  const double lr = 2.0 * (pla - s->hla[c]);
  bool acc = (lr >= 0.0) ? true : (log(nqs::xs_uniform(st)) < lr);
  s->htot++;
  if (acc) {
    s->hacc++;
    memcpy(x, s->hp.data(), sizeof(int) * N);
    s->hla[c] = pla;
  }
}

int ensure_model_for_sampler(nqs_sampler* s) {
  CKR(ansatz_upload(s->src.a));
  return 0;
}

// rounds of the learned proposal: without local moves all rounds run in
// one launch; with opt.ar_local > 0 each round is ar_local exchange moves
// followed by a one-round learned launch
template <typename R>
int learned_rounds(nqs_sampler* s, int rounds, int n_out, int* d_out) {
  const int C = s->cs.C;
  ModelView<R> mv;
  if (s->opt.ar_local > 0) CKR(ratio_view<R>(s->src.a, mv));   // exchange moves between
  else CKR(model_view<R>(s->src.a, mv, 0));   // amplitudes only: no wait for the range verdict
  nqs::ArRounds ar;
  ar.constrained = 1;
  ar.lacc = s->lacc.as<unsigned long long>();
  if (s->opt.ar_local <= 0) {
    // proposals of every round at once, their amplitudes in one forward
    // pass, then the per-chain accept / reject rounds
    const size_t nprop = (size_t)rounds * C;
    CKR(s->prop_bits.ensure(sizeof(uint64_t) * nprop));
    CKR(s->prop_spins.ensure(sizeof(int) * nprop * s->N));
    CKR(s->lq.ensure(sizeof(double) * (nprop + C)));
    CKR(s->la_prop.ensure(sizeof(double) * nprop));
    nqs::ArBulk b;
    b.rounds = rounds;
    b.prop = s->prop_bits.as<uint64_t>();
    b.spins = s->prop_spins.as<int>();
    b.lq = s->lq.as<double>();
    b.lp = s->la_prop.as<double>();
    b.n_out = n_out;
    b.out = d_out;
    b.lacc = s->lacc.as<unsigned long long>();
    // the accept rounds also write the training window's entries (what
    // ar_window_append would pack from the samples)
    if (s->win.p && n_out > 0) {
      b.win = s->win.as<uint64_t>();
      b.win_cap = std::max(1, s->opt.ar_train_window);
      b.win_pos = s->win_pos;
      b.win_first = n_out - std::min<long>(n_out, b.win_cap);
      s->win_written = true;
    }
    CKR(nqs::launch_ar_propose(s->ar->dev, s->cs, b, g_stream));
    // log|psi| only (no tanh theta)
    CKR(nqs::launch_forward<R>(mv, b.spins, (int)nprop, nullptr, nullptr, b.lp, nullptr, g_stream));
    CKR(nqs::launch_ar_accept_rounds(s->cs, b, g_stream));
  } else {
    nqs::MHLaunch ml;
    ml.move = 1;
    ml.steps = s->opt.ar_local;
    ml.rounds = 1;
    ar.rounds = 1;
    for (int r = 0; r < rounds; r++) {
      CKR(nqs::launch_mh<R>(mv, s->cs, ml, g_stream));
      const long first = (long)r * C;
      ar.out = first < n_out ? d_out + first * s->N : nullptr;
      ar.n_out = std::max(0L, (long)n_out - first);
      CKR(nqs::launch_ar_mh<R>(s->ar->dev, mv, s->cs, ar, g_stream));
    }
  }
  s->ar_props += (long long)rounds * C;
  return 0;
}

// run `steps` moves per round for `rounds` rounds on every chain, optionally
// recording samples r*C + c < n_out into d_out (device) / h_out (host).
// A learned-proposal sampler trains when due before a recorded batch, runs
// learned rounds once trained, and appends the batch to its window.
int sampler_advance(nqs_sampler* s, int steps, int rounds, int n_out, int* d_out, int* h_out) {
  if (rounds <= 0) return 0;
  if (s->device) {
    CKR(ensure_model_for_sampler(s));
    const bool learned_batch = s->ar && d_out;
    if (learned_batch) CKR(ar_maybe_train(s));
    if (s->ar) CKR(ar_join(s->ar));
    if (learned_batch && s->trained) {
      if (low_prec(s->opt.precision)) CKR(learned_rounds<float>(s, rounds, n_out, d_out));
      else CKR(learned_rounds<double>(s, rounds, n_out, d_out));
    } else {
      nqs::MHLaunch ml;
      ml.move = s->opt.move == NQS_HIP_MOVE_FLIP ? 0 : 1;
      ml.steps = steps;
      ml.rounds = rounds;
      ml.record = d_out ? 1 : 0;
      ml.n_out = n_out;
      ml.out = d_out;
      if (low_prec(s->opt.precision)) {
        ModelView<float> v;
        CKR(ratio_view<float>(s->src.a, v));
        CKR(nqs::launch_mh<float>(v, s->cs, ml, g_stream));
      } else {
        ModelView<double> v;
        CKR(ratio_view<double>(s->src.a, v));
        CKR(nqs::launch_mh<double>(v, s->cs, ml, g_stream));
      }
    }
    if (learned_batch) {
      CKR(ar_window_append(s, n_out, d_out));
      if (ar_overlap_on()) CKR(ar_train_ahead(s));
    }
    if (h_out && d_out) {
      CK(hipMemcpyAsync(h_out, d_out, sizeof(int) * (size_t)n_out * s->N, hipMemcpyDeviceToHost,
                        g_stream));
      return sampler_sync_cur(s);
    }
    // chain 0's configuration is fetched when a caller asks for it
    s->cur_dirty = 1;
    return 0;
  }
  const int C = s->opt.num_chains;
  for (int r = 0; r < rounds; r++) {
    for (int c = 0; c < C; c++)
      for (int k = 0; k < steps; k++) host_step(s, c);
    if (h_out) {
      for (int c = 0; c < C; c++) {
        const long b = (long)r * C + c;
        if (b < n_out) memcpy(h_out + b * s->N, s->hc.data() + (size_t)c * s->N, sizeof(int) * s->N);
      }
    }
  }
  if (d_out && h_out)
    CK(hipMemcpyAsync(d_out, h_out, sizeof(int) * (size_t)n_out * s->N, hipMemcpyHostToDevice,
                      g_stream));
  return sampler_sync_cur(s);
}

// log|psi| of every chain with the current parameters (the value the
// reference computes at sampler creation, nqs_sampler.c:73-74)
int chains_eval_la(nqs_sampler* s) {
  DBuf rows;
  CKR(rows.ensure(sizeof(int) * (size_t)s->cs.C * s->N));
  CKR(nqs::launch_unpack_all(s->cs, rows.as<int>(), g_stream));
  CKR(ansatz_upload(s->src.a));
  CKR(forward_t<double>(s->src.a, rows.as<int>(), s->cs.C, s->cs.la, nullptr));
  CK(hipStreamSynchronize(g_stream));
  rows.release();
  return 0;
}

nqs_sampler* sampler_new(int N, const nqs_config_t* cfg, const nqs_hip_options_t* opt,
                         nqs_log_amp_fn_t fn, void* user) {
  if (N <= 0 || !cfg || !fn) return nullptr;
  if (stream_ready()) return nullptr;
  nqs_sampler* s = new nqs_sampler();
  s->N = N;
  s->cfg = *cfg;
  s->opt = opt ? *opt : nqs_hip_options_defaults();
  if (!opt) {
    const char* env = getenv("NQS_HIP_CHAINS");
    if (env && atoi(env) > 0) s->opt.num_chains = atoi(env);
  }
  if (s->opt.num_chains < 1) s->opt.num_chains = 1;
  s->fn = fn;
  s->user = user;
  s->h_cur.assign(N, 1);
  const int C = s->opt.num_chains;
  s->device = recognise(fn, user, s->src) ? 1 : 0;
  if (s->device && s->src.a->N != N) s->device = 0;
  if (s->device && (N > nqs::kMaxSites || s->src.a->M > nqs::kMaxHidden)) s->device = 0;
  if (s->device) {
    const int nw = (N + 63) / 64;
    if (s->d_spins.ensure(sizeof(uint64_t) * (size_t)C * nw) ||
        s->d_rng.ensure(sizeof(uint64_t) * C) || s->d_acc.ensure(sizeof(uint64_t) * C) ||
        s->d_tot.ensure(sizeof(uint64_t) * C) || s->d_la.ensure(sizeof(double) * C) ||
        s->d_cur.ensure(sizeof(int) * N)) {
      delete s;
      return nullptr;
    }
    s->cs.C = C;
    s->cs.N = N;
    s->cs.nw = nw;
    s->cs.spins = s->d_spins.as<uint64_t>();
    s->cs.rng = s->d_rng.as<uint64_t>();
    s->cs.acc = s->d_acc.as<unsigned long long>();
    s->cs.tot = s->d_tot.as<unsigned long long>();
    s->cs.la = s->d_la.as<double>();
    std::vector<uint64_t> seeds(C);
    for (int c = 0; c < C; c++) seeds[c] = chain_seed(cfg->rng_seed, c);
    DBuf tmp;
    if (tmp.ensure(sizeof(uint64_t) * C) ||
        hipMemcpyAsync(tmp.p, seeds.data(), sizeof(uint64_t) * C, hipMemcpyHostToDevice, g_stream) ||
        nqs::launch_chain_init(s->cs, tmp.as<uint64_t>(), s->opt.sz_zero_start, g_stream) ||
        sampler_sync_cur(s) || chains_eval_la(s)) {
      delete s;
      return nullptr;
    }
    (void)hipStreamSynchronize(g_stream);
    tmp.release();
    if (s->opt.move == NQS_HIP_MOVE_LEARNED) {
      const uint64_t seed = s->opt.ar_seed ? (uint64_t)s->opt.ar_seed : chain_seed(cfg->rng_seed, 0);
      s->ar = ar_new(N, seed, 1.0);
      const size_t cap = (size_t)std::max(1, s->opt.ar_train_window);
      if (!s->ar || s->win.ensure(sizeof(uint64_t) * cap) || s->lacc.ensure(sizeof(uint64_t) * C) ||
          hipMemsetAsync(s->lacc.p, 0, sizeof(uint64_t) * C, g_stream) != hipSuccess ||
          hipStreamSynchronize(g_stream) != hipSuccess) {
        delete s;
        return nullptr;
      }
    }
    return s;
  }
  if (s->opt.move == NQS_HIP_MOVE_LEARNED) {
    fprintf(stderr, "nqs_hip: the learned proposal needs a device amplitude (nqs_ansatz_log_amp "
                    "or nqs_marshall_log_amp over it)\n");
    delete s;
    return nullptr;
  }
  // host-callback chains: the reference algorithm against the caller's function
  s->hc.assign((size_t)C * N, 1);
  s->hp.assign(N, 1);
  s->hla.assign(C, 0.0);
  s->hrng.assign(C, 0);
  for (int c = 0; c < C; c++) {
    uint64_t st = chain_seed(cfg->rng_seed, c);
    int* x = s->hc.data() + (size_t)c * N;
    if (s->opt.sz_zero_start) {
      for (int i = 0; i < N; i++) x[i] = (i < N / 2) ? 1 : -1;
      for (int i = N - 1; i > 0; i--) {
        const int j = (int)(nqs::xs_next(st) % (uint64_t)(i + 1));
        std::swap(x[i], x[j]);
      }
    } else {
      for (int i = 0; i < N; i++) x[i] = (nqs::xs_next(st) & 1) ? 1 : -1;
    }
    s->hrng[c] = st;
    double ar = 0.0;
    fn(x, N, user, &s->hla[c], &ar);
    g_host_calls++;
  }
  memcpy(s->h_cur.data(), s->hc.data(), sizeof(int) * N);
  return s;
}

}  // namespace

NQS_EXPORT nqs_hip_options_t nqs_hip_options_defaults(void) {
  nqs_hip_options_t o;
  memset(&o, 0, sizeof(o));
  o.num_chains = 1;
  o.move = NQS_HIP_MOVE_FLIP;
  o.sz_zero_start = 0;
  o.periodic = 0;
  o.precision = NQS_HIP_FP64;
  o.gram = NQS_HIP_GRAM_DIRECT;
  o.ar_hidden = 64;
  o.ar_train_every = 10;
  o.ar_train_window = 16384;
  o.ar_epochs = 2;
  o.ar_batch = 2048;
  o.ar_lr = 3e-3;
  o.ar_local = 0;
  o.ar_seed = 0;
  return o;
}

NQS_EXPORT nqs_sampler_t* nqs_sampler_create(int num_sites, const nqs_config_t* cfg,
                                             nqs_log_amp_fn_t log_amp, void* log_amp_user) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  return sampler_new(num_sites, cfg, nullptr, log_amp, log_amp_user);
}

NQS_EXPORT nqs_sampler_t* nqs_hip_sampler_create_ex(int num_sites, const nqs_config_t* cfg,
                                                    const nqs_hip_options_t* opt,
                                                    nqs_log_amp_fn_t log_amp, void* user) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!opt) return nullptr;
  if (opt->move == NQS_HIP_MOVE_LEARNED) {
    if (num_sites < 2 || num_sites > nqs::ar::kMaxN ||
        (opt->ar_hidden != 0 && opt->ar_hidden != nqs::ar::H) || opt->ar_local < 0) {
      fprintf(stderr, "nqs_hip: learned proposal needs 2 <= N <= %d, ar_hidden = %d and "
                      "ar_local >= 0\n", nqs::ar::kMaxN, nqs::ar::H);
      return nullptr;
    }
  }
  return sampler_new(num_sites, cfg, opt, log_amp, user);
}

NQS_EXPORT int nqs_hip_sampler_get_options(const nqs_sampler_t* s, nqs_hip_options_t* out) {
  if (!s || !out) return -1;
  *out = s->opt;
  return 0;
}

NQS_EXPORT void nqs_sampler_free(nqs_sampler_t* s) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!s) return;
  (void)hipStreamSynchronize(g_stream);
  delete s;
}

NQS_EXPORT void nqs_sampler_thermalize(nqs_sampler_t* s) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!s) return;
  (void)sampler_advance(s, s->cfg.num_thermalize, s->cfg.num_thermalize > 0 ? 1 : 0, 0, nullptr,
                        nullptr);
}

NQS_EXPORT const int* nqs_sampler_next(nqs_sampler_t* s) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!s) return nullptr;
  const int k = s->cfg.num_decorrelate > 0 ? s->cfg.num_decorrelate : 1;
  if (sampler_advance(s, k, 1, 0, nullptr, nullptr) || sampler_sync_cur(s)) return nullptr;
  return s->h_cur.data();
}

NQS_EXPORT int nqs_sampler_batch(nqs_sampler_t* s, int n, int* out) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!s || !out || n <= 0) return -1;
  const int C = s->opt.num_chains;
  const int k = s->cfg.num_decorrelate > 0 ? s->cfg.num_decorrelate : 1;
  const int rounds = (n + C - 1) / C;
  if (s->d_samples.ensure(sizeof(int) * (size_t)n * s->N)) return -1;
  return sampler_advance(s, k, rounds, n, s->d_samples.as<int>(), out);
}

NQS_EXPORT double nqs_sampler_acceptance_ratio(const nqs_sampler_t* s) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!s) return 0.0;
  if (!s->device) return s->htot ? (double)s->hacc / (double)s->htot : 0.0;
  const int C = s->cs.C;
  std::vector<unsigned long long> a(C), t(C);
  if (hipMemcpyAsync(a.data(), s->cs.acc, sizeof(uint64_t) * C, hipMemcpyDeviceToHost, g_stream) ||
      hipMemcpyAsync(t.data(), s->cs.tot, sizeof(uint64_t) * C, hipMemcpyDeviceToHost, g_stream) ||
      hipStreamSynchronize(g_stream))
    return 0.0;
  unsigned long long sa = 0, stt = 0;
  for (int c = 0; c < C; c++) { sa += a[c]; stt += t[c]; }
  return stt ? (double)sa / (double)stt : 0.0;
}

NQS_EXPORT const int* nqs_sampler_current(const nqs_sampler_t* s) {
  if (!s) return nullptr;
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  nqs_sampler* m = const_cast<nqs_sampler*>(s);
  if (m->cur_dirty && sampler_sync_cur(m)) return nullptr;
  return m->h_cur.data();
}

NQS_EXPORT int nqs_sampler_num_sites(const nqs_sampler_t* s) { return s ? s->N : 0; }

// ===========================================================================
// local energy
// ===========================================================================
namespace {

int local_energy_any(const nqs_config_t* cfg, int Lx, int Ly, int per, int prec, const int* spins,
                     int B, nqs_log_amp_fn_t fn, void* user, double* re, double* im) {
  AmpSource src;
  const int N = nqs::ham_sites(cfg, Lx, Ly);
  CKR(upload_spins(spins, (long)B * N));
  CKR(g_ws.e_re.ensure(sizeof(double) * B));
  CKR(g_ws.e_im.ensure(sizeof(double) * B));
  if (recognise(fn, user, src) && src.a->N == N) {
    CKR(ansatz_upload(src.a));
    CKR(eloc_device(cfg, Lx, Ly, per, prec, src, g_ws.spins.as<int>(), B, g_ws.e_re.as<double>(),
                    g_ws.e_im.as<double>()));
  } else {
    CKR(eloc_callback(cfg, Lx, Ly, per, spins, B, fn, user, g_ws.e_re.as<double>(),
                      g_ws.e_im.as<double>(), g_ws.spins.as<int>()));
  }
  if (re) CK(hipMemcpyAsync(re, g_ws.e_re.p, sizeof(double) * B, hipMemcpyDeviceToHost, g_stream));
  if (im) CK(hipMemcpyAsync(im, g_ws.e_im.p, sizeof(double) * B, hipMemcpyDeviceToHost, g_stream));
  CK(hipStreamSynchronize(g_stream));
  return 0;
}

}  // namespace

NQS_EXPORT double nqs_local_energy(const nqs_config_t* cfg, int size_x, int size_y, const int* spins,
                                   nqs_log_amp_fn_t log_amp, void* log_amp_user) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!cfg || !spins || !log_amp) return 0.0;
  if (stream_ready()) return 0.0;
  double re = 0.0;
  if (local_energy_any(cfg, size_x, size_y, 0, NQS_HIP_FP64, spins, 1, log_amp, log_amp_user, &re,
                       nullptr))
    return 0.0;
  return re;
}

NQS_EXPORT void nqs_local_energy_batch(const nqs_config_t* cfg, int size_x, int size_y,
                                       const int* spins_batch, int batch_size,
                                       nqs_log_amp_fn_t log_amp, void* log_amp_user,
                                       double* out_energies) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!cfg || !spins_batch || !out_energies || batch_size <= 0 || !log_amp) return;
  if (stream_ready()) return;
  if (local_energy_any(cfg, size_x, size_y, 0, NQS_HIP_FP64, spins_batch, batch_size, log_amp,
                       log_amp_user, out_energies, nullptr))
    for (int i = 0; i < batch_size; i++) out_energies[i] = 0.0;
}

NQS_EXPORT void nqs_local_energy_complex(const nqs_config_t* cfg, int size_x, int size_y,
                                         const int* spins, nqs_log_amp_fn_t log_amp,
                                         void* log_amp_user, double* out_re, double* out_im) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!cfg || !spins || !log_amp || !out_re || !out_im) {
    if (out_re) *out_re = 0.0;
    if (out_im) *out_im = 0.0;
    return;
  }
  if (stream_ready() || local_energy_any(cfg, size_x, size_y, 0, NQS_HIP_FP64, spins, 1, log_amp,
                                         log_amp_user, out_re, out_im)) {
    *out_re = 0.0;
    *out_im = 0.0;
  }
}

NQS_EXPORT void nqs_local_energy_batch_complex(const nqs_config_t* cfg, int size_x, int size_y,
                                               const int* spins_batch, int batch_size,
                                               nqs_log_amp_fn_t log_amp, void* log_amp_user,
                                               double* out_re, double* out_im) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!cfg || !spins_batch || !out_re || !out_im || batch_size <= 0 || !log_amp) return;
  if (stream_ready()) return;
  if (local_energy_any(cfg, size_x, size_y, 0, NQS_HIP_FP64, spins_batch, batch_size, log_amp,
                       log_amp_user, out_re, out_im))
    for (int i = 0; i < batch_size; i++) out_re[i] = out_im[i] = 0.0;
}

NQS_EXPORT int nqs_hip_local_energy_batch_ex(const nqs_config_t* cfg, int size_x, int size_y,
                                             const nqs_hip_options_t* opt, const int* spins, int B,
                                             nqs_log_amp_fn_t fn, void* user, double* re,
                                             double* im) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!cfg || !spins || B <= 0 || !fn || !re) return -1;
  if (stream_ready()) return -1;
  const int per = opt ? opt->periodic : 0;
  const int prec = opt ? opt->precision : NQS_HIP_FP64;
  return local_energy_any(cfg, size_x, size_y, per, prec, spins, B, fn, user, re, im);
}

// ===========================================================================
// stochastic reconfiguration
// ===========================================================================
int solo_cg_step(const nqs_config_t* cfg, int Lx, int Ly, nqs_ansatz* a, nqs_sampler* s,
                 nqs_sr_step_info_t* info, bool in_run);

namespace {

enum { SR_CG = 0, SR_HOLO = 1, SR_MINSR = 2 };

struct StepTimer {
  hipEvent_t ev[7];
  int n = 0;
  bool on = false;
  void init() {
    on = true;
    for (int i = 0; i < 7; i++)
      if (hipEventCreate(&ev[i]) != hipSuccess) on = false;
  }
  void mark() {
    if (on && n < 7) (void)hipEventRecord(ev[n++], g_stream);
  }
  void finish(nqs_sampler* s) {
    if (!on) return;
    (void)hipEventSynchronize(ev[n - 1]);
    double tot = 0.0;
    for (int i = 0; i + 1 < n && i < 6; i++) {
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, ev[i], ev[i + 1]);
      s->stage_ms[i] = ms;
      tot += ms;
    }
    s->stage_ms[6] = tot;
    s->have_stage = 1;
    for (int i = 0; i < 7; i++) (void)hipEventDestroy(ev[i]);
  }
};

// CG solve of (A + eps I) x = F on the device; A applied through `apply`.
template <typename Apply>
int cg_device(long P, const double* F, double* x, int max_iters, double tol, Apply apply,
              int& iters, int& converged, int& breakdown, double& pAp_out, double& rs_ratio) {
  CKR(g_ws.cg.ensure(sizeof(nqs::CGScalars)));
  nqs::CGScalars* sc = g_ws.cg.as<nqs::CGScalars>();
  double* base = g_ws.vecs.as<double>();
  double* r = base + 2 * P;
  double* p = base + 3 * P;
  double* Ap = base + 4 * P;
  CKR(nqs::launch_cg_init(F, P, x, r, p, sc, g_stream));
  nqs::CGScalars h;
  int it = 0;
  for (; it < max_iters; it++) {
    CKR(apply(p, Ap));
    CKR(nqs::launch_cg_alpha(sc, p, Ap, P, g_stream));
    CKR(nqs::launch_cg_update(sc, x, r, p, Ap, P, tol, g_stream));
    CKR(nqs::launch_cg_direction(sc, p, r, P, g_stream));
    if ((it & 7) == 7 || it + 1 == max_iters) {
      CK(hipMemcpyAsync(&h, sc, sizeof(h), hipMemcpyDeviceToHost, g_stream));
      CK(hipStreamSynchronize(g_stream));
      if (h.done) break;
    }
  }
  CK(hipMemcpyAsync(&h, sc, sizeof(h), hipMemcpyDeviceToHost, g_stream));
  CK(hipStreamSynchronize(g_stream));
  iters = h.iters;
  converged = h.converged;
  breakdown = h.breakdown;
  pAp_out = h.pAp;
  rs_ratio = h.rs0 > 0.0 ? sqrt(h.rs / h.rs0) : 0.0;
  if (max_iters <= 0) { iters = 0; converged = 0; breakdown = 0; }
  return 0;
}

// MinSR solve (src/nqs/nqs_optimizer.c:1119-1177) on device data, shared by
// nqs_sr_step_minsr_full and nqs_hip_minsr_solve.  Inputs: the samples, the
// uncentred O [B][P] with its column means (null / unused when the factored
// Gram is used without O), the energies E with their mean in sc[0].  The
// sample-space solve: eps_i = E_i - <E>, T = O_c O_c^T / B + eps I (direct
// over O, or factored from the RBM's t = tanh theta left in the workspace by
// the caller), Cholesky with the reference's retry rule, delta = O_c^T y / B.
// The parameter-space solve (dual): (O_c^T O_c / B + eps I) delta =
// O_c^T eps / B.  Optional host copies of T (before factorisation) and y.
struct MinsrArgs {
  nqs_ansatz* a = nullptr;
  const int* d_samples = nullptr;
  int B = 0;
  double* O = nullptr;        // [B][P], centred in place (null: factored, no O)
  const double* Om = nullptr; // column means of O
  const double* E = nullptr;  // [B]
  double* sc = nullptr;       // device scalars, sc[0] = <E>
  double* work = nullptr;     // colsum / factored work
  double* rhs = nullptr;      // [P] scratch (parameter space)
  double eps = 0.0;
  bool factored = false, dual = false;
  bool mixed = false;         // dense Gram on the split-BF16 MFMA (FP32 O)
  bool fdual = false;         // mixed parameter-space solve of an RBM without O:
                              // O_c^T and the force from the spins and FP32 factors
  double* delta = nullptr;    // [P] out (device)
  double* h_T = nullptr;      // optional [B][B] host copy of T (sample space)
  double* h_y = nullptr;      // optional [B] host copy of y (sample space)
};

// NQS_HIP_FACTORED_DUAL=0: the mixed parameter-space MinSR of an RBM goes
// through the materialised O (read at each step); so does an RBM with more
// than 256 staged factor rows per sample
bool factored_dual_on(const nqs_ansatz* a) {
  const char* e = getenv("NQS_HIP_FACTORED_DUAL");
  if (e && e[0] == '0') return false;
  // the generator stages N + M (complex: N + 2M) factor rows per sample
  return a->N + (a->kind == NQS_ANSATZ_COMPLEX_RBM ? 2 : 1) * a->M <= 256;
}

// NQS_HIP_GRAM_SPLIT=0: the mixed parameter-space S is formed in one
// product before the factorisation starts (read at each step)
bool gram_split_on() {
  const char* e = getenv("NQS_HIP_GRAM_SPLIT");
  return !(e && e[0] == '0');
}

int minsr_core(const MinsrArgs& m, int& converged) {
  nqs_ansatz* a = m.a;
  const int B = m.B, N = a->N;
  const long P = a->P;
  const double eps = m.eps;
  double* O = m.O;
  double* sc = m.sc;
  double* work = m.work;
  double* delta = m.delta;
  const bool ofree = m.factored && !m.dual;
  converged = 0;
  double* u = g_ws.u1.as<double>();
  // eps_i = E_i - <E>, O_c = O - <O>  (nqs_optimizer.c:1119-1126)
  CKR(nqs::launch_sub_scalar(u, m.E, B, sc + 0, g_stream));
  if (O) CKR(nqs::launch_center(O, B, P, m.Om, g_stream));
  CKR(g_ws.diag.ensure(sizeof(double) * B));
  CKR(g_ws.y.ensure(sizeof(double) * B * 2));
  CKR(g_ws.fail.ensure(sizeof(int) * 4));
  double* T = nullptr;
  const int cpl = a->kind == NQS_ANSATZ_COMPLEX_RBM;
  auto gram_factored = [&](double* out) -> int {
    CKR(g_ws.Sd.ensure(sizeof(double) * (size_t)B * (N + 2)));
    return nqs::launch_gram_factored(m.d_samples, B, N, g_ws.t_re.as<double>(),
                                     cpl ? g_ws.t_im.as<double>() : nullptr, a->M, eps, out,
                                     work, g_ws.Sd.as<double>(), g_ws.u2.as<double>(), sc + 6,
                                     g_stream);
  };
  if (m.dual) {
    // parameter-space MinSR: delta = S^-1 (O_c^T eps / Ns)
    CKR(g_ws.T.ensure(sizeof(double) * (size_t)P * P));
    T = g_ws.T.as<double>();
    double* rhs = m.rhs;
    // S = O_c^T O_c / Ns (lower triangle): FP64 MFMA over the transposed O,
    // or the split-BF16 MFMA over its FP32 transpose (K padded with zeros)
    const long Kp = nqs::x3_pad(B);
    double* OT = nullptr;
    float* OT32 = nullptr;
    if (m.mixed) {
      CKR(g_ws.o32.ensure(sizeof(float) * (size_t)P * Kp));
      CKR(g_ws.x3max.ensure(sizeof(float) * nqs::x3_max_elems(B, P, Kp, 1)));
      OT32 = g_ws.o32.as<float>();
      if (m.fdual)
        CKR(nqs::launch_oct_factored(m.d_samples, B, N, a->M, g_ws.t_re.as<float>(),
                                     cpl ? g_ws.t_im.as<float>() : nullptr, m.Om, P, OT32, Kp,
                                     g_ws.x3max.as<float>(), g_stream));
      else
        CKR(nqs::launch_to_f32_t(O, B, P, OT32, Kp, g_ws.x3max.as<float>(), g_stream));
    } else {
      CKR(g_ws.T2.ensure(sizeof(double) * (size_t)P * B));
      OT = g_ws.T2.as<double>();
      CKR(nqs::launch_transpose(O, B, P, OT, g_stream));
    }
    if (m.mixed)
      CKR(g_ws.x3part.ensure(sizeof(float) * nqs::x3_part_elems((int)P, (int)P, Kp) + 16));
    auto gram_param = [&](double* out) -> int {
      if (m.mixed)
        CKR(nqs::launch_gemm_x3((int)P, (int)P, (int)Kp, 1.0 / B, OT32, Kp, OT32, Kp,
                                g_ws.x3max.as<float>(), g_ws.x3max.as<float>(), out, P, 1,
                                g_ws.x3part.as<float>(), g_stream));
      else
        CKR(nqs::launch_gemm_nt((int)P, (int)P, B, 1.0 / B, OT, B, OT, B, 0.0, out, P, 1,
                                g_stream));
      return nqs::launch_add_diag(out, (int)P, P, eps, g_stream);
    };
    // The Cholesky's first wide panel reads only columns [0, w) of S: those
    // are formed first, and the trailing block of S on the factorisation's
    // second stream, beside that panel (the same sums per element;
    // NQS_HIP_GRAM_SPLIT=0: one product, then the factorisation)
    const int w0 = nqs::potrf_first_width((int)P);
    hipEvent_t gram_done = nullptr;
    if (m.mixed && w0 < P && gram_split_on()) {
      static hipEvent_t ev_cols = nullptr, ev_rest = nullptr;
      if (!ev_cols) CK(hipEventCreateWithFlags(&ev_cols, hipEventDisableTiming));
      if (!ev_rest) CK(hipEventCreateWithFlags(&ev_rest, hipEventDisableTiming));
      hipStream_t side = nqs::potrf_side_stream();
      if (!side) return -1;
      const long Pr = P - w0;
      float* part = g_ws.x3part.as<float>();
      const float* mx = g_ws.x3max.as<float>();
      const size_t p1 = nqs::x3_part_elems(w0, w0, Kp), p2 = nqs::x3_part_elems((int)Pr, w0, Kp);
      CKR(nqs::launch_gemm_x3(w0, w0, (int)Kp, 1.0 / B, OT32, Kp, OT32, Kp, mx, mx, T, P, 1, part,
                              g_stream));
      CKR(nqs::launch_gemm_x3((int)Pr, w0, (int)Kp, 1.0 / B, OT32 + (size_t)w0 * Kp, Kp, OT32, Kp, mx,
                              mx, T + (size_t)w0 * P, P, 0, part + p1, g_stream));
      CKR(nqs::launch_add_diag(T, w0, P, eps, g_stream));
      CK(hipEventRecord(ev_cols, g_stream));
      CK(hipStreamWaitEvent(side, ev_cols, 0));
      double* Tr = T + (size_t)w0 * P + w0;
      CKR(nqs::launch_gemm_x3((int)Pr, (int)Pr, (int)Kp, 1.0 / B, OT32 + (size_t)w0 * Kp, Kp,
                              OT32 + (size_t)w0 * Kp, Kp, mx, mx, Tr, P, 1, part + p1 + p2, side));
      CKR(nqs::launch_add_diag(Tr, (int)Pr, P, eps, side));
      CK(hipEventRecord(ev_rest, side));
      gram_done = ev_rest;
    } else {
      CKR(gram_param(T));
    }
    if (m.fdual) {
      // O_c^T eps / Ns = O^T (eps - <eps>) / Ns from the factors
      double* uc = g_ws.u2.as<double>();
      CKR(nqs::launch_sum(u, B, nullptr, 1.0 / B, sc + 7, g_stream));
      CKR(nqs::launch_sub_scalar(uc, u, B, sc + 7, g_stream));
      CKR(nqs::launch_factored_delta(m.d_samples, B, N, a->M, g_ws.t_re.as<float>(),
                                     cpl ? g_ws.t_im.as<float>() : nullptr, uc, 1.0 / B, work,
                                     rhs, g_stream));
    } else {
      CKR(nqs::launch_colsum(O, B, P, u, 1.0 / B, rhs, work, g_stream));
    }
    CKR(g_ws.linv.ensure(sizeof(double) * nqs::potrf_linv_elems((int)P)));
    CKR(g_ws.y.ensure(sizeof(double) * 2 * (size_t)P));
    double* Linv = g_ws.linv.as<double>();
    int f = nqs::launch_potrf(T, (int)P, Linv, g_ws.fail.as<int>(), g_stream, gram_done);
    if (f < 0) return -1;
    if (f > 0) {
      // as in the sample-space solve: the partially factored matrix with
      // +1e-6 on its diagonal is factored once more (input columns from a
      // second evaluation of S)
      CKR(g_ws.Oi.ensure(sizeof(double) * (size_t)P * P));
      double* S2 = g_ws.Oi.as<double>();
      CKR(gram_param(S2));
      CKR(nqs::launch_potrf_repair(T, (int)P, f - 1, S2, g_stream));
      CKR(nqs::launch_add_diag(T, (int)P, P, 1e-6, g_stream));
      f = nqs::launch_potrf(T, (int)P, Linv, g_ws.fail.as<int>(), g_stream);
      if (f < 0) return -1;
    }
    if (f == 0) {
      CKR(nqs::launch_potrs(T, Linv, (int)P, rhs, delta, g_ws.y.as<double>(), g_stream));
      converged = 1;
    } else {
      fprintf(stderr, "nqs_sr_step_minsr: Cholesky breakdown after retry; increase "
                      "sr_diag_shift or num_samples. Update set to 0.\n");
      CKR(nqs::launch_fill(delta, 0.0, P, g_stream));
    }
    return 0;
  }
  CKR(g_ws.T.ensure(sizeof(double) * (size_t)B * B));
  T = g_ws.T.as<double>();
  // T = O_c O_c^T / Ns + eps I over the materialised O: FP64 MFMA, or the
  // split-BF16 MFMA over its FP32 copy (columns padded with zeros)
  const long Pp = nqs::x3_pad(P);
  if (!m.factored && m.mixed) {
    CKR(g_ws.o32.ensure(sizeof(float) * (size_t)B * Pp));
    CKR(g_ws.x3part.ensure(sizeof(float) * nqs::x3_part_elems(B, B, Pp) + 16));
    CKR(g_ws.x3max.ensure(sizeof(float) * nqs::x3_max_elems(B, P, Pp, 0)));
    CKR(nqs::launch_to_f32_pad(O, B, P, g_ws.o32.as<float>(), Pp, g_ws.x3max.as<float>(),
                               g_stream));
  }
  auto gram_dense = [&](double* out) -> int {
    if (m.mixed)
      CKR(nqs::launch_gemm_x3(B, B, (int)Pp, 1.0 / B, g_ws.o32.as<float>(), Pp,
                              g_ws.o32.as<float>(), Pp, g_ws.x3max.as<float>(),
                              g_ws.x3max.as<float>(), out, B, 1, g_ws.x3part.as<float>(),
                              g_stream));
    else
      CKR(nqs::launch_gemm_nt(B, B, (int)P, 1.0 / B, O, P, O, P, 0.0, out, B, 1, g_stream));
    return nqs::launch_add_diag(out, B, B, eps, g_stream);
  };
  if (m.factored) {
    CKR(gram_factored(T));
  } else {
    CKR(gram_dense(T));
  }
  if (m.h_T) {
    CK(hipMemcpyAsync(m.h_T, T, sizeof(double) * (size_t)B * B, hipMemcpyDeviceToHost, g_stream));
    CK(hipStreamSynchronize(g_stream));
  }
  CKR(g_ws.linv.ensure(sizeof(double) * nqs::potrf_linv_elems(B)));
  double* Linv = g_ws.linv.as<double>();
  int f = nqs::launch_potrf(T, B, Linv, g_ws.fail.as<int>(), g_stream);
  if (f < 0) return -1;
  if (f > 0) {
    // the reference retries on its partially factored matrix with +1e-6 on
    // the diagonal (nqs_optimizer.c:1156-1160): rebuild that matrix (input
    // columns from a second Gram evaluation) and factor again
    CKR(g_ws.T2.ensure(sizeof(double) * (size_t)B * B));
    double* T2 = g_ws.T2.as<double>();
    if (m.factored) {
      CKR(gram_factored(T2));
    } else {
      CKR(gram_dense(T2));
    }
    CKR(nqs::launch_potrf_repair(T, B, f - 1, T2, g_stream));
    CKR(nqs::launch_add_diag(T, B, B, 1e-6, g_stream));
    f = nqs::launch_potrf(T, B, Linv, g_ws.fail.as<int>(), g_stream);
    if (f < 0) return -1;
  }
  double* y = g_ws.y.as<double>();
  if (f == 0) {
    CKR(nqs::launch_potrs(T, Linv, B, u, y, y + B, g_stream));
    converged = 1;
  } else {
    fprintf(stderr, "nqs_sr_step_minsr: Cholesky breakdown after retry; increase "
                    "sr_diag_shift or num_samples. Update set to 0.\n");
    CKR(nqs::launch_fill(y, 0.0, B, g_stream));
  }
  if (m.h_y) {
    CK(hipMemcpyAsync(m.h_y, y, sizeof(double) * B, hipMemcpyDeviceToHost, g_stream));
    CK(hipStreamSynchronize(g_stream));
  }
  // delta_k = (1/Ns) sum_i O_c[i][k] y_i = (1/Ns) sum_i O[i][k] (y_i - mean y)
  if (ofree) {
    double* yc = y + B;
    CKR(nqs::launch_sum(y, B, nullptr, 1.0 / B, sc + 7, g_stream));
    CKR(nqs::launch_sub_scalar(yc, y, B, sc + 7, g_stream));
    CKR(nqs::launch_factored_delta(m.d_samples, B, N, a->M, g_ws.t_re.as<double>(),
                                   cpl ? g_ws.t_im.as<double>() : nullptr, yc, 1.0 / B, work,
                                   delta, g_stream));
  } else {
    CKR(nqs::launch_colsum(O, B, P, y, 1.0 / B, delta, work, g_stream));
  }
  return 0;
}

// Variants of the holomorphic step (src/nqs/nqs_optimizer.c:562-993):
// excited-state SR (local energies augmented by the orthogonality penalty
// mu r conj(<r>), r = psi_ref / psi) and real-time tVMC (force
// 2[Cov(R, E_im) - Cov(I, E_re)], theta += dt delta).  apply = false leaves
// the parameters alone and copies delta to keep_delta (Heun's two stages).
struct DynStep {
  int kind = 0;                 // 1 excited, 2 tVMC
  double mu = 0.0;
  nqs_log_amp_fn_t ref_fn = nullptr;
  void* ref_user = nullptr;
  double dt = 0.0;
  bool apply = true;
  double* keep_delta = nullptr;
};

// log|psi| and arg psi of the B samples under a callback: on the device for
// a recognised ansatz (plus the Marshall sign), else through the host
// callback on the host copies
int logamp_samples(nqs_log_amp_fn_t fn, void* user, int N, const int* d_samples,
                   const int* h_samples, int B, double* d_la, double* d_ar) {
  AmpSource src;
  if (recognise(fn, user, src) && src.a->N == N) {
    CKR(ansatz_upload(src.a));
    CKR(forward_t<double>(src.a, d_samples, B, d_la, d_ar));
    if (src.marshall) CKR(nqs::launch_marshall_arg(d_samples, B, N, src.mlx, src.mly, d_ar, g_stream));
    return 0;
  }
  if (!h_samples) return -1;
  std::vector<double> la(B), ar(B);
  for (int b = 0; b < B; b++) {
    fn(h_samples + (size_t)b * N, N, user, &la[b], &ar[b]);
    g_host_calls++;
  }
  CK(hipMemcpyAsync(d_la, la.data(), sizeof(double) * B, hipMemcpyHostToDevice, g_stream));
  CK(hipMemcpyAsync(d_ar, ar.data(), sizeof(double) * B, hipMemcpyHostToDevice, g_stream));
  CK(hipStreamSynchronize(g_stream));
  return 0;
}

int sr_step_impl(int method, const nqs_config_t* cfg, int Lx, int Ly, nqs_ansatz* a,
                 nqs_sampler* s, nqs_log_amp_fn_t fn, void* user, nqs_gradient_fn_t gfn,
                 nqs_holomorphic_gradient_fn_t hfn, void* guser, nqs_sr_step_info_t* info,
                 bool in_run, const DynStep* dyn = nullptr) {
  const int N = s->N;
  const long P = a->P;
  const int B = cfg->num_samples;
  if (P <= 0 || B <= 0) return -1;
  if (!fn) { fn = nqs_ansatz_log_amp; user = a; }
  const int holo = (method == SR_HOLO);
  const int prec = s->opt.precision;
  {
    // CG-SR of a real RBM whose sampler and local energies use the ansatz
    // itself: the batched path with one problem (factored QGT products, no
    // host round trip inside the CG, one graph per step);
    // NQS_HIP_SOLO_BATCH=0 keeps the general path
    static const bool solo_on = [] {
      const char* e = getenv("NQS_HIP_SOLO_BATCH");
      return !(e && e[0] == '0');
    }();
    AmpSource es;
    if (solo_on && method == SR_CG && !gfn && s->device && a->kind == NQS_ANSATZ_RBM && a->M >= 1 &&
        s->src.a == a && !s->src.marshall && prec == NQS_HIP_FP64 &&
        s->opt.move != NQS_HIP_MOVE_LEARNED && recognise(fn, user, es) && es.a == a &&
        !es.marshall && (a->M + 1) * (a->N + 1) <= nqs::kMultiVecBlock * nqs::kMultiMaxPart)
      return solo_cg_step(cfg, Lx, Ly, a, s, info, in_run);
  }
  StepTimer tm;
  if (s->device) tm.init();
  if (!in_run) a->dev_fresh = 0;
  CKR(ansatz_upload(a));
  a->dev_fresh = 1;
  tm.mark();
  // ---- 1. samples ----------------------------------------------------------
  const int C = s->opt.num_chains;
  const int k = s->cfg.num_decorrelate > 0 ? s->cfg.num_decorrelate : 1;
  const int rounds = (B + C - 1) / C;
  CKR(s->d_samples.ensure(sizeof(int) * (size_t)B * N));
  std::vector<int> h_samples;
  AmpSource esrc, rsrc;
  const bool excited = dyn && dyn->kind == 1, tvmc = dyn && dyn->kind == 2;
  const bool need_host_samples =
      !s->device || gfn || hfn ||
      (excited && !(recognise(fn, user, rsrc) && recognise(dyn->ref_fn, dyn->ref_user, rsrc)));
  const bool e_dev = recognise(fn, user, esrc) && esrc.a->N == N;
  if (need_host_samples || !e_dev) h_samples.resize((size_t)B * N);
  CKR(sampler_advance(s, k, rounds, B, s->d_samples.as<int>(),
                      h_samples.empty() ? nullptr : h_samples.data()));
  if (!h_samples.empty()) CK(hipStreamSynchronize(g_stream));
  const int* d_samples = s->d_samples.as<int>();
  tm.mark();
  // ---- 2. local energies -----------------------------------------------------
  CKR(g_ws.e_re.ensure(sizeof(double) * B));
  CKR(g_ws.e_im.ensure(sizeof(double) * B));
  double* E = g_ws.e_re.as<double>();
  double* Ei = g_ws.e_im.as<double>();
  if (e_dev) {
    CKR(ansatz_upload(esrc.a));
    CKR(eloc_device(cfg, Lx, Ly, s->opt.periodic, prec, esrc, d_samples, B, E, Ei));
  } else {
    CKR(eloc_callback(cfg, Lx, Ly, s->opt.periodic, h_samples.data(), B, fn, user, E, Ei,
                      d_samples));
  }
  CKR(g_ws.scal.ensure(sizeof(double) * 16));
  double* sc = g_ws.scal.as<double>();      // [0] <E_re> [1] <E_re^2> [2] <E_im> [3,4] u means
  if (excited) {
    // physical statistics first (they are what info reports), then
    // r = psi_ref / psi, its means, the augmented energies and their means
    // in closed form (sc[10], sc[11]), nqs_optimizer.c:605-657
    CKR(nqs::launch_sum(E, B, nullptr, 1.0 / B, sc + 0, g_stream));
    CKR(nqs::launch_sum(E, B, E, 1.0 / B, sc + 1, g_stream));
    CKR(nqs::launch_sum(Ei, B, nullptr, 1.0 / B, sc + 2, g_stream));
    CKR(g_ws.xr.ensure(sizeof(double) * 6 * (size_t)B));
    double* xr = g_ws.xr.as<double>();
    const int* hs = h_samples.empty() ? nullptr : h_samples.data();
    CKR(logamp_samples(fn, user, N, d_samples, hs, B, xr, xr + B));
    CKR(logamp_samples(dyn->ref_fn, dyn->ref_user, N, d_samples, hs, B, xr + 2 * B, xr + 3 * B));
    CKR(nqs::launch_excited_ratio(xr, xr + B, xr + 2 * B, xr + 3 * B, B, xr + 4 * B, xr + 5 * B,
                                  g_stream));
    CKR(nqs::launch_sum(xr + 4 * B, B, nullptr, 1.0 / B, sc + 8, g_stream));
    CKR(nqs::launch_sum(xr + 5 * B, B, nullptr, 1.0 / B, sc + 9, g_stream));
    CKR(nqs::launch_excited_augment(E, Ei, xr + 4 * B, xr + 5 * B, sc + 8, dyn->mu, B, g_stream));
    CKR(nqs::launch_excited_means(sc + 0, sc + 2, sc + 8, dyn->mu, sc + 10, g_stream));
  }
  tm.mark();
  // MinSR path: sample space (T = O_c O_c^T / Ns + eps I, Ns x Ns) or, with
  // NQS_HIP_GRAM_AUTO when it costs less, parameter space
  // (S = O_c^T O_c / Ns + eps I, P x P): O_c^T (O_c O_c^T/Ns + eps)^-1 =
  // (O_c^T O_c/Ns + eps)^-1 O_c^T, so both give the same delta.  A
  // sample-space solve with the factored Gram never forms O: the update is
  // taken from the RBM factors as well.
  const bool rbm_like = (a->kind == NQS_ANSATZ_RBM || a->kind == NQS_ANSATZ_COMPLEX_RBM) && !gfn;
  const bool factored = method == SR_MINSR && rbm_like &&
                        (s->opt.gram == NQS_HIP_GRAM_FACTORED || s->opt.gram == NQS_HIP_GRAM_AUTO);
  bool dual = method == SR_MINSR && s->opt.gram == NQS_HIP_GRAM_PARAMETER;
  if (method == SR_MINSR && s->opt.gram == NQS_HIP_GRAM_AUTO) {
    const double b = (double)B, p = (double)P;
    const double gram_b = factored ? b * b * (N + 2.0 * a->M) : b * b * p;
    const double primal = gram_b + b * b * b / 3.0;
    // the parameter-space Gram of the mixed mode runs on the split-BF16
    // MFMA at about an eighth of the FP64 MFMA's time per product
    const double dualc = (prec == NQS_HIP_MIXED ? 0.125 : 1.0) * b * p * p + p * p * p / 3.0;
    dual = dualc < primal;
  }
  const bool ofree = factored && !dual;
  // the mixed parameter-space solve of an RBM forms O_c^T in FP32 straight
  // from the spins and the FP32 factors, and the means and force from factor
  // sums, so O itself is never written (NQS_HIP_FACTORED_DUAL=0: through O)
  const bool fdual = method == SR_MINSR && dual && prec == NQS_HIP_MIXED && rbm_like && !holo &&
                     factored_dual_on(a);
  // ---- 3. O matrix -----------------------------------------------------------
  const size_t ob = sizeof(double) * (size_t)B * (size_t)P;
  if (!ofree && !fdual) CKR(g_ws.O.ensure(ob));
  if (holo) CKR(g_ws.Oi.ensure(ob));
  double* O = (ofree || fdual) ? nullptr : g_ws.O.as<double>();
  double* Oi = holo ? g_ws.Oi.as<double>() : nullptr;
  if (ofree) {
    // t in FP64 for the factored Gram and update (the FP64 local energies of
    // this ansatz have just left it in the workspace)
    if (!(e_dev && esrc.a == a && prec == NQS_HIP_FP64))
      CKR(forward_t<double>(a, d_samples, B, nullptr, nullptr));
  } else if (fdual) {
    CKR(forward_t<float>(a, d_samples, B, nullptr, nullptr, false, true));
  } else if (OrbitSource og; orbit_gradient_cb(holo, gfn, hfn, guser, og)) {
    // gradients of a symmetry-projected amplitude: orbit rows of every
    // sample, weights and base gradients on the device
    CKR(orbit_gradients(og, a, d_samples, B, holo, O, Oi));
  } else if ((!holo && gfn) || (holo && hfn)) {
    std::vector<double> hr((size_t)B * P), hi(holo ? (size_t)B * P : 0);
    for (int b = 0; b < B; b++) {
      const int* sp = h_samples.data() + (size_t)b * N;
      if (holo) hfn(guser, a, sp, N, hr.data() + (size_t)b * P, hi.data() + (size_t)b * P);
      else gfn(guser, a, sp, N, hr.data() + (size_t)b * P);
      g_host_calls++;
    }
    CK(hipMemcpyAsync(O, hr.data(), ob, hipMemcpyHostToDevice, g_stream));
    if (holo) CK(hipMemcpyAsync(Oi, hi.data(), ob, hipMemcpyHostToDevice, g_stream));
    CK(hipStreamSynchronize(g_stream));
  } else {
    if (low_prec(prec)) {
      CKR(forward_t<float>(a, d_samples, B, nullptr, nullptr, false, prec == NQS_HIP_MIXED));
      ModelView<float> v;
      CKR(model_view<float>(a, v));
      CKR(nqs::launch_gen_O<float>(v, a->kind, d_samples, B, g_ws.t_re.as<float>(),
                                   g_ws.t_im.as<float>(), holo, O, Oi, P, g_stream));
    } else {
      CKR(forward_t<double>(a, d_samples, B, nullptr, nullptr));
      ModelView<double> v;
      CKR(model_view<double>(a, v));
      CKR(nqs::launch_gen_O<double>(v, a->kind, d_samples, B, g_ws.t_re.as<double>(),
                                    g_ws.t_im.as<double>(), holo, O, Oi, P, g_stream));
    }
  }
  // ---- 4. means and force ----------------------------------------------------
  CKR(g_ws.vecs.ensure(sizeof(double) * 9 * (size_t)P));
  if (ofree) {
    const int nblk = a->kind == NQS_ANSATZ_COMPLEX_RBM ? 2 : 1;
    CKR(g_ws.colwork.ensure(sizeof(double) * nqs::gram_factored_work_elems(B, N, a->M, nblk)));
  } else if (fdual) {
    const int nblk = a->kind == NQS_ANSATZ_COMPLEX_RBM ? 2 : 1;
    CKR(g_ws.colwork.ensure(sizeof(double) * nqs::factored_delta_work_elems(B, N, a->M, nblk)));
  } else {
    CKR(g_ws.colwork.ensure(sizeof(double) * nqs::colsum_work_elems(B, P)));
  }
  CKR(g_ws.u1.ensure(sizeof(double) * B));
  CKR(g_ws.u2.ensure(sizeof(double) * B));
  double* vb = g_ws.vecs.as<double>();
  double* delta = vb;                        // [0, P)
  double* F = vb + P;                        // [P, 2P)
  double* Om = vb + 5 * P;                   // means of O (R)
  double* Im = vb + 6 * P;                   // means of I
  double* c1 = vb + 7 * P;
  double* c2 = vb + 8 * P;
  double* work = g_ws.colwork.as<double>();
  if (!excited) {   // (the excited step summed the physical energies before augmenting)
    CKR(nqs::launch_sum(E, B, nullptr, 1.0 / B, sc + 0, g_stream));
    CKR(nqs::launch_sum(E, B, E, 1.0 / B, sc + 1, g_stream));
    CKR(nqs::launch_sum(Ei, B, nullptr, 1.0 / B, sc + 2, g_stream));
  }
  if (fdual)
    CKR(nqs::launch_factored_delta(d_samples, B, N, a->M, g_ws.t_re.as<float>(),
                                   a->kind == NQS_ANSATZ_COMPLEX_RBM ? g_ws.t_im.as<float>()
                                                                     : nullptr,
                                   nullptr, 1.0 / B, work, Om, g_stream));
  else if (!ofree)
    CKR(nqs::launch_colsum(O, B, P, nullptr, 1.0 / B, Om, work, g_stream));
  if (holo && tvmc) {
    // F = 2[(R^T E_im - I^T E_re)/Ns - <R><E_im> + <I><E_re>]
    // (nqs_optimizer.c:797-810): second term and mean carry the minus sign
    CKR(nqs::launch_colsum(Oi, B, P, nullptr, 1.0 / B, Im, work, g_stream));
    CKR(nqs::launch_colsum(O, B, P, Ei, 1.0 / B, c1, work, g_stream));
    CKR(nqs::launch_colsum(Oi, B, P, E, -1.0 / B, c2, work, g_stream));
    CKR(nqs::launch_sum(E, B, nullptr, -1.0 / B, sc + 12, g_stream));
    CKR(nqs::launch_qgt_finish(F, c1, Om, sc + 2, c2, Im, sc + 12, 2.0, 0.0, c1, P, g_stream));
  } else if (holo) {
    // holomorphic force; the excited step's with the augmented energies and
    // their closed-form means
    CKR(nqs::launch_colsum(Oi, B, P, nullptr, 1.0 / B, Im, work, g_stream));
    CKR(nqs::launch_colsum(O, B, P, E, 1.0 / B, c1, work, g_stream));
    CKR(nqs::launch_colsum(Oi, B, P, Ei, 1.0 / B, c2, work, g_stream));
    CKR(nqs::launch_qgt_finish(F, c1, Om, sc + (excited ? 10 : 0), c2, Im, sc + (excited ? 11 : 2),
                               2.0, 0.0, c1, P, g_stream));
  } else if (method == SR_CG) {
    CKR(nqs::launch_colsum(O, B, P, E, 1.0 / B, c1, work, g_stream));
    CKR(nqs::launch_qgt_finish(F, c1, Om, sc + 0, nullptr, nullptr, nullptr, 1.0, 0.0, c1, P,
                               g_stream));
  }
  tm.mark();
  // ---- 5. solve ----------------------------------------------------------------
  int iters = 0, converged = 0;
  const double eps = cfg->sr_diag_shift;
  if (method == SR_MINSR) {
    MinsrArgs m;
    m.a = a;
    m.d_samples = d_samples;
    m.B = B;
    m.O = O;
    m.Om = Om;
    m.E = E;
    m.sc = sc;
    m.work = work;
    m.rhs = c1;
    m.eps = eps;
    m.factored = factored;
    m.dual = dual;
    m.mixed = prec == NQS_HIP_MIXED;
    m.fdual = fdual;
    m.delta = delta;
    CKR(minsr_core(m, converged));
  } else {
    double* u1 = g_ws.u1.as<double>();
    double* u2 = g_ws.u2.as<double>();
    auto apply = [&](const double* v, double* out) -> int {
      CKR(nqs::launch_rowdot(O, B, P, v, u1, g_stream));
      CKR(nqs::launch_sum(u1, B, nullptr, 1.0 / B, sc + 3, g_stream));
      CKR(nqs::launch_colsum(O, B, P, u1, 1.0 / B, c1, work, g_stream));
      if (holo) {
        CKR(nqs::launch_rowdot(Oi, B, P, v, u2, g_stream));
        CKR(nqs::launch_sum(u2, B, nullptr, 1.0 / B, sc + 4, g_stream));
        CKR(nqs::launch_colsum(Oi, B, P, u2, 1.0 / B, c2, work, g_stream));
        return nqs::launch_qgt_finish(out, c1, Om, sc + 3, c2, Im, sc + 4, 2.0, eps, v, P,
                                      g_stream);
      }
      return nqs::launch_qgt_finish(out, c1, Om, sc + 3, nullptr, nullptr, nullptr, 1.0, eps, v,
                                    P, g_stream);
    };
    int breakdown = 0;
    double pAp = 0.0, rr = 0.0;
    CKR(cg_device(P, F, delta, cfg->sr_cg_max_iters, cfg->sr_cg_tol, apply, iters, converged,
                  breakdown, pAp, rr));
    if (breakdown && !holo)
      fprintf(stderr, "nqs_sr cg_solve: Krylov breakdown at iter=%d, pAp=%.3e, residual/rhs=%.3e "
                      "(tol=%.3e). Returning unconverged partial step.\n",
              iters, pAp, rr, cfg->sr_cg_tol);
  }
  tm.mark();
  // ---- 6. update -----------------------------------------------------------------
  const double step = tvmc ? dyn->dt : -cfg->learning_rate;
  if (!dyn || dyn->apply) {
    CKR(nqs::launch_axpy(a->d_par.as<double>(), delta, step, P, g_stream));
    a->m64.valid = 0;
    a->m32.valid = 0;
  }
  if (dyn && dyn->keep_delta)
    CK(hipMemcpyAsync(dyn->keep_delta, delta, sizeof(double) * P, hipMemcpyDeviceToDevice, g_stream));
  CKR(nqs::launch_sum(delta, P, delta, 1.0, sc + 5, g_stream));
  // scalars, parameters and energies through one pinned staging buffer:
  // three copies, one synchronisation
  static double* stage = nullptr;
  static size_t stage_n = 0;
  const size_t need = 8 + (size_t)P + (size_t)B;
  if (stage_n < need) {
    if (stage) (void)hipHostFree(stage);
    stage = nullptr;
    stage_n = 0;
    CK(hipHostMalloc((void**)&stage, sizeof(double) * need, hipHostMallocDefault));
    stage_n = need;
  }
  CK(hipMemcpyAsync(stage, sc, sizeof(double) * 6, hipMemcpyDeviceToHost, g_stream));
  CK(hipMemcpyAsync(stage + 8, a->d_par.p, sizeof(double) * P, hipMemcpyDeviceToHost, g_stream));
  CK(hipMemcpyAsync(stage + 8 + P, E, sizeof(double) * B, hipMemcpyDeviceToHost, g_stream));
  CK(hipStreamSynchronize(g_stream));
  double hs[6];
  memcpy(hs, stage, sizeof(hs));
  memcpy(a->h.data(), stage + 8, sizeof(double) * P);
  tm.mark();
  if (info) {
    info->mean_energy = hs[0];
    info->variance_energy = hs[1] - hs[0] * hs[0];
    info->update_norm = sqrt(hs[5]) * fabs(step);
    info->acceptance_ratio = nqs_sampler_acceptance_ratio(s);
    info->cg_iterations = (method == SR_MINSR) ? 0 : iters;
    info->converged = converged;
  }
  // per-chain energy series for tau_int
  s->last_E.assign(stage + 8 + P, stage + 8 + P + B);
  s->last_rounds = rounds;
  tm.finish(s);
  if (!in_run) a->dev_fresh = 0;
  return 0;
}

int sr_run_impl(int method, const nqs_config_t* cfg, int Lx, int Ly, nqs_ansatz* a, nqs_sampler* s,
                nqs_log_amp_fn_t fn, void* user, nqs_gradient_fn_t gfn,
                nqs_holomorphic_gradient_fn_t hfn, void* guser, double* trace,
                const DynStep* dyn = nullptr) {
  // a host callback may read the parameters between steps: keep the host
  // copy authoritative then
  AmpSource src;
  AmpSource rsrc;
  const bool pure_device = s->device && (!fn || recognise(fn, user, src)) && !gfn && !hfn &&
                           (!dyn || dyn->kind != 1 || recognise(dyn->ref_fn, dyn->ref_user, rsrc));
  a->dev_fresh = 0;
  sampler_advance(s, s->cfg.num_thermalize, s->cfg.num_thermalize > 0 ? 1 : 0, 0, nullptr, nullptr);
  for (int it = 0; it < cfg->num_iterations; it++) {
    nqs_sr_step_info_t info;
    int rc = sr_step_impl(method, cfg, Lx, Ly, a, s, fn, user, gfn, hfn, guser, &info, pure_device,
                          dyn);
    if (rc) {
      a->dev_fresh = 0;
      return rc;
    }
    if (trace) trace[it] = info.mean_energy;
  }
  a->dev_fresh = 0;
  return 0;
}

}  // namespace

NQS_EXPORT int nqs_sr_step_custom_full(const nqs_config_t* cfg, int size_x, int size_y,
                                       nqs_ansatz_t* ansatz, nqs_sampler_t* sampler,
                                       nqs_log_amp_fn_t log_amp_fn, void* log_amp_user,
                                       nqs_gradient_fn_t gradient_fn, void* grad_user,
                                       nqs_sr_step_info_t* out_info) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!cfg || !ansatz || !sampler) return -1;
  return sr_step_impl(SR_CG, cfg, size_x, size_y, ansatz, sampler, log_amp_fn, log_amp_user,
                      gradient_fn, nullptr, grad_user, out_info, false);
}

NQS_EXPORT int nqs_sr_step_custom(const nqs_config_t* cfg, int size_x, int size_y,
                                  nqs_ansatz_t* ansatz, nqs_sampler_t* sampler,
                                  nqs_log_amp_fn_t log_amp_fn, void* log_amp_user,
                                  nqs_sr_step_info_t* out_info) {
  return nqs_sr_step_custom_full(cfg, size_x, size_y, ansatz, sampler, log_amp_fn, log_amp_user,
                                 nullptr, nullptr, out_info);
}

NQS_EXPORT int nqs_sr_step(const nqs_config_t* cfg, int size_x, int size_y, nqs_ansatz_t* ansatz,
                           nqs_sampler_t* sampler, nqs_sr_step_info_t* out_info) {
  return nqs_sr_step_custom_full(cfg, size_x, size_y, ansatz, sampler, nullptr, nullptr, nullptr,
                                 nullptr, out_info);
}

NQS_EXPORT int nqs_sr_run_custom_full(const nqs_config_t* cfg, int size_x, int size_y,
                                      nqs_ansatz_t* ansatz, nqs_sampler_t* sampler,
                                      nqs_log_amp_fn_t log_amp_fn, void* log_amp_user,
                                      nqs_gradient_fn_t gradient_fn, void* grad_user,
                                      double* out_energy_trace) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!cfg || !ansatz || !sampler) return -1;
  return sr_run_impl(SR_CG, cfg, size_x, size_y, ansatz, sampler, log_amp_fn, log_amp_user,
                     gradient_fn, nullptr, grad_user, out_energy_trace);
}

NQS_EXPORT int nqs_sr_run_custom(const nqs_config_t* cfg, int size_x, int size_y,
                                 nqs_ansatz_t* ansatz, nqs_sampler_t* sampler,
                                 nqs_log_amp_fn_t log_amp_fn, void* log_amp_user,
                                 double* out_energy_trace) {
  return nqs_sr_run_custom_full(cfg, size_x, size_y, ansatz, sampler, log_amp_fn, log_amp_user,
                                nullptr, nullptr, out_energy_trace);
}

NQS_EXPORT int nqs_sr_run(const nqs_config_t* cfg, int size_x, int size_y, nqs_ansatz_t* ansatz,
                          nqs_sampler_t* sampler, double* out_energy_trace) {
  return nqs_sr_run_custom_full(cfg, size_x, size_y, ansatz, sampler, nullptr, nullptr, nullptr,
                                nullptr, out_energy_trace);
}

NQS_EXPORT int nqs_sr_step_holomorphic_full(const nqs_config_t* cfg, int size_x, int size_y,
                                            nqs_ansatz_t* ansatz, nqs_sampler_t* sampler,
                                            nqs_log_amp_fn_t log_amp_fn, void* log_amp_user,
                                            nqs_holomorphic_gradient_fn_t grad_fn,
                                            void* grad_user, nqs_sr_step_info_t* out_info) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!cfg || !ansatz || !sampler) return -1;
  return sr_step_impl(SR_HOLO, cfg, size_x, size_y, ansatz, sampler, log_amp_fn, log_amp_user,
                      nullptr, grad_fn, grad_user, out_info, false);
}

NQS_EXPORT int nqs_sr_step_holomorphic(const nqs_config_t* cfg, int size_x, int size_y,
                                       nqs_ansatz_t* ansatz, nqs_sampler_t* sampler,
                                       nqs_log_amp_fn_t log_amp_fn, void* log_amp_user,
                                       nqs_sr_step_info_t* out_info) {
  return nqs_sr_step_holomorphic_full(cfg, size_x, size_y, ansatz, sampler, log_amp_fn,
                                      log_amp_user, nullptr, nullptr, out_info);
}

NQS_EXPORT int nqs_sr_run_holomorphic_full(const nqs_config_t* cfg, int size_x, int size_y,
                                           nqs_ansatz_t* ansatz, nqs_sampler_t* sampler,
                                           nqs_log_amp_fn_t log_amp_fn, void* log_amp_user,
                                           nqs_holomorphic_gradient_fn_t grad_fn, void* grad_user,
                                           double* out_energy_trace) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!cfg || !ansatz || !sampler) return -1;
  return sr_run_impl(SR_HOLO, cfg, size_x, size_y, ansatz, sampler, log_amp_fn, log_amp_user,
                     nullptr, grad_fn, grad_user, out_energy_trace);
}

NQS_EXPORT int nqs_sr_run_holomorphic(const nqs_config_t* cfg, int size_x, int size_y,
                                      nqs_ansatz_t* ansatz, nqs_sampler_t* sampler,
                                      nqs_log_amp_fn_t log_amp_fn, void* log_amp_user,
                                      double* out_energy_trace) {
  return nqs_sr_run_holomorphic_full(cfg, size_x, size_y, ansatz, sampler, log_amp_fn,
                                     log_amp_user, nullptr, nullptr, out_energy_trace);
}

NQS_EXPORT int nqs_sr_step_minsr_full(const nqs_config_t* cfg, int size_x, int size_y,
                                      nqs_ansatz_t* ansatz, nqs_sampler_t* sampler,
                                      nqs_log_amp_fn_t log_amp_fn, void* log_amp_user,
                                      nqs_gradient_fn_t gradient_fn, void* grad_user,
                                      nqs_sr_step_info_t* out_info) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!cfg || !ansatz || !sampler) return -1;
  return sr_step_impl(SR_MINSR, cfg, size_x, size_y, ansatz, sampler, log_amp_fn, log_amp_user,
                      gradient_fn, nullptr, grad_user, out_info, false);
}

NQS_EXPORT int nqs_sr_run_minsr(const nqs_config_t* cfg, int size_x, int size_y,
                                nqs_ansatz_t* ansatz, nqs_sampler_t* sampler,
                                nqs_log_amp_fn_t log_amp_fn, void* log_amp_user,
                                double* out_energy_trace) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!cfg || !ansatz || !sampler) return -1;
  return sr_run_impl(SR_MINSR, cfg, size_x, size_y, ansatz, sampler, log_amp_fn, log_amp_user,
                     nullptr, nullptr, nullptr, out_energy_trace);
}

NQS_EXPORT int nqs_sr_step_excited(const nqs_config_t* cfg, int size_x, int size_y,
                                   nqs_ansatz_t* ansatz, nqs_sampler_t* sampler,
                                   nqs_log_amp_fn_t log_amp_fn, void* log_amp_user,
                                   nqs_log_amp_fn_t ref_log_amp_fn, void* ref_log_amp_user,
                                   double penalty_mu, nqs_sr_step_info_t* out_info) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!cfg || !ansatz || !sampler || !ref_log_amp_fn) return -1;
  DynStep d;
  d.kind = 1;
  d.mu = penalty_mu;
  d.ref_fn = ref_log_amp_fn;
  d.ref_user = ref_log_amp_user;
  return sr_step_impl(SR_HOLO, cfg, size_x, size_y, ansatz, sampler, log_amp_fn, log_amp_user,
                      nullptr, nullptr, nullptr, out_info, false, &d);
}

NQS_EXPORT int nqs_sr_run_excited(const nqs_config_t* cfg, int size_x, int size_y,
                                  nqs_ansatz_t* ansatz, nqs_sampler_t* sampler,
                                  nqs_log_amp_fn_t log_amp_fn, void* log_amp_user,
                                  nqs_log_amp_fn_t ref_log_amp_fn, void* ref_log_amp_user,
                                  double penalty_mu, double* out_energy_trace) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!cfg || !ansatz || !sampler || !ref_log_amp_fn) return -1;
  DynStep d;
  d.kind = 1;
  d.mu = penalty_mu;
  d.ref_fn = ref_log_amp_fn;
  d.ref_user = ref_log_amp_user;
  return sr_run_impl(SR_HOLO, cfg, size_x, size_y, ansatz, sampler, log_amp_fn, log_amp_user,
                     nullptr, nullptr, nullptr, out_energy_trace, &d);
}

NQS_EXPORT int nqs_tvmc_step_real_time(const nqs_config_t* cfg, double dt, int size_x, int size_y,
                                       nqs_ansatz_t* ansatz, nqs_sampler_t* sampler,
                                       nqs_log_amp_fn_t log_amp_fn, void* log_amp_user,
                                       nqs_sr_step_info_t* out_info) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!cfg || !ansatz || !sampler) return -1;
  DynStep d;
  d.kind = 2;
  d.dt = dt;
  return sr_step_impl(SR_HOLO, cfg, size_x, size_y, ansatz, sampler, log_amp_fn, log_amp_user,
                      nullptr, nullptr, nullptr, out_info, false, &d);
}

// Heun (nqs_optimizer.c:830-902): k1 = delta(theta) and k2 = delta(theta +
// dt k1), each from its own batch, then theta <- theta + dt (k1 + k2) / 2.
// The parameters stay on the device between the stages; the batch has the
// sampler's site count (the reference sizes its Heun batch Lx * Ly, the
// same number on square lattices).
NQS_EXPORT int nqs_tvmc_step_heun(const nqs_config_t* cfg, double dt, int size_x, int size_y,
                                  nqs_ansatz_t* ansatz, nqs_sampler_t* sampler,
                                  nqs_log_amp_fn_t log_amp_fn, void* log_amp_user,
                                  nqs_sr_step_info_t* out_info) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!cfg || !ansatz || !sampler) return -1;
  nqs_ansatz* a = ansatz;
  const long P = a->P;
  if (P <= 0 || stream_ready()) return -1;
  CKR(g_ws.heun.ensure(sizeof(double) * 3 * (size_t)P));
  double* th0 = g_ws.heun.as<double>();
  double* k1 = th0 + P;
  double* k2 = th0 + 2 * P;
  a->dev_fresh = 0;
  CKR(ansatz_upload(a));
  a->dev_fresh = 1;
  CK(hipMemcpyAsync(th0, a->d_par.p, sizeof(double) * P, hipMemcpyDeviceToDevice, g_stream));
  DynStep d;
  d.kind = 2;
  d.dt = dt;
  d.apply = false;
  d.keep_delta = k1;
  nqs_sr_step_info_t i1, i2;
  int rc = sr_step_impl(SR_HOLO, cfg, size_x, size_y, a, sampler, log_amp_fn, log_amp_user,
                        nullptr, nullptr, nullptr, &i1, true, &d);
  if (rc == 0) {
    CKR(nqs::launch_axpy(a->d_par.as<double>(), k1, dt, P, g_stream));
    a->m64.valid = 0;
    a->m32.valid = 0;
    d.keep_delta = k2;
    rc = sr_step_impl(SR_HOLO, cfg, size_x, size_y, a, sampler, log_amp_fn, log_amp_user, nullptr,
                      nullptr, nullptr, &i2, true, &d);
  }
  // theta0 back, then + dt (k1 + k2) / 2 (k1 + k2 formed first, scaled by
  // dt / 2: the same rounding as dt ((k1 + k2) / 2))
  CK(hipMemcpyAsync(a->d_par.p, th0, sizeof(double) * P, hipMemcpyDeviceToDevice, g_stream));
  a->m64.valid = 0;
  a->m32.valid = 0;
  if (rc == 0) {
    CKR(nqs::launch_axpy(k1, k2, 1.0, P, g_stream));
    CKR(nqs::launch_axpy(a->d_par.as<double>(), k1, 0.5 * dt, P, g_stream));
    double* sc = g_ws.scal.as<double>();
    CKR(nqs::launch_sum(k1, P, k1, 1.0, sc + 13, g_stream));
    double n2 = 0.0;
    CK(hipMemcpyAsync(&n2, sc + 13, sizeof(double), hipMemcpyDeviceToHost, g_stream));
    CKR(ansatz_download(a));
    if (out_info) {
      *out_info = i1;
      out_info->update_norm = 0.5 * sqrt(n2) * fabs(dt);
      out_info->acceptance_ratio = nqs_sampler_acceptance_ratio(sampler);
      out_info->cg_iterations = i1.cg_iterations + i2.cg_iterations;
      out_info->converged = i1.converged && i2.converged;
    }
  } else {
    (void)ansatz_download(a);
  }
  a->dev_fresh = 0;
  return rc;
}

// ===========================================================================
// extensions
// ===========================================================================
NQS_EXPORT int nqs_hip_sampler_last_energies(const nqs_sampler_t* s, double* out, int capacity) {
  if (!s || !out || s->last_E.empty()) return -1;
  const int C = s->opt.num_chains, R = s->last_rounds, B = (int)s->last_E.size();
  if (capacity < C * R) return -1;
  for (int c = 0; c < C; c++)
    for (int r = 0; r < R; r++) {
      const long b = (long)r * C + c;
      out[(size_t)c * R + r] = b < B ? s->last_E[b] : NAN;
    }
  return R;
}

NQS_EXPORT int nqs_hip_sampler_rounds(const nqs_sampler_t* s) { return s ? s->last_rounds : -1; }

NQS_EXPORT double nqs_hip_tau_int(const double* x, int n) {
  // tau = 1/2 + sum_{t=1}^{W} rho(t), smallest W >= 5 tau(W) (Sokal)
  if (!x || n < 4) return 0.5;
  double m = 0.0;
  for (int i = 0; i < n; i++) m += x[i];
  m /= n;
  double c0 = 0.0;
  for (int i = 0; i < n; i++) c0 += (x[i] - m) * (x[i] - m);
  c0 /= n;
  if (c0 <= 0.0) return 0.5;
  double tau = 0.5;
  for (int t = 1; t < n; t++) {
    double c = 0.0;
    for (int i = 0; i + t < n; i++) c += (x[i] - m) * (x[i + t] - m);
    c /= n;
    tau += c / c0;
    if (t >= 5.0 * tau) break;
  }
  return tau;
}

NQS_EXPORT int nqs_hip_last_stage_ms(const nqs_sampler_t* s, double out[7]) {
  if (!s || !out || !s->have_stage) return -1;
  for (int i = 0; i < 7; i++) out[i] = s->stage_ms[i];
  return 0;
}

NQS_EXPORT int nqs_hip_gram(const double* O, int B, long P, double scale, double* T) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!O || !T || B <= 0 || P <= 0 || stream_ready()) return -1;
  const size_t ob = sizeof(double) * (size_t)B * P, tb = sizeof(double) * (size_t)B * B;
  if (g_ws.O.ensure(ob) || g_ws.T.ensure(tb)) return -1;
  CK(hipMemcpyAsync(g_ws.O.p, O, ob, hipMemcpyHostToDevice, g_stream));
  CKR(nqs::launch_gemm_nt(B, B, (int)P, scale, g_ws.O.as<double>(), P, g_ws.O.as<double>(), P, 0.0,
                          g_ws.T.as<double>(), B, 1, g_stream));
  CKR(nqs::launch_mirror_lower(g_ws.T.as<double>(), B, g_stream));
  CK(hipMemcpyAsync(T, g_ws.T.p, tb, hipMemcpyDeviceToHost, g_stream));
  CK(hipStreamSynchronize(g_stream));
  return 0;
}

NQS_EXPORT int nqs_hip_cholesky_solve(double* T, int n, const double* b, double* x) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!T || !b || !x || n <= 0 || stream_ready()) return -1;
  const size_t tb = sizeof(double) * (size_t)n * n;
  if (g_ws.T.ensure(tb) || g_ws.T2.ensure(tb) || g_ws.y.ensure(sizeof(double) * 3 * n) ||
      g_ws.fail.ensure(16) || g_ws.linv.ensure(sizeof(double) * nqs::potrf_linv_elems(n)))
    return -1;
  double* dT = g_ws.T.as<double>();
  double* y = g_ws.y.as<double>();
  double* Linv = g_ws.linv.as<double>();
  CK(hipMemcpyAsync(dT, T, tb, hipMemcpyHostToDevice, g_stream));
  CK(hipMemcpyAsync(g_ws.T2.p, dT, tb, hipMemcpyDeviceToDevice, g_stream));
  CK(hipMemcpyAsync(y + 2 * n, b, sizeof(double) * n, hipMemcpyHostToDevice, g_stream));
  int f = nqs::launch_potrf(dT, n, Linv, g_ws.fail.as<int>(), g_stream);
  if (f < 0) return -1;
  if (f > 0) {
    CKR(nqs::launch_potrf_repair(dT, n, f - 1, g_ws.T2.as<double>(), g_stream));
    CKR(nqs::launch_add_diag(dT, n, n, 1e-6, g_stream));
    f = nqs::launch_potrf(dT, n, Linv, g_ws.fail.as<int>(), g_stream);
    if (f < 0) return -1;
  }
  if (f == 0) CKR(nqs::launch_potrs(dT, Linv, n, y + 2 * n, y, y + n, g_stream));
  else CKR(nqs::launch_fill(y, 0.0, n, g_stream));
  CK(hipMemcpyAsync(T, dT, tb, hipMemcpyDeviceToHost, g_stream));
  CK(hipMemcpyAsync(x, y, sizeof(double) * n, hipMemcpyDeviceToHost, g_stream));
  CK(hipStreamSynchronize(g_stream));
  return f == 0 ? 1 : 0;
}

// ---- allocation-free Cholesky solve ------------------------------------------
// Device memory of nqs_hip_cholesky_solve_noalloc: one file-scope array that
// every GPU the library runs on holds (its own copy, until the process ends).
// Layout for an n x n system: the failure flag (first 256 bytes), the copy of
// the input the reference's retry repairs from (n^2 doubles), the block
// inverses of the factorisation (potrf_linv_elems(n)), and the solve's two
// vectors (2n), each region 256-byte aligned.  64 MiB serve every n <= 2319.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];

namespace {
constexpr int kMaxDevices = 64;
double* g_scratch_dev[kMaxDevices] = {};

// the array's device address on the current device, looked up once per
// device with that device current (the name itself is a host-side stand-in)
double* scratch_device_ptr() {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    fprintf(stderr, "nqs_hip: no usable current device for the allocation-free scratch\n");
    return nullptr;
  }
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      fprintf(stderr, "nqs_hip: hipGetSymbolAddress(scratch) failed on device %d: %s\n", dev,
              hipGetErrorString(e));
      return nullptr;
    }
    g_scratch_dev[dev] = (double*)p;
  }
  return g_scratch_dev[dev];
}

// offsets (in doubles) of the regions, each 256-byte aligned as hipMalloc's
size_t noalloc_align(size_t d) { return (d + 31) & ~(size_t)31; }
size_t noalloc_linv(int n) { return noalloc_align(32 + (size_t)n * n); }
size_t noalloc_w(int n) { return noalloc_align(noalloc_linv(n) + nqs::potrf_linv_elems(n)); }
size_t noalloc_doubles(int n) { return noalloc_w(n) + 2 * (size_t)n; }
}  // namespace

NQS_EXPORT int nqs_hip_cholesky_solve_noalloc(double* d_T, int n, const double* d_b, double* d_x) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!d_T || !d_b || !d_x || n <= 0) return -1;
  const size_t need = noalloc_doubles(n) * sizeof(double);
  if (need > kNoAllocScratchBytes) {
    int nmax = 1;
    while (noalloc_doubles(nmax + 1) * sizeof(double) <= kNoAllocScratchBytes) nmax++;
    fprintf(stderr, "nqs_hip_cholesky_solve_noalloc: n = %d needs %zu bytes of scratch; the "
                    "library holds %zu (n <= %d)\n", n, need, kNoAllocScratchBytes, nmax);
    return -1;
  }
  if (stream_ready()) return -1;
  double* base = scratch_device_ptr();
  if (!base) return -1;
  int* fail = reinterpret_cast<int*>(base);          // zeroed by launch_potrf's own kernel
  double* orig = base + 32;
  double* Linv = base + noalloc_linv(n);
  double* w = base + noalloc_w(n);
  // the input, for the retry's repair of the partially factored columns
  CK(hipMemcpyAsync(orig, d_T, sizeof(double) * (size_t)n * n, hipMemcpyDeviceToDevice, g_stream));
  int f = nqs::launch_potrf(d_T, n, Linv, fail, g_stream);
  if (f < 0) return -1;
  if (f > 0) {
    CKR(nqs::launch_potrf_repair(d_T, n, f - 1, orig, g_stream));
    CKR(nqs::launch_add_diag(d_T, n, n, 1e-6, g_stream));
    f = nqs::launch_potrf(d_T, n, Linv, fail, g_stream);
    if (f < 0) return -1;
  }
  if (f == 0) CKR(nqs::launch_potrs(d_T, Linv, n, d_b, d_x, w, g_stream));
  else CKR(nqs::launch_fill(d_x, 0.0, n, g_stream));
  CK(hipStreamSynchronize(g_stream));
  return f == 0 ? 1 : 0;
}

// ===========================================================================
// sample diagnostics (src/nqs/nqs_diagnostics.c)
// ===========================================================================
namespace {

// kagome bonds in the reference's enumeration (nqs_diagnostics.c:141-197,
// nqs_lanczos.c:343-371): per cell (cx, cy), cx-major, the up triangle
// (A,B), (A,C), (B,C), then, with periodic cells or cx, cy >= 1, the down
// triangle (A,Bm), (A,Cm), (Bm,Cm) with Bm the B site of cell (cx-1, cy) and
// Cm the C site of cell (cx, cy-1)
std::vector<std::array<int, 2>> kagome_bonds(int Lx, int Ly, int pbc) {
  std::vector<std::array<int, 2>> out;
  for (int cx = 0; cx < Lx; cx++)
    for (int cy = 0; cy < Ly; cy++) {
      const int A = 3 * (cx * Ly + cy);
      out.push_back({A, A + 1});
      out.push_back({A, A + 2});
      out.push_back({A + 1, A + 2});
      if (!pbc && (cx == 0 || cy == 0)) continue;
      const int Bm = 3 * (((cx - 1 + Lx) % Lx) * Ly + cy) + 1;
      const int Cm = 3 * (cx * Ly + (cy - 1 + Ly) % Ly) + 2;
      out.push_back({A, Bm});
      out.push_back({A, Cm});
      out.push_back({Bm, Cm});
    }
  return out;
}

// the sampler's next cfg->num_samples configurations, as nqs_sampler_batch
// draws them, into s->d_samples (and h_out when given)
int diag_batch(nqs_sampler* s, int B, int* h_out) {
  const int C = s->opt.num_chains;
  const int k = s->cfg.num_decorrelate > 0 ? s->cfg.num_decorrelate : 1;
  const int rounds = (B + C - 1) / C;
  CKR(s->d_samples.ensure(sizeof(int) * (size_t)B * s->N));
  std::vector<int> tmp;
  if (!h_out && !s->device) {
    tmp.resize((size_t)B * s->N);
    h_out = tmp.data();
  }
  CKR(sampler_advance(s, k, rounds, B, s->d_samples.as<int>(), h_out));
  return 0;
}

}  // namespace

// Tr(S) of the quantum geometric tensor over a fresh batch
// (nqs_diagnostics.c:14-92): holomorphic O rows of every sample on the
// device, column moments in a fixed order.
NQS_EXPORT int nqs_compute_chi_F(const nqs_config_t* cfg, int size_x, int size_y,
                                 nqs_ansatz_t* a, nqs_sampler_t* s, double* out_trace_S,
                                 double* out_per_param) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  (void)size_x;
  (void)size_y;
  if (!cfg || !a || !s || !out_trace_S) return -1;
  const long P = a->P;
  const int B = cfg->num_samples;
  if (P <= 0 || B <= 0) return -1;
  if (a->N != s->N) {
    fprintf(stderr, "nqs_compute_chi_F: the ansatz has %d sites and the sampler %d\n", a->N, s->N);
    return -1;
  }
  if (stream_ready()) return -1;
  CKR(diag_batch(s, B, nullptr));
  a->dev_fresh = 0;
  CKR(ansatz_upload(a));
  const int* d_samples = s->d_samples.as<int>();
  CKR(forward_t<double>(a, d_samples, B, nullptr, nullptr));
  const bool cplx = a->kind == NQS_ANSATZ_COMPLEX_RBM;
  CKR(g_ws.O.ensure(sizeof(double) * (size_t)B * (size_t)P));
  if (cplx) CKR(g_ws.Oi.ensure(sizeof(double) * (size_t)B * (size_t)P));
  double* Or = g_ws.O.as<double>();
  double* Oi = cplx ? g_ws.Oi.as<double>() : nullptr;
  ModelView<double> v;
  CKR(model_view<double>(a, v));
  CKR(nqs::launch_gen_O<double>(v, a->kind, d_samples, B, g_ws.t_re.as<double>(),
                                g_ws.t_im.as<double>(), cplx ? 1 : 0, Or, Oi, P, g_stream));
  CKR(g_ws.colwork.ensure(sizeof(double) * nqs::chi_work_elems(P)));
  CKR(g_ws.scal.ensure(sizeof(double) * 16));
  CKR(nqs::launch_chi_trace(Or, Oi, B, P, g_ws.colwork.as<double>(), g_ws.scal.as<double>(),
                            g_stream));
  double tr = 0.0;
  CK(hipMemcpyAsync(&tr, g_ws.scal.p, sizeof(double), hipMemcpyDeviceToHost, g_stream));
  CK(hipStreamSynchronize(g_stream));
  *out_trace_S = tr;
  if (out_per_param) *out_per_param = tr / (double)P;
  return 0;
}

// Per-bond-class mean of psi(s_ij)/psi(s) over the opposite-spin kagome
// bonds of a fresh batch (nqs_diagnostics.c:113-222).  Bonds in the
// reference's order: per cell (cx, cy), cx-major, the up triangle (A,B),
// (A,C), (B,C), then, with periodic cells or cx, cy >= 1, the down triangle
// (A,Bm), (A,Cm), (Bm,Cm) with Bm the B site of cell (cx-1, cy) and Cm the C
// site of cell (cx, cy-1).  Class of a bond: A-B 0, A-C 1, B-C 2.  The
// ratios of all (bond, sample) pairs come from the E_loc pair kernel; they
// are added per class in the reference's order (sample, then bond).
NQS_EXPORT int nqs_compute_kagome_bond_phase(const nqs_config_t* cfg, int Lx_cells, int Ly_cells,
                                             nqs_ansatz_t* a, nqs_sampler_t* s,
                                             double* out_mean_re, double* out_mean_im,
                                             long* out_counts) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!cfg || !a || !s || !out_mean_re || !out_mean_im) return -1;
  if (cfg->hamiltonian != NQS_HAM_KAGOME_HEISENBERG) return -2;
  if (Lx_cells < 1 || Ly_cells < 1) return -1;
  const int N = 3 * Lx_cells * Ly_cells, pbc = cfg->kagome_pbc;
  if (s->N != N) return -1;
  const int B = cfg->num_samples;
  if (B <= 0) return -1;
  if (a->N != N) {
    fprintf(stderr, "nqs_compute_kagome_bond_phase: the ansatz has %d sites, the lattice %d\n",
            a->N, N);
    return -1;
  }
  if (stream_ready()) return -1;
  nqs::HamHost hh;
  hh.n_sites = N;
  std::vector<int> cls;
  auto sub_class = [](int u, int w) {
    const int lo = std::min(u % 3, w % 3), hi = std::max(u % 3, w % 3);
    return lo == 0 ? (hi == 1 ? 0 : 1) : 2;
  };
  for (const auto& e : kagome_bonds(Lx_cells, Ly_cells, pbc)) {
    hh.pair_a.push_back(e[0]);
    hh.pair_b.push_back(e[1]);
  }
  const int T = (int)hh.pair_a.size();
  for (int t = 0; t < T; t++) cls.push_back(sub_class(hh.pair_a[t], hh.pair_b[t]));
  hh.pair_c.assign(T, 1.0);
  std::vector<int> h_samples((size_t)B * N);
  CKR(diag_batch(s, B, h_samples.data()));
  AmpSource src;
  src.a = a;
  nqs::HamDev H;
  CKR(ham_upload(hh, src, g_ws.bond_ham, g_ws.bond_sub, H));
  a->dev_fresh = 0;
  CKR(ansatz_upload(a));
  ModelView<double> v;
  CKR(ratio_view<double>(a, v));
  const int* d_samples = s->d_samples.as<int>();
  CKR(forward_t<double>(a, d_samples, B, nullptr, nullptr, true));
  const size_t wb = nqs::eloc_ws_bytes<double>(v, H, B);
  CKR(g_ws.elw.ensure(wb));
  double* res[2] = {nullptr, nullptr};
  CKR(nqs::launch_pair_ratios<double>(v, H, d_samples, B, g_ws.t_re.as<double>(),
                                      g_ws.q_re.as<double>(), g_ws.q_im.as<double>(), g_ws.elw.p,
                                      wb, res, g_stream));
  std::vector<double> rr((size_t)T * B), ri((size_t)T * B);
  CK(hipMemcpyAsync(rr.data(), res[0], sizeof(double) * rr.size(), hipMemcpyDeviceToHost,
                    g_stream));
  CK(hipMemcpyAsync(ri.data(), res[1], sizeof(double) * ri.size(), hipMemcpyDeviceToHost,
                    g_stream));
  CK(hipStreamSynchronize(g_stream));
  double sr[3] = {0, 0, 0}, si[3] = {0, 0, 0};
  long cnt[3] = {0, 0, 0};
  for (int b = 0; b < B; b++) {
    const int* sp = h_samples.data() + (size_t)b * N;
    for (int t = 0; t < T; t++) {
      if (sp[hh.pair_a[t]] == sp[hh.pair_b[t]]) continue;
      sr[cls[t]] += rr[(size_t)t * B + b];
      si[cls[t]] += ri[(size_t)t * B + b];
      cnt[cls[t]]++;
    }
  }
  for (int c = 0; c < 3; c++) {
    out_mean_re[c] = cnt[c] ? sr[c] / (double)cnt[c] : 0.0;
    out_mean_im[c] = cnt[c] ? si[c] / (double)cnt[c] : 0.0;
    if (out_counts) out_counts[c] = cnt[c];
  }
  return 0;
}

// ===========================================================================
// symmetry projection (src/nqs/nqs_symproj.c, src/nqs/nqs_translation.c)
// ===========================================================================
namespace {

// device buffer released at scope exit
struct ScopedBuf {
  DBuf b;
  ~ScopedBuf() { b.release(); }
};

// (lp, arg) of the G orbit members g.s of one configuration: one device
// batch when the base amplitude is a device source, else G callback calls
int orbit_member_amps(nqs_log_amp_fn_t bfn, void* buser, const int* perm, int G, int N,
                      const int* spins, std::vector<int>& rows, std::vector<double>& lp,
                      std::vector<double>& ar) {
  rows.resize((size_t)G * N);
  for (int g = 0; g < G; g++)
    for (int i = 0; i < N; i++) {
      const int j = perm[(size_t)g * N + i];
      if (j < 0 || j >= N) return -1;
      rows[(size_t)g * N + i] = spins[j];
    }
  lp.assign(G, 0.0);
  ar.assign(G, 0.0);
  AmpSource bs;
  if (recognise(bfn, buser, bs) && bs.a->N == N) {
    if (stream_ready()) return -1;
    CKR(ansatz_upload(bs.a));
    ModelView<double> v;
    CKR(model_view<double>(bs.a, v));
    CKR(g_ws.orb_cfg.ensure(sizeof(int) * rows.size()));
    CKR(g_ws.orb_lp.ensure(sizeof(double) * 2 * G));
    double* d = g_ws.orb_lp.as<double>();
    CK(hipMemcpyAsync(g_ws.orb_cfg.p, rows.data(), sizeof(int) * rows.size(),
                      hipMemcpyHostToDevice, g_stream));
    CKR(nqs::launch_forward<double>(v, g_ws.orb_cfg.as<int>(), G, nullptr, nullptr, d, d + G,
                                    g_stream));
    if (bs.marshall)
      CKR(nqs::launch_marshall_arg(g_ws.orb_cfg.as<int>(), G, N, bs.mlx, bs.mly, d + G, g_stream));
    CK(hipMemcpyAsync(lp.data(), d, sizeof(double) * G, hipMemcpyDeviceToHost, g_stream));
    CK(hipMemcpyAsync(ar.data(), d + G, sizeof(double) * G, hipMemcpyDeviceToHost, g_stream));
    CK(hipStreamSynchronize(g_stream));
    return 0;
  }
  for (int g = 0; g < G; g++) {
    bfn(rows.data() + (size_t)g * N, N, buser, &lp[g], &ar[g]);
    g_host_calls++;
  }
  return 0;
}

// the real combination (nqs_symproj.c:56-79)
void orbit_combine_host(const std::vector<double>& lp, const std::vector<double>& ar,
                        const double* chi, int G, double* out_la, double* out_arg) {
  double m = -1e300;
  for (int g = 0; g < G; g++) m = lp[g] > m ? lp[g] : m;
  double S = 0.0;
  for (int g = 0; g < G; g++) S += chi[g] * cos(ar[g]) * exp(lp[g] - m);
  if (fabs(S) > 0.0) {
    if (out_la) *out_la = log(fabs(S)) + m - 0.5 * log((double)G);
    if (out_arg) *out_arg = S < 0.0 ? M_PI : 0.0;
  } else {
    if (out_la) *out_la = -1e300;
    if (out_arg) *out_arg = 0.0;
  }
}

// sum_g w_g grad log psi_a(g.s): weights from the orbit members' base
// amplitudes (real or complex), base gradients of the G rows on the device,
// accumulated on the host in g order (nqs_symproj.c:120-150, 220-262)
int orbit_gradient_one(nqs_log_amp_fn_t bfn, void* buser, const int* perm, const double* chi,
                       int G, nqs_ansatz* a, const int* spins, int N, bool cplx, double* out_re,
                       double* out_im) {
  std::vector<int> rows;
  std::vector<double> lp, ar;
  CKR(orbit_member_amps(bfn, buser, perm, G, N, spins, rows, lp, ar));
  const long P = a->P;
  double m = -1e300;
  for (int g = 0; g < G; g++) m = lp[g] > m ? lp[g] : m;
  std::vector<double> cr(G), ci(G, 0.0);
  double dr = 0.0, di = 0.0;
  for (int g = 0; g < G; g++) {
    if (cplx) {
      const double mag = chi[g] * exp(lp[g] - m);
      cr[g] = mag * cos(ar[g]);
      ci[g] = mag * sin(ar[g]);
    } else {
      cr[g] = chi[g] * cos(ar[g]) * exp(lp[g] - m);
    }
    dr += cr[g];
    di += ci[g];
  }
  std::fill(out_re, out_re + P, 0.0);
  if (cplx) std::fill(out_im, out_im + P, 0.0);
  const double d2 = dr * dr + di * di;
  if ((cplx && d2 == 0.0) || (!cplx && dr == 0.0)) return 0;
  // base gradients of all G rows in one device batch
  if (stream_ready()) return -1;
  CKR(ansatz_upload(a));
  CKR(upload_spins(rows.data(), (long)rows.size()));
  CKR(forward_t<double>(a, g_ws.spins.as<int>(), G, nullptr, nullptr));
  ModelView<double> v;
  CKR(model_view<double>(a, v));
  CKR(g_ws.orb_O.ensure(sizeof(double) * (size_t)G * P));
  if (cplx) CKR(g_ws.orb_Oi.ensure(sizeof(double) * (size_t)G * P));
  CKR(nqs::launch_gen_O<double>(v, a->kind, g_ws.spins.as<int>(), G, g_ws.t_re.as<double>(),
                                g_ws.t_im.as<double>(), cplx ? 1 : 0, g_ws.orb_O.as<double>(),
                                cplx ? g_ws.orb_Oi.as<double>() : nullptr, P, g_stream));
  std::vector<double> gr((size_t)G * P), gi(cplx ? (size_t)G * P : 0);
  CK(hipMemcpyAsync(gr.data(), g_ws.orb_O.p, sizeof(double) * gr.size(), hipMemcpyDeviceToHost,
                    g_stream));
  if (cplx)
    CK(hipMemcpyAsync(gi.data(), g_ws.orb_Oi.p, sizeof(double) * gi.size(), hipMemcpyDeviceToHost,
                      g_stream));
  CK(hipStreamSynchronize(g_stream));
  for (int g = 0; g < G; g++) {
    double wr, wi = 0.0;
    if (cplx) {
      wr = (cr[g] * dr + ci[g] * di) / d2;
      wi = (ci[g] * dr - cr[g] * di) / d2;
    } else {
      wr = cr[g] / dr;
    }
    if (wr == 0.0 && wi == 0.0) continue;
    const double* r = gr.data() + (size_t)g * P;
    if (!cplx) {
      for (long k = 0; k < P; k++) out_re[k] += wr * r[k];
      continue;
    }
    const double* q = gi.data() + (size_t)g * P;
    for (long k = 0; k < P; k++) {
      out_re[k] += wr * r[k] - wi * q[k];
      out_im[k] += wr * q[k] + wi * r[k];
    }
  }
  return 0;
}

}  // namespace

NQS_EXPORT void nqs_symproj_log_amp(const int* spins, int num_sites, void* user,
                                    double* out_log_abs, double* out_arg) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  const nqs_symproj_wrapper_t* w = (const nqs_symproj_wrapper_t*)user;
  auto zero = [&] {
    if (out_log_abs) *out_log_abs = 0.0;
    if (out_arg) *out_arg = 0.0;
  };
  if (!w || !w->base_log_amp || !w->perm || !w->characters || !spins) return zero();
  const int N = w->num_sites, G = w->num_group_elements;
  if (num_sites != N || N <= 0 || G <= 0) return zero();
  std::vector<int> rows;
  std::vector<double> lp, ar;
  if (orbit_member_amps(w->base_log_amp, w->base_user, w->perm, G, N, spins, rows, lp, ar))
    return zero();
  orbit_combine_host(lp, ar, w->characters, G, out_log_abs, out_arg);
}

NQS_EXPORT int nqs_symproj_gradient(void* grad_user, nqs_ansatz_t* ansatz, const int* spins,
                                    int num_sites, double* out_grad) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!grad_user || !ansatz || !spins || !out_grad) return -1;
  const nqs_symproj_wrapper_t* w = (const nqs_symproj_wrapper_t*)grad_user;
  const int N = w->num_sites, G = w->num_group_elements;
  if (num_sites != N || N <= 0 || G <= 0 || ansatz->N != N) return -1;
  return orbit_gradient_one(w->base_log_amp, w->base_user, w->perm, w->characters, G, ansatz,
                            spins, N, false, out_grad, nullptr);
}

NQS_EXPORT int nqs_symproj_gradient_complex(void* grad_user, nqs_ansatz_t* ansatz,
                                            const int* spins, int num_sites, double* out_grad_re,
                                            double* out_grad_im) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!grad_user || !ansatz || !spins || !out_grad_re || !out_grad_im) return -1;
  const nqs_symproj_wrapper_t* w = (const nqs_symproj_wrapper_t*)grad_user;
  const int N = w->num_sites, G = w->num_group_elements;
  if (num_sites != N || N <= 0 || G <= 0 || ansatz->N != N) return -1;
  return orbit_gradient_one(w->base_log_amp, w->base_user, w->perm, w->characters, G, ansatz,
                            spins, N, true, out_grad_re, out_grad_im);
}

NQS_EXPORT void nqs_translation_log_amp(const int* spins, int num_sites, void* user,
                                        double* out_log_abs, double* out_arg) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  const nqs_translation_wrapper_t* w = (const nqs_translation_wrapper_t*)user;
  auto zero = [&] {
    if (out_log_abs) *out_log_abs = 0.0;
    if (out_arg) *out_arg = 0.0;
  };
  if (!w || !w->base_log_amp || !spins || w->size_x <= 0 || w->size_y <= 0) return zero();
  const int N = w->size_x * w->size_y;
  if (num_sites != N) return zero();
  const TranslationTable& t = translation_table(w->size_x, w->size_y);
  std::vector<int> rows;
  std::vector<double> lp, ar;
  if (orbit_member_amps(w->base_log_amp, w->base_user, t.perm.data(), N, N, spins, rows, lp, ar))
    return zero();
  orbit_combine_host(lp, ar, t.chi.data(), N, out_log_abs, out_arg);
}

NQS_EXPORT int nqs_translation_gradient(void* grad_user, nqs_ansatz_t* ansatz, const int* spins,
                                        int num_sites, double* out_grad) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!grad_user || !ansatz || !spins || !out_grad) return -1;
  const nqs_translation_wrapper_t* w = (const nqs_translation_wrapper_t*)grad_user;
  if (w->size_x <= 0 || w->size_y <= 0) return -1;
  const int N = w->size_x * w->size_y;
  if (num_sites != N || ansatz->N != N) return -1;
  const TranslationTable& t = translation_table(w->size_x, w->size_y);
  return orbit_gradient_one(w->base_log_amp, w->base_user, t.perm.data(), t.chi.data(), N, ansatz,
                            spins, N, false, out_grad, nullptr);
}

namespace {

// the projector's bit tables: bit i of T(g) s is bit perm[g][i] of s, so
// bit j of s lands at position pinv_g[j]; tab [G][nchunk][8] holds, for
// every 3-bit chunk c of s and its value v, the OR of the bits it sets in
// T(g) s.  strict: every row must be a permutation of 0..N-1.
int proj_tables(const int* perm, int G, int N, bool strict, const char* who,
                std::vector<unsigned long long>& tab, bool* identity) {
  const int nchunk = (N + 2) / 3;
  tab.assign((size_t)G * nchunk * 8, 0ULL);
  bool id = true;
  for (int g = 0; g < G; g++) {
    std::vector<int> pinv(N, -1);
    for (int i = 0; i < N; i++) {
      const int j = perm[(size_t)g * N + i];
      if (j < 0 || j >= N || (strict && pinv[j] >= 0)) {
        fprintf(stderr, "%s: perm[%d] is not a permutation of 0..%d\n", who, g, N - 1);
        return -1;
      }
      pinv[j] = i;
      if (j != i) id = false;
    }
    for (int c = 0; c < nchunk; c++)
      for (int v = 0; v < 8; v++) {
        unsigned long long m = 0;
        for (int b = 0; b < 3; b++)
          if (((v >> b) & 1) && 3 * c + b < N && pinv[3 * c + b] >= 0)
            m |= 1ULL << pinv[3 * c + b];
        tab[((size_t)g * nchunk + c) * 8 + v] = m;
      }
  }
  if (identity) *identity = id;
  return 0;
}

}  // namespace

// psi <- (1/G) sum_g chi_g T(g) psi on the device (the ED projector)
NQS_EXPORT void nqs_kagome_p6m_project_inplace(double* psi, int N, int G, const int* perm,
                                               const double* characters) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!psi || !perm || !characters || N <= 0 || N > 30 || G <= 0) return;
  const int nchunk = (N + 2) / 3;
  std::vector<unsigned long long> tab;
  if (proj_tables(perm, G, N, false, "nqs_kagome_p6m_project_inplace", tab, nullptr)) return;
  if (stream_ready()) return;
  const long dim = 1L << N;
  ScopedBuf in, out, part, scal, dtab, dchi;
  if (in.b.ensure(sizeof(double) * dim) || out.b.ensure(sizeof(double) * dim) ||
      part.b.ensure(sizeof(double) * nqs::ed_partials()) || scal.b.ensure(64) ||
      dtab.b.ensure(sizeof(unsigned long long) * tab.size()) || dchi.b.ensure(sizeof(double) * G))
    return;
  if (hipMemcpyAsync(in.b.p, psi, sizeof(double) * dim, hipMemcpyHostToDevice, g_stream) ||
      hipMemcpyAsync(dtab.b.p, tab.data(), sizeof(unsigned long long) * tab.size(),
                     hipMemcpyHostToDevice, g_stream) ||
      hipMemcpyAsync(dchi.b.p, characters, sizeof(double) * G, hipMemcpyHostToDevice, g_stream))
    return;
  if (nqs::launch_ed_project(in.b.as<double>(), out.b.as<double>(), dim,
                             dtab.b.as<unsigned long long>(), dchi.b.as<double>(), G, nchunk,
                             part.b.as<double>(), scal.b.as<double>(), g_stream))
    return;
  if (hipMemcpyAsync(psi, out.b.p, sizeof(double) * dim, hipMemcpyDeviceToHost, g_stream) ||
      hipStreamSynchronize(g_stream))
    return;
}

// ===========================================================================
// exact diagonalisation on the 2^N basis (src/nqs/nqs_lanczos.c,
// src/mps/lanczos.c)
// ===========================================================================
namespace {

// number of eigenvalues below x of the symmetric tridiagonal matrix with
// diagonal d[0..K-1] and off-diagonal e[0..K-2]: the negative pivots of the
// LDL^T factorisation of T - x I (Sturm count)
int sturm_below(const double* d, const double* e, int K, double x) {
  int c = 0;
  double q = d[0] - x;
  for (int i = 0;;) {
    if (q < 0.0) c++;
    if (++i >= K) break;
    if (q == 0.0) q = 1e-300;
    q = d[i] - x - e[i - 1] * e[i - 1] / q;
  }
  return c;
}

// the idx-th smallest eigenvalue (0-based) of that matrix: bisection on the
// Sturm count inside the Gershgorin interval, down to adjacent doubles
double tridiag_eig(const double* d, const double* e, int K, int idx) {
  double lo = 1e300, hi = -1e300;
  for (int i = 0; i < K; i++) {
    const double r = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i + 1 < K ? fabs(e[i]) : 0.0);
    lo = std::min(lo, d[i] - r);
    hi = std::max(hi, d[i] + r);
  }
  const double pad = 1e-12 * (hi - lo) + 1e-300;
  lo -= pad;
  hi += pad;
  for (int it = 0; it < 2000; it++) {
    const double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    if (sturm_below(d, e, K, mid) >= idx + 1) hi = mid;
    else lo = mid;
  }
  return 0.5 * (lo + hi);
}

double tridiag_smallest(const double* d, const double* e, int K) {
  return tridiag_eig(d, e, K, 0);
}

// unit eigenvector z of that matrix for its eigenvalue lam by inverse
// iteration: Gaussian elimination of T - lam I with partial pivoting (U
// with two superdiagonals; a zero pivot becomes eps |T|), three solves from
// a fixed start vector, each rescaled by its largest entry
void tridiag_vec(const double* d, const double* e, int K, double lam, double* z) {
  if (K == 1) {
    z[0] = 1.0;
    return;
  }
  double nrm = 0.0;
  for (int i = 0; i < K; i++)
    nrm = std::max(nrm, fabs(d[i] - lam) + (i > 0 ? fabs(e[i - 1]) : 0.0) +
                            (i + 1 < K ? fabs(e[i]) : 0.0));
  const double tiny = nrm > 0.0 ? 2.220446049250313e-16 * nrm : 1e-300;
  std::vector<double> u0(K), u1(K, 0.0), u2(K, 0.0), l(K, 0.0);
  std::vector<char> sw(K, 0);
  // p: the pivot candidate row (columns i, i+1, i+2), q: row i+1 as stored
  double p0 = d[0] - lam, p1 = e[0], p2 = 0.0;
  for (int i = 0; i + 1 < K; i++) {
    double q0 = e[i], q1 = d[i + 1] - lam, q2 = i + 2 < K ? e[i + 1] : 0.0;
    if (fabs(q0) > fabs(p0)) {
      std::swap(p0, q0);
      std::swap(p1, q1);
      std::swap(p2, q2);
      sw[i] = 1;
    }
    if (p0 == 0.0) p0 = tiny;
    const double m = q0 / p0;
    l[i] = m;
    u0[i] = p0;
    u1[i] = p1;
    u2[i] = p2;
    p0 = q1 - m * p1;
    p1 = q2 - m * p2;
    p2 = 0.0;
  }
  u0[K - 1] = p0 == 0.0 ? tiny : p0;
  std::vector<double> b(K);
  uint64_t st = 0x9E3779B97F4A7C15ULL;
  for (int i = 0; i < K; i++) {
    st ^= st << 13;
    st ^= st >> 7;
    st ^= st << 17;
    b[i] = 0.5 + (double)(st >> 11) / 9007199254740992.0;
  }
  for (int it = 0; it < 3; it++) {
    for (int i = 0; i + 1 < K; i++) {
      if (sw[i]) std::swap(b[i], b[i + 1]);
      b[i + 1] -= l[i] * b[i];
    }
    for (int i = K - 1; i >= 0; i--) {
      double x = b[i];
      if (i + 1 < K) x -= u1[i] * b[i + 1];
      if (i + 2 < K) x -= u2[i] * b[i + 2];
      b[i] = x / u0[i];
    }
    double mx = 0.0;
    for (int i = 0; i < K; i++) mx = std::max(mx, fabs(b[i]));
    if (!(mx > 0.0) || !std::isfinite(mx)) break;
    for (int i = 0; i < K; i++) b[i] /= mx;
  }
  double n2 = 0.0;
  for (int i = 0; i < K; i++) n2 += b[i] * b[i];
  const double inv = n2 > 0.0 ? 1.0 / sqrt(n2) : 0.0;
  for (int i = 0; i < K; i++) z[i] = b[i] * inv;
}

// the reference's deterministic start vector (src/mps/lanczos.c:160-176):
// xorshift64 from seed ^ dim, element i = (x_i >> 11) / 2^53 - 1/2;
// generated on the host in slices and uploaded
int ed_seed_vector(double* d_v, long dim, unsigned long long seed) {
  const long slice = 1L << 22;
  std::vector<double> h((size_t)std::min(dim, slice));
  unsigned long long x = seed ^ (unsigned long long)dim;
  for (long s0 = 0; s0 < dim; s0 += slice) {
    const long n = std::min(slice, dim - s0);
    for (long i = 0; i < n; i++) {
      x ^= x << 13;
      x ^= x >> 7;
      x ^= x << 17;
      h[(size_t)i] = (double)(x >> 11) / 9007199254740992.0 - 0.5;
    }
    CK(hipMemcpyAsync(d_v + s0, h.data(), sizeof(double) * n, hipMemcpyHostToDevice, g_stream));
    CK(hipStreamSynchronize(g_stream));
  }
  return 0;
}

double ed_read(const double* d) {
  double v = 0.0;
  if (hipMemcpyAsync(&v, d, sizeof(double), hipMemcpyDeviceToHost, g_stream) != hipSuccess ||
      hipStreamSynchronize(g_stream) != hipSuccess)
    return NAN;
  return v;
}

// H on the device: bond list and coefficients
struct EdHam {
  ScopedBuf bonds;
  nqs::EdOp op;
};

int ed_ham_bonds(const std::vector<std::array<int, 2>>& bl, int N, EdHam& h) {
  std::vector<int> flat;
  for (const auto& e : bl) {
    flat.push_back(e[0]);
    flat.push_back(e[1]);
  }
  CKR(h.bonds.b.ensure(sizeof(int) * flat.size() + 16));
  if (!flat.empty())
    CK(hipMemcpyAsync(h.bonds.b.p, flat.data(), sizeof(int) * flat.size(), hipMemcpyHostToDevice,
                      g_stream));
  CK(hipStreamSynchronize(g_stream));
  h.op.bonds = h.bonds.b.as<int>();
  h.op.nb = (int)bl.size();
  h.op.N = N;
  return 0;
}

// kagome Heisenberg J S_u . S_v (nqs_lanczos.c:336-432)
int ed_ham_kagome(int Lx, int Ly, double J, int pbc, EdHam& h) {
  CKR(ed_ham_bonds(kagome_bonds(Lx, Ly, pbc), 3 * Lx * Ly, h));
  h.op.jd = 0.25 * J;
  h.op.jo = 0.5 * J;
  h.op.field = 0.0;
  return 0;
}

// open Lx x Ly square lattice, site x Ly + y, bonds to (x + 1, y), (x, y + 1)
std::vector<std::array<int, 2>> square_open_bonds(int Lx, int Ly) {
  std::vector<std::array<int, 2>> bl;
  for (int x = 0; x < Lx; x++)
    for (int y = 0; y < Ly; y++) {
      if (x + 1 < Lx) bl.push_back({x * Ly + y, (x + 1) * Ly + y});
      if (y + 1 < Ly) bl.push_back({x * Ly + y, x * Ly + y + 1});
    }
  return bl;
}

// XXZ J/2 (S+S- + S-S+) + Jz Sz Sz on the open square lattice
// (nqs_lanczos.c:224-281)
int ed_ham_heisenberg(int Lx, int Ly, double J, double Jz, EdHam& h) {
  CKR(ed_ham_bonds(square_open_bonds(Lx, Ly), Lx * Ly, h));
  h.op.jd = 0.25 * Jz;
  h.op.jo = 0.5 * J;
  h.op.field = 0.0;
  return 0;
}

// TFIM -J sz sz - Gamma sx on the open square lattice (nqs_lanczos.c:150-191)
int ed_ham_tfim(int Lx, int Ly, double J, double Gamma, EdHam& h) {
  CKR(ed_ham_bonds(square_open_bonds(Lx, Ly), Lx * Ly, h));
  h.op.jd = -J;
  h.op.jo = 0.0;
  h.op.field = -Gamma;
  return 0;
}

// the sector projector P = (1/G) sum_g chi_g T(g) on the device; identity
// when G = 1 with the identity permutation and chi = 1
struct EdProj {
  bool identity = true;
  int G = 1, nchunk = 0;
  ScopedBuf tab, chi;
};

int ed_proj_build(const int* perm, const double* chars, int G, int N, const char* who,
                  EdProj& p) {
  std::vector<unsigned long long> tab;
  bool id = true;
  if (proj_tables(perm, G, N, true, who, tab, &id)) return -1;
  p.identity = id && G == 1 && chars[0] == 1.0;
  p.G = G;
  p.nchunk = (N + 2) / 3;
  if (stream_ready()) return -1;
  CKR(p.tab.b.ensure(sizeof(unsigned long long) * tab.size()));
  CKR(p.chi.b.ensure(sizeof(double) * G));
  CK(hipMemcpyAsync(p.tab.b.p, tab.data(), sizeof(unsigned long long) * tab.size(),
                    hipMemcpyHostToDevice, g_stream));
  CK(hipMemcpyAsync(p.chi.b.p, chars, sizeof(double) * G, hipMemcpyHostToDevice, g_stream));
  CK(hipStreamSynchronize(g_stream));
  return 0;
}

// x <- P x (out of place into tmp, then the pointers swap; in place for the
// identity); returns |x|^2
double ed_project_norm2(const EdProj& P, double*& x, double*& tmp, long dim, double* pt,
                        double* sc) {
  if (P.identity) {
    if (nqs::launch_ed_scale(x, x, 1.0, dim, pt, sc + 1, g_stream)) return NAN;
  } else {
    if (nqs::launch_ed_project(x, tmp, dim, P.tab.b.as<unsigned long long>(),
                               P.chi.b.as<double>(), P.G, P.nchunk, pt, sc + 1, g_stream))
      return NAN;
    std::swap(x, tmp);
  }
  return ed_read(sc + 1);
}

double ed_norm2(const double* x, long dim, double* pt, double* sc) {
  if (nqs::launch_ed_scale(x, const_cast<double*>(x), 1.0, dim, pt, sc + 1, g_stream)) return NAN;
  return ed_read(sc + 1);
}

// log|psi| and arg of every basis state (bit i set <-> s_i = -1) into the
// device arrays la, ar: a device ansatz source, an orbit source, or one
// host call per state
int ed_log_amps(nqs_log_amp_fn_t fn, void* user, int N, double* la, double* ar) {
  const long dim = 1L << N;
  ScopedBuf spins;
  AmpSource src;
  if (recognise(fn, user, src) && src.a->N == N) {
    nqs_ansatz* a = src.a;
    a->dev_fresh = 0;
    CKR(ansatz_upload(a));
    ModelView<double> v;
    CKR(model_view<double>(a, v));
    const long chunk = std::min(dim, 1L << 20);
    CKR(spins.b.ensure(sizeof(int) * (size_t)chunk * N));
    for (long s0 = 0; s0 < dim; s0 += chunk) {
      const long n = std::min(chunk, dim - s0);
      CKR(nqs::launch_ed_state_spins(s0, n, N, spins.b.as<int>(), g_stream));
      CKR(nqs::launch_forward<double>(v, spins.b.as<int>(), (int)n, nullptr, nullptr, la + s0,
                                      ar + s0, g_stream));
      if (src.marshall)
        CKR(nqs::launch_marshall_arg(spins.b.as<int>(), (int)n, N, src.mlx, src.mly, ar + s0,
                                     g_stream));
    }
    return 0;
  }
  if (OrbitSource orb; recognise_orbit(fn, user, orb) && orb.base.a->N == N) {
    const long chunk = std::min(dim, 1L << 16);
    CKR(spins.b.ensure(sizeof(int) * (size_t)chunk * N));
    for (long s0 = 0; s0 < dim; s0 += chunk) {
      const long n = std::min(chunk, dim - s0);
      CKR(nqs::launch_ed_state_spins(s0, n, N, spins.b.as<int>(), g_stream));
      CKR(orbit_log_amp(orb, spins.b.as<int>(), n, la + s0, ar + s0));
    }
    return 0;
  }
  std::vector<double> hl(dim), ha(dim);
  std::vector<int> sp(N);
  for (long s = 0; s < dim; s++) {
    for (int i = 0; i < N; i++) sp[i] = ((s >> i) & 1) ? -1 : 1;
    fn(sp.data(), N, user, &hl[s], &ha[s]);
    g_host_calls++;
  }
  CK(hipMemcpyAsync(la, hl.data(), sizeof(double) * dim, hipMemcpyHostToDevice, g_stream));
  CK(hipMemcpyAsync(ar, ha.data(), sizeof(double) * dim, hipMemcpyHostToDevice, g_stream));
  CK(hipStreamSynchronize(g_stream));
  return 0;
}

// psi = Re psi / |Re psi| over the 2^N states into d_psi, with
// Re psi(s) = cos(arg) exp(log|psi| - max log|psi|) (nqs_lanczos.c:115-149);
// a warning on stderr when Im psi holds more than 1e-6 of the norm (:39-62).
// -1 when N is outside 1..24.
int ed_materialise(nqs_log_amp_fn_t fn, void* user, int N, double* d_psi) {
  if (!fn || N <= 0 || N > 24) return -1;
  const long dim = 1L << N;
  ScopedBuf la, ar, part, scal;
  CKR(la.b.ensure(sizeof(double) * dim));
  CKR(ar.b.ensure(sizeof(double) * dim));
  CKR(part.b.ensure(sizeof(double) * nqs::ed_partials()));
  CKR(scal.b.ensure(sizeof(double) * 8));
  CKR(ed_log_amps(fn, user, N, la.b.as<double>(), ar.b.as<double>()));
  double* sc = scal.b.as<double>();
  CKR(nqs::launch_ed_psi(la.b.as<double>(), ar.b.as<double>(), dim, d_psi, part.b.as<double>(),
                         sc, g_stream));
  const double n2 = ed_read(sc + 1), tot = ed_read(sc + 2);
  if (!std::isfinite(n2) || !std::isfinite(tot)) return -1;
  const double frac = tot > 0.0 ? (tot - n2) / tot : 0.0;
  if (frac > 1e-6)
    fprintf(stderr, "nqs_lanczos: warning: Im(psi) holds %.3e of the wavefunction norm; the "
                    "materialised state is Re(psi), which is exact only for stoquastic ground "
                    "states\n", frac);
  const double inv = n2 > 0.0 ? 1.0 / sqrt(n2) : 1.0;
  CKR(nqs::launch_ed_scale(d_psi, d_psi, inv, dim, part.b.as<double>(), nullptr, g_stream));
  CK(hipStreamSynchronize(g_stream));
  return 0;
}

// <psi|H|psi> / <psi|psi> of the materialised state
int ed_exact_energy(nqs_log_amp_fn_t fn, void* user, int N, const EdHam& H, double* out) {
  const long dim = 1L << N;
  ScopedBuf psi, hpsi, part, scal;
  CKR(psi.b.ensure(sizeof(double) * dim));
  CKR(hpsi.b.ensure(sizeof(double) * dim));
  CKR(part.b.ensure(sizeof(double) * nqs::ed_partials()));
  CKR(scal.b.ensure(sizeof(double) * 8));
  CKR(ed_materialise(fn, user, N, psi.b.as<double>()));
  double* sc = scal.b.as<double>();
  const double den = ed_norm2(psi.b.as<double>(), dim, part.b.as<double>(), sc);
  CKR(nqs::launch_ed_matvec(H.op, psi.b.as<double>(), hpsi.b.as<double>(), dim,
                            part.b.as<double>(), sc + 2, g_stream));
  const double num = ed_read(sc + 2);
  if (!(den > 0.0) || !std::isfinite(num)) return -1;
  *out = num / den;
  return 0;
}

// Full-reorthogonalisation Lanczos (src/mps/lanczos.c:119-306: smallest Ritz
// value, stop when |beta_{k+1} z_{k}| < tol after at least two steps;
// :308-433: the k smallest Ritz values after max_iters steps; :435-711: the
// same with a projector applied to every new Krylov vector).  The basis V
// [max_iters][dim] stays in HBM.  The reference's modified Gram-Schmidt
// sweep over V_0..V_k is done as two passes of classical Gram-Schmidt
// (h = V^T w, w -= V h, twice), which leaves w orthogonal to the basis to
// working precision in two sweeps over V instead of k + 1 dependent ones.
// The tridiagonal's eigenvalues come from Sturm bisection, the Ritz
// vector's coefficients from inverse iteration.
struct FullLanczos {
  const EdHam* H = nullptr;
  const EdProj* P = nullptr;     // null: no projection
  int N = 0;
  const double* seed = nullptr;  // device start vector; null: the deterministic one
  int max_iters = 0;
  double tol = 0.0;
  int k = 0;                     // 0: smallest with convergence test, > 0: k smallest
};

int full_lanczos(const FullLanczos& q, double* out_vec, double* out_vals, lanczos_result_t* out) {
  if (q.N <= 0 || q.N > 30 || q.max_iters <= 0 || q.k < 0) return -1;
  const long dim = 1L << q.N;
  if (out) *out = lanczos_result_t{0.0, 0, 0, 0.0};
  int max_iters = q.max_iters;
  if (max_iters > dim) max_iters = (int)dim;
  int kk = std::min(q.k, max_iters);
  const bool proj = q.P && !q.P->identity;
  ScopedBuf vb, wb, tb, part, hb, scal;
  if (vb.b.ensure(sizeof(double) * (size_t)max_iters * dim)) {
    fprintf(stderr, "nqs_lanczos: the Krylov basis of %d vectors of 2^%d doubles does not fit\n",
            max_iters, q.N);
    return -1;
  }
  CKR(wb.b.ensure(sizeof(double) * dim));
  if (proj) CKR(tb.b.ensure(sizeof(double) * dim));
  CKR(part.b.ensure(sizeof(double) *
                    std::max<size_t>(nqs::ed_partials(), nqs::ed_mdot_partials(dim, max_iters))));
  CKR(hb.b.ensure(sizeof(double) * (max_iters + 1)));
  CKR(scal.b.ensure(sizeof(double) * 8));
  double* V = vb.b.as<double>();
  double *wp = wb.b.as<double>(), *tp = tb.b.as<double>();
  double *pt = part.b.as<double>(), *sc = scal.b.as<double>(), *h = hb.b.as<double>();
  auto start_norm2 = [&]() -> double {
    return q.P ? ed_project_norm2(*q.P, wp, tp, dim, pt, sc) : ed_norm2(wp, dim, pt, sc);
  };
  // start vector in w, then V_0 = w / |w|; a caller's start vector is not
  // projected
  double n2;
  if (q.seed) {
    CK(hipMemcpyAsync(wp, q.seed, sizeof(double) * dim, hipMemcpyDeviceToDevice, g_stream));
    n2 = ed_norm2(wp, dim, pt, sc);
  } else {
    CKR(ed_seed_vector(wp, dim, 0xA5A5A5A5A5A5A5A5ULL));
    n2 = start_norm2();
  }
  if (!(sqrt(n2) >= 1e-14)) {
    CKR(ed_seed_vector(wp, dim, 0xBEEFBABEULL));
    n2 = start_norm2();
  }
  if (!std::isfinite(n2)) return -1;
  CKR(nqs::launch_ed_scale(wp, V, n2 > 0.0 ? 1.0 / sqrt(n2) : 1.0, dim, pt, nullptr, g_stream));
  std::vector<double> alpha(max_iters), beta(max_iters + 1, 0.0), d(max_iters), e(max_iters),
      z(max_iters);
  auto load_t = [&](int K) {
    for (int i = 0; i < K; i++) d[i] = alpha[i];
    for (int i = 0; i + 1 < K; i++) e[i] = beta[i + 1];
  };
  // Ritz vector sum_j z_j V_j / |.| into out_vec (host)
  auto ritz = [&](int K) -> int {
    std::vector<double> nz(K);
    for (int i = 0; i < K; i++) nz[i] = -z[i];
    CK(hipMemcpyAsync(h, nz.data(), sizeof(double) * K, hipMemcpyHostToDevice, g_stream));
    CKR(nqs::launch_ed_mupdate(V, dim, K, h, wp, 0, pt, sc + 1, g_stream));
    const double r2 = ed_read(sc + 1);
    if (!std::isfinite(r2)) return -1;
    if (r2 > 0.0) CKR(nqs::launch_ed_scale(wp, wp, 1.0 / sqrt(r2), dim, pt, nullptr, g_stream));
    CK(hipMemcpyAsync(out_vec, wp, sizeof(double) * dim, hipMemcpyDeviceToHost, g_stream));
    CK(hipStreamSynchronize(g_stream));
    return 0;
  };
  double resid = 0.0;
  int kb;
  for (kb = 0; kb < max_iters; kb++) {
    double* vk = V + (size_t)kb * dim;
    CKR(nqs::launch_ed_matvec(q.H->op, vk, wp, dim, pt, sc, g_stream));
    alpha[kb] = ed_read(sc);
    CKR(nqs::launch_ed_axpy2(wp, vk, alpha[kb], kb > 0 ? vk - dim : nullptr, beta[kb], dim, pt,
                             nullptr, g_stream));
    for (int pass = 0; pass < 2; pass++) {
      CKR(nqs::launch_ed_mdot(V, dim, kb + 1, wp, pt, h, g_stream));
      CKR(nqs::launch_ed_mupdate(V, dim, kb + 1, h, wp, 1, pt,
                                 (pass == 1 && !q.P) ? sc + 1 : nullptr, g_stream));
    }
    const double b2 = q.P ? ed_project_norm2(*q.P, wp, tp, dim, pt, sc) : ed_read(sc + 1);
    if (!std::isfinite(alpha[kb]) || !std::isfinite(b2)) return -1;
    beta[kb + 1] = sqrt(b2);
    if (beta[kb + 1] < 1e-14) {
      kb++;
      break;
    }
    if (q.k == 0) {
      const int K = kb + 1;
      load_t(K);
      const double lam = tridiag_eig(d.data(), e.data(), K, 0);
      tridiag_vec(d.data(), e.data(), K, lam, z.data());
      resid = fabs(beta[kb + 1] * z[K - 1]);
      if (resid < q.tol && kb >= 1) {
        if (out_vec) CKR(ritz(K));
        out_vals[0] = lam;
        if (out) *out = lanczos_result_t{lam, K, 1, resid};
        return 0;
      }
    }
    if (kb + 1 < max_iters)
      CKR(nqs::launch_ed_scale(wp, V + (size_t)(kb + 1) * dim, 1.0 / beta[kb + 1], dim, pt,
                               nullptr, g_stream));
  }
  const int K = std::max(kb, 1);
  load_t(K);
  if (q.k == 0) {
    const double lam = tridiag_eig(d.data(), e.data(), K, 0);
    if (out_vec) {
      tridiag_vec(d.data(), e.data(), K, lam, z.data());
      CKR(ritz(K));
    }
    out_vals[0] = lam;
    if (out) *out = lanczos_result_t{lam, kb, 0, resid};
    return 0;
  }
  kk = std::min(kk, K);
  for (int i = 0; i < kk; i++) out_vals[i] = tridiag_eig(d.data(), e.data(), K, i);
  if (out)
    *out = lanczos_result_t{tridiag_eig(d.data(), e.data(), K, 0), K, beta[K] < 1e-10 ? 1 : 0,
                            beta[K]};
  return 0;
}

// the reference's seeded drivers: start from the materialised Re psi when
// it exists (N <= 24), else from the deterministic start vector
int seeded_lanczos(nqs_log_amp_fn_t fn, void* user, const EdHam& H, const EdProj* P, int N,
                   int max_iters, double tol, int k, double* out_vec, double* out_vals,
                   lanczos_result_t* out) {
  if (N <= 0 || N > 30) return -1;
  ScopedBuf seed;
  const double* sp = nullptr;
  if (fn && N <= 24 && max_iters > 0) {
    CKR(seed.b.ensure(sizeof(double) * (1L << N)));
    if (ed_materialise(fn, user, N, seed.b.as<double>()) == 0) sp = seed.b.as<double>();
  }
  FullLanczos q;
  q.H = &H;
  q.P = P;
  q.N = N;
  q.seed = sp;
  q.max_iters = max_iters;
  q.tol = tol;
  q.k = k;
  return full_lanczos(q, out_vec, out_vals, out);
}

// Memory-lean projecting Lanczos (src/mps/lanczos.c:824-970): three vectors
// and the projector's output in HBM; alpha and beta are device reductions
// read back every step, the tridiagonal's smallest eigenvalue is taken on
// the host.  Stops when the smallest Ritz value moves by less than tol after
// more than 5 steps or when beta < 1e-14 (then the previous step's Ritz
// value is returned, as in the reference).
//
// With out_vec (src/mps/lanczos.c:972-1130): the same first pass, then the
// Ritz value and its coefficients z from the final tridiagonal, then a
// replay of the recurrence from the same start vector accumulating
// sum_k z_k v_k; the kernels reduce in a fixed order, so the replay
// reproduces the first pass's vectors bit for bit.
int lean_lanczos(const EdHam& H, const EdProj& P, int N, int max_iters, double tol,
                 double* out_val, double* out_vec, lanczos_result_t* out) {
  const long dim = 1L << N;
  if (out) *out = lanczos_result_t{0.0, 0, 0, 0.0};
  if (max_iters > dim) max_iters = (int)dim;
  const int nvec = out_vec ? 6 : 4;
  ScopedBuf vb[6], part, scal;
  for (int i = 0; i < nvec; i++) CKR(vb[i].b.ensure(sizeof(double) * dim));
  CKR(part.b.ensure(sizeof(double) * nqs::ed_partials()));
  CKR(scal.b.ensure(sizeof(double) * 8));
  double* sc = scal.b.as<double>();
  double* pt = part.b.as<double>();
  // v_cur, v_prev, w, projector output (rotated by pointer); start vector
  // and Ritz vector for the two-pass variant
  double *vc = vb[0].b.as<double>(), *vp = vb[1].b.as<double>(), *w = vb[2].b.as<double>(),
         *tmp = vb[3].b.as<double>();
  CKR(ed_seed_vector(vc, dim, 0xA5A5A5A5A5A5A5A5ULL));
  double n2 = ed_project_norm2(P, vc, tmp, dim, pt, sc);
  if (!(sqrt(n2) >= 1e-14)) {
    // the projector removed the start vector: the reference's second seed
    CKR(ed_seed_vector(vc, dim, 0xBEEFBABEULL));
    n2 = ed_project_norm2(P, vc, tmp, dim, pt, sc);
  }
  if (!std::isfinite(n2)) return -1;
  if (n2 > 0.0) CKR(nqs::launch_ed_scale(vc, vc, 1.0 / sqrt(n2), dim, pt, nullptr, g_stream));
  double* seed = out_vec ? vb[4].b.as<double>() : nullptr;
  if (seed)
    CK(hipMemcpyAsync(seed, vc, sizeof(double) * dim, hipMemcpyDeviceToDevice, g_stream));
  std::vector<double> alpha(max_iters), beta(max_iters + 1, 0.0), d(max_iters), e(max_iters);
  // w = P (H v_cur - alpha v_cur - beta_k v_prev); alpha from the matvec's
  // dot when fresh, |w|^2 returned
  auto step = [&](int k, bool fresh) -> double {
    if (nqs::launch_ed_matvec(H.op, vc, w, dim, pt, sc, g_stream)) return NAN;
    if (fresh) alpha[k] = ed_read(sc);
    if (nqs::launch_ed_axpy2(w, vc, alpha[k], k > 0 ? vp : nullptr, beta[k], dim, pt,
                             P.identity ? sc + 1 : nullptr, g_stream))
      return NAN;
    return P.identity ? ed_read(sc + 1) : ed_project_norm2(P, w, tmp, dim, pt, sc);
  };
  double lambda = 0.0, prev = 0.0;
  int k = 0, converged = 0, K = 0;
  for (k = 0; k < max_iters; k++) {
    const double b2 = step(k, true);
    if (!std::isfinite(alpha[k]) || !std::isfinite(b2)) return -1;
    beta[k + 1] = sqrt(b2);
    K = k + 1;
    if (beta[k + 1] < 1e-14) {
      k++;
      break;
    }
    for (int i = 0; i < K; i++) d[i] = alpha[i];
    for (int i = 0; i + 1 < K; i++) e[i] = beta[i + 1];
    prev = lambda;
    lambda = tridiag_smallest(d.data(), e.data(), K);
    if (k > 5 && fabs(lambda - prev) < tol) {
      converged = 1;
      k++;
      break;
    }
    // v_prev <- v_cur, v_cur <- w / beta (into the old v_prev's buffer)
    std::swap(vp, vc);
    CKR(nqs::launch_ed_scale(w, vc, 1.0 / beta[k + 1], dim, pt, nullptr, g_stream));
  }
  if (!out_vec) {
    *out_val = lambda;
    if (out) *out = lanczos_result_t{lambda, k, converged, k > 0 ? fabs(lambda - prev) : 0.0};
    return 0;
  }
  // second pass: Ritz value and coefficients of the final tridiagonal
  for (int i = 0; i < K; i++) d[i] = alpha[i];
  for (int i = 0; i + 1 < K; i++) e[i] = beta[i + 1];
  lambda = tridiag_eig(d.data(), e.data(), K, 0);
  std::vector<double> z(K);
  tridiag_vec(d.data(), e.data(), K, lambda, z.data());
  // third pass: replay from the start vector, acc = sum_k z_k v_k
  double* acc = vb[5].b.as<double>();
  CK(hipMemcpyAsync(vc, seed, sizeof(double) * dim, hipMemcpyDeviceToDevice, g_stream));
  CKR(nqs::launch_ed_scale(seed, acc, z[0], dim, pt, nullptr, g_stream));
  for (int j = 0; j + 1 < K; j++) {
    if (!std::isfinite(step(j, false))) return -1;
    if (beta[j + 1] < 1e-14) break;
    std::swap(vp, vc);
    CKR(nqs::launch_ed_scale(w, vc, 1.0 / beta[j + 1], dim, pt, nullptr, g_stream));
    CKR(nqs::launch_ed_axpy2(acc, vc, -z[j + 1], nullptr, 0.0, dim, pt, nullptr, g_stream));
  }
  const double a2 = ed_norm2(acc, dim, pt, sc);
  if (!std::isfinite(a2)) return -1;
  if (a2 > 0.0) CKR(nqs::launch_ed_scale(acc, acc, 1.0 / sqrt(a2), dim, pt, nullptr, g_stream));
  CK(hipMemcpyAsync(out_vec, acc, sizeof(double) * dim, hipMemcpyDeviceToHost, g_stream));
  CK(hipStreamSynchronize(g_stream));
  *out_val = lambda;
  const double dl = fabs(lambda - prev);
  if (out) *out = lanczos_result_t{lambda, K, (K >= 6 && dl < tol) ? 1 : 0, K >= 2 ? dl : 0.0};
  return 0;
}

// the materialised state as a malloc'd host vector (the reference's
// nqs_materialise_state* contract: the caller frees it)
int materialise_host(nqs_log_amp_fn_t fn, void* user, int N, double** out_psi, long* out_dim) {
  if (!fn || !out_psi || !out_dim || N <= 0 || N > 24) return -1;
  if (stream_ready()) return -1;
  const long dim = 1L << N;
  ScopedBuf psi;
  CKR(psi.b.ensure(sizeof(double) * dim));
  CKR(ed_materialise(fn, user, N, psi.b.as<double>()));
  double* h = (double*)malloc(sizeof(double) * dim);
  if (!h) return -1;
  if (hipMemcpyAsync(h, psi.b.p, sizeof(double) * dim, hipMemcpyDeviceToHost, g_stream) ||
      hipStreamSynchronize(g_stream)) {
    free(h);
    return -1;
  }
  *out_psi = h;
  *out_dim = dim;
  return 0;
}

bool kagome_size_ok(int Lx, int Ly, int max_n, const char* who) {
  if (Lx >= 1 && Ly >= 1 && 3 * Lx * Ly <= max_n) return true;
  fprintf(stderr, "%s: N = 3 Lx Ly must be 3..%d\n", who, max_n);
  return false;
}

bool square_size_ok(int Lx, int Ly, int max_n, const char* who) {
  if (Lx >= 1 && Ly >= 1 && Lx * Ly <= max_n) return true;
  fprintf(stderr, "%s: N = Lx Ly must be 1..%d\n", who, max_n);
  return false;
}

}  // namespace

NQS_EXPORT int nqs_materialise_state(nqs_ansatz_t* a, int Lx, int Ly, double** out_psi,
                                     long* out_dim) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!a || Lx <= 0 || Ly <= 0) return -1;
  return materialise_host(nqs_ansatz_log_amp, a, Lx * Ly, out_psi, out_dim);
}

NQS_EXPORT int nqs_materialise_state_with_cb(nqs_log_amp_fn_t log_amp, void* user, int Lx, int Ly,
                                             double** out_psi, long* out_dim) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (Lx <= 0 || Ly <= 0) return -1;
  return materialise_host(log_amp, user, Lx * Ly, out_psi, out_dim);
}

NQS_EXPORT int nqs_materialise_state_with_cb_N(nqs_log_amp_fn_t log_amp, void* user, int N,
                                               double** out_psi, long* out_dim) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  return materialise_host(log_amp, user, N, out_psi, out_dim);
}

NQS_EXPORT int nqs_exact_energy_tfim(nqs_ansatz_t* a, int Lx, int Ly, double J, double Gamma,
                                     double* out_energy) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!a || !out_energy || !square_size_ok(Lx, Ly, 24, "nqs_exact_energy_tfim")) return -1;
  if (stream_ready()) return -1;
  EdHam H;
  CKR(ed_ham_tfim(Lx, Ly, J, Gamma, H));
  return ed_exact_energy(nqs_ansatz_log_amp, a, Lx * Ly, H, out_energy);
}

NQS_EXPORT int nqs_lanczos_refine_tfim(nqs_ansatz_t* a, int Lx, int Ly, double J, double Gamma,
                                       int max_iters, double tol, double* out_eigenvalue,
                                       double* out_eigenvector, lanczos_result_t* out_result) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!a || !out_eigenvalue || !square_size_ok(Lx, Ly, 30, "nqs_lanczos_refine_tfim")) return -1;
  if (stream_ready()) return -1;
  EdHam H;
  CKR(ed_ham_tfim(Lx, Ly, J, Gamma, H));
  // the reference starts the TFIM refiner from its deterministic vector
  return seeded_lanczos(nullptr, nullptr, H, nullptr, Lx * Ly, max_iters, tol, 0, out_eigenvector,
                        out_eigenvalue, out_result);
}

NQS_EXPORT int nqs_exact_energy_heisenberg(nqs_ansatz_t* a, int Lx, int Ly, double J, double Jz,
                                           double* out_energy) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!a || !out_energy || !square_size_ok(Lx, Ly, 24, "nqs_exact_energy_heisenberg")) return -1;
  if (stream_ready()) return -1;
  EdHam H;
  CKR(ed_ham_heisenberg(Lx, Ly, J, Jz, H));
  return ed_exact_energy(nqs_ansatz_log_amp, a, Lx * Ly, H, out_energy);
}

NQS_EXPORT int nqs_lanczos_refine_heisenberg(nqs_ansatz_t* a, int Lx, int Ly, double J, double Jz,
                                             int max_iters, double tol, double* out_eigenvalue,
                                             double* out_eigenvector,
                                             lanczos_result_t* out_result) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!a || !out_eigenvalue || !square_size_ok(Lx, Ly, 30, "nqs_lanczos_refine_heisenberg"))
    return -1;
  if (stream_ready()) return -1;
  EdHam H;
  CKR(ed_ham_heisenberg(Lx, Ly, J, Jz, H));
  return seeded_lanczos(nqs_ansatz_log_amp, a, H, nullptr, Lx * Ly, max_iters, tol, 0,
                        out_eigenvector, out_eigenvalue, out_result);
}

NQS_EXPORT int nqs_exact_energy_kagome_heisenberg(nqs_ansatz_t* a, int Lx_cells, int Ly_cells,
                                                  double J, int pbc, double* out_energy) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!a || !out_energy ||
      !kagome_size_ok(Lx_cells, Ly_cells, 24, "nqs_exact_energy_kagome_heisenberg"))
    return -1;
  if (stream_ready()) return -1;
  EdHam H;
  CKR(ed_ham_kagome(Lx_cells, Ly_cells, J, pbc, H));
  return ed_exact_energy(nqs_ansatz_log_amp, a, 3 * Lx_cells * Ly_cells, H, out_energy);
}

NQS_EXPORT int nqs_exact_energy_kagome_heisenberg_with_cb(nqs_log_amp_fn_t log_amp, void* user,
                                                          int Lx_cells, int Ly_cells, double J,
                                                          int pbc, double* out_energy) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!log_amp || !out_energy ||
      !kagome_size_ok(Lx_cells, Ly_cells, 24, "nqs_exact_energy_kagome_heisenberg"))
    return -1;
  if (stream_ready()) return -1;
  EdHam H;
  CKR(ed_ham_kagome(Lx_cells, Ly_cells, J, pbc, H));
  return ed_exact_energy(log_amp, user, 3 * Lx_cells * Ly_cells, H, out_energy);
}

NQS_EXPORT int nqs_lanczos_refine_kagome_heisenberg_with_cb(
    nqs_log_amp_fn_t log_amp, void* user, int Lx_cells, int Ly_cells, double J, int pbc,
    int max_iters, double tol, double* out_eigenvalue, double* out_eigenvector,
    lanczos_result_t* out_result) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!log_amp || !out_eigenvalue ||
      !kagome_size_ok(Lx_cells, Ly_cells, 30, "nqs_lanczos_refine_kagome_heisenberg"))
    return -1;
  if (stream_ready()) return -1;
  EdHam H;
  CKR(ed_ham_kagome(Lx_cells, Ly_cells, J, pbc, H));
  return seeded_lanczos(log_amp, user, H, nullptr, 3 * Lx_cells * Ly_cells, max_iters, tol, 0,
                        out_eigenvector, out_eigenvalue, out_result);
}

NQS_EXPORT int nqs_lanczos_refine_kagome_heisenberg(nqs_ansatz_t* a, int Lx_cells, int Ly_cells,
                                                    double J, int pbc, int max_iters, double tol,
                                                    double* out_eigenvalue,
                                                    double* out_eigenvector,
                                                    lanczos_result_t* out_result) {
  if (!a) return -1;
  return nqs_lanczos_refine_kagome_heisenberg_with_cb(nqs_ansatz_log_amp, a, Lx_cells, Ly_cells,
                                                      J, pbc, max_iters, tol, out_eigenvalue,
                                                      out_eigenvector, out_result);
}

NQS_EXPORT int nqs_lanczos_k_lowest_kagome_heisenberg_with_cb(
    nqs_log_amp_fn_t log_amp, void* user, int Lx_cells, int Ly_cells, double J, int pbc,
    int max_iters, int k, double* out_eigenvalues, lanczos_result_t* out_result) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!log_amp || !out_eigenvalues || k <= 0 ||
      !kagome_size_ok(Lx_cells, Ly_cells, 30, "nqs_lanczos_k_lowest_kagome_heisenberg"))
    return -1;
  if (stream_ready()) return -1;
  EdHam H;
  CKR(ed_ham_kagome(Lx_cells, Ly_cells, J, pbc, H));
  return seeded_lanczos(log_amp, user, H, nullptr, 3 * Lx_cells * Ly_cells, max_iters, 0.0, k,
                        nullptr, out_eigenvalues, out_result);
}

NQS_EXPORT int nqs_lanczos_k_lowest_kagome_heisenberg(nqs_ansatz_t* a, int Lx_cells, int Ly_cells,
                                                      double J, int pbc, int max_iters, int k,
                                                      double* out_eigenvalues,
                                                      lanczos_result_t* out_result) {
  if (!a) return -1;
  return nqs_lanczos_k_lowest_kagome_heisenberg_with_cb(nqs_ansatz_log_amp, a, Lx_cells, Ly_cells,
                                                        J, pbc, max_iters, k, out_eigenvalues,
                                                        out_result);
}

NQS_EXPORT int nqs_lanczos_k_lowest_kagome_heisenberg_projected(
    nqs_log_amp_fn_t log_amp, void* user, int Lx_cells, int Ly_cells, double J, int pbc,
    const int* perm, const double* characters, int G, int max_iters, int k,
    double* out_eigenvalues, lanczos_result_t* out_result) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!log_amp || !out_eigenvalues || !perm || !characters || G <= 0 || k <= 0 ||
      !kagome_size_ok(Lx_cells, Ly_cells, 30, "nqs_lanczos_k_lowest_kagome_heisenberg_projected"))
    return -1;
  const int N = 3 * Lx_cells * Ly_cells;
  EdProj P;
  CKR(ed_proj_build(perm, characters, G, N, "nqs_lanczos_k_lowest_kagome_heisenberg_projected",
                    P));
  EdHam H;
  CKR(ed_ham_kagome(Lx_cells, Ly_cells, J, pbc, H));
  return seeded_lanczos(log_amp, user, H, &P, N, max_iters, 0.0, k, nullptr, out_eigenvalues,
                        out_result);
}

NQS_EXPORT int nqs_lanczos_refine_kagome_heisenberg_projected(
    nqs_log_amp_fn_t log_amp, void* user, int Lx_cells, int Ly_cells, double J, int pbc,
    const int* perm, const double* characters, int G, int max_iters, double tol,
    double* out_eigenvalue, double* out_eigenvector, lanczos_result_t* out_result) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!log_amp || !out_eigenvalue || !perm || !characters || G <= 0 ||
      !kagome_size_ok(Lx_cells, Ly_cells, 30, "nqs_lanczos_refine_kagome_heisenberg_projected"))
    return -1;
  const int N = 3 * Lx_cells * Ly_cells;
  EdProj P;
  CKR(ed_proj_build(perm, characters, G, N, "nqs_lanczos_refine_kagome_heisenberg_projected", P));
  EdHam H;
  CKR(ed_ham_kagome(Lx_cells, Ly_cells, J, pbc, H));
  return seeded_lanczos(log_amp, user, H, &P, N, max_iters, tol, 0, out_eigenvector,
                        out_eigenvalue, out_result);
}

NQS_EXPORT int nqs_lanczos_e0_kagome_heisenberg_projected_lean(
    int Lx_cells, int Ly_cells, double J, int pbc, const int* perm, const double* characters,
    int G, int max_iters, double tol, double* out_eigenvalue, lanczos_result_t* out_result) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!out_eigenvalue || !perm || !characters || G <= 0) return -1;
  if (Lx_cells < 1 || Ly_cells < 1 || 3 * Lx_cells * Ly_cells > 30 || max_iters <= 0) return -1;
  const int N = 3 * Lx_cells * Ly_cells;
  EdProj P;
  CKR(ed_proj_build(perm, characters, G, N, "nqs_lanczos_e0_kagome_heisenberg_projected_lean", P));
  EdHam H;
  CKR(ed_ham_kagome(Lx_cells, Ly_cells, J, pbc, H));
  return lean_lanczos(H, P, N, max_iters, tol, out_eigenvalue, nullptr, out_result);
}

NQS_EXPORT int nqs_lanczos_e0_kagome_heisenberg_projected_lean_eigvec(
    int Lx_cells, int Ly_cells, double J, int pbc, const int* perm, const double* characters,
    int G, int max_iters, double tol, double* out_eigenvalue, double* out_eigenvector,
    lanczos_result_t* out_result) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!out_eigenvalue || !out_eigenvector || !perm || !characters || G <= 0) return -1;
  if (Lx_cells < 1 || Ly_cells < 1 || 3 * Lx_cells * Ly_cells > 30 || max_iters <= 0) return -1;
  const int N = 3 * Lx_cells * Ly_cells;
  EdProj P;
  CKR(ed_proj_build(perm, characters, G, N,
                    "nqs_lanczos_e0_kagome_heisenberg_projected_lean_eigvec", P));
  EdHam H;
  CKR(ed_ham_kagome(Lx_cells, Ly_cells, J, pbc, H));
  return lean_lanczos(H, P, N, max_iters, tol, out_eigenvalue, out_eigenvector, out_result);
}

namespace {

int minsr_solve_impl(nqs_ansatz_t* a, const int* spins, int B, const double* e_loc, double shift,
                     int gram, int prec, double* out_delta, double* out_y, double* out_T,
                     const char* who) {
  if (!a || !spins || !e_loc || !out_delta || B <= 0) return -1;
  const bool rbm_like = a->kind == NQS_ANSATZ_RBM || a->kind == NQS_ANSATZ_COMPLEX_RBM;
  if (gram != NQS_HIP_GRAM_DIRECT && gram != NQS_HIP_GRAM_FACTORED &&
      gram != NQS_HIP_GRAM_PARAMETER) {
    fprintf(stderr, "%s: gram must be DIRECT, FACTORED or PARAMETER\n", who);
    return -1;
  }
  if ((gram == NQS_HIP_GRAM_FACTORED && !rbm_like) ||
      (gram == NQS_HIP_GRAM_PARAMETER && (out_y || out_T))) {
    fprintf(stderr, "%s: unsupported combination (factored Gram needs an RBM; "
                    "y and T exist in sample space only)\n", who);
    return -1;
  }
  if (stream_ready()) return -1;
  const int N = a->N;
  const long P = a->P;
  const bool factored = gram == NQS_HIP_GRAM_FACTORED, dual = gram == NQS_HIP_GRAM_PARAMETER;
  const bool fdual = dual && prec == NQS_HIP_MIXED && rbm_like && factored_dual_on(a);
  a->dev_fresh = 0;
  CKR(ansatz_upload(a));
  CKR(upload_spins(spins, (long)B * N));
  const int* d_samples = g_ws.spins.as<int>();
  CKR(g_ws.e_re.ensure(sizeof(double) * B));
  double* E = g_ws.e_re.as<double>();
  CK(hipMemcpyAsync(E, e_loc, sizeof(double) * B, hipMemcpyHostToDevice, g_stream));
  double* O = nullptr;
  if (factored) {
    CKR(forward_t<double>(a, d_samples, B, nullptr, nullptr));
  } else if (fdual) {
    CKR(forward_t<float>(a, d_samples, B, nullptr, nullptr, false, true));
  } else {
    CKR(g_ws.O.ensure(sizeof(double) * (size_t)B * (size_t)P));
    O = g_ws.O.as<double>();
    if (low_prec(prec)) {
      CKR(forward_t<float>(a, d_samples, B, nullptr, nullptr, false, prec == NQS_HIP_MIXED));
      ModelView<float> v;
      CKR(model_view<float>(a, v));
      CKR(nqs::launch_gen_O<float>(v, a->kind, d_samples, B, g_ws.t_re.as<float>(),
                                   g_ws.t_im.as<float>(), 0, O, nullptr, P, g_stream));
    } else {
      CKR(forward_t<double>(a, d_samples, B, nullptr, nullptr));
      ModelView<double> v;
      CKR(model_view<double>(a, v));
      CKR(nqs::launch_gen_O<double>(v, a->kind, d_samples, B, g_ws.t_re.as<double>(),
                                    g_ws.t_im.as<double>(), 0, O, nullptr, P, g_stream));
    }
  }
  CKR(g_ws.vecs.ensure(sizeof(double) * 9 * (size_t)P));
  if (factored) {
    const int nblk = a->kind == NQS_ANSATZ_COMPLEX_RBM ? 2 : 1;
    CKR(g_ws.colwork.ensure(sizeof(double) * nqs::gram_factored_work_elems(B, N, a->M, nblk)));
  } else if (fdual) {
    const int nblk = a->kind == NQS_ANSATZ_COMPLEX_RBM ? 2 : 1;
    CKR(g_ws.colwork.ensure(sizeof(double) * nqs::factored_delta_work_elems(B, N, a->M, nblk)));
  } else {
    CKR(g_ws.colwork.ensure(sizeof(double) * nqs::colsum_work_elems(B, P)));
  }
  CKR(g_ws.scal.ensure(sizeof(double) * 16));
  CKR(g_ws.u1.ensure(sizeof(double) * B));
  CKR(g_ws.u2.ensure(sizeof(double) * B));
  double* sc = g_ws.scal.as<double>();
  double* vb = g_ws.vecs.as<double>();
  double* work = g_ws.colwork.as<double>();
  CKR(nqs::launch_sum(E, B, nullptr, 1.0 / B, sc + 0, g_stream));
  if (O) CKR(nqs::launch_colsum(O, B, P, nullptr, 1.0 / B, vb + 5 * P, work, g_stream));
  if (fdual)
    CKR(nqs::launch_factored_delta(d_samples, B, N, a->M, g_ws.t_re.as<float>(),
                                   a->kind == NQS_ANSATZ_COMPLEX_RBM ? g_ws.t_im.as<float>()
                                                                     : nullptr,
                                   nullptr, 1.0 / B, work, vb + 5 * P, g_stream));
  MinsrArgs m;
  m.a = a;
  m.d_samples = d_samples;
  m.B = B;
  m.O = O;
  m.Om = vb + 5 * P;
  m.E = E;
  m.sc = sc;
  m.work = work;
  m.rhs = vb + 7 * P;
  m.eps = shift;
  m.factored = factored;
  m.dual = dual;
  m.mixed = prec == NQS_HIP_MIXED;
  m.fdual = fdual;
  m.delta = vb;
  m.h_T = out_T;
  m.h_y = out_y;
  int converged = 0;
  CKR(minsr_core(m, converged));
  CK(hipMemcpyAsync(out_delta, vb, sizeof(double) * P, hipMemcpyDeviceToHost, g_stream));
  CK(hipStreamSynchronize(g_stream));
  return converged ? 1 : 0;
}

}  // namespace

NQS_EXPORT int nqs_hip_minsr_solve(nqs_ansatz_t* a, const int* spins, int B, const double* e_loc,
                                   double shift, int gram, double* out_delta, double* out_y,
                                   double* out_T) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  return minsr_solve_impl(a, spins, B, e_loc, shift, gram, NQS_HIP_FP64, out_delta, out_y, out_T,
                          "nqs_hip_minsr_solve");
}

NQS_EXPORT int nqs_hip_minsr_solve_ex(nqs_ansatz_t* a, const int* spins, int B,
                                      const double* e_loc, double shift,
                                      const nqs_hip_options_t* opt, double* out_delta,
                                      double* out_y, double* out_T) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!opt) return -1;
  if (opt->precision < NQS_HIP_FP64 || opt->precision > NQS_HIP_MIXED) {
    fprintf(stderr, "nqs_hip_minsr_solve_ex: unknown precision %d\n", opt->precision);
    return -1;
  }
  return minsr_solve_impl(a, spins, B, e_loc, shift, opt->gram, opt->precision, out_delta, out_y,
                          out_T, "nqs_hip_minsr_solve_ex");
}

// Streaming copy of `bytes` bytes between two device buffers, `reps` times;
// returns the average time per copy in milliseconds (HIP events) or < 0.
NQS_EXPORT double nqs_hip_copy_probe(long bytes, int reps) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (bytes <= 0 || reps <= 0 || stream_ready()) return -1.0;
  const long n = (bytes / 32) * 4;
  DBuf a, b;
  if (a.ensure(sizeof(double) * n) || b.ensure(sizeof(double) * n)) return -1.0;
  if (nqs::launch_fill(a.as<double>(), 1.0, n, g_stream)) return -1.0;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) || hipEventCreate(&e1)) return -1.0;
  (void)nqs::launch_copy_probe(a.as<double>(), b.as<double>(), n, g_stream);
  (void)hipEventRecord(e0, g_stream);
  for (int r = 0; r < reps; r++) (void)nqs::launch_copy_probe(a.as<double>(), b.as<double>(), n, g_stream);
  (void)hipEventRecord(e1, g_stream);
  (void)hipEventSynchronize(e1);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  a.release();
  b.release();
  return ms / reps;
}

// ---- learned-proposal network API -------------------------------------------
namespace {
int ar_rows_to_device(nqs_hip_ar* a, const int* spins, int B, uint64_t*& d_bits) {
  CKR(a->spins.ensure(sizeof(int) * (size_t)B * a->N));
  CKR(a->bits.ensure(sizeof(uint64_t) * (size_t)B));
  CK(hipMemcpyAsync(a->spins.p, spins, sizeof(int) * (size_t)B * a->N, hipMemcpyHostToDevice,
                    g_stream));
  CKR(nqs::launch_spins_to_bits(a->spins.as<int>(), B, a->N, a->bits.as<uint64_t>(), g_stream));
  d_bits = a->bits.as<uint64_t>();
  return 0;
}
}  // namespace

NQS_EXPORT nqs_hip_ar_t* nqs_hip_ar_create(int num_sites, unsigned long long seed, double init_scale) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  return ar_new(num_sites, seed, init_scale);
}

NQS_EXPORT void nqs_hip_ar_free(nqs_hip_ar_t* a) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!a) return;
  (void)hipStreamSynchronize(g_stream);
  delete a;
}

NQS_EXPORT long nqs_hip_ar_num_params(const nqs_hip_ar_t* a) { return a ? a->d.P : 0; }

NQS_EXPORT int nqs_hip_ar_get_params(nqs_hip_ar_t* a, float* out) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!a || !out) return -1;
  CKR(ar_join(a));
  CK(hipMemcpyAsync(out, a->w.p, sizeof(float) * a->d.P, hipMemcpyDeviceToHost, g_stream));
  CK(hipStreamSynchronize(g_stream));
  return 0;
}

NQS_EXPORT int nqs_hip_ar_set_params(nqs_hip_ar_t* a, const float* in) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!a || !in) return -1;
  CKR(ar_join(a));
  CKR(ar_set(a, in));
  CK(hipStreamSynchronize(g_stream));
  return 0;
}

NQS_EXPORT int nqs_hip_ar_log_prob(nqs_hip_ar_t* a, const int* spins, int B, int constrained,
                                   double* out) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!a || !spins || !out || B <= 0) return -1;
  CKR(ar_join(a));
  uint64_t* bits = nullptr;
  CKR(ar_rows_to_device(a, spins, B, bits));
  CKR(a->lq.ensure(sizeof(double) * B));
  nqs::ArWalk w;
  w.rows = B;
  w.n_sample = 0;
  w.constrained = constrained ? 1 : 0;
  w.bits = bits;
  w.lq = a->lq.as<double>();
  CKR(nqs::launch_ar_walk(a->dev, w, g_stream));
  CK(hipMemcpyAsync(out, a->lq.p, sizeof(double) * B, hipMemcpyDeviceToHost, g_stream));
  CK(hipStreamSynchronize(g_stream));
  return 0;
}

NQS_EXPORT int nqs_hip_ar_sample(nqs_hip_ar_t* a, const int* ref_spins, int B, int constrained,
                                 const unsigned long long* seeds, int* out_spins, double* out_logq) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!a || !ref_spins || !seeds || !out_spins || B <= 0) return -1;
  CKR(ar_join(a));
  std::vector<int> ref((size_t)B * a->N);
  for (int b = 0; b < B; b++) memcpy(ref.data() + (size_t)b * a->N, ref_spins, sizeof(int) * a->N);
  uint64_t* bits = nullptr;
  CKR(ar_rows_to_device(a, ref.data(), B, bits));
  std::vector<uint64_t> st(B);
  for (int b = 0; b < B; b++) st[b] = seeds[b] ? seeds[b] : 1ULL;
  CKR(a->rng.ensure(sizeof(uint64_t) * B));
  CKR(a->lq.ensure(sizeof(double) * B));
  CKR(a->grad.ensure(sizeof(int) * (size_t)B * a->N));
  CK(hipMemcpyAsync(a->rng.p, st.data(), sizeof(uint64_t) * B, hipMemcpyHostToDevice, g_stream));
  nqs::ArWalk w;
  w.rows = B;
  w.n_sample = B;
  w.constrained = constrained ? 1 : 0;
  w.bits = bits;
  w.rng = a->rng.as<uint64_t>();
  w.out_spins = a->grad.as<int>();
  w.lq = a->lq.as<double>();
  CKR(nqs::launch_ar_walk(a->dev, w, g_stream));
  CK(hipMemcpyAsync(out_spins, a->grad.p, sizeof(int) * (size_t)B * a->N, hipMemcpyDeviceToHost,
                    g_stream));
  if (out_logq)
    CK(hipMemcpyAsync(out_logq, a->lq.p, sizeof(double) * B, hipMemcpyDeviceToHost, g_stream));
  CK(hipStreamSynchronize(g_stream));
  return 0;
}

NQS_EXPORT int nqs_hip_ar_gradient(nqs_hip_ar_t* a, const int* spins, int B, int constrained,
                                   float* grad, double* mean_nll) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!a || !spins || !grad || B <= 0) return -1;
  CKR(ar_join(a));
  uint64_t* bits = nullptr;
  CKR(ar_rows_to_device(a, spins, B, bits));
  const int nb = nqs::ar_grad_blocks(B);
  CKR(a->partial.ensure(sizeof(float) * (size_t)nb * a->d.P));
  CKR(a->loss.ensure(sizeof(double) * nb));
  CKR(a->grad.ensure(sizeof(float) * a->d.P));
  nqs::ArGrad g;
  g.rows = bits;
  g.n_avail = B;
  g.start = 0;
  g.count = B;
  g.constrained = constrained ? 1 : 0;
  g.scale = 1.0f / (float)B;
  g.partial = a->partial.as<float>();
  g.loss = a->loss.as<double>();
  CKR(nqs::launch_ar_grad(a->dev, g, g_stream));
  CKR(nqs::launch_ar_sum_grad(a->dev, a->partial.as<float>(), nb, a->grad.as<float>(), g_stream));
  std::vector<double> l(nb);
  CK(hipMemcpyAsync(grad, a->grad.p, sizeof(float) * a->d.P, hipMemcpyDeviceToHost, g_stream));
  CK(hipMemcpyAsync(l.data(), a->loss.p, sizeof(double) * nb, hipMemcpyDeviceToHost, g_stream));
  CK(hipStreamSynchronize(g_stream));
  if (mean_nll) {
    double t = 0.0;
    for (double x : l) t += x;
    *mean_nll = t / (double)B;
  }
  return 0;
}

NQS_EXPORT int nqs_hip_ar_train(nqs_hip_ar_t* a, const int* spins, int n, int constrained,
                                int epochs, int minibatch, double lr, double* mean_nll) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!a || !spins || n <= 0 || epochs <= 0 || minibatch <= 0) return -1;
  uint64_t* bits = nullptr;
  CKR(ar_rows_to_device(a, spins, n, bits));
  return ar_train(a, bits, n, constrained ? 1 : 0, epochs, minibatch, lr, mean_nll);
}

NQS_EXPORT nqs_hip_ar_t* nqs_hip_sampler_ar(nqs_sampler_t* s) { return s ? s->ar : nullptr; }

NQS_EXPORT int nqs_hip_sampler_learned_stats(const nqs_sampler_t* s, double* acceptance,
                                             long long* proposals, int* trainings,
                                             double* last_nll) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!s || !s->ar) return -1;
  const int C = s->cs.C;
  std::vector<unsigned long long> acc(C);
  CK(hipMemcpyAsync(acc.data(), s->lacc.p, sizeof(uint64_t) * C, hipMemcpyDeviceToHost, g_stream));
  CK(hipStreamSynchronize(g_stream));
  unsigned long long t = 0;
  for (unsigned long long x : acc) t += x;
  if (acceptance) *acceptance = s->ar_props ? (double)t / (double)s->ar_props : 0.0;
  if (proposals) *proposals = s->ar_props;
  if (trainings) *trainings = s->n_train;
  if (last_nll) CKR(ar_last_loss(s->ar, last_nll));
  return 0;
}

NQS_EXPORT long long nqs_hip_device_launches(void) { return g_launches.load(); }
NQS_EXPORT long long nqs_hip_host_callback_calls(void) { return g_host_calls.load(); }
NQS_EXPORT const char* nqs_hip_version(void) { return "nqs_hip 0.1.0 gfx950"; }

// ABI layout check: sizeof(nqs_config_t), then offsets of its members in
// declaration order; returns the number of values written.
NQS_EXPORT int nqs_hip_config_layout(long* out, int cap) {
  const long v[] = {
      (long)sizeof(nqs_config_t),
      (long)offsetof(nqs_config_t, ansatz), (long)offsetof(nqs_config_t, patch_size),
      (long)offsetof(nqs_config_t, depth), (long)offsetof(nqs_config_t, width),
      (long)offsetof(nqs_config_t, heads), (long)offsetof(nqs_config_t, rbm_hidden_units),
      (long)offsetof(nqs_config_t, rbm_init_scale), (long)offsetof(nqs_config_t, symmetries),
      (long)offsetof(nqs_config_t, hamiltonian), (long)offsetof(nqs_config_t, j_coupling),
      (long)offsetof(nqs_config_t, j2_coupling), (long)offsetof(nqs_config_t, j_z_coupling),
      (long)offsetof(nqs_config_t, transverse_field), (long)offsetof(nqs_config_t, kh_K),
      (long)offsetof(nqs_config_t, kh_J), (long)offsetof(nqs_config_t, kagome_pbc),
      (long)offsetof(nqs_config_t, num_samples), (long)offsetof(nqs_config_t, num_thermalize),
      (long)offsetof(nqs_config_t, num_decorrelate), (long)offsetof(nqs_config_t, cluster_moves),
      (long)offsetof(nqs_config_t, rng_seed), (long)offsetof(nqs_config_t, num_iterations),
      (long)offsetof(nqs_config_t, learning_rate), (long)offsetof(nqs_config_t, sr_diag_shift),
      (long)offsetof(nqs_config_t, sr_cg_max_iters), (long)offsetof(nqs_config_t, sr_cg_tol),
      (long)sizeof(nqs_sr_step_info_t), (long)sizeof(nqs_marshall_wrapper_t),
      (long)sizeof(nqs_hip_options_t)};
  const int n = (int)(sizeof(v) / sizeof(v[0]));
  if (!out || cap < n) return -1;
  for (int i = 0; i < n; i++) out[i] = v[i];
  return n;
}

// ===========================================================================
// batched independent problems: n (real RBM ansatz, device sampler) pairs,
// one CG-SR iteration each per nqs_hip_multi_step (nqs_sr_step semantics,
// nqs_optimizer.c:255-262).  Per problem, on one of kMultiSide streams:
// tables, Metropolis batch, forward, E_loc, energy sums; then for all
// problems together (k_multi.hip): operands, force, a fixed number of CG
// iterations (finished problems skip theirs), update and a gather of the
// step's scalars.  The first step runs eagerly, times the split-K factor
// candidates of the CG product on that step's data, and captures the whole
// step into one hipGraph that every later step replays.
// ===========================================================================
struct nqs_hip_multi {
  static constexpr int kSide = 4;
  struct Prob {
    nqs_ansatz* a = nullptr;
    nqs_sampler* s = nullptr;
    nqs_config_t cfg;
    nqs::HamDev H;
    DBuf ham, sublat, t, E, S1, S1t, T1t, U, Gp, vec, part, ticket, sc, stats;
    int B = 0, rounds = 0, steps = 1, max_split = 1;
  };
  std::vector<Prob> p;
  std::vector<nqs::MultiProb> mp;
  DBuf d_mp, d_out;
  nqs::MultiBatch mb;
  int max_iters = 0;
  int target = 0;                       // autotuned: G tiles x splits per problem
  std::vector<std::pair<int, double>> tuning;   // (target, matvec us)
  hipStream_t side[kSide] = {};
  hipEvent_t ev_fork = nullptr, ev_join[kSide] = {};
  hipEvent_t ev_s0 = nullptr, ev_s1 = nullptr;   // around the Metropolis launches
  hipEvent_t ev_t0 = nullptr, ev_t1 = nullptr;   // around the whole step
  double sample_ms = -1.0, step_ms = -1.0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  double* h_out = nullptr;              // pinned, kMultiOut per problem
  long long steps_done = 0;
  int tuned = 0, use_graph = 1;
  // the CG iterations in one cooperative launch, the captured graph then
  // holding the step up to the CG start: chosen by the first step's
  // autotuning against the launch sequence, unless NQS_HIP_MULTI_PERSIST at
  // creation or nqs_hip_multi_use_persistent chose (persist_set)
  int persist = 0, persist_set = 0;
  // Metropolis jobs of all problems, grouped by kernel variant:
  // (variant, first job, jobs, blocks)
  std::vector<nqs::MHJob> jobs;
  std::vector<std::array<int, 4>> groups;
  DBuf d_jobs;
  ~nqs_hip_multi() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (int k = 0; k < kSide; k++) {
      if (side[k]) (void)hipStreamDestroy(side[k]);
      if (ev_join[k]) (void)hipEventDestroy(ev_join[k]);
    }
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    hipEvent_t evs[] = {ev_s0, ev_s1, ev_t0, ev_t1};
    for (hipEvent_t e : evs)
      if (e) (void)hipEventDestroy(e);
    if (h_out) (void)hipHostFree(h_out);
    for (Prob& q : p) {
      DBuf* all[] = {&q.ham, &q.sublat, &q.t, &q.E, &q.S1, &q.S1t, &q.T1t, &q.U,
                     &q.Gp, &q.vec, &q.part, &q.ticket, &q.sc, &q.stats};
      for (DBuf* b : all) b->release();
    }
    d_mp.release();
    d_out.release();
    d_jobs.release();
  }
};

namespace {

// split factors from the target count of G tiles x splits per problem
int multi_set_split(nqs_hip_multi* m, int target) {
  // the previous descriptor copy reads m->mp
  CK(hipStreamSynchronize(g_stream));
  for (size_t i = 0; i < m->p.size(); i++) {
    const nqs_hip_multi::Prob& q = m->p[i];
    const int M1 = q.a->M + 1, N1 = q.a->N + 1;
    const int tiles = ((M1 + 63) / 64) * ((N1 + 63) / 64);
    int ns = 1;
    while (ns * 2 <= q.max_split && tiles * ns * 2 <= target) ns *= 2;
    m->mp[i].nsplit = ns;
  }
  nqs::multi_layout(m->mp.data(), (int)m->mp.size(), m->mb);
  // the host table is rewritten by the next call (the autotuner sets several
  // targets in a row): the copy must have read it before this returns
  CK(hipMemcpyAsync(m->d_mp.p, m->mp.data(), sizeof(nqs::MultiProb) * m->mp.size(),
                    hipMemcpyHostToDevice, g_stream));
  CK(hipStreamSynchronize(g_stream));
  m->mb.probs = m->d_mp.as<nqs::MultiProb>();
  m->target = target;
  return 0;
}

int multi_fork(nqs_hip_multi* m) {
  CK(hipEventRecord(m->ev_fork, g_stream));
  for (int k = 0; k < nqs_hip_multi::kSide; k++) CK(hipStreamWaitEvent(m->side[k], m->ev_fork, 0));
  return 0;
}

int multi_join(nqs_hip_multi* m) {
  for (int k = 0; k < nqs_hip_multi::kSide; k++) {
    CK(hipEventRecord(m->ev_join[k], m->side[k]));
    CK(hipStreamWaitEvent(g_stream, m->ev_join[k], 0));
  }
  return 0;
}

// the Metropolis jobs of all problems (the tables' addresses are fixed once
// the first preparation has allocated them)
int multi_build_jobs(nqs_hip_multi* m) {
  std::vector<std::pair<int, int>> order;   // (variant, problem)
  for (size_t i = 0; i < m->p.size(); i++)
    order.push_back({nqs::mh_variant(m->p[i].a->m64.view), (int)i});
  std::stable_sort(order.begin(), order.end());
  m->jobs.clear();
  m->groups.clear();
  for (size_t k = 0; k < order.size(); k++) {
    const int v = order[k].first;
    nqs_hip_multi::Prob& q = m->p[order[k].second];
    nqs_sampler* s = q.s;
    if (m->groups.empty() || m->groups.back()[0] != v) m->groups.push_back({v, (int)k, 0, 0});
    std::array<int, 4>& g = m->groups.back();
    nqs::MHJob jb;
    jb.m = q.a->m64.view;
    jb.cs = s->cs;
    jb.ml.move = s->opt.move == NQS_HIP_MOVE_FLIP ? 0 : 1;
    jb.ml.steps = q.steps;
    jb.ml.rounds = q.rounds;
    jb.ml.record = 1;
    jb.ml.n_out = q.B;
    jb.ml.out = s->d_samples.as<int>();
    jb.blk0 = g[3];
    g[2]++;
    g[3] += nqs::mh_blocks(v, s->cs.C);
    m->jobs.push_back(jb);
  }
  CKR(m->d_jobs.ensure(sizeof(nqs::MHJob) * m->jobs.size()));
  CK(hipMemcpyAsync(m->d_jobs.p, m->jobs.data(), sizeof(nqs::MHJob) * m->jobs.size(),
                    hipMemcpyHostToDevice, g_stream));
  CK(hipStreamSynchronize(g_stream));
  return 0;
}

// per-problem tables on the side streams; every problem's Metropolis batch
// in one launch per kernel variant; per-problem forward, E_loc and energy
// sums on the side streams, joined into g_stream
int multi_issue_front(nqs_hip_multi* m) {
  CKR(multi_fork(m));
  for (size_t i = 0; i < m->p.size(); i++) {
    nqs_ansatz* a = m->p[i].a;
    // inside the captured graph: the range verdict is read after the step
    CKR(nqs::launch_prep<double>(a->d_par.as<double>(), a->kind, a->N, a->M, a->m64,
                                 m->side[i % nqs_hip_multi::kSide], 0));
  }
  CKR(multi_join(m));
  if (m->jobs.empty()) CKR(multi_build_jobs(m));
  const nqs::MHJob* dj = m->d_jobs.as<nqs::MHJob>();
  CK(hipEventRecord(m->ev_s0, g_stream));
  for (const std::array<int, 4>& g : m->groups)
    CKR(nqs::launch_mh_batch(g[0], dj + g[1], g[2], g[3], g_stream));
  CK(hipEventRecord(m->ev_s1, g_stream));
  CKR(multi_fork(m));
  for (size_t i = 0; i < m->p.size(); i++) {
    nqs_hip_multi::Prob& q = m->p[i];
    hipStream_t st = m->side[i % nqs_hip_multi::kSide];
    nqs_ansatz* a = q.a;
    const ModelView<double> v = a->m64.view;
    const int* smp = q.s->d_samples.as<int>();
    const size_t tb = (size_t)q.B * a->M;
    double* t = q.t.as<double>();
    // real RBM: t in the first half, the small logistic members in the second
    CKR(nqs::launch_forward<double>(v, smp, q.B, t, nullptr, nullptr, nullptr, st, t + tb,
                                    nullptr));
    double* E = q.E.as<double>();
    CKR(nqs::launch_eloc<double>(v, q.H, smp, q.B, t, nullptr, t + tb, nullptr, E, E + q.B, st));
    double* stats = q.stats.as<double>();
    CKR(nqs::launch_sum(E, q.B, nullptr, 1.0 / q.B, stats + 0, st));
    CKR(nqs::launch_sum(E, q.B, E, 1.0 / q.B, stats + 1, st));
  }
  return multi_join(m);
}

int multi_issue_back(nqs_hip_multi* m) {
  if (m->persist && nqs::launch_multi_cg_persist(m->mb, m->max_iters, g_stream) != 0) {
    fprintf(stderr, "nqs_hip_multi_step: the CG runs as a launch sequence from here on\n");
    m->persist = 0;
    if (m->exec) {
      (void)hipGraphExecDestroy(m->exec);
      m->exec = nullptr;
    }
    if (m->graph) {
      (void)hipGraphDestroy(m->graph);
      m->graph = nullptr;
    }
  }
  if (!m->persist)
    for (int it = 0; it < m->max_iters; it++) CKR(nqs::launch_multi_cg_iter(m->mb, 0, g_stream));
  CKR(nqs::launch_multi_apply(m->mb, g_stream));
  CKR(nqs::launch_multi_gather(m->mb, m->d_out.as<double>(), g_stream));
  CK(hipMemcpyAsync(m->h_out, m->d_out.p, sizeof(double) * nqs::kMultiOut * m->p.size(),
                    hipMemcpyDeviceToHost, g_stream));
  return 0;
}

// the step up to the CG start
int multi_issue_pre(nqs_hip_multi* m) {
  CKR(multi_issue_front(m));
  CKR(nqs::launch_multi_prep(m->mb, g_stream));
  return nqs::launch_multi_force(m->mb, g_stream);
}

int multi_issue_step(nqs_hip_multi* m) {
  CKR(multi_issue_pre(m));
  return multi_issue_back(m);
}

// time the CG product (matvec and alpha, idempotent at the CG start) for
// each target and keep the fastest
int multi_autotune(nqs_hip_multi* m) {
  // NQS_HIP_MULTI_SPLIT=t: split target t without timing, so that the
  // summation order (and with it every result) does not depend on timing
  const char* fixed = getenv("NQS_HIP_MULTI_SPLIT");
  if (fixed && atoi(fixed) > 0) {
    m->tuning.clear();
    return multi_set_split(m, atoi(fixed));
  }
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  const int targets[] = {8, 16, 32, 64, 128, 256};
  double best = 1e300;
  int best_t = 32;
  m->tuning.clear();
  for (int t : targets) {
    CKR(multi_set_split(m, t));
    CKR(nqs::launch_multi_cg_iter(m->mb, 1, g_stream));   // warm
    CK(hipEventRecord(e0, g_stream));
    const int reps = 5;
    for (int r = 0; r < reps; r++) CKR(nqs::launch_multi_cg_iter(m->mb, 1, g_stream));
    CK(hipEventRecord(e1, g_stream));
    CK(hipEventSynchronize(e1));
    float ms = 0.f;
    CK(hipEventElapsedTime(&ms, e0, e1));
    const double us = 1e3 * ms / reps;
    m->tuning.push_back({t, us});
    if (us < best) {
      best = us;
      best_t = t;
    }
  }
  CKR(multi_set_split(m, best_t));
  // launch shape: the persistent cooperative launch against the launch
  // sequence on the same product at the chosen split (bit-equal results, so
  // the choice is timing only), unless the caller or NQS_HIP_MULTI_PERSIST
  // chose; recorded in the tuning list as target 0
  if (!m->persist_set && nqs::launch_multi_cg_persist(m->mb, 1, g_stream, 1) == 0) {
    CK(hipEventRecord(e0, g_stream));
    const int reps = 5;
    for (int r = 0; r < reps; r++) CKR(nqs::launch_multi_cg_persist(m->mb, 1, g_stream, 1));
    CK(hipEventRecord(e1, g_stream));
    CK(hipEventSynchronize(e1));
    float ms = 0.f;
    CK(hipEventElapsedTime(&ms, e0, e1));
    const double us = 1e3 * ms / 5;
    m->tuning.push_back({0, us});
    m->persist = us < best ? 1 : 0;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return 0;
}

void multi_fill_info(nqs_hip_multi* m, nqs_sr_step_info_t* infos) {
  for (size_t i = 0; i < m->p.size(); i++) {
    const double* o = m->h_out + i * nqs::kMultiOut;
    const nqs_config_t& c = m->p[i].cfg;
    if (o[5] != 0.0)
      fprintf(stderr, "nqs_sr cg_solve: Krylov breakdown at iter=%d, pAp=%.3e, residual/rhs=%.3e "
                      "(tol=%.3e). Returning unconverged partial step.\n",
              (int)o[3], o[6], o[8] > 0.0 ? sqrt(o[7] / o[8]) : 0.0, c.sr_cg_tol);
    if (!infos) continue;
    nqs_sr_step_info_t& f = infos[i];
    f.mean_energy = o[0];
    f.variance_energy = o[1] - o[0] * o[0];
    f.update_norm = sqrt(o[2]) * c.learning_rate;
    f.acceptance_ratio = o[9];
    f.cg_iterations = (int)o[3];
    f.converged = (int)o[4];
  }
}

int multi_capture(nqs_hip_multi* m) {
  const int kt = g_kt_on.exchange(0);
  hipError_t e = hipStreamBeginCapture(g_stream, hipStreamCaptureModeRelaxed);
  // the cooperative CG launch stays outside the graph
  int rc = e == hipSuccess ? (m->persist ? multi_issue_pre(m) : multi_issue_step(m)) : -1;
  hipGraph_t g = nullptr;
  hipError_t e2 = e == hipSuccess ? hipStreamEndCapture(g_stream, &g) : e;
  g_kt_on = kt;
  if (rc || e2 != hipSuccess || !g) {
    fprintf(stderr, "nqs_hip: graph capture of the batched step failed (%s)\n",
            hipGetErrorString(e2));
    if (g) (void)hipGraphDestroy(g);
    return -1;
  }
  m->graph = g;
  CK(hipGraphInstantiate(&m->exec, m->graph, nullptr, nullptr, 0));
  return 0;
}

int multi_step(nqs_hip_multi* m, nqs_sr_step_info_t* infos) {
  CK(hipEventRecord(m->ev_t0, g_stream));
  if (!m->tuned) {
    // first step: eager, with the split factors tuned on its own data
    CKR(multi_issue_front(m));
    CKR(multi_set_split(m, 32));
    CKR(nqs::launch_multi_prep(m->mb, g_stream));
    CKR(nqs::launch_multi_force(m->mb, g_stream));
    CKR(multi_autotune(m));
    CKR(multi_issue_back(m));
    m->tuned = 1;
  } else if (!m->use_graph) {
    CKR(multi_issue_step(m));
  } else {
    if (!m->exec) CKR(multi_capture(m));
    CK(hipGraphLaunch(m->exec, g_stream));
    if (m->persist) CKR(multi_issue_back(m));
  }
  CK(hipEventRecord(m->ev_t1, g_stream));
  CK(hipStreamSynchronize(g_stream));
  float sms = 0.f, tms = 0.f;
  m->sample_ms = hipEventElapsedTime(&sms, m->ev_s0, m->ev_s1) == hipSuccess ? sms : -1.0;
  m->step_ms = hipEventElapsedTime(&tms, m->ev_t0, m->ev_t1) == hipSuccess ? tms : -1.0;
  multi_fill_info(m, infos);
  for (nqs_hip_multi::Prob& q : m->p) q.a->m64.valid = 0;
  // weights beyond the ratio tables' range (flagged by the tables of this
  // step): the step's moves and energies are not trusted
  for (size_t i = 0; i < m->p.size(); i++)
    if (m->h_out[i * nqs::kMultiOut + 10] != 0.0) {
      fprintf(stderr, "nqs_hip_multi_step: problem %zu has a weight |Re W| beyond the range of "
                      "the FP64 amplitude-ratio tables (|Re W| <= %g); the step is refused\n", i,
              nqs::RatioRange<double>::max_w);
      return -1;
    }
  m->steps_done++;
  return 0;
}

}  // namespace

void multi_delete(nqs_hip_multi* m) { delete m; }

nqs_hip_multi* multi_new(int n, const nqs_config_t* cfgs, const int* size_x, const int* size_y,
                         nqs_ansatz* const* ansatz, nqs_sampler* const* samplers) {
  if (n <= 0 || !cfgs || !size_x || !size_y || !ansatz || !samplers) return nullptr;
  if (stream_ready()) return nullptr;
  nqs_hip_multi* m = new nqs_hip_multi();
  {
    const char* e = getenv("NQS_HIP_MULTI_PERSIST");
    m->persist = (e && e[0] == '1') ? 1 : 0;
    m->persist_set = (e && (e[0] == '0' || e[0] == '1')) ? 1 : 0;
  }
  m->p.resize(n);
  m->mp.resize(n);
  auto fail = [&](int i, const char* why) -> nqs_hip_multi_t* {
    fprintf(stderr, "nqs_hip_multi_create: problem %d: %s\n", i, why);
    delete m;
    return nullptr;
  };
  for (int i = 0; i < n; i++) {
    nqs_hip_multi::Prob& q = m->p[i];
    q.a = ansatz[i];
    q.s = samplers[i];
    q.cfg = cfgs[i];
    nqs_ansatz* a = q.a;
    nqs_sampler* s = q.s;
    if (!a || !s) return fail(i, "null ansatz or sampler");
    for (int j = 0; j < i; j++)
      if (m->p[j].a == a || m->p[j].s == s) return fail(i, "ansatz or sampler used twice");
    if (a->kind != NQS_ANSATZ_RBM || a->M < 1) return fail(i, "the ansatz is not a real RBM");
    if (!s->device || s->src.a != a || s->src.marshall)
      return fail(i, "the sampler does not sample this ansatz on the device");
    if (s->opt.precision != NQS_HIP_FP64 || s->opt.move == NQS_HIP_MOVE_LEARNED)
      return fail(i, "the sampler needs FP64 and flip or exchange moves");
    const int N = a->N, M = a->M, N1 = N + 1, M1 = M + 1;
    if ((M1 * N1 + nqs::kMultiVecBlock - 1) / nqs::kMultiVecBlock > nqs::kMultiMaxPart)
      return fail(i, "problem too large for the batched CG");
    q.B = q.cfg.num_samples;
    if (q.B <= 0) return fail(i, "num_samples must be positive");
    const int C = s->cs.C;
    q.rounds = (q.B + C - 1) / C;
    q.steps = s->cfg.num_decorrelate > 0 ? s->cfg.num_decorrelate : 1;
    nqs::HamHost hh;
    if (nqs::build_hamiltonian(&q.cfg, size_x[i], size_y[i], s->opt.periodic, hh) != 0 ||
        hh.n_sites != N)
      return fail(i, "Hamiltonian unavailable or of another size");
    if (ham_upload(hh, s->src, q.ham, q.sublat, q.H)) return fail(i, "device memory");
    q.max_split = 1;
    while (q.max_split * 2 <= 32 && q.max_split * 2 * 64 <= q.B) q.max_split *= 2;
    const size_t L = (size_t)M1 * N1, B = (size_t)q.B, d = sizeof(double);
    if (q.t.ensure(d * 2 * B * M) || q.E.ensure(d * 2 * B) || q.S1.ensure(d * B * N1) ||
        q.S1t.ensure(d * B * N1) || q.T1t.ensure(d * B * M1) || q.U.ensure(d * B * ((M1 + 63) / 64)) ||
        q.Gp.ensure(d * 2 * q.max_split * L) ||
        q.vec.ensure(d * 6 * L) || q.part.ensure(d * nqs::kMultiMaxPart) ||
        q.ticket.ensure(64) || q.sc.ensure(sizeof(nqs::CGScalars)) || q.stats.ensure(d * 8) ||
        s->d_samples.ensure(sizeof(int) * B * N))
      return fail(i, "device memory");
    if (hipMemsetAsync(q.ticket.p, 0, 64, g_stream) != hipSuccess ||
        hipMemsetAsync(q.sc.p, 0, sizeof(nqs::CGScalars), g_stream) != hipSuccess)
      return fail(i, "device memory");
    a->dev_fresh = 0;
    if (ansatz_upload(a)) return fail(i, "parameter upload");
    a->dev_fresh = 1;
    ModelView<double> mv0;
    if (ratio_view<double>(a, mv0)) return fail(i, "amplitude-ratio tables");
    nqs::MultiProb& mp = m->mp[i];
    mp.range_flag = a->m64.range_flag;
    mp.B = q.B;
    mp.N = N;
    mp.M = M;
    mp.nsplit = 1;
    mp.max_it = q.cfg.sr_cg_max_iters;
    mp.C = C;
    mp.eps = q.cfg.sr_diag_shift;
    mp.tol = q.cfg.sr_cg_tol;
    mp.lr = q.cfg.learning_rate;
    mp.acc = s->cs.acc;
    mp.tot = s->cs.tot;
    mp.spins = s->d_samples.as<int>();
    mp.t = q.t.as<double>();
    mp.E = q.E.as<double>();
    mp.S1 = q.S1.as<double>();
    mp.S1t = q.S1t.as<double>();
    mp.T1t = q.T1t.as<double>();
    mp.u = q.U.as<double>();
    mp.Gp = q.Gp.as<double>();
    double* v = q.vec.as<double>();
    mp.x = v;
    mp.r = v + L;
    mp.p = v + 2 * L;
    mp.Ap = v + 3 * L;
    mp.Om = v + 4 * L;
    mp.F = v + 5 * L;
    mp.part = q.part.as<double>();
    mp.ticket = q.ticket.as<unsigned>();
    mp.sc = q.sc.as<nqs::CGScalars>();
    mp.stats = q.stats.as<double>();
    mp.norm2 = q.stats.as<double>() + 5;
    mp.par = a->d_par.as<double>();
    m->max_iters = std::max(m->max_iters, q.cfg.sr_cg_max_iters);
  }
  if (m->d_mp.ensure(sizeof(nqs::MultiProb) * n) || m->d_out.ensure(sizeof(double) * nqs::kMultiOut * n) ||
      hipHostMalloc((void**)&m->h_out, sizeof(double) * nqs::kMultiOut * n, hipHostMallocDefault) !=
          hipSuccess)
    return fail(0, "memory");
  if (hipEventCreateWithFlags(&m->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreate(&m->ev_s0) != hipSuccess || hipEventCreate(&m->ev_s1) != hipSuccess ||
      hipEventCreate(&m->ev_t0) != hipSuccess || hipEventCreate(&m->ev_t1) != hipSuccess)
    return fail(0, "event");
  for (int k = 0; k < nqs_hip_multi::kSide; k++)
    if (hipStreamCreateWithFlags(&m->side[k], hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&m->ev_join[k], hipEventDisableTiming) != hipSuccess)
      return fail(0, "stream");
  if (multi_set_split(m, 32) || hipStreamSynchronize(g_stream) != hipSuccess)
    return fail(0, "setup");
  return m;
}

NQS_EXPORT nqs_hip_multi_t* nqs_hip_multi_create(int n, const nqs_config_t* cfgs, const int* size_x,
                                                 const int* size_y, nqs_ansatz_t* const* ansatz,
                                                 nqs_sampler_t* const* samplers) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  return multi_new(n, cfgs, size_x, size_y, ansatz, samplers);
}

// the fields of a configuration that a CG-SR step reads
static bool same_step_cfg(const nqs_config_t& x, const nqs_config_t& y) {
  return x.hamiltonian == y.hamiltonian && x.j_coupling == y.j_coupling &&
         x.j2_coupling == y.j2_coupling && x.j_z_coupling == y.j_z_coupling &&
         x.transverse_field == y.transverse_field && x.kh_K == y.kh_K && x.kh_J == y.kh_J &&
         x.kagome_pbc == y.kagome_pbc && x.num_samples == y.num_samples &&
         x.learning_rate == y.learning_rate && x.sr_diag_shift == y.sr_diag_shift &&
         x.sr_cg_max_iters == y.sr_cg_max_iters && x.sr_cg_tol == y.sr_cg_tol;
}

// One CG-SR step of a real RBM through the batched path with a single
// problem (the sampler keeps the batch and its captured graph between
// steps).  Host parameters are uploaded before and downloaded after, as in
// the general path.
int solo_cg_step(const nqs_config_t* cfg, int Lx, int Ly, nqs_ansatz* a, nqs_sampler* s,
                 nqs_sr_step_info_t* info, bool in_run) {
  if (!in_run) a->dev_fresh = 0;
  CKR(ansatz_upload(a));
  a->dev_fresh = 1;
  const int B = cfg->num_samples;
  CKR(s->d_samples.ensure(sizeof(int) * (size_t)B * s->N));
  const bool same = s->solo && s->solo_ansatz == a->id && s->solo_Lx == Lx && s->solo_Ly == Ly &&
                    s->solo_samples == s->d_samples.p && same_step_cfg(s->solo_cfg, *cfg);
  if (!same) {
    multi_delete(s->solo);
    s->solo = nullptr;
    nqs_ansatz* av[1] = {a};
    nqs_sampler* sv[1] = {s};
    s->solo = multi_new(1, cfg, &Lx, &Ly, av, sv);
    if (!s->solo) return -1;
    s->solo_ansatz = a->id;
    s->solo_cfg = *cfg;
    s->solo_Lx = Lx;
    s->solo_Ly = Ly;
    s->solo_samples = s->d_samples.p;
  }
  CKR(multi_step(s->solo, info));
  // the step's local energies for the per-chain series
  s->last_E.resize(B);
  CK(hipMemcpyAsync(s->last_E.data(), s->solo->p[0].E.p, sizeof(double) * B, hipMemcpyDeviceToHost,
                    g_stream));
  CKR(ansatz_download(a));
  s->last_rounds = (B + s->opt.num_chains - 1) / s->opt.num_chains;
  // stage timings: the sampling launches and the whole step (the rest of
  // the step is one graph; its stages are not timed separately)
  s->have_stage = s->solo->sample_ms >= 0.0 && s->solo->step_ms >= 0.0 ? 1 : 0;
  for (int i = 0; i < 7; i++) s->stage_ms[i] = 0.0;
  if (s->have_stage) {
    s->stage_ms[0] = s->solo->sample_ms;
    s->stage_ms[6] = s->solo->step_ms;
  }
  s->cur_dirty = 1;
  if (!in_run) a->dev_fresh = 0;
  return 0;
}

NQS_EXPORT int nqs_hip_multi_step(nqs_hip_multi_t* m, nqs_sr_step_info_t* out_infos) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!m) return -1;
  return multi_step(m, out_infos);
}

NQS_EXPORT int nqs_hip_multi_run(nqs_hip_multi_t* m, int num_iterations, double* out_traces) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!m || num_iterations < 0) return -1;
  // nqs_sr_run: thermalize once, then the iterations (nqs_optimizer.c:264-305)
  for (nqs_hip_multi::Prob& q : m->p)
    CKR(sampler_advance(q.s, q.s->cfg.num_thermalize, q.s->cfg.num_thermalize > 0 ? 1 : 0, 0,
                        nullptr, nullptr));
  std::vector<nqs_sr_step_info_t> info(m->p.size());
  for (int it = 0; it < num_iterations; it++) {
    CKR(multi_step(m, info.data()));
    if (out_traces)
      for (size_t i = 0; i < m->p.size(); i++)
        out_traces[i * (size_t)num_iterations + it] = info[i].mean_energy;
  }
  return 0;
}

NQS_EXPORT int nqs_hip_multi_sync(nqs_hip_multi_t* m) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!m) return -1;
  for (nqs_hip_multi::Prob& q : m->p) {
    CKR(ansatz_download(q.a));
    CKR(sampler_sync_cur(q.s));
  }
  return 0;
}

NQS_EXPORT int nqs_hip_multi_tuning(const nqs_hip_multi_t* m, int* targets, double* matvec_us,
                                    int capacity) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!m) return -1;
  const int n = (int)m->tuning.size();
  for (int i = 0; i < n && i < capacity; i++) {
    if (targets) targets[i] = m->tuning[i].first;
    if (matvec_us) matvec_us[i] = m->tuning[i].second;
  }
  return m->target;
}

NQS_EXPORT int nqs_hip_multi_last_sample_ms(const nqs_hip_multi_t* m, double* ms) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!m || !ms || m->sample_ms < 0.0) return -1;
  *ms = m->sample_ms;
  return 0;
}

NQS_EXPORT int nqs_hip_multi_use_graph(nqs_hip_multi_t* m, int on) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!m) return -1;
  m->use_graph = on ? 1 : 0;
  return 0;
}

NQS_EXPORT int nqs_hip_multi_use_persistent(nqs_hip_multi_t* m, int on) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!m) return -1;
  if ((on ? 1 : 0) != m->persist) {
    // the captured graph's extent depends on the mode
    if (m->exec) (void)hipGraphExecDestroy(m->exec);
    if (m->graph) (void)hipGraphDestroy(m->graph);
    m->exec = nullptr;
    m->graph = nullptr;
  }
  m->persist = on ? 1 : 0;
  m->persist_set = 1;
  return 0;
}

NQS_EXPORT int nqs_hip_multi_last_energies(const nqs_hip_multi_t* m, int i, double* out,
                                           int capacity) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!m || i < 0 || i >= (int)m->p.size() || !out || !m->steps_done) return -1;
  const nqs_hip_multi::Prob& q = m->p[i];
  const int n = std::min(capacity, q.B);
  CK(hipMemcpyAsync(out, q.E.p, sizeof(double) * n, hipMemcpyDeviceToHost, g_stream));
  CK(hipStreamSynchronize(g_stream));
  return n;
}

NQS_EXPORT void nqs_hip_multi_free(nqs_hip_multi_t* m) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!m) return;
  (void)hipStreamSynchronize(g_stream);
  for (nqs_hip_multi::Prob& q : m->p) {
    (void)ansatz_download(q.a);
    (void)sampler_sync_cur(q.s);
    q.a->dev_fresh = 0;
    q.a->m64.valid = 0;
  }
  delete m;
}
