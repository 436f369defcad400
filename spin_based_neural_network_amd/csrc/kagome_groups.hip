// kagome_groups.hip — site-permutation tables of the kagome wallpaper
// groups (host code), the reference's include/nqs/nqs_symproj.h builders.
//
// Table format (the reference's, nqs_symproj.h:40-47): row g of perm
// [G][N] lists, for every output site i, the site pi_g(i) whose spin g.s
// carries there: (g.s)_i = s_{pi_g(i)}.  Site index 3 (cx L + cy) + sub,
// sub A = 0 at the cell origin, B = 1 at a1/2, C = 2 at a2/2, with
// a1 = (1, 0), a2 = (1/2, sqrt(3)/2).
//
// Geometry is done exactly in integers: a point u a1 + v a2 is held as
// (U, V) = (6u, 6v), so every site, the up-triangle centroid (1/6, 1/6) and
// the hexagon centre (1/2, 1/2) have integer coordinates.  A rotation by 60
// degrees maps a1 -> a2, a2 -> a2 - a1, i.e. (U, V) -> (-V, U + V); its
// inverse is (U, V) -> (U + V, -U).  The mirror y -> sqrt(3)/2 - y through
// the hexagon centre is (U, V) -> (U + V - 3, 6 - V).
//
// Element order and composition follow the reference builders
// (src/nqs/nqs_symproj.c:269-785) so that row g means the same element:
//   translations (tx, ty), tx-major:  source cell (cx + tx, cy + ty);
//   p2: translations, then inversion composed with translation, source
//       -p + t (nqs_symproj.c:341-364);
//   p3 / p6: rotation k, tx, ty (k-major), source R^-k (p - c) + c - t with c
//       the up-triangle centroid (p3) or the hexagon centre (p6);
//   p6m: ops 0..5 rotations R^k, ops 6..11 M R^k, then tx, ty; source
//       R^-k (M^m (p - t) - h) + h;
// characters of the C6v irreps at Gamma per op (nqs_symproj.c:655-735),
// 2D irreps carrying d_E chi(g).
#include <stdio.h>
#include <stdlib.h>

#include "../../include/nqs/nqs_symproj.h"

#define NQS_EXPORT extern "C" __attribute__((visibility("default")))

namespace {

struct P2i {
  long U, V;
};

inline long mod(long a, long m) { return ((a % m) + m) % m; }

P2i site_pos(int cx, int cy, int sub) {
  return {6L * cx + (sub == 1 ? 3 : 0), 6L * cy + (sub == 2 ? 3 : 0)};
}

// site of an integer lattice point on the L x L torus, -1 if none
int pos_site(P2i p, int L) {
  const long ru = mod(p.U, 6), rv = mod(p.V, 6);
  int sub;
  if (ru == 0 && rv == 0) sub = 0;
  else if (ru == 3 && rv == 0) sub = 1;
  else if (ru == 0 && rv == 3) sub = 2;
  else return -1;
  const long cx = mod((p.U - ru) / 6, L), cy = mod((p.V - rv) / 6, L);
  return (int)(3 * (cx * L + cy) + sub);
}

P2i rot_inv(P2i p, int k) {     // R^-k
  for (int i = 0; i < k; i++) p = {p.U + p.V, -p.U};
  return p;
}

P2i mirror(P2i p) { return {p.U + p.V - 3, 6 - p.V}; }

// C6v character of op (0..5 rotations R^k, 6..11 M R^k) in an irrep
double c6v_char(int op, nqs_symproj_kagome_irrep_t irrep) {
  const bool mir = op >= 6;
  const int k = mir ? op - 6 : op;
  const bool sv = mir && k % 2 == 0, sd = mir && k % 2 == 1;
  const bool c6 = !mir && (k == 1 || k == 5), c3 = !mir && (k == 2 || k == 4);
  const bool c2 = !mir && k == 3, e = !mir && k == 0;
  switch (irrep) {
    case NQS_SYMPROJ_KAGOME_GAMMA_A1: return 1.0;
    case NQS_SYMPROJ_KAGOME_GAMMA_A2: return mir ? -1.0 : 1.0;
    case NQS_SYMPROJ_KAGOME_GAMMA_B1: return (c6 || c2 || sd) ? -1.0 : 1.0;
    case NQS_SYMPROJ_KAGOME_GAMMA_B2: return (c6 || c2 || sv) ? -1.0 : 1.0;
    case NQS_SYMPROJ_KAGOME_GAMMA_E1:
      return mir ? 0.0 : e ? 4.0 : c6 ? 2.0 : c3 ? -2.0 : c2 ? -4.0 : 0.0;
    case NQS_SYMPROJ_KAGOME_GAMMA_E2:
      return mir ? 0.0 : e ? 4.0 : c6 ? -2.0 : c3 ? -2.0 : c2 ? 4.0 : 0.0;
  }
  return 1.0;
}

// fill one row: src(p) for every output site; false when a source is not a
// lattice point (cannot happen for the groups below)
template <typename F>
bool fill_row(int L, int* row, F src) {
  for (int cx = 0; cx < L; cx++)
    for (int cy = 0; cy < L; cy++)
      for (int sub = 0; sub < 3; sub++) {
        const int s = pos_site(src(site_pos(cx, cy, sub)), L);
        if (s < 0) return false;
        row[3 * (cx * L + cy) + sub] = s;
      }
  return true;
}

int alloc_tables(int G, int N, int** perm, double** chars) {
  *perm = (int*)malloc(sizeof(int) * (size_t)G * N);
  if (chars) *chars = (double*)malloc(sizeof(double) * (size_t)G);
  if (!*perm || (chars && !*chars)) {
    free(*perm);
    if (chars) free(*chars);
    return -1;
  }
  return 0;
}

}  // namespace

NQS_EXPORT int nqs_kagome_translation_perm(int Lx, int Ly, int** out_perm, int* out_num_elements) {
  if (Lx <= 0 || Ly <= 0 || !out_perm || !out_num_elements) return -1;
  const int N = 3 * Lx * Ly, G = Lx * Ly;
  int* perm;
  if (alloc_tables(G, N, &perm, nullptr)) return -1;
  int g = 0;
  for (int tx = 0; tx < Lx; tx++)
    for (int ty = 0; ty < Ly; ty++, g++)
      for (int cx = 0; cx < Lx; cx++)
        for (int cy = 0; cy < Ly; cy++)
          for (int sub = 0; sub < 3; sub++)
            perm[(size_t)g * N + 3 * (cx * Ly + cy) + sub] =
                3 * (((cx + tx) % Lx) * Ly + (cy + ty) % Ly) + sub;
  *out_perm = perm;
  *out_num_elements = G;
  return 0;
}

NQS_EXPORT int nqs_kagome_p2_perm(int Lx, int Ly, int** out_perm, double** out_characters,
                                  int* out_num_elements) {
  if (Lx <= 0 || Ly <= 0 || !out_perm || !out_characters || !out_num_elements) return -1;
  const int N = 3 * Lx * Ly, G = 2 * Lx * Ly;
  int* perm;
  double* chars;
  if (alloc_tables(G, N, &perm, &chars)) return -1;
  int g = 0;
  for (int inv = 0; inv < 2; inv++)
    for (int tx = 0; tx < Lx; tx++)
      for (int ty = 0; ty < Ly; ty++, g++) {
        for (int cx = 0; cx < Lx; cx++)
          for (int cy = 0; cy < Ly; cy++)
            for (int sub = 0; sub < 3; sub++) {
              // inversion through the A site: B at cx + 1/2 -> -cx - 1/2
              // lies in cell -cx - 1, C likewise in the second direction
              const int xs = inv ? (int)mod(-cx - (sub == 1) + tx, Lx) : (cx + tx) % Lx;
              const int ys = inv ? (int)mod(-cy - (sub == 2) + ty, Ly) : (cy + ty) % Ly;
              perm[(size_t)g * N + 3 * (cx * Ly + cy) + sub] = 3 * (xs * Ly + ys) + sub;
            }
        chars[g] = 1.0;
      }
  *out_perm = perm;
  *out_characters = chars;
  *out_num_elements = G;
  return 0;
}

// rotations about a centre c (in 6x lattice units), order k, tx, ty
static int rotation_tables(int L, int nrot, P2i c, const char* name, int** out_perm,
                           double** out_characters, int* out_num_elements) {
  if (L <= 0 || !out_perm || !out_characters || !out_num_elements) return -1;
  const int N = 3 * L * L, G = nrot * L * L, step = 6 / nrot;
  int* perm;
  double* chars;
  if (alloc_tables(G, N, &perm, &chars)) return -1;
  int g = 0;
  for (int k = 0; k < nrot; k++)
    for (int tx = 0; tx < L; tx++)
      for (int ty = 0; ty < L; ty++, g++) {
        const bool ok = fill_row(L, perm + (size_t)g * N, [&](P2i p) {
          P2i r = rot_inv({p.U - c.U, p.V - c.V}, k * step);
          return P2i{r.U + c.U - 6L * tx, r.V + c.V - 6L * ty};
        });
        if (!ok) {
          fprintf(stderr, "%s: row %d is not a lattice map\n", name, g);
          free(perm);
          free(chars);
          return -1;
        }
        chars[g] = 1.0;
      }
  *out_perm = perm;
  *out_characters = chars;
  *out_num_elements = G;
  return 0;
}

NQS_EXPORT int nqs_kagome_p3_perm(int L, int** out_perm, double** out_characters,
                                  int* out_num_elements) {
  // C3 about the centroid of up triangle (0, 0): (1/6, 1/6) in lattice units
  return rotation_tables(L, 3, P2i{1, 1}, "nqs_kagome_p3_perm", out_perm, out_characters,
                         out_num_elements);
}

NQS_EXPORT int nqs_kagome_p6_perm(int L, int** out_perm, double** out_characters,
                                  int* out_num_elements) {
  // C6 about the hexagon centre (a1 + a2) / 2
  return rotation_tables(L, 6, P2i{3, 3}, "nqs_kagome_p6_perm", out_perm, out_characters,
                         out_num_elements);
}

NQS_EXPORT int nqs_kagome_p6m_perm_irrep(int L, nqs_symproj_kagome_irrep_t irrep, int** out_perm,
                                         double** out_characters, int* out_num_elements) {
  if (L <= 0 || !out_perm || !out_characters || !out_num_elements) return -1;
  const int N = 3 * L * L, G = 12 * L * L;
  int* perm;
  double* chars;
  if (alloc_tables(G, N, &perm, &chars)) return -1;
  const P2i h{3, 3};
  int g = 0;
  for (int op = 0; op < 12; op++) {
    const double chi = c6v_char(op, irrep);
    for (int tx = 0; tx < L; tx++)
      for (int ty = 0; ty < L; ty++, g++) {
        const bool ok = fill_row(L, perm + (size_t)g * N, [&](P2i p) {
          P2i q{p.U - 6L * tx, p.V - 6L * ty};
          if (op >= 6) q = mirror(q);
          P2i r = rot_inv({q.U - h.U, q.V - h.V}, op % 6);
          return P2i{r.U + h.U, r.V + h.V};
        });
        if (!ok) {
          fprintf(stderr, "nqs_kagome_p6m_perm_irrep: row %d is not a lattice map\n", g);
          free(perm);
          free(chars);
          return -1;
        }
        chars[g] = chi;
      }
  }
  *out_perm = perm;
  *out_characters = chars;
  *out_num_elements = G;
  return 0;
}

NQS_EXPORT int nqs_kagome_p6m_perm(int L, int** out_perm, double** out_characters,
                                   int* out_num_elements) {
  return nqs_kagome_p6m_perm_irrep(L, NQS_SYMPROJ_KAGOME_GAMMA_A1, out_perm, out_characters,
                                   out_num_elements);
}
