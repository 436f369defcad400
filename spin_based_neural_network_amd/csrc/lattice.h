// lattice.h — Hamiltonian term lists built on the host, in the bond order of
// the reference kernels (src/nqs/nqs_gradient.c: TFIM :50-82, XXZ /
// Heisenberg :87-139, J1-J2 :146-205, kagome :374-445), plus periodic
// square lattices (a dimension of length >= 3 wraps).
#pragma once

#include <vector>

#include "../../include/nqs/nqs_config.h"

namespace nqs {

struct HamHost {
  std::vector<int> diag_a, diag_b, flip_i, pair_a, pair_b;
  std::vector<double> diag_c, flip_c, pair_c;
  int n_sites = 0;
};

// 0 on success; -1 for Hamiltonians the device path does not provide.
inline int ham_sites(const nqs_config_t* cfg, int Lx, int Ly) {
  return cfg->hamiltonian == NQS_HAM_KAGOME_HEISENBERG ? 3 * Lx * Ly : Lx * Ly;
}

inline int build_hamiltonian(const nqs_config_t* cfg, int Lx, int Ly, int periodic, HamHost& H) {
  H = HamHost();
  H.n_sites = ham_sites(cfg, Lx, Ly);
  auto nxt = [&](int x, int L) { return (x + 1 < L) ? x + 1 : ((periodic && L >= 3) ? 0 : -1); };
  auto pair = [&](int a, int b, double dc, double oc) {
    H.diag_a.push_back(a); H.diag_b.push_back(b); H.diag_c.push_back(dc);
    H.pair_a.push_back(a); H.pair_b.push_back(b); H.pair_c.push_back(oc);
  };
  const int ham = cfg->hamiltonian;
  if (ham == NQS_HAM_TFIM) {
    const double J = cfg->j_coupling, G = cfg->transverse_field;
    for (int x = 0; x < Lx; x++)
      for (int y = 0; y < Ly; y++) {
        const int xn = nxt(x, Lx), yn = nxt(y, Ly);
        if (xn >= 0) { H.diag_a.push_back(x * Ly + y); H.diag_b.push_back(xn * Ly + y); H.diag_c.push_back(-J); }
        if (yn >= 0) { H.diag_a.push_back(x * Ly + y); H.diag_b.push_back(x * Ly + yn); H.diag_c.push_back(-J); }
      }
    for (int i = 0; i < Lx * Ly; i++) { H.flip_i.push_back(i); H.flip_c.push_back(-G); }
    return 0;
  }
  if (ham == NQS_HAM_HEISENBERG || ham == NQS_HAM_XXZ || ham == NQS_HAM_J1_J2) {
    const double Jxy = cfg->j_coupling;
    const double Jz = (ham == NQS_HAM_XXZ) ? cfg->j_z_coupling : cfg->j_coupling;
    for (int x = 0; x < Lx; x++)
      for (int y = 0; y < Ly; y++) {
        const int xn = nxt(x, Lx), yn = nxt(y, Ly);
        if (xn >= 0) pair(x * Ly + y, xn * Ly + y, 0.25 * Jz, 0.5 * Jxy);
        if (yn >= 0) pair(x * Ly + y, x * Ly + yn, 0.25 * Jz, 0.5 * Jxy);
      }
    if (ham == NQS_HAM_J1_J2) {
      const double J2 = cfg->j2_coupling;
      const int xe = (periodic && Lx >= 3) ? Lx : Lx - 1;
      const int ye = (periodic && Ly >= 3) ? Ly : Ly - 1;
      for (int x = 0; x < xe; x++)
        for (int y = 0; y < ye; y++) {
          const int x1 = (x + 1) % Lx, y1 = (y + 1) % Ly;
          pair(x * Ly + y, x1 * Ly + y1, 0.25 * J2, 0.5 * J2);
          pair(x * Ly + y1, x1 * Ly + y, 0.25 * J2, 0.5 * J2);
        }
    }
    return 0;
  }
  if (ham == NQS_HAM_KAGOME_HEISENBERG) {
    const double J = cfg->j_coupling;
    for (int cx = 0; cx < Lx; cx++)
      for (int cy = 0; cy < Ly; cy++) {
        const int A = 3 * (cx * Ly + cy);
        pair(A, A + 1, 0.25 * J, 0.5 * J);
        pair(A, A + 2, 0.25 * J, 0.5 * J);
        pair(A + 1, A + 2, 0.25 * J, 0.5 * J);
        int cxm, cym;
        if (cfg->kagome_pbc) {
          cxm = (cx - 1 + Lx) % Lx;
          cym = (cy - 1 + Ly) % Ly;
        } else {
          if (cx == 0 || cy == 0) continue;
          cxm = cx - 1;
          cym = cy - 1;
        }
        const int Bm = 3 * (cxm * Ly + cy) + 1, Cm = 3 * (cx * Ly + cym) + 2;
        pair(A, Bm, 0.25 * J, 0.5 * J);
        pair(A, Cm, 0.25 * J, 0.5 * J);
        pair(Bm, Cm, 0.25 * J, 0.5 * J);
      }
    return 0;
  }
  return -1;
}

}  // namespace nqs
