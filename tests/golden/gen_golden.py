"""Write tests/golden/golden.npz from the compiled reference.

Run where the reference checkout exists (it does not on the GPU machine):

    make -C oracle ref && python tests/golden/gen_golden.py

``oracle/_ref/libnqs_ref.so`` is the reference's own src/nqs + src/mps
sources compiled unmodified by oracle/Makefile.  Every scenario in
tests/scenarios.py is executed through its ``nqs_*`` entry points and the
outputs (inputs included) are stored as arrays; nothing from the reference
source is copied.  ED anchors come from the reference's Lanczos entry
points (src/nqs/nqs_lanczos.c:210-224, 300-321, 477-502).
"""
import ctypes
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from oracle import load_reference  # noqa: E402
from spin_based_neural_network_amd.abi import config_defaults  # noqa: E402
import scenarios  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.npz")


def ed_anchors(lib):
    """E0 from the reference Lanczos refiners (seeded by an RBM)."""
    c_d = ctypes.c_double
    out = {}
    res_buf = (ctypes.c_char * 4096)()

    def ans(N, ham):
        cfg = config_defaults(ansatz=5, hamiltonian=ham, rbm_init_scale=0.1)
        return lib.nqs_ansatz_create(ctypes.byref(cfg), N)

    f = lib.nqs_lanczos_refine_tfim
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, c_d, c_d, ctypes.c_int, c_d,
                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    for (Lx, Ly, J, G) in [(2, 2, 1.0, 1.0), (4, 1, 1.0, 1.0), (3, 3, 1.0, 0.5), (8, 1, 1.0, 1.0)]:
        a = ans(Lx * Ly, 0)
        e = c_d(0)
        rc = f(a, Lx, Ly, J, G, 200, 1e-12, ctypes.byref(e), None, res_buf)
        assert rc == 0
        out[f"ed/tfim_{Lx}x{Ly}_J{J}_G{G}"] = np.array([e.value])
        lib.nqs_ansatz_free(a)
    f = lib.nqs_lanczos_refine_heisenberg
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, c_d, c_d, ctypes.c_int, c_d,
                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    for (Lx, Ly, J, Jz) in [(4, 1, 1.0, 1.0), (8, 1, 1.0, 1.0), (4, 2, 1.0, 1.0), (3, 3, 1.0, 0.5)]:
        a = ans(Lx * Ly, 1)
        e = c_d(0)
        rc = f(a, Lx, Ly, J, Jz, 200, 1e-12, ctypes.byref(e), None, res_buf)
        assert rc == 0
        out[f"ed/heis_{Lx}x{Ly}_J{J}_Jz{Jz}"] = np.array([e.value])
        lib.nqs_ansatz_free(a)
    f = lib.nqs_lanczos_refine_kagome_heisenberg
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, c_d, ctypes.c_int, ctypes.c_int, c_d,
                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    for (Lx, Ly, pbc) in [(2, 2, 1), (2, 2, 0), (1, 1, 1)]:
        a = ans(3 * Lx * Ly, 6)
        e = c_d(0)
        rc = f(a, Lx, Ly, 1.0, pbc, 300, 1e-12, ctypes.byref(e), None, res_buf)
        assert rc == 0
        out[f"ed/kagome_{Lx}x{Ly}_pbc{pbc}"] = np.array([e.value])
        lib.nqs_ansatz_free(a)
    # the reference's lean Lanczos (identity group): E_0, iterations, converged
    from spin_based_neural_network_amd.abi import LanczosResult
    f = lib.nqs_lanczos_e0_kagome_heisenberg_projected_lean
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int, ctypes.c_int, c_d, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                  ctypes.c_int, ctypes.c_int, c_d, ctypes.c_void_p, ctypes.POINTER(LanczosResult)]
    for (Lx, Ly, pbc, iters, tol) in [(1, 1, 1, 50, 1e-9), (2, 2, 1, 300, 1e-11),
                                      (2, 2, 0, 300, 1e-11), (1, 2, 1, 300, 1e-11)]:
        N = 3 * Lx * Ly
        perm = np.arange(N, dtype=np.int32)
        chi = np.ones(1)
        e = c_d(0)
        lr = LanczosResult()
        assert f(Lx, Ly, 1.0, pbc, perm.ctypes.data, chi.ctypes.data, 1, iters, tol,
                 ctypes.byref(e), ctypes.byref(lr)) == 0
        out[f"lean/kagome_{Lx}x{Ly}_pbc{pbc}"] = np.array([e.value, lr.iterations, lr.converged,
                                                          lr.residual_norm])
    # deterministic variational energies <psi|H|psi> of a real RBM
    f = lib.nqs_exact_energy_tfim
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, c_d, c_d, ctypes.c_void_p]
    cfg = config_defaults(ansatz=5, rbm_init_scale=0.4, rng_seed=21)
    a = lib.nqs_ansatz_create(ctypes.byref(cfg), 9)
    e = c_d(0)
    assert f(a, 3, 3, 1.0, 0.8, ctypes.byref(e)) == 0
    out["exact/tfim_3x3_rbm_s04_seed21_G08"] = np.array([e.value])
    lib.nqs_ansatz_free(a)
    return out


def main():
    lib = load_reference()
    if lib is None:
        raise SystemExit("oracle/_ref/libnqs_ref.so missing: run `make -C oracle ref`")
    flat = {}
    res = scenarios.run_all(lib, "nqs_")
    for key, d in res.items():
        for k, v in d.items():
            flat[f"{key}/{k}"] = np.asarray(v)
    flat.update(ed_anchors(lib))
    np.savez_compressed(OUT, **flat)
    print(f"wrote {OUT}: {len(flat)} arrays, {os.path.getsize(OUT) / 1e6:.2f} MB")


if __name__ == "__main__":
    main()
