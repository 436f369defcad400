"""Parity at large weights: RBM / complex-RBM inits of scale 3 to 10, where
|theta| reaches 20-40 and tanh(theta) rounds to +-1.

The device prices a connected configuration as prod_h (p_h e^{d_h} +
q_h e^{-d_h}) with the logistic pair (p, q) of theta_h (cancellation-free);
the oracle, like the reference (src/nqs/nqs_ansatz.c:277-294, 309-318),
differences two log-domain amplitudes.  Local energies must agree to 1e-10
relative to the largest |E_loc| of the batch, and Metropolis chains (whose
acceptance tests use the same ratios) must draw the oracle's samples bit for
bit.  Weights beyond the served range (|Re W| > 20) are refused loudly.
"""
import ctypes

import numpy as np
import pytest

import scenarios
from spin_based_neural_network_amd.abi import config_defaults, fnptr, options_defaults, ptr

pytestmark = pytest.mark.gpu


def _close(a, b, tol):
    a, b = np.asarray(a, float), np.asarray(b, float)
    s = max(1.0, float(np.max(np.abs(b))))
    err = float(np.max(np.abs(a - b)))
    assert err <= tol * s, f"max err {err:.3e} > {tol * s:.3e}"


def _pair(hip, orc, cfg, N):
    ah = hip.nqs_ansatz_create(ctypes.byref(cfg), N)
    ao = orc.orc_ansatz_create(ctypes.byref(cfg), N)
    assert ah and ao
    return ah, ao


def _theta_max(hip, a, cfg, N, spins):
    """max |Re theta| over the batch (host restatement, for the record)."""
    P = hip.nqs_ansatz_num_params(a)
    par = np.ctypeslib.as_array(ctypes.cast(hip.nqs_ansatz_params_raw(a), ctypes.POINTER(ctypes.c_double)), (P,))
    M = cfg.rbm_hidden_units
    b, W = par[N:N + M], par[N + M:N + M + M * N].reshape(M, N)
    return float(np.abs(spins @ W.T + b).max()), float(np.abs(W).max())


@pytest.mark.parametrize("ham,kind,Lx,Ly,M,scale,per,marsh", [
    (1, 6, 6, 6, 144, 3.0, 1, True),     # C2 shape: 6x6 Heisenberg + Marshall, cRBM
    (1, 6, 6, 6, 144, 10.0, 1, True),
    (6, 6, 2, 4, 96, 6.0, 1, False),     # C5 shape: kagome 2x4, cRBM M=96
    (6, 6, 2, 4, 96, 10.0, 1, False),
    (0, 5, 16, 1, 32, 10.0, 1, False),   # C1 shape: TFIM chain, real RBM
    (2, 5, 10, 10, 800, 3.0, 1, False),  # C3 shape: J1-J2 10x10, real RBM alpha=8
])
def test_large_weight_local_energy_vs_oracle(hip, orc, ham, kind, Lx, Ly, M, scale, per, marsh):
    N = 3 * Lx * Ly if ham == 6 else Lx * Ly
    cfg = config_defaults(ansatz=kind, hamiltonian=ham, rbm_hidden_units=M, rbm_init_scale=scale,
                          j2_coupling=0.5, transverse_field=0.7, rng_seed=77)
    ah, ao = _pair(hip, orc, cfg, N)
    B = 16 if M >= 800 else 64
    s = scenarios.spins_random(B, N, 5, sz0=ham != 0)
    th, wm = _theta_max(hip, ah, cfg, N, s)
    print(f"ham={ham} N={N} M={M} scale={scale}: max|theta| = {th:.1f}, max|W| = {wm:.2f}", flush=True)
    fh, fo = fnptr(hip, "nqs_ansatz_log_amp"), fnptr(orc, "orc_ansatz_log_amp")
    uh, uo, keep = ah, ao, []
    if marsh:
        from spin_based_neural_network_amd.abi import MarshallWrapper
        wh, wo = MarshallWrapper(fh, ah, Lx, Ly), MarshallWrapper(fo, ao, Lx, Ly)
        keep += [wh, wo]
        fh, uh = fnptr(hip, "nqs_marshall_log_amp"), ctypes.addressof(wh)
        fo, uo = fnptr(orc, "orc_marshall_log_amp"), ctypes.addressof(wo)
    opt = options_defaults(periodic=per)
    eh, ehi, eo, eoi = (np.zeros(B) for _ in range(4))
    assert hip.nqs_hip_local_energy_batch_ex(ctypes.byref(cfg), Lx, Ly, ctypes.byref(opt), ptr(s), B,
                                             fh, uh, ptr(eh), ptr(ehi)) == 0
    orc.orc_local_energy_batch_ex(ctypes.byref(cfg), Lx, Ly, per, ptr(s), B, fo, uo, ptr(eo), ptr(eoi))
    scale_e = max(np.abs(eo).max(), np.abs(eoi).max())
    print(f"  max|E_loc| = {scale_e:.3e}, err re {np.abs(eh - eo).max():.2e}, "
          f"im {np.abs(ehi - eoi).max():.2e}", flush=True)
    assert np.all(np.isfinite(eh)) and np.all(np.isfinite(ehi))
    assert np.abs(eh - eo).max() <= 1e-10 * max(1.0, scale_e)
    assert np.abs(ehi - eoi).max() <= 1e-10 * max(1.0, scale_e)
    hip.nqs_ansatz_free(ah)
    orc.orc_ansatz_free(ao)


@pytest.mark.parametrize("kind,ham,L,M,scale,move", [
    (5, 0, (16, 1), 32, 4.0, 0),      # flips, real RBM
    (6, 1, (6, 6), 144, 3.0, 1),      # exchange, C2 shape
    (6, 6, (2, 4), 96, 5.0, 1),       # exchange, C5 shape
])
def test_large_weight_sampler_bit_exact_vs_oracle(hip, orc, kind, ham, L, M, scale, move):
    """Multi-chain Metropolis at large weights draws the oracle's samples."""
    Lx, Ly = L
    N = 3 * Lx * Ly if ham == 6 else Lx * Ly
    cfg = config_defaults(ansatz=kind, hamiltonian=ham, rbm_hidden_units=M, rbm_init_scale=scale,
                          num_thermalize=3 * N, num_decorrelate=N, rng_seed=31)
    opt = options_defaults(num_chains=16, move=move, sz_zero_start=move, periodic=1)
    out = {}
    for name, lib, pre in (("hip", hip, "nqs_"), ("orc", orc, "orc_")):
        a = getattr(lib, pre + "ansatz_create")(ctypes.byref(cfg), N)
        fn = fnptr(lib, pre + "ansatz_log_amp")
        create = lib.nqs_hip_sampler_create_ex if name == "hip" else lib.orc_sampler_create_ex
        s = create(N, ctypes.byref(cfg), ctypes.byref(opt), fn, a)
        getattr(lib, pre + "sampler_thermalize")(s)
        smp = np.zeros((64, N), dtype=np.int32)
        assert getattr(lib, pre + "sampler_batch")(s, 64, ptr(smp)) == 0
        out[name] = (smp, getattr(lib, pre + "sampler_acceptance_ratio")(s))
        getattr(lib, pre + "sampler_free")(s)
        getattr(lib, pre + "ansatz_free")(a)
    print(f"kind={kind} ham={ham} scale={scale}: acceptance {out['orc'][1]:.3f}", flush=True)
    np.testing.assert_array_equal(out["hip"][0], out["orc"][0])
    assert out["hip"][1] == out["orc"][1]
    assert 0.0 < out["orc"][1] < 1.0


def test_weights_beyond_table_range_refused(hip, capfd):
    """|Re W| above 20 is refused with a message: moves and local energies
    return an error instead of unreliable ratios; amplitudes still work."""
    cfg = config_defaults(ansatz=5, hamiltonian=0, rbm_hidden_units=8, rng_seed=3)
    N = 8
    a = hip.nqs_ansatz_create(ctypes.byref(cfg), N)
    P = hip.nqs_ansatz_num_params(a)
    par = np.ctypeslib.as_array(ctypes.cast(hip.nqs_ansatz_params_raw(a), ctypes.POINTER(ctypes.c_double)), (P,))
    par[N + 8 + 5] = 25.0
    s = scenarios.spins_random(4, N, 1)
    fn = fnptr(hip, "nqs_ansatz_log_amp")
    e, ei = np.zeros(4), np.zeros(4)
    opt = options_defaults()
    assert hip.nqs_hip_local_energy_batch_ex(ctypes.byref(cfg), N, 1, ctypes.byref(opt), ptr(s), 4, fn, a,
                                             ptr(e), ptr(ei)) == -1
    assert "beyond the range" in capfd.readouterr().err
    la, ar = np.zeros(4), np.zeros(4)
    assert hip.nqs_hip_log_amp_batch(a, ptr(s), 4, ptr(la), ptr(ar)) == 0
    assert np.all(np.isfinite(la))
    hip.nqs_ansatz_free(a)
