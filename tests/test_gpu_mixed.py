"""Mixed precision (NQS_HIP_MIXED): the dense MinSR Gram on the split
16-bit MFMA against the FP64 Gram on the same samples
(nqs_hip_minsr_solve_ex).

Bound (include/nqs_hip/nqs_hip.h): the FP32 entries are scaled by a power of
two and split into hi + lo halves, each product taken as hi hi + hi lo +
lo hi, error at most about 3 * 2^-22 |a b|; the products accumulate in FP32
over at most 2048 terms (at most 3 * 2048 / 32 = 192 roundings, each at most
2^-24 of the running sum, so at most 1.2e-5 of the sum of |products|), and
the 2048-term pieces add in FP64; the O entries come from FP32 amplitudes
(relative error ~1e-7).  So

    |T_mixed[i][j] - T_fp64[i][j]| <= 2e-5 * (|O_c| |O_c|^T)[i][j] / Ns

elementwise, with |O_c| from the oracle's FP64 gradients of the same ansatz.
The MinSR update delta inherits the Gram's error through the shifted solve;
its relative deviation from the FP64 update is measured and bounded here at
the shifts the engine's MinSR configurations use.
"""
import ctypes

import numpy as np
import pytest

from spin_based_neural_network_amd.abi import (FP32, FP64, GRAM_DIRECT, GRAM_PARAMETER, MIXED,
                                               config_defaults, options_defaults, ptr)

pytestmark = pytest.mark.gpu

CRBM, KAGOME = 6, 6


def _setup(hip, orc, Lx, Ly, M, B, seed, ansatz=CRBM):
    N = 3 * Lx * Ly
    cfg = config_defaults(ansatz=ansatz, hamiltonian=KAGOME, kagome_pbc=1, rbm_hidden_units=M,
                          rbm_init_scale=0.1, rng_seed=seed)
    a = hip.nqs_ansatz_create(ctypes.byref(cfg), N)
    o = orc.orc_ansatz_create(ctypes.byref(cfg), N)
    P = hip.nqs_ansatz_num_params(a)
    rng = np.random.default_rng(seed)
    spins = np.where(rng.random((B, N)) < 0.5, 1, -1).astype(np.int32)
    e = rng.normal(-0.4 * N, 1.0, B)
    O = np.zeros((B, P))
    for b in range(B):
        assert orc.orc_ansatz_logpsi_gradient(o, ptr(np.ascontiguousarray(spins[b])), N,
                                              ptr(O[b])) == 0
    orc.orc_ansatz_free(o)
    return a, P, spins, e, O - O.mean(0)


def _solve(hip, a, spins, e, shift, prec, gram, want_T):
    B = spins.shape[0]
    P = hip.nqs_ansatz_num_params(a)
    d = np.zeros(P)
    T = np.zeros((B, B)) if want_T else None
    opt = options_defaults(precision=prec, gram=gram)
    rc = hip.nqs_hip_minsr_solve_ex(a, ptr(spins), B, ptr(e), shift, ctypes.byref(opt), ptr(d),
                                    None, ptr(T) if want_T else None)
    assert rc == 1, f"solve (precision {prec}, gram {gram}) returned {rc}"
    return d, T


@pytest.mark.parametrize("Lx,Ly,M,B", [(2, 2, 24, 1000), (2, 4, 96, 2048)])
def test_mixed_gram_sample_space_within_bound(hip, orc, Lx, Ly, M, B):
    """T = O_c O_c^T / Ns + shift I (K = P, padded to 32; B = 1000 leaves
    partial 128-row tiles) against the FP64 Gram, elementwise, and the
    FP32-amplitude Gram for the split of the error."""
    a, P, spins, e, Oc = _setup(hip, orc, Lx, Ly, M, B, 700 + M)
    shift = 1e-3
    print(f"[mixed] sample-space Gram: N={3 * Lx * Ly} M={M} P={P} B={B}", flush=True)
    d64, T64 = _solve(hip, a, spins, e, shift, FP64, GRAM_DIRECT, True)
    d32, T32 = _solve(hip, a, spins, e, shift, FP32, GRAM_DIRECT, True)
    dmx, Tmx = _solve(hip, a, spins, e, shift, MIXED, GRAM_DIRECT, True)
    hip.nqs_ansatz_free(a)
    low = np.tril_indices(B)
    ref = (Oc @ Oc.T / B + shift * np.eye(B))[low]
    np.testing.assert_allclose(T64[low], ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    bound = 2e-5 * (np.abs(Oc) @ np.abs(Oc).T / B)[low]
    err32 = np.abs(T32[low] - T64[low])
    errmx = np.abs(Tmx[low] - T64[low])
    scale = np.abs(T64[low]).max()
    print(f"[mixed] max |T32 - T64| / max|T| = {err32.max() / scale:.3e}; "
          f"max |Tmx - T64| / max|T| = {errmx.max() / scale:.3e}; "
          f"worst ratio to the bound {np.max(errmx / bound):.3f}", flush=True)
    assert np.all(errmx <= bound)
    assert np.all(err32 <= 1e-5 * bound / 2e-5)
    rel32 = np.linalg.norm(d32 - d64) / np.linalg.norm(d64)
    relmx = np.linalg.norm(dmx - d64) / np.linalg.norm(d64)
    print(f"[mixed] |d32 - d64| / |d64| = {rel32:.3e}, |dmx - d64| / |d64| = {relmx:.3e}",
          flush=True)
    assert relmx < 1e-3


@pytest.mark.parametrize("shift", [1e-2, 1e-4])
def test_mixed_parameter_space_update(hip, orc, shift):
    """The parameter-space solve (S = O_c^T O_c / Ns + shift I, K = Ns =
    4096 samples, P = 4848: the C5 ansatz) in the mixed mode against FP64 on
    the same samples, and against the FP64 sample-space update (the same
    delta by the push-through identity)."""
    a, P, spins, e, Oc = _setup(hip, orc, 2, 4, 96, 4096, 901)
    print(f"[mixed] parameter-space update: P={P} B=4096 shift={shift}", flush=True)
    d64, _ = _solve(hip, a, spins, e, shift, FP64, GRAM_PARAMETER, False)
    dmx, _ = _solve(hip, a, spins, e, shift, MIXED, GRAM_PARAMETER, False)
    dss, _ = _solve(hip, a, spins, e, shift, FP64, GRAM_DIRECT, False)
    hip.nqs_ansatz_free(a)
    rel_dual = np.linalg.norm(d64 - dss) / np.linalg.norm(dss)
    relmx = np.linalg.norm(dmx - d64) / np.linalg.norm(d64)
    cos = float(dmx @ d64) / (np.linalg.norm(dmx) * np.linalg.norm(d64))
    print(f"[mixed] |d64(param) - d64(sample)| / |d| = {rel_dual:.3e}; "
          f"|dmx - d64| / |d64| = {relmx:.3e}, cos = {cos:.9f}", flush=True)
    assert rel_dual < 1e-6
    assert relmx < (1e-4 if shift >= 1e-2 else 5e-3)
    assert cos > 0.99999


@pytest.mark.parametrize("ansatz,M,B", [(CRBM, 96, 4096), (5, 40, 1000), (5, 3, 200)])
def test_mixed_parameter_space_without_o_matrix(hip, orc, monkeypatch, ansatz, M, B):
    """The mixed parameter-space solve of an RBM takes O_c^T (FP32) from the
    spins and factors and the means and force from FP64 factor sums, without
    forming O: against the path through the materialised O on the same
    samples (same FP32 entries up to the rounding of the column means, so
    the updates agree far inside the mixed mode's bound), and against the
    FP64 update.  Cases: the C5 complex RBM and a real RBM with B = 1000
    (samples padded to 1024), and a small real RBM (P = 99, fewer parameter
    rows than the generator's hidden-unit slices)."""
    a, P, spins, e, Oc = _setup(hip, orc, 2, 4, M, B, 905, ansatz=ansatz)
    shift = 1e-2
    print(f"[mixed] parameter space without O: ansatz {ansatz} P={P} B={B}", flush=True)
    monkeypatch.setenv("NQS_HIP_FACTORED_DUAL", "0")
    d_mat, _ = _solve(hip, a, spins, e, shift, MIXED, GRAM_PARAMETER, False)
    monkeypatch.setenv("NQS_HIP_FACTORED_DUAL", "1")
    d_fac, _ = _solve(hip, a, spins, e, shift, MIXED, GRAM_PARAMETER, False)
    d64, _ = _solve(hip, a, spins, e, shift, FP64, GRAM_PARAMETER, False)
    hip.nqs_ansatz_free(a)
    rel = np.linalg.norm(d_fac - d_mat) / np.linalg.norm(d_mat)
    rel64 = np.linalg.norm(d_fac - d64) / np.linalg.norm(d64)
    print(f"[mixed] |d(no O) - d(O)| / |d| = {rel:.3e}; against FP64 {rel64:.3e}", flush=True)
    assert rel < 1e-5
    assert rel64 < 1e-4


@pytest.mark.parametrize("ansatz,M,B,split_dual", [(CRBM, 96, 4096, "1"), (5, 40, 1000, "1"),
                                                   (5, 40, 1000, "0")])
def test_mixed_gram_split_bit_equal(hip, orc, monkeypatch, ansatz, M, B, split_dual):
    """S formed in two parts (the first wide panel's columns, then the
    trailing block on the factorisation's second stream, beside that panel)
    gives the update of the single product bit for bit.  Cases: the C5
    complex RBM (P = 4848: ten wide panels), a real RBM with P = 1024 (two
    wide panels), with and without the materialised O."""
    a, P, spins, e, Oc = _setup(hip, orc, 2, 4, M, B, 907, ansatz=ansatz)
    monkeypatch.setenv("NQS_HIP_FACTORED_DUAL", split_dual)
    print(f"[mixed] Gram split: ansatz {ansatz} P={P} B={B} factored {split_dual}", flush=True)
    out = {}
    for sp in ("0", "1", "0"):
        monkeypatch.setenv("NQS_HIP_GRAM_SPLIT", sp)
        d, _ = _solve(hip, a, spins, e, 1e-3, MIXED, GRAM_PARAMETER, False)
        if sp in out:
            np.testing.assert_array_equal(d, out[sp])
        out[sp] = d
    hip.nqs_ansatz_free(a)
    assert P > 512
    np.testing.assert_array_equal(out["1"], out["0"])
    assert np.all(np.isfinite(out["1"])) and np.linalg.norm(out["1"]) > 0
