"""The algebra of the factored MinSR Gram matrix, checked against explicit O rows.

For an RBM the log-derivative row of a sample is O = (s, t, t (x) s) with
t = tanh(theta) (src/nqs/nqs_ansatz.c gradient, restated in
oracle/nqs_oracle.c), so the kernels never form O and build the Gram matrix
from the factors instead:

- product:   O_i . O_j = (1 + s_i . s_j)(1 + u_i . u_j) - 1, where u = t for
  a real RBM and u = [Re t | Im t] for a complex RBM in the real
  (Re-projected) gradient the reference's MinSR uses
  (src/nqs/nqs_optimizer.c:325-470);
- row sums:  sum_j G_ij = s_i . (sum s + Y_i) + u_i . sum u, with K = S^T U
  and Y = U K^T;
- centring:  T = (G - r_i - r_j + g) / B with r the row means and g the grand
  mean of G equals Oc Oc^T / B for the column-centred O.

Here O comes from the oracle's own gradient routine on random spins, and s
and u are read out of O's a- and b-blocks, so a gradient with a dropped or
mis-indexed term, or a wrong factored formula, fails.
"""
import numpy as np
import pytest

import scenarios
from spin_based_neural_network_amd.abi import ANSATZ_CRBM, ANSATZ_RBM, config_defaults, ptr


def _rows(orc, kind, N, M, B, seed):
    r = scenarios.Runner(orc, "orc_")
    cfg = config_defaults(ansatz=kind, rbm_hidden_units=M, rbm_init_scale=0.3, rng_seed=seed)
    h = r.ansatz(cfg, N)
    P = r.f("ansatz_num_params")(h)
    spins = scenarios.spins_random(B, N, seed)
    O = np.zeros((B, P))
    for b in range(B):
        row = np.ascontiguousarray(spins[b])
        assert r.f("ansatz_logpsi_gradient")(h, ptr(row), N, ptr(O[b])) == 0
    r.f("ansatz_free")(h)
    return spins, O


@pytest.mark.parametrize("kind", [ANSATZ_RBM, ANSATZ_CRBM], ids=["rbm", "crbm"])
def test_factored_gram_identities(orc, kind):
    N, M, B = 12, 24, 40
    spins, O = _rows(orc, kind, N, M, B, seed=5)
    P = O.shape[1]
    half = P // 2 if kind == ANSATZ_CRBM else P
    assert half == N + M + M * N

    S = O[:, :N]
    np.testing.assert_array_equal(S, spins.astype(float))
    U = O[:, N:N + M]
    if kind == ANSATZ_CRBM:
        U = np.hstack([U, O[:, half + N:half + N + M]])
    assert np.abs(U).max() > 0.05, "hidden activations too small to test anything"

    G = O @ O.T
    F = (1.0 + S @ S.T) * (1.0 + U @ U.T) - 1.0
    np.testing.assert_allclose(F, G, rtol=1e-12, atol=1e-12 * np.abs(G).max())

    K = S.T @ U
    Y = U @ K.T
    rows = (S * (S.sum(0) + Y)).sum(1) + U @ U.sum(0)
    np.testing.assert_allclose(rows, G.sum(1), rtol=1e-12, atol=1e-12 * np.abs(G).max() * B)

    r = G.mean(1)
    g = G.mean()
    T = (G - r[:, None] - r[None, :] + g) / B
    Oc = O - O.mean(0)
    np.testing.assert_allclose(T, Oc @ Oc.T / B, rtol=1e-10, atol=1e-12 * np.abs(G).max())
