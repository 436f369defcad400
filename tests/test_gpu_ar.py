"""GPU tests of the learned proposal (k_ar.hip) through the C ABI.

* the network initialisation equals the oracle's bit for bit;
* the device log q sums to one over the support (autoregressive masks and
  the walk's site order on the device) and matches the oracle's BF16 forward
  within BF16 rounding;
* sampling and evaluation are the same function: a sample's log q equals
  the evaluation of that configuration exactly, and sample frequencies match
  q (chi-square);
* the training gradient matches the oracle's FP64 backward within the BF16
  bound, and Adam's first step follows its closed form;
* a learned-proposal sampler leaves |psi|^2 invariant: chain states after
  learned moves with a deliberately poor network follow the exact
  distribution by enumeration (chi-square); on a multimodal |psi|^2 the
  trained network reaches the exact <E_loc> where exchange moves at the same
  sample count do not, at a higher effective-sample rate.
"""
import ctypes
import itertools

import numpy as np
import pytest
from scipy import stats

from spin_based_neural_network_amd import abi
from spin_based_neural_network_amd.abi import ptr
from spin_based_neural_network_amd.stats import pooled_tau

pytestmark = pytest.mark.gpu


def _configs(N, n_down=None):
    rows = [c for c in itertools.product([1, -1], repeat=N)
            if n_down is None or sum(1 for x in c if x < 0) == n_down]
    return np.array(rows, dtype=np.int32)


def _sector(rng, B, N, nd):
    sp = np.ones((B, N), dtype=np.int32)
    for b in range(B):
        sp[b, rng.permutation(N)[:nd]] = -1
    return sp


class Net:
    def __init__(self, hip, N, seed, scale):
        self.hip, self.N = hip, N
        self.h = hip.nqs_hip_ar_create(N, seed, scale)
        assert self.h
        self.P = hip.nqs_hip_ar_num_params(self.h)

    def params(self):
        w = np.zeros(self.P, dtype=np.float32)
        assert self.hip.nqs_hip_ar_get_params(self.h, ptr(w)) == 0
        return w

    def logq(self, sp, constrained):
        sp = np.ascontiguousarray(sp, dtype=np.int32)
        out = np.zeros(len(sp))
        assert self.hip.nqs_hip_ar_log_prob(self.h, ptr(sp), len(sp), constrained, ptr(out)) == 0
        return out

    def sample(self, ref, B, constrained, seed):
        seeds = (np.arange(B, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15) + np.uint64(seed)) | np.uint64(1)
        ref = np.ascontiguousarray(ref, dtype=np.int32)
        out = np.zeros((B, self.N), dtype=np.int32)
        lq = np.zeros(B)
        assert self.hip.nqs_hip_ar_sample(self.h, ptr(ref), B, constrained, ptr(seeds), ptr(out), ptr(lq)) == 0
        return out, lq

    def free(self):
        self.hip.nqs_hip_ar_free(self.h)


@pytest.mark.parametrize("N", [5, 24, 40, 64])
def test_init_matches_oracle(hip, orc, N):
    net = Net(hip, N, 1234 + N, 0.7)
    w = np.zeros(orc.orc_ar_num_params(N), dtype=np.float32)
    orc.orc_ar_init(N, 1234 + N, 0.7, ptr(w))
    assert net.P == len(w)
    assert np.array_equal(net.params(), w)
    net.free()


@pytest.mark.parametrize("N", [6, 10, 17])
def test_log_prob_normalised_and_close_to_oracle(hip, orc, N):
    net = Net(hip, N, 11, 2.0)
    w = net.params()
    full = _configs(N)
    lq = net.logq(full, 0)
    assert abs(np.exp(lq).sum() - 1.0) < 1e-9
    sec = _configs(N, N // 2)
    lqs = net.logq(sec, 1)
    assert abs(np.exp(lqs).sum() - 1.0) < 1e-9
    ref = np.zeros(len(sec))
    orc.orc_ar_log_prob(N, ptr(w), ptr(sec), len(sec), 1, 1, ptr(ref))
    # BF16 activations: FP32 vs FP64 sums can move an activation by one BF16
    # ulp (2^-8 relative); bound on log q of a whole configuration
    assert np.max(np.abs(lqs - ref)) < 2e-2 * N / 10
    net.free()


def test_log_prob_large_n_close_to_oracle(hip, orc):
    N = 64
    net = Net(hip, N, 3, 1.0)
    w = net.params()
    sp = _sector(np.random.default_rng(2), 256, N, N // 2)
    for c in (0, 1):
        a = net.logq(sp, c)
        b = np.zeros(len(sp))
        orc.orc_ar_log_prob(N, ptr(w), ptr(sp), len(sp), c, 1, ptr(b))
        assert np.max(np.abs(a - b)) < 0.1
    net.free()


@pytest.mark.parametrize("constrained", [0, 1])
def test_sample_logq_is_evaluation_and_frequencies_match(hip, constrained):
    N = 8
    net = Net(hip, N, 21, 2.0)
    ref = np.array([1, -1] * (N // 2), dtype=np.int32)
    B = 200000
    smp, lq = net.sample(ref, B, constrained, 99)
    # same arithmetic: the walk's log q of its own draws is the evaluation
    assert np.array_equal(lq, net.logq(smp, constrained))
    if constrained:
        assert np.all((smp < 0).sum(axis=1) == N // 2)
    space = _configs(N, N // 2 if constrained else None)
    q = np.exp(net.logq(space, constrained))
    key = {tuple(r): i for i, r in enumerate(space)}
    counts = np.bincount([key[tuple(r)] for r in smp], minlength=len(space))
    expct = q * B
    ok = expct >= 5
    chi = stats.chisquare(counts[ok], expct[ok] * counts[ok].sum() / expct[ok].sum())
    assert chi.pvalue > 1e-4, chi
    net.free()


def test_gradient_matches_oracle(hip, orc):
    N, B = 24, 512
    net = Net(hip, N, 5, 1.0)
    w = net.params()
    sp = _sector(np.random.default_rng(4), B, N, N // 2)
    for c in (0, 1):
        g = np.zeros(net.P, dtype=np.float32)
        nll = ctypes.c_double()
        assert hip.nqs_hip_ar_gradient(net.h, ptr(sp), B, c, ptr(g), ctypes.byref(nll)) == 0
        ref = np.zeros(net.P)
        rnll = ctypes.c_double()
        orc.orc_ar_gradient(N, ptr(w), ptr(sp), B, c, 1, ptr(ref), ctypes.byref(rnll))
        assert abs(nll.value - rnll.value) < 1e-3 * abs(rnll.value)
        # BF16 staging of activations and back-propagated errors: 2^-9
        # relative per element, accumulated in FP32
        err = np.linalg.norm(g - ref) / np.linalg.norm(ref)
        assert err < 2e-2, err
        assert np.all(g[ref == 0] == 0)
    net.free()


def test_adam_first_step_and_training_lowers_nll(hip):
    N, B = 12, 256
    net = Net(hip, N, 8, 1.0)
    w0 = net.params()
    sp = _sector(np.random.default_rng(6), B, N, N // 2)
    g = np.zeros(net.P, dtype=np.float32)
    nll0 = ctypes.c_double()
    hip.nqs_hip_ar_gradient(net.h, ptr(sp), B, 1, ptr(g), ctypes.byref(nll0))
    last = ctypes.c_double()
    assert hip.nqs_hip_ar_train(net.h, ptr(sp), B, 1, 1, B, 1e-3, ctypes.byref(last)) == 0
    w1 = net.params()
    big = np.abs(g) > 1e-4
    expect = w0.astype(np.float64) - 1e-3 * g / (np.abs(g) + 1e-8)
    assert np.allclose(w1[big], expect[big], rtol=0, atol=2e-6)
    assert np.all(w1[g == 0] == w0[g == 0])
    # a few more epochs: the training loss falls
    hip.nqs_hip_ar_train(net.h, ptr(sp), B, 1, 30, 64, 3e-3, ctypes.byref(last))
    nll1 = ctypes.c_double()
    hip.nqs_hip_ar_gradient(net.h, ptr(sp), B, 1, ptr(g), ctypes.byref(nll1))
    assert nll1.value < nll0.value - 0.5
    net.free()


def _heis_chain_setup(hip, N, chains, seed, scale, learned, ar_local=0, train_every=1000000,
                      ar_epochs=2):
    cfg = abi.config_defaults(ansatz=5, hamiltonian=1, rbm_hidden_units=2 * N, j_coupling=1.0,
                              num_samples=chains, num_thermalize=16 * N, num_decorrelate=N,
                              rng_seed=seed)
    a = hip.nqs_ansatz_create(ctypes.byref(cfg), N)
    P = hip.nqs_ansatz_num_params(a)
    par = np.ctypeslib.as_array(ctypes.cast(hip.nqs_ansatz_params_raw(a), ctypes.POINTER(ctypes.c_double)), (P,))
    par[:] = np.random.default_rng(seed).normal(0.0, scale, P)
    opt = abi.options_defaults(num_chains=chains, move=abi.MOVE_LEARNED if learned else abi.MOVE_EXCHANGE,
                               sz_zero_start=1, ar_local=ar_local, ar_train_every=train_every,
                               ar_epochs=ar_epochs, ar_batch=256)
    fn = abi.fnptr(hip, "nqs_ansatz_log_amp")
    s = hip.nqs_hip_sampler_create_ex(N, ctypes.byref(cfg), ctypes.byref(opt), fn, a)
    assert s
    return cfg, a, s


def _exact_sector_probs(hip, a, N):
    space = _configs(N, N // 2)
    la = np.zeros(len(space))
    ar = np.zeros(len(space))
    assert hip.nqs_hip_log_amp_batch(a, ptr(space), len(space), ptr(la), ptr(ar)) == 0
    p = np.exp(2 * (la - la.max()))
    return space, p / p.sum()


def test_training_ahead_gives_the_same_chains(hip, monkeypatch):
    """A due retraining launched right after a batch on the side stream
    (beside the caller's work) instead of at the start of the next batch:
    same window, same network, same Adam steps, so identical samples and
    statistics batch after batch (three retrainings used, the first one as
    soon as the window holds samples).  After the last batch the ahead
    schedule has already run the next due retraining, which the other runs
    only if another batch comes: one training more, and its network."""
    N, C = 8, 512
    runs = {}
    for ov in ("0", "1"):
        monkeypatch.setenv("NQS_HIP_AR_OVERLAP", ov)
        cfg, a, s = _heis_chain_setup(hip, N, C, 91, 0.3, learned=True, train_every=3)
        hip.nqs_sampler_thermalize(s)
        outs = []
        for _ in range(10):
            out = np.zeros((C, N), dtype=np.int32)
            assert hip.nqs_sampler_batch(s, C, ptr(out)) == 0
            outs.append(out)
        acc, props, ntr, nll = ctypes.c_double(), ctypes.c_longlong(), ctypes.c_int(), ctypes.c_double()
        assert hip.nqs_hip_sampler_learned_stats(s, ctypes.byref(acc), ctypes.byref(props),
                                                 ctypes.byref(ntr), ctypes.byref(nll)) == 0
        runs[ov] = (outs, acc.value, props.value, ntr.value)
        hip.nqs_sampler_free(s)
        hip.nqs_ansatz_free(a)
    (o0, a0, p0, t0), (o1, a1, p1, t1) = runs["0"], runs["1"]
    print(f"trainings {t0}, acceptance {a0:.4f}", flush=True)
    assert t0 >= 3 and t1 == t0 + 1 and p0 == p1 and a0 == a1
    for b0, b1 in zip(o0, o1):
        np.testing.assert_array_equal(b0, b1)


def test_learned_sampler_exact_stationary_distribution(hip):
    # independent chains: the chi-square uses one state per chain, taken
    # after many learned moves, so chain autocorrelation cannot inflate it
    N, C = 8, 32768
    cfg, a, s = _heis_chain_setup(hip, N, C, 77, 0.35, learned=True)
    hip.nqs_sampler_thermalize(s)
    out = np.zeros((C, N), dtype=np.int32)
    assert hip.nqs_sampler_batch(s, C, ptr(out)) == 0          # exchange batch; window
    # trains once (train_every is huge), then a deliberately poor network:
    # the trained weights times 3, far from |psi|^2
    net = hip.nqs_hip_sampler_ar(s)
    P = hip.nqs_hip_ar_num_params(net)
    assert hip.nqs_sampler_batch(s, C, ptr(out)) == 0          # trains, learned rounds
    w = np.zeros(P, dtype=np.float32)
    hip.nqs_hip_ar_get_params(net, ptr(w))
    w2 = (w * 3.0).astype(np.float32)
    hip.nqs_hip_ar_set_params(net, ptr(w2))
    space, p = _exact_sector_probs(hip, a, N)
    key = {tuple(r): i for i, r in enumerate(space)}
    rounds = 48
    buf = np.zeros((rounds * C, N), dtype=np.int32)
    assert hip.nqs_sampler_batch(s, rounds * C, ptr(buf)) == 0
    counts = np.zeros(len(space))
    for r in buf[(rounds - 1) * C:]:
        counts[key[tuple(r)]] += 1
    acc, props, ntr = ctypes.c_double(), ctypes.c_longlong(), ctypes.c_int()
    nll = ctypes.c_double()
    assert hip.nqs_hip_sampler_learned_stats(s, ctypes.byref(acc), ctypes.byref(props),
                                             ctypes.byref(ntr), ctypes.byref(nll)) == 0
    assert ntr.value == 1 and props.value == (1 + rounds) * C
    assert 0.01 < acc.value < 0.99
    # the proposal is far from the target: without the q(s)/q(s') factor the
    # chains would follow a different distribution
    lq = np.zeros(len(space))
    assert hip.nqs_hip_ar_log_prob(net, ptr(space), len(space), 1, ptr(lq)) == 0
    assert 0.5 * np.abs(np.exp(lq) - p).sum() > 0.2
    expct = p * counts.sum()
    ok = expct >= 5
    chi = stats.chisquare(counts[ok], expct[ok] * counts[ok].sum() / expct[ok].sum())
    assert chi.pvalue > 1e-4, (chi, acc.value)
    hip.nqs_sampler_free(s)
    hip.nqs_ansatz_free(a)


def _timed_batch(hip, cfg, a, s, N, C, rounds):
    """E_loc of one recorded batch [rounds][C], its tau_int and wall time."""
    import time
    B = rounds * C
    buf = np.zeros((B, N), dtype=np.int32)
    e = np.zeros(B)
    t0 = time.perf_counter()
    assert hip.nqs_sampler_batch(s, B, ptr(buf)) == 0
    dt = time.perf_counter() - t0
    fn = abi.fnptr(hip, "nqs_ansatz_log_amp")
    hip.nqs_local_energy_batch(ctypes.byref(cfg), N, 1, ptr(buf), B, fn, a, ptr(e))
    return e, pooled_tau(e.reshape(rounds, C).T), dt


def _exact_energy(hip, cfg, a, N):
    space = np.array([[(-1 if i in idx else 1) for i in range(N)]
                      for idx in itertools.combinations(range(N), N // 2)], dtype=np.int32)
    la = np.zeros(len(space))
    ar = np.zeros(len(space))
    hip.nqs_hip_log_amp_batch(a, ptr(space), len(space), ptr(la), ptr(ar))
    p = np.exp(2 * (la - la.max()))
    p /= p.sum()
    e = np.zeros(len(space))
    hip.nqs_local_energy_batch(ctypes.byref(cfg), N, 1, ptr(space), len(space),
                               abi.fnptr(hip, "nqs_ansatz_log_amp"), a, ptr(e))
    return float(np.dot(p, e))


def test_trained_proposal_gains_over_exchange(hip):
    # A frozen psi (random RBM, 16-site Heisenberg chain, Sz = 0) whose
    # |psi|^2 has well separated modes.  Same chains, thermalisation and
    # sample count for both samplers; plain: one sweep (N exchange moves)
    # per sample, learned: one global move per sample.  Exact <E_loc> by
    # enumeration of the 12870 sector states.  Exchange chains stay in the
    # modes they thermalised into and miss the exact value by many standard
    # errors; learned moves cross between modes, reach it, and give more
    # effective samples per second.
    N, C, rounds, nb = 16, 2048, 32, 4
    cfg, a, s_plain = _heis_chain_setup(hip, N, C, 5, 0.3, learned=False)
    cfg2, a2, s_learn = _heis_chain_setup(hip, N, C, 5, 0.3, learned=True, ar_epochs=8)
    exact = _exact_energy(hip, cfg, a, N)
    for s in (s_plain, s_learn):
        hip.nqs_sampler_thermalize(s)
    _timed_batch(hip, cfg2, a2, s_learn, N, C, rounds)      # exchange; fills the window
    _timed_batch(hip, cfg2, a2, s_learn, N, C, rounds)      # trains once, learned from here
    res = {}
    for name, cf, aa, s in (("plain", cfg, a, s_plain), ("learned", cfg2, a2, s_learn)):
        means, taus, dts = [], [], []
        for _ in range(nb):
            e, tau, dt = _timed_batch(hip, cf, aa, s, N, C, rounds)
            means.append(e.mean())
            taus.append(tau)
            dts.append(dt)
        m = np.mean(means)
        err = np.std(means) / np.sqrt(nb - 1)
        ess = rounds * C / (2 * np.median(taus)) / np.median(dts)
        res[name] = (m, err, ess)
        print(f"{name}: <E> {m:.5f} +- {err:.5f} (exact {exact:.5f}) tau {np.median(taus):.3f} "
              f"batch {np.median(dts) * 1e3:.2f} ms ESS/s {ess:.3e}")
    acc = ctypes.c_double()
    hip.nqs_hip_sampler_learned_stats(s_learn, ctypes.byref(acc), None, None, None)
    (mp, ep, essp), (ml, el, essl) = res["plain"], res["learned"]
    assert acc.value > 0.3, acc.value
    assert abs(ml - exact) < 5 * el + 0.01 * abs(exact), (ml, el, exact)
    assert abs(mp - exact) > 3 * abs(ml - exact), (mp, ml, exact)
    assert essl > essp, (essl, essp)
    for s, aa in ((s_plain, a), (s_learn, a2)):
        hip.nqs_sampler_free(s)
        hip.nqs_ansatz_free(aa)
