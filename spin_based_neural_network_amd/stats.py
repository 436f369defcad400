"""Sampling statistics of the metric: integrated autocorrelation time and
effective sample size of per-chain local-energy series."""
import numpy as np


def pooled_tau(series):
    """tau_int from the chain-averaged autocorrelation (Sokal window, c = 5).

    series: [C][n] per-chain series (each chain's mean removed).  Returns at
    least 0.5 (independent samples)."""
    x = np.asarray(series, dtype=float)
    x = x - x.mean(axis=1, keepdims=True)
    C, n = x.shape
    var = np.mean(x * x)
    if var <= 0 or n < 4:
        return 0.5
    tau = 0.5
    for t in range(1, n):
        rho = np.mean(x[:, :-t] * x[:, t:]) * n / (n - t) / var
        tau += rho
        if t >= 5.0 * tau:
            break
    return max(tau, 0.5)


def ess(series):
    """Effective sample size sum_c n_c / (2 tau) with the pooled tau."""
    x = np.asarray(series, dtype=float)
    return x.size / (2.0 * pooled_tau(x))


def tau_ci(series, groups=8, z=1.96):
    """tau_int with a confidence interval from a delete-one-group jackknife
    over chains: the chains are split into `groups` interleaved groups, tau
    is recomputed with each group left out, and the interval is
    tau +- z * sqrt((G - 1) / G * sum (tau_(-g) - mean)^2).  Returns
    (tau, lo, hi)."""
    x = np.asarray(series, dtype=float)
    C = x.shape[0]
    tau = pooled_tau(x)
    G = min(groups, C)
    if G < 2:
        return tau, tau, tau
    idx = np.arange(C)
    loo = np.array([pooled_tau(x[idx % G != g]) for g in range(G)])
    se = np.sqrt((G - 1) / G * np.sum((loo - loo.mean()) ** 2))
    return tau, max(0.5, tau - z * se), tau + z * se
