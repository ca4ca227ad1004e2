"""Error analysis of Monte Carlo time series (host side, numpy).

jackknife      binned jackknife error of the mean
gamma_method   Wolff's Gamma method (U. Wolff, Comput. Phys. Commun. 156
               (2004) 143, hep-lat/0306017): mean, error including
               autocorrelations, integrated autocorrelation time, window.
               The same algorithm as include/u1/actime.h (the drop-in for the
               reference's actime.cpp).
"""
import math

import numpy as np


def jackknife(x, nbins=20):
    """Mean and binned-jackknife error over axis 0 (any trailing shape)."""
    x = np.asarray(x, dtype=np.float64)
    n = len(x) // nbins * nbins
    if n < nbins:
        return x.mean(axis=0), np.full(x.shape[1:], np.nan) if x.ndim > 1 else float("nan")
    b = x[:n].reshape(nbins, -1, *x.shape[1:]).mean(axis=1)
    jk = (b.sum(axis=0) - b) / (nbins - 1)
    err = np.sqrt((nbins - 1) / nbins * ((jk - jk.mean(axis=0)) ** 2).sum(axis=0))
    mean = x[:n].mean(axis=0)
    if x.ndim == 1:
        return float(mean), float(err)
    return mean, err


def gamma_method(a, S=1.5):
    """Mean, error, tau_int and summation window W of a 1-D series."""
    a = np.asarray(a, dtype=np.float64)
    N = len(a)
    d = a - a.mean()
    tmax = N // 2
    G = [float(np.dot(d, d) / N)]
    sum_rho, W, found = 0.0, 0, False
    t = 1
    while t <= tmax:
        G.append(float(np.dot(d[:N - t], d[t:]) / (N - t)))
        if not found:
            if G[0] <= 0:
                break
            sum_rho += G[t] / G[0]
            tau = 1e-300 if sum_rho <= 0 else S / math.log((sum_rho + 1) / sum_rho)
            if math.exp(-t / tau) - tau / math.sqrt(t * N) < 0:
                W, found = t, True
                tmax = min(tmax, 2 * t)
        t += 1
    if not found:
        W = len(G) - 1
    if G[0] <= 0:
        return float(a.mean()), 0.0, 0.5, W
    C = G[0] + 2 * sum(G[1:W + 1])
    G = [g + C / N for g in G]
    C = G[0] + 2 * sum(G[1:W + 1])
    return float(a.mean()), math.sqrt(C / N), C / (2 * G[0]), W


def mean_error(x, nbins=20):
    """Mean and the larger of the binned-jackknife and Gamma-method errors."""
    m, ej = jackknife(x, nbins)
    _, eg, _, _ = gamma_method(x)
    return m, max(ej, eg)
