"""FP32-bulk vs FP64 dynamics: ensemble means from independent chains.

Errors come from the spread of independent replicas (different seeds), not
from a single chain's jackknife, so long autocorrelations cannot hide a bias.
beta = 0.8 (confined) and 1.3 (Coulomb phase) are far from the U(1) phase
transition near beta = 1.011, where single chains are metastable.
"""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


def replica_means(u1, prec, beta, nrep, dims, ntherm, nmeas, seed0):
    geoms = [u1.geom(dims, beta) for _ in range(nrep)]
    seeds = [seed0 + 7919 * i for i in range(nrep)]
    e = u1.Ensemble(geoms, seeds, [1] * nrep, prec, 1, 3, u1.BENCH_METRO)
    e.sweep(ntherm, graph=True)
    acc = np.zeros(nrep)
    for _ in range(nmeas):
        e.sweep(2, graph=True)
        acc += np.array([u1.mean_plaquette(r) for r in e.plaquette()])
    e.destroy()
    return acc / nmeas


@pytest.mark.parametrize("beta", [0.8, 1.3])
def test_fp32_fp64_ensemble_means_agree(u1, beta):
    nrep, dims = 32, (12, 12, 12, 12)
    m32 = replica_means(u1, u1.F32, beta, nrep, dims, 300, 250, 1000)
    m64 = replica_means(u1, u1.F64, beta, nrep, dims, 300, 250, 5000)
    a, ea = m32.mean(), m32.std(ddof=1) / math.sqrt(nrep)
    b, eb = m64.mean(), m64.std(ddof=1) / math.sqrt(nrep)
    z = (a - b) / math.hypot(ea, eb)
    say(f"beta={beta}: <P> fp32 {a:.7f} +- {ea:.1e}, fp64 {b:.7f} +- {eb:.1e}, diff {a - b:+.2e} = {z:+.2f} sigma")
    assert abs(z) < 3.5
