"""FP32-bulk vs FP64 dynamics: ensemble means from independent chains.

Errors come from the spread of independent replicas (different seeds), not
from a single chain's jackknife, so long autocorrelations cannot hide a bias.
beta = 0.8 (confined) and 1.3 (Coulomb phase) are far from the U(1) phase
transition near beta = 1.011, where single chains are metastable; beta = 1.0
and 0.95 sit next to it, with many replicas instead of long chains.
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


def test_fp32_fp64_independent_chains_48x4(u1):
    """The 48^4 benchmark configuration itself (beta = 1, hot start, 1 Metropolis
    with 4 hits + 3 OR, 500 thermalisation + 1000 measured sweeps with the
    plaquette every sweep), four independent chains per precision (different
    seeds): the mean over chains, with Gamma-method errors per chain, agrees
    between precisions.  beta = 1 is close to the transition, so the
    chain-to-chain spread is reported beside the per-chain errors."""
    from latticeu1_amd.analysis import mean_error

    res = {}
    for prec, name in ((u1.F32, "fp32"), (u1.F64, "fp64")):
        means, errs = [], []
        for c in range(4):
            seed = 20241 + 1000 * c + (0 if prec == u1.F32 else 500000)
            lat = u1.Lattice.cold(u1.geom((48, 48, 48, 48), 1.0), prec)
            lat.hotstart(seed)
            lat.bench(1, 3, u1.BENCH_METRO, seed, 0, 500, 0)
            _, _, series = lat.bench(1, 3, u1.BENCH_METRO, seed, 500, 1000, 1)
            m, e = mean_error([u1.mean_plaquette(r) for r in series])
            means.append(m)
            errs.append(e)
            del lat
        means, errs = np.array(means), np.array(errs)
        mean = means.mean()
        err = max(math.sqrt(np.sum(errs ** 2)) / len(means), means.std(ddof=1) / math.sqrt(len(means)))
        res[name] = (mean, err)
        say(f"48^4 {name}: chain means {np.array2string(means, precision=6)} (errors "
            f"{np.array2string(errs, precision=1)}), mean {mean:.7f} +- {err:.1e}")
    d = res["fp32"][0] - res["fp64"][0]
    z = d / math.hypot(res["fp32"][1], res["fp64"][1])
    say(f"48^4 <P> FP32 - FP64 over independent chains: {d:+.2e} = {z:+.2f} combined sigma")
    assert abs(z) < 3.0


@pytest.mark.parametrize("dims,nrep,beta", [((12, 12, 12, 12), 128, 1.0), ((16, 16, 16, 16), 128, 1.0),
                                            ((24, 24, 24, 24), 64, 1.0), ((16, 16, 16, 16), 128, 0.95)])
def test_fp32_fp64_ensemble_means_near_transition(u1, dims, nrep, beta):
    """Near the transition, where <P> moves fastest with beta (a shift of the
    effective coupling would show most): independent replicas through the
    batched path, 400 thermalisation sweeps, then 1200 sweeps with the
    plaquette summed inside every sweep."""
    def means(prec, seed0):
        geoms = [u1.geom(dims, beta) for _ in range(nrep)]
        e = u1.Ensemble(geoms, [seed0 + 7919 * i for i in range(nrep)], [1] * nrep, prec, 1, 3, u1.BENCH_METRO)
        e.sweep(400, graph=True)
        s = e.sweep_plaquette(1200)
        e.destroy()
        return (1.0 - 0.5 * (s[:, :, 0] + s[:, :, 2])).mean(axis=0)  # mean_plaquette per sweep and replica

    m32, m64 = means(u1.F32, 1000), means(u1.F64, 5000)
    a, ea = m32.mean(), m32.std(ddof=1) / math.sqrt(nrep)
    b, eb = m64.mean(), m64.std(ddof=1) / math.sqrt(nrep)
    z = (a - b) / math.hypot(ea, eb)
    say(f"{dims[0]}^4 x {nrep} beta={beta}: <P> fp32 {a:.7f} +- {ea:.1e}, fp64 {b:.7f} +- {eb:.1e}, "
        f"diff {a - b:+.2e} = {z:+.2f} sigma")
    assert abs(z) < 3.5
