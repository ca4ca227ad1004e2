"""Pins of the FP64 oracle's lattice algorithms (CPU).

The reference ships no tests, fixtures or golden data, so each function of
the oracle is pinned by something the mathematics fixes:
  geometry      brute-force bijection / neighbour checks from coordinates
  staple        read off the total Wilson action (independent numpy statement)
  OR            exact action conservation and involution
  Metropolis    closed-form first-phase acceptance (e^{-6b} I0(6b)) and 1-D
                quadrature for the theta+delta proposal; von Mises single-link
                law through the stochastic -> analytic multihit limit
  measurements  cold-start closed forms, numpy restatements, gauge invariance
  multilevel    exact reduction to Poly2(multihit) for k4 = k2 = 0, n4 = n2 = 1
  ensembles     strong- and weak-coupling limits of <cos P>
"""
import math

import numpy as np
import pytest

from tests import lattice_np as L


def rand_lat(dims, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(-np.pi, np.pi, 4 * int(np.prod(dims)))


# ---------------------------------------------------------------- geometry
@pytest.mark.parametrize("dims", [(2, 2, 2, 2), (4, 4, 4, 4), (4, 6, 8, 2), (6, 2, 4, 10)])
def test_eo_coords_bijection(orc, dims):
    g = orc.geom(dims)
    V = int(np.prod(dims))
    seen = set()
    for parity in (0, 1):
        for i in range(V // 2):
            x = orc.eo_coords(g, i, parity)
            assert all(0 <= x[k] < dims[k] for k in range(4))
            lex = x[0] + dims[0] * (x[1] + dims[1] * (x[2] + dims[2] * x[3]))
            assert lex >> 1 == i and (sum(x) & 1) == parity
            seen.add(tuple(x))
    assert len(seen) == V


@pytest.mark.parametrize("dims", [(4, 4, 4, 4), (6, 4, 2, 8)])
def test_eo_neighbors_bruteforce(orc, dims):
    g = orc.geom(dims)
    V = int(np.prod(dims))
    rng = np.random.default_rng(0)
    for _ in range(300):
        parity = int(rng.integers(2))
        i = int(rng.integers(V // 2))
        mu, nu = [int(v) for v in rng.choice(4, 2, replace=False)]
        lmu, lnu = int(rng.choice([-1, 1])), int(rng.choice([-1, 0, 1]))
        x = orc.eo_coords(g, i, parity)
        y = list(x)
        y[mu] = (y[mu] + lmu) % dims[mu]
        y[nu] = (y[nu] + lnu) % dims[nu]
        lex = y[0] + dims[0] * (y[1] + dims[1] * (y[2] + dims[2] * y[3]))
        assert orc.eo_neighbor(g, i, parity, mu, lmu, nu, lnu) == (lex >> 1) + (sum(y) & 1) * (V // 2)


# ---------------------------------------------------------------- staple
@pytest.mark.parametrize("xi", [1.0, 1.3])
def test_staple_from_action(orc, xi):
    dims = (4, 4, 6, 4)
    g = orc.geom(dims, beta=1.0, aniso=xi)
    lat = rand_lat(dims, 5)
    th = L.to_lex(lat, dims)
    rng = np.random.default_rng(1)
    for _ in range(12):
        mu = int(rng.integers(4))
        site = tuple(int(rng.integers(n)) for n in (dims[3], dims[2], dims[1], dims[0]))  # (t, z, y, x)
        t, z, y, x = site
        lex = x + dims[0] * (y + dims[1] * (z + dims[2] * t))
        parity = (x + y + z + t) & 1
        S = orc.staple(g, lat, lex >> 1, parity, mu)
        ref = L.staple_from_action(th, xi, site, mu)
        assert abs(S - ref) < 1e-11 * max(1.0, abs(ref))


def test_gauge_invariance(orc):
    dims = (4, 4, 4, 6)
    g = orc.geom(dims, aniso=1.2)
    lat = rand_lat(dims, 9)
    th = L.to_lex(lat, dims)
    alpha = np.random.default_rng(3).uniform(-np.pi, np.pi, th[0].shape)
    th2 = [th[mu] + alpha - L.shift(alpha, mu) for mu in range(4)]
    lat2 = L.from_lex(th2, dims)
    assert np.allclose(orc.plaquette(g, lat), orc.plaquette(g, lat2), atol=1e-13)
    V = int(np.prod(dims))
    for parity in (0, 1):
        for i in (0, 7, V // 2 - 1):
            for mu in range(4):
                a = abs(orc.staple(g, lat, i, parity, mu))
                b = abs(orc.staple(g, lat2, i, parity, mu))
                assert abs(a - b) < 1e-12


# ---------------------------------------------------------------- updates
def weighted_action(orc, g, lat):
    p = orc.plaquette(g, lat)
    V = int(np.prod(list(g.dims)))
    xi = g.aniso
    return (3 * V * (1 - p[0])) / xi + 3 * V * (1 - p[2]) * xi


@pytest.mark.parametrize("xi", [1.0, 1.7])
def test_overrelaxation_preserves_action(orc, xi):
    dims = (4, 6, 4, 4)
    g = orc.geom(dims, beta=1.0, aniso=xi)
    lat = orc.hotstart(g, 21)
    a0 = weighted_action(orc, g, lat)
    for s in range(3):
        orc.update(g, lat, 0, 3, orc.metro(), 21, s)
    a1 = weighted_action(orc, g, lat)
    assert abs(a1 - a0) <= 1e-12 * abs(a0) + 1e-10
    assert lat.min() >= -np.pi and lat.max() < np.pi


def test_overrelaxation_involution(orc):
    dims = (4, 4, 4, 4)
    g = orc.geom(dims)
    lat = orc.hotstart(g, 3)
    ref = lat.copy()
    orc.overrelax_phase(g, lat, 1, 2)
    assert not np.array_equal(lat, ref)
    orc.overrelax_phase(g, lat, 1, 2)
    assert np.max(np.abs(np.angle(np.exp(1j * (lat - ref))))) < 1e-12


def test_phase_write_set(orc):
    """A (parity, mu) phase writes only its own half-plane."""
    dims = (4, 4, 6, 4)
    g = orc.geom(dims)
    V = int(np.prod(dims))
    lat = orc.hotstart(g, 8)
    ref = lat.copy()
    orc.metropolis_phase(g, lat, 1, 2, orc.metro(1, 4, 0.5), 8)
    changed = np.nonzero(lat != ref)[0]
    lo, hi = 2 * V + V // 2, 3 * V
    assert changed.size > 0 and changed.min() >= lo and changed.max() < hi


@pytest.mark.parametrize("beta", [0.5, 1.0])
def test_first_phase_acceptance_independent(orc, beta):
    """Cold start, xi = 1: S = 6 for every link, so an independent uniform
    proposal is accepted with probability E[exp(6 b (cos t - 1))] = e^{-6b} I0(6b)."""
    sp = pytest.importorskip("scipy.special")
    dims = (16, 16, 16, 16)
    g = orc.geom(dims, beta=beta)
    lat = orc.cold(g)
    n = int(np.prod(dims)) // 2
    acc = orc.metropolis_phase(g, lat, 0, 0, orc.metro(0, 1, 0.0), 1234)
    p = math.exp(-6 * beta) * sp.i0(6 * beta)
    sigma = math.sqrt(p * (1 - p) / n)
    assert abs(acc / n - p) < 4.5 * sigma


def test_first_phase_acceptance_delta(orc):
    """theta + delta, delta ~ U[-0.5, 0.5): acceptance = E_delta[min(1, exp(6b(cos delta - 1)))]."""
    integrate = pytest.importorskip("scipy.integrate")
    beta = 1.0
    dims = (16, 16, 16, 16)
    g = orc.geom(dims, beta=beta)
    lat = orc.cold(g)
    n = int(np.prod(dims)) // 2
    acc = orc.metropolis_phase(g, lat, 0, 1, orc.metro(1, 1, 0.5), 99)
    p = integrate.quad(lambda d: math.exp(6 * beta * (math.cos(d) - 1)), -0.5, 0.5)[0] / 1.0
    sigma = math.sqrt(p * (1 - p) / n)
    assert abs(acc / n - p) < 4.5 * sigma


def test_stochastic_multihit_converges_to_analytic(orc):
    """Metropolis with a fixed staple samples the von Mises law, whose mean of
    e^{i theta} is conj(S)/|S| I1(b|S|)/I0(b|S|) -- the analytic multihit."""
    dims = (4, 4, 4, 4)
    g = orc.geom(dims, beta=1.0)
    lat = orc.hotstart(g, 4)
    for s in range(5):
        orc.update(g, lat, 1, 3, orc.metro(1, 4, 0.5), 4, s)
    exact = orc.temporal_links(g, lat, 1)
    mc = orc.temporal_links(g, lat, 2, orc.metro(0, 4000, 0.0), seed=77)
    # independent proposals: per-hit variance <= 1, acceptance >~ 0.2
    assert np.max(np.abs(mc - exact)) < 0.08
    assert np.mean(np.abs(mc - exact)) < 0.02


# ---------------------------------------------------------------- measurements
def test_cold_start_closed_forms(orc):
    sp = pytest.importorskip("scipy.special")
    dims = (6, 6, 6, 8)
    beta = 0.8
    g = orc.geom(dims, beta=beta)
    lat = orc.cold(g)
    assert np.array_equal(orc.plaquette(g, lat), np.zeros(4))
    assert orc.polyakov(g, lat) == 1.0
    assert np.allclose(orc.poly2(g, lat, 0), 1.0, atol=0)
    u = sp.i1(6 * beta) / sp.i0(6 * beta)
    assert np.allclose(orc.poly2(g, lat, 1), u ** (2 * dims[3]), rtol=1e-13)
    pc = orc.plaquette_host(g, lat)
    assert pc[0] == 1.0 and pc[1] == 0.0


def test_measurements_match_numpy(orc):
    dims = (4, 6, 4, 8)
    g = orc.geom(dims)
    lat = rand_lat(dims, 12)
    th = L.to_lex(lat, dims)
    pa = L.plaquette_angles(th)
    ss = sum(np.sum(1 - np.exp(1j * pa[k])) for k in [(0, 1), (0, 2), (1, 2)])
    st = sum(np.sum(1 - np.exp(1j * pa[k])) for k in [(0, 3), (1, 3), (2, 3)])
    V = int(np.prod(dims))
    p = orc.plaquette(g, lat)
    assert np.allclose(p, [ss.real / (3 * V), ss.imag / (3 * V), st.real / (3 * V), st.imag / (3 * V)], atol=1e-14)
    hp = orc.plaquette_host(g, lat)
    assert abs(hp[0] - (1 - (p[0] + p[2]) / 2)) < 1e-14
    P = L.polyakov_field(th)
    assert abs(orc.polyakov(g, lat) - P.mean()) < 1e-14
    Pv = orc.poly_volume(g, lat, 0)
    assert np.allclose(Pv, P.ravel(), atol=1e-13)
    assert np.allclose(orc.pp(g, Pv, dims[0] // 2), L.correlator(P, dims[0] // 2), atol=1e-13)


def test_hot_start_statistics(orc):
    dims = (8, 8, 8, 8)
    g = orc.geom(dims)
    lat = orc.hotstart(g, 2024)
    assert lat.min() >= -np.pi and lat.max() < np.pi
    assert abs(lat.mean()) < 4 * np.pi / math.sqrt(3 * lat.size)
    p = orc.plaquette_host(g, lat)
    assert abs(p[0]) < 5 / math.sqrt(2 * 6 * lat.size / 4)
    assert not np.array_equal(lat, orc.hotstart(g, 2025))


def test_metropolis_weight_identity():
    """|S| cos(theta + atan2(Im S, Re S)) == Re(S e^{i theta}): the reference's
    weight (lattice_functions.h:68-80) and the form evaluated here."""
    rng = np.random.default_rng(0)
    S = rng.normal(size=1000) * 4 + 1j * rng.normal(size=1000) * 4
    th = rng.uniform(-np.pi, np.pi, 1000)
    a = np.abs(S) * np.cos(th + np.arctan2(S.imag, S.real))
    b = S.real * np.cos(th) - S.imag * np.sin(th)
    assert np.allclose(a, b, atol=1e-12)


# ---------------------------------------------------------------- multilevel
def test_multilevel_reduces_to_poly2_multihit(orc):
    dims = (4, 4, 4, 8)
    g = orc.geom(dims, beta=1.0)
    lat = orc.hotstart(g, 31)
    for s in range(3):
        orc.update(g, lat, 1, 3, orc.metro(1, 4, 0.5), 31, s)
    ml = orc.multilevel(g, lat, 1, 0, 1, 0, 1, 3, orc.metro(1, 4, 0.5), 31)
    p2 = orc.poly2(g, lat, 1)
    assert np.allclose(ml, p2, rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("n4,n2", [(3, 4), (5, 2)])
def test_multilevel_level_averages_of_frozen_field(orc, n4, n2):
    """With no sublattice updates (k4 = k2 = 0) every level-2 measurement sees
    the same field, so the 1/n2 and 1/n4 averages (multilevel.cu:665, 668) and
    the T2 clear before each level-4 step (multilevel.cu:674) must give back
    Poly2 with the multihit links.  A missing norm or reset scales C(R) by a
    power of n2 or n4."""
    dims = (4, 4, 4, 8)
    g = orc.geom(dims, beta=1.0)
    lat = orc.hotstart(g, 44)
    for s in range(3):
        orc.update(g, lat, 1, 3, orc.metro(1, 4, 0.5), 44, s)
    ml = orc.multilevel(g, lat, n4, 0, n2, 0, 1, 3, orc.metro(1, 4, 0.5), 44, mh_mode=1)
    p2 = orc.poly2(g, lat, 1)
    assert np.allclose(ml, p2, rtol=1e-12, atol=1e-16)


def _multilevel_numpy(orc, g, lat, n4, k4, n2, k2, metrop, ovrn, m, seed, call_id, mh_mode, mh):
    """The two-level algorithm (multilevel.cu:639-700, T2 kernel :331-360, T2 -> T4
    :439-456, final :532-553) recomposed in numpy from the oracle's pinned
    primitives: masked updates (Philox streams 3 and 4, sweep index = step
    counter, c3 = call_id) and the temporal-link field (stream 5, c1 = level-2
    step).  Axis order of the lexicographic field: (t, z, y, x)."""
    Nx, Ny, Nz, Nt = g.dims[0], g.dims[1], g.dims[2], g.dims[3]
    R = Nx // 2
    lat = lat.copy()
    T4 = np.zeros((Nt // 4, 3, R, Nz, Ny, Nx), complex)
    for i in range(n4):
        for j in range(k4):
            orc.update(g, lat, metrop, ovrn, m, seed, i * k4 + j, stream=3, c3=call_id, mask_ml=4)
        T2 = np.zeros((Nt // 2, 3, R, Nz, Ny, Nx), complex)
        for k in range(n2):
            for l in range(k2):
                orc.update(g, lat, metrop, ovrn, m, seed, (i * n2 + k) * k2 + l, stream=4, c3=call_id, mask_ml=2)
            tl = orc.temporal_links(g, lat, mh_mode, mh, seed, 5, i * n2 + k, call_id).reshape(Nt, Nz, Ny, Nx)
            two = tl[0::2] * tl[1::2]  # line of two temporal links from t = 2l
            for d, ax in enumerate((3, 2, 1)):  # x, y, z in the (l, z, y, x) array
                for r in range(1, R + 1):
                    T2[:, d, r - 1] += two * np.conj(np.roll(two, -r, axis=ax))
        T2 /= n2
        T4 += T2[0::2] * T2[1::2]
    T4 /= n4
    return np.prod(T4, axis=0).mean(axis=(0, 2, 3, 4))


@pytest.mark.parametrize("mh_mode", [1, 2])
def test_multilevel_matches_numpy_recomposition(orc, mh_mode):
    """Level averages over fields that do change (k4, k2 > 0, n4 = 2, n2 = 3),
    against the numpy recomposition above: pins the T2/T4 products, the level
    norms and the per-level reset where the measured values differ."""
    dims = (4, 4, 4, 8)
    g = orc.geom(dims, beta=1.0)
    lat = orc.hotstart(g, 71)
    m = orc.metro(1, 4, 0.5)
    mh = orc.metro(1, 10, 0.5)
    for s in range(3):
        orc.update(g, lat, 1, 3, m, 71, s)
    ml = orc.multilevel(g, lat, 2, 1, 3, 1, 1, 2, m, 71, call_id=5, mh_mode=mh_mode, mh=mh)
    ref = _multilevel_numpy(orc, g, lat, 2, 1, 3, 1, 1, 2, m, 71, 5, mh_mode, mh)
    assert np.allclose(ml, ref, rtol=1e-11, atol=1e-15)
    # the level averaging matters here: a single level-2 measurement differs
    one = _multilevel_numpy(orc, g, lat, 1, 1, 1, 1, 1, 2, m, 71, 5, mh_mode, mh)
    assert not np.allclose(ml, one, rtol=1e-3)


def test_multilevel_masks_freeze_boundaries(orc):
    dims = (4, 4, 4, 8)
    g = orc.geom(dims)
    V = int(np.prod(dims))
    lat = orc.hotstart(g, 5)
    ref = lat.copy()
    orc.update(g, lat, 1, 3, orc.metro(1, 4, 0.5), 5, 0, stream=3, mask_ml=4)
    th0, th1 = L.to_lex(ref, dims), L.to_lex(lat, dims)
    t = np.arange(dims[3])
    for mu in range(4):
        same = np.all(th0[mu] == th1[mu], axis=(1, 2, 3))
        if mu == 3:
            assert not same.any()
        else:
            assert np.array_equal(same, t % 4 == 0)


def test_multilevel_small_run_is_finite_and_real_dominated(orc):
    dims = (4, 4, 4, 8)
    g = orc.geom(dims, beta=1.0)
    lat = orc.hotstart(g, 8)
    res, field = orc.multilevel(g, lat, 2, 1, 2, 1, 1, 1, orc.metro(1, 4, 0.5), 8, field=True)
    assert np.all(np.isfinite(res)) and np.all(np.isfinite(field))
    SV = dims[0] * dims[1] * dims[2]
    r = dims[0] // 2
    # the field summed over sites and directions reproduces the correlator
    f = field.reshape(3, r, SV)
    assert np.allclose(f.sum(axis=(0, 2)) / (3 * SV), res, rtol=1e-12, atol=1e-15)


# ---------------------------------------------------------------- layout / fields
def test_convert_roundtrip_and_permutation(orc):
    dims = (4, 6, 4, 2)
    g = orc.geom(dims)
    lat = rand_lat(dims, 1)
    no = orc.convert(g, lat, True)
    th = L.to_lex(lat, dims)
    V = int(np.prod(dims))
    for mu in range(4):
        assert np.array_equal(no[mu * V:(mu + 1) * V], th[mu].ravel())
    assert np.array_equal(orc.convert(g, no, False), lat)


def test_plaquette_fields_consistent(orc):
    dims = (4, 4, 6, 4)
    g = orc.geom(dims)
    lat = rand_lat(dims, 2)
    p = orc.plaquette(g, lat)
    F, means = orc.plaquette_fields(g, lat, spacetime=True, eo_order=True)
    # SixPlaquette order: (0,3) (1,3) (2,3) (1,2) (2,0) (0,1)
    assert abs(means[:3].mean() - complex(p[2], p[3])) < 1e-14
    # (2,0) is the (0,2) plaquette with reversed orientation: 1 - e^{-iP}
    ss = np.array([means[3], np.conj(means[4]), means[5]])
    assert abs(ss.mean() - complex(p[0], p[1])) < 1e-14
    assert np.allclose(F.mean(axis=1), means, atol=1e-14)
    th = L.to_lex(lat, dims)
    pa = L.plaquette_angles(th)
    assert np.allclose(F[0], (1 - np.exp(1j * pa[(0, 3)])).ravel(), atol=1e-14)
    assert np.allclose(F[4], (1 - np.exp(1j * (th[2] + L.shift(th[0], 2) - L.shift(th[2], 0) - th[0]))).ravel(),
                       atol=1e-14)
    Fs, ms = orc.plaquette_fields(g, lat, spacetime=False, eo_order=True)
    assert np.allclose(Fs, F.reshape(6, dims[3], -1).mean(axis=1), atol=1e-14)
    no = orc.convert(g, lat, True)
    Fn, mn = orc.plaquette_fields(g, no, spacetime=True, eo_order=False)
    assert np.allclose(Fn, F, atol=1e-14) and np.allclose(mn, means, atol=1e-14)
    Fsn, _ = orc.plaquette_fields(g, no, spacetime=False, eo_order=False)
    assert np.allclose(Fsn, Fs, atol=1e-14)


# ---------------------------------------------------------------- ensemble limits
def run_chain(orc, dims, beta, ntherm, nmeas, seed, hot=True):
    g = orc.geom(dims, beta=beta)
    lat = orc.hotstart(g, seed) if hot else orc.cold(g)
    m = orc.metro(1, 4, 0.5)
    vals = []
    for s in range(ntherm + nmeas):
        orc.update(g, lat, 1, 3, m, seed, s)
        if s >= ntherm:
            vals.append(orc.plaquette_host(g, lat)[0])
    return np.array(vals)


def test_strong_coupling_limit(orc):
    """beta -> 0: <cos P> = I1(b)/I0(b) + O(b^5)."""
    sp = pytest.importorskip("scipy.special")
    beta = 0.2
    v = run_chain(orc, (8, 8, 8, 8), beta, 10, 40, 42)
    u = sp.i1(beta) / sp.i0(beta)
    err = v.std(ddof=1) / math.sqrt(len(v))
    assert abs(v.mean() - u) < 5 * err + 2e-4


def test_weak_coupling_limit(orc):
    """beta >> 1 (Coulomb phase): 1 - <cos P> = 1/(4 beta) + O(1/beta^2)."""
    beta = 4.0
    v = run_chain(orc, (6, 6, 6, 6), beta, 30, 30, 7, hot=False)
    assert abs((1 - v.mean()) / (1 / (4 * beta)) - 1) < 0.10
