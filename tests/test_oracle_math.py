"""Pins of the oracle's elementary functions and random numbers (CPU).

* Philox4x32-10: the Random123 known-answer vectors (Salmon et al., SC'11) and
  an independent pure-Python restatement of the published round function.
* sin/cos/exp/atan2: ulp error against glibc (math module) over wide samples
  and special values; the libm variant of the oracle uses glibc directly.
* Bessel I0/I1: against scipy.special.
* The generated constants header must equal what its generator derives.
"""
import math
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

KAT = [
    # (ctr, key, expected) -- Random123 kat_vectors, philox4x32_10
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


def philox_py(ctr, key):
    """Philox4x32-10 from the paper's definition (independent of u1_philox.h)."""
    M0, M1, W0, W1 = 0xD2511F53, 0xCD9E8D57, 0x9E3779B9, 0xBB67AE85
    x = list(ctr)
    k0, k1 = key
    mask = 0xFFFFFFFF
    for r in range(10):
        p0 = M0 * x[0]
        p1 = M1 * x[2]
        x = [((p1 >> 32) ^ x[1] ^ k0) & mask, p1 & mask, ((p0 >> 32) ^ x[3] ^ k1) & mask, p0 & mask]
        k0 = (k0 + W0) & mask
        k1 = (k1 + W1) & mask
    return tuple(x)


def test_philox_kat(orc):
    ctr = np.array([c for c, _, _ in KAT], dtype=np.uint32)
    key = np.array([k for _, k, _ in KAT], dtype=np.uint32)
    out = orc.philox(ctr, key)
    for i, (_, _, exp) in enumerate(KAT):
        assert tuple(int(v) for v in out[i]) == exp


def test_philox_matches_python_restatement(orc):
    rng = np.random.default_rng(7)
    ctr = rng.integers(0, 2**32, size=(500, 4), dtype=np.uint64).astype(np.uint32)
    key = rng.integers(0, 2**32, size=(500, 2), dtype=np.uint64).astype(np.uint32)
    out = orc.philox(ctr, key)
    for i in range(500):
        assert tuple(int(v) for v in out[i]) == philox_py(tuple(int(v) for v in ctr[i]), tuple(int(v) for v in key[i]))


def ulp_err(a, ref):
    ref = np.asarray(ref, dtype=np.float64)
    sp = np.spacing(np.abs(ref))
    sp = np.where(sp == 0, np.finfo(np.float64).tiny, sp)
    return np.abs(np.asarray(a) - ref) / sp


def test_sincos_ulp(orc):
    rng = np.random.default_rng(1)
    x = np.concatenate([
        rng.uniform(-200, 200, 60000),
        rng.uniform(-4 * math.pi, 4 * math.pi, 60000),
        rng.uniform(-1e-3, 1e-3, 5000),
        np.array([0.0, -0.0, 1e-300, math.pi / 4, math.pi / 2, math.pi, -math.pi, 3 * math.pi / 4, 1e5, -1e5]),
        np.arange(-64, 64) * (math.pi / 4),
    ])
    s, c = orc.sincos(x)
    rs = np.array([math.sin(v) for v in x])
    rc = np.array([math.cos(v) for v in x])
    # absolute error scaled by ulp(1) near zeros of sin/cos, relative elsewhere
    es = np.minimum(ulp_err(s, rs), np.abs(s - rs) / np.spacing(1.0))
    ec = np.minimum(ulp_err(c, rc), np.abs(c - rc) / np.spacing(1.0))
    assert es.max() <= 1.0 and ec.max() <= 1.0
    s0, c0 = orc.sincos(np.array([0.0, -0.0]))
    assert s0[0] == 0.0 and c0[0] == 1.0 and np.signbit(s0[1])


def test_exp_ulp(orc):
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.uniform(-60, 60, 100000), rng.uniform(-700, 700, 20000),
                        np.array([0.0, -0.0, 1.0, -1.0, 709.0, -708.0, 1e-12, -1e-12])])
    y = orc.exp(x)
    ref = np.array([math.exp(v) for v in x])
    assert ulp_err(y, ref).max() <= 1.0
    assert orc.exp(np.array([0.0]))[0] == 1.0
    big = orc.exp(np.array([800.0, -800.0]))
    assert np.isinf(big[0]) and big[1] == 0.0


def test_exp_range_ends(orc):
    """exp at the ends of its range, where the scaling by 2^k leaves the
    normal numbers: subnormal results (x in [-745.14, -708.3]) within one unit
    of the subnormal spacing of glibc's exp, results near the overflow
    threshold within one ulp, infinity where glibc overflows, and 0, inf and
    NaN at -inf, inf and NaN."""
    rng = np.random.default_rng(12)
    lo = np.concatenate([rng.uniform(-745.14, -708.3, 50000), np.linspace(-745.1332191019412, -745.13, 1001)])
    y = orc.exp(lo)
    ref = np.array([math.exp(v) for v in lo])
    assert np.max(np.abs(y - ref)) <= 2.0 ** -1074
    hi = np.concatenate([rng.uniform(700.0, 709.78, 50000), np.linspace(709.78, 709.782712893384, 1001)])
    assert ulp_err(orc.exp(hi), [math.exp(v) for v in hi]).max() <= 1.0
    over = np.array([709.7828, 709.79, 710.0, 1e300])
    assert np.all(np.isinf(orc.exp(over)))
    e = orc.exp(np.array([-np.inf, np.inf, np.nan]))
    assert e[0] == 0.0 and np.isinf(e[1]) and np.isnan(e[2])


def test_atan2_ulp_and_quadrants(orc):
    rng = np.random.default_rng(3)
    yv = np.concatenate([rng.uniform(-10, 10, 100000), rng.normal(size=20000) * 1e-6])
    xv = np.concatenate([rng.uniform(-10, 10, 100000), rng.normal(size=20000)])
    a = orc.atan2(yv, xv)
    ref = np.array([math.atan2(p, q) for p, q in zip(yv, xv)])
    assert ulp_err(a, ref).max() <= 4.0
    # signed zeros and axes (IEEE atan2 special values)
    sy = np.array([0.0, -0.0, 0.0, -0.0, 1.0, -1.0, 1.0, 1.0, -1.0])
    sx = np.array([0.0, 0.0, -0.0, -0.0, 0.0, 0.0, -1.0, 1.0, -1.0])
    out = orc.atan2(sy, sx)
    exp = np.array([math.atan2(p, q) for p, q in zip(sy, sx)])
    assert np.array_equal(np.signbit(out), np.signbit(exp))
    assert np.allclose(out, exp, rtol=0, atol=1e-15)


def test_bessel_against_scipy(orc):
    sp = pytest.importorskip("scipy.special")
    x = np.concatenate([np.linspace(0, 50, 2001), np.array([1e-8, 1.5, 6.0, 12.0, 24.0])])
    i0, i1 = orc.bessel_i0i1(x)
    assert np.allclose(i0, sp.i0(x), rtol=1e-14, atol=0)
    assert np.allclose(i1, sp.i1(x), rtol=1e-14, atol=0)
    # known values: libstdc++ std::cyl_bessel_i(0 and 1, 1.5)
    a, b = orc.bessel_i0i1(np.array([1.5]))
    assert abs(a[0] - 1.6467231897728907) < 4e-16 and abs(b[0] - 0.98166642857790731) < 4e-16


def test_constants_header_is_generated():
    import importlib.util

    spec = importlib.util.spec_from_file_location("gen", os.path.join(ROOT, "latticeu1_amd", "csrc",
                                                                      "gen_math_constants.py"))
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    with open(os.path.join(ROOT, "latticeu1_amd", "csrc", "u1_math_constants.h")) as f:
        assert f.read() == gen.render()
    c, _ = gen.constants()
    assert c["PI"] == math.pi and c["TWO_PI"] == 2 * math.pi and c["LN2_HI"] == math.log(2)
    # Cody-Waite split: P1 + P2 + P3 reproduces pi/2 far beyond binary64
    from decimal import Decimal, getcontext
    getcontext().prec = 100
    err = Decimal(c["PIO2_1"]) + Decimal(c["PIO2_2"]) + Decimal(c["PIO2_3"]) - gen.pi_dec() / 2
    assert abs(err) < Decimal(10) ** -45


def test_atan_minimax_table():
    """The binary32 minimax table of the FP32 update kernels' atan2
    (atan z = z + z^3 A(z^2) on [0, 1]): the rounded polynomial's relative
    error on a dense grid, the equioscillation level of a fresh Remez run, and
    the table being that run's result rounded to binary32 (within one unit of
    the last place, to allow for another LAPACK)."""
    import importlib.util

    spec = importlib.util.spec_from_file_location("gen", os.path.join(ROOT, "latticeu1_amd", "csrc",
                                                                      "gen_math_constants.py"))
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    table = np.array([float.fromhex(h) for h in gen.ATAN01_MM32_HEX])  # highest order first
    z = np.linspace(0.0, 1.0, 1_000_001)[1:]
    w = z * z
    rel = np.abs(z + z**3 * np.polyval(table, w) - np.arctan(z)) / np.arctan(z)
    assert rel.max() < 2e-8
    # the Taylor series truncated at the same degree is far worse on [0, 1]
    taylor = np.array([(-1.0) ** (k + 1) / (2 * k + 3) for k in range(7, -1, -1)])
    assert (np.abs(z + z**3 * np.polyval(taylor, w) - np.arctan(z)) / np.arctan(z)).max() > 1e-2
    cw, lev = gen.remez_atan01(8)
    assert 1.6e-8 < lev < 1.75e-8
    bits_fit = cw[::-1].astype(np.float32).view(np.int32).astype(np.int64)
    bits_tab = table.astype(np.float32).view(np.int32).astype(np.int64)
    assert np.max(np.abs(bits_fit - bits_tab)) <= 1


def test_fp64_minimax_tables():
    """The binary64 minimax tables of the shared sin, cos, exp and atan: the
    committed (rounded) polynomials' relative error, measured exactly at 2000
    points of each interval, stays below 0.12 units of 2^-53; the truncated
    Taylor series of the same degrees is worse by orders of magnitude; and a
    fresh Remez exchange reproduces the committed sin table to within a few
    units in the last place."""
    import importlib.util

    import mpmath as mp

    spec = importlib.util.spec_from_file_location("gen", os.path.join(ROOT, "latticeu1_amd", "csrc",
                                                                      "gen_math_constants.py"))
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    ulp = 2.0**-53
    for name, n in gen.MM64.items():
        assert len(gen.MM64_HEX[name]) == n
        assert gen.mm64_table_error(name) < 0.12 * ulp, name
    # Taylor series of sin through r^13 on |r| <= pi/4: the minimax table is
    # needed to stay under one unit
    g, W, a, b = gen._mm64_problem("SIN64")
    with mp.workdps(40):
        x = b
        taylor = sum(mp.mpf(-1) ** (k + 1) / mp.factorial(2 * k + 3) * x**k for k in range(6))
        assert abs(W(x) * (g(x) - taylor)) > 100 * ulp
    c, lev = gen.remez_mp("SIN64", 6)
    tab = np.array([float.fromhex(h) for h in reversed(gen.MM64_HEX["SIN64"])])
    assert np.all(np.abs(np.array(c) - tab) <= 8 * np.spacing(np.abs(tab)))
    assert 3.5e-18 < lev < 3.7e-18


def test_libm_variant_agrees_short_time(orc, orc_libm):
    """Same algorithms with glibc math.  One sweep must agree to rounding level
    (identical accept decisions, angles within 1e-7): this pins that the shared
    elementary functions introduce no systematic error.  Longer runs are not
    comparable -- overrelaxation amplifies a 1e-16 difference to O(1) within
    about three sweeps on a hot start, which is why the GPU path shares the
    oracle's arithmetic bit for bit."""
    g = orc.geom((4, 4, 4, 4), beta=1.0)
    m = orc.metro(1, 4, 0.5)
    a = orc.hotstart(g, 11)
    b = a.copy()
    na = orc.update(g, a, 1, 0, m, 11, 0, count_accepts=True)
    nb = orc_libm.update(g, b, 1, 0, m, 11, 0, count_accepts=True)
    assert na == nb
    orc.update(g, a, 0, 3, m, 11, 0)
    orc_libm.update(g, b, 0, 3, m, 11, 0)
    pa, pb = orc.plaquette(g, a), orc_libm.plaquette(g, b)
    assert np.max(np.abs(pa - pb)) < 1e-10
    assert np.max(np.abs(np.angle(np.exp(1j * (a - b))))) < 1e-7


def test_gamma_method_on_ar1(tmp_path):
    """Drop-in actime.h: tau_int of AR(1) series is (1 + rho) / (2 (1 - rho))."""
    import subprocess

    src = os.path.join(ROOT, "tests", "cpu_checks", "check_gamma.cpp")
    exe = tmp_path / "check_gamma"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include", "u1"), src, "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    for line in filter(None, out):
        rho, exact, tau, dtau, dval, dexact, w, ok = (float(v) for v in line.split())
        assert ok == 1
        assert abs(tau - exact) < 3 * dtau + 0.02, line
        assert abs(dval / dexact - 1) < 0.1, line


def test_python_gamma_method_on_ar1():
    """latticeu1_amd.analysis.gamma_method (used by bench.py and the tests):
    tau_int of AR(1) series is (1 + rho) / (2 (1 - rho)); the error of the
    mean matches sqrt(2 tau_int / N); jackknife agrees for uncorrelated data."""
    from latticeu1_amd.analysis import gamma_method, jackknife

    rng = np.random.default_rng(12345)
    for rho in (0.0, 0.5, 0.8, 0.9):
        n = 100000
        e = rng.standard_normal(n)
        x = np.empty(n)
        v = 0.0
        for i in range(n):
            v = rho * v + math.sqrt(1 - rho * rho) * e[i]
            x[i] = v + 3.0
        mean, err, tau, w = gamma_method(x)
        exact = (1 + rho) / (2 * (1 - rho))
        assert abs(tau - exact) < 0.12 * exact + 0.03, (rho, tau, exact)
        assert abs(err / math.sqrt(2 * exact / n) - 1) < 0.12, (rho, err)
        assert abs(mean - 3.0) < 5 * err
    y = rng.standard_normal(20000)
    m, ej = jackknife(y, 20)
    assert abs(ej / (1 / math.sqrt(len(y))) - 1) < 0.35
