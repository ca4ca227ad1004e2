"""GPU parity of the gfx950 path against the FP64 host oracle, through the C ABI.

Bars (written in each test):
  FP64 updates      bit-identical lattices and identical accept counts after
                    every sweep; plaquette within 1e-12 (only the summation
                    order of the FP64 reductions differs)
  FP64 measurements relative 1e-12 (Polyakov, correlators, multihit,
                    multilevel: 1e-11)
  FP32 bulk         fixed configurations: relative 1e-6 (plaquette, correlators)
  layout / fields   exact / 1e-13
"""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def make_pair(u1, orc, dims, beta=1.0, aniso=1.0, hot=True, seed=1, prec=0):
    g = u1.geom(dims, beta, aniso)
    og = orc.geom(dims, beta, aniso)
    lat = u1.Lattice.cold(g, prec)
    if hot:
        lat.hotstart(seed)
        ref = orc.hotstart(og, seed)
    else:
        ref = orc.cold(og)
    return g, og, lat, ref


# ------------------------------------------------------------ shared arithmetic
def test_device_math_bit_identical(u1, orc):
    say("device math self test")
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(-60, 60, 20000), rng.uniform(-1e-6, 1e-6, 100), [0.0, -0.0, 1e-300, np.pi]])
    y = rng.uniform(-5, 5, x.size)
    d = u1.selftest_math(0, x)
    s, c = orc.sincos(x)
    assert np.array_equal(bits(d[:, 0]), bits(s)) and np.array_equal(bits(d[:, 1]), bits(c))
    xe = x / 2
    assert np.array_equal(bits(u1.selftest_math(1, xe)), bits(orc.exp(xe)))
    # exp at the ends of its range (subnormal results, overflow, NaN, inf),
    # and the binary64 Metropolis hits' exp (fn 10), whose wavefronts with
    # every |x| < 708 skip the range tests: in shuffled (mixed) wavefronts
    # and in wavefronts wholly inside
    xr = rng.permutation(np.concatenate([xe[:6000], rng.uniform(-750, -700, 3000), rng.uniform(700, 712, 3000),
                                         [np.nan, np.inf, -np.inf, 709.782712893384, -745.1332191019412, 708.0]]))
    for fn in (1, 10):
        assert np.array_equal(bits(u1.selftest_math(fn, xr)), bits(orc.exp(xr)))
    assert np.array_equal(bits(u1.selftest_math(10, xe)), bits(orc.exp(xe)))
    assert np.array_equal(bits(u1.selftest_math(2, x, y)), bits(orc.atan2(x, y)))
    xb = np.abs(x[:2000])
    b = u1.selftest_math(5, xb)
    i0, i1 = orc.bessel_i0i1(xb)
    assert np.array_equal(bits(b[:, 0]), bits(i0)) and np.array_equal(bits(b[:, 1]), bits(i1))
    w = rng.uniform(-3 * np.pi, 3 * np.pi, 5000)
    wd = u1.selftest_math(6, w)
    assert wd.min() >= -np.pi and wd.max() < np.pi
    # FP32 sincos / atan2 accuracy (these are not bit-matched to the host)
    xf = rng.uniform(-20, 20, 20000).astype(np.float32).astype(np.float64)
    f = u1.selftest_math(3, xf)
    assert np.max(np.abs(f[:, 0] - np.sin(xf))) < 2.5e-7 and np.max(np.abs(f[:, 1] - np.cos(xf))) < 2.5e-7
    yf = rng.uniform(-5, 5, 20000).astype(np.float32).astype(np.float64)
    assert np.max(np.abs(u1.selftest_math(4, xf, yf) - np.arctan2(xf, yf))) < 5e-7


def test_fp32_update_math_accuracy(u1):
    """Hardware transcendentals used by the FP32 update kernels (not by the
    measurement kernels, which keep the polynomial sincos_f)."""
    rng = np.random.default_rng(4)
    x = rng.uniform(-3 * np.pi, 3 * np.pi, 50000).astype(np.float32).astype(np.float64)
    f = u1.selftest_math(7, x)
    es = np.max(np.abs(f[:, 0] - np.sin(x)))
    ec = np.max(np.abs(f[:, 1] - np.cos(x)))
    y = rng.uniform(-6, 6, 50000).astype(np.float32).astype(np.float64)
    ea = np.max(np.abs(u1.selftest_math(8, x, y) - np.arctan2(x, y)))
    # binary32 wrap: the angle of the binary64 wrap_d within one binary32 ulp
    # of pi, in [-pi, pi] up to a few ulps (inputs up to 3 pi + 1, as the OR angle)
    edges = np.add.outer(np.pi * np.array([-3.0, -1.0, 1.0, 3.0]), 1e-7 * np.arange(-40, 41)).ravel()
    xw = np.concatenate([rng.uniform(-3 * np.pi - 1, 3 * np.pi + 1, 50000), edges]).astype(np.float32)
    xw = xw.astype(np.float64)
    wf = u1.selftest_math(9, xw)
    ulp_pi = float(np.spacing(np.float32(np.pi)))
    dw = wf - np.arctan2(np.sin(xw), np.cos(xw))
    dw -= 2 * np.pi * np.rint(dw / (2 * np.pi))
    ew = np.max(np.abs(dw))
    say(f"fp32 update math: max |err| sin {es:.2e} cos {ec:.2e} atan2 {ea:.2e} wrap {ew:.2e}")
    assert es < 5e-6 and ec < 5e-6 and ea < 2e-6
    assert ew <= ulp_pi and np.max(np.abs(wf)) <= np.pi + 4 * ulp_pi


def test_device_philox_kat(u1):
    from tests.test_oracle_math import KAT

    ctr = np.array([c for c, _, _ in KAT], dtype=np.uint32)
    key = np.array([k for _, k, _ in KAT], dtype=np.uint32)
    out = u1.selftest_philox(ctr, key)
    for i, (_, _, exp) in enumerate(KAT):
        assert tuple(int(v) for v in out[i]) == exp


# ------------------------------------------------------------ FP64 update parity
@pytest.mark.parametrize("start", ["cold", "hot"])
@pytest.mark.parametrize("mode", ["baseline", "reference"])
def test_fp64_sweep_parity_8x4(u1, orc, start, mode):
    """8^4, beta = 1, 200 sweeps of (1 Metropolis + 3 OR); baseline proposal
    theta + U[-0.5,0.5) with 4 hits, or the reference's independent proposal."""
    dims = (8, 8, 8, 8)
    g, og, lat, ref = make_pair(u1, orc, dims, 1.0, 1.0, start == "hot", seed=1234)
    m = u1.BENCH_METRO if mode == "baseline" else u1.REFERENCE_METRO
    om = orc.metro(m.proposal, m.nhit, m.step)
    say(f"fp64 parity 8^4 start={start} mode={mode}")
    assert np.array_equal(bits(lat.to_host()), bits(ref))
    worst = 0.0
    for s in range(200):
        a_gpu = lat.update(1, 3, m, 1234, s, count_accepts=True)
        a_orc = orc.update(og, ref, 1, 3, om, 1234, s, count_accepts=True)
        assert a_gpu == a_orc, f"sweep {s}: accepts {a_gpu} != {a_orc}"
        host = lat.to_host()
        if not np.array_equal(bits(host), bits(ref)):
            bad = np.nonzero(bits(host) != bits(ref))[0]
            pytest.fail(f"sweep {s}: {bad.size} links differ, first {bad[:5]}")
        p_gpu, p_orc = lat.plaquette(), orc.plaquette(og, ref)
        worst = max(worst, float(np.max(np.abs(p_gpu - p_orc))))
        assert worst <= 1e-12, f"sweep {s}: plaquette {p_gpu} vs {p_orc}"
        if s % 50 == 49:
            say(f"  sweep {s + 1}: <P> = {u1.mean_plaquette(p_gpu):.12f}, max |dplaq| = {worst:.2e}")


@pytest.mark.parametrize("nhit", [2, 3, 5, 6])
def test_fp64_hit_counts_parity(u1, orc, nhit):
    """The kernels run the hits in blocks of four (two Philox draws): hit
    counts that end inside a block (one pair, a pair and a single hit, a
    second block of one or two hits) give the oracle's links and accept
    counts bit for bit, 8^4, hot start, 10 sweeps of 2 Metropolis + 1 OR."""
    g, og, lat, ref = make_pair(u1, orc, (8, 8, 8, 8), 1.0, 1.0, True, seed=91)
    m = u1.metro(1, nhit, 0.5)
    om = orc.metro(1, nhit, 0.5)
    say(f"fp64 parity nhit={nhit}")
    for s in range(10):
        a_gpu = lat.update(2, 1, m, 91, s, count_accepts=True)
        a_orc = orc.update(og, ref, 2, 1, om, 91, s, count_accepts=True)
        assert a_gpu == a_orc, f"nhit {nhit} sweep {s}: accepts {a_gpu} != {a_orc}"
        assert np.array_equal(bits(lat.to_host()), bits(ref)), f"nhit {nhit} sweep {s}: links differ"


@pytest.mark.parametrize("dims,beta,aniso,metrop,ovrn,nsw", [
    ((4, 4, 4, 8), 1.5, 2.0, 1, 3, 30),      # anisotropic (space-time staples weighted xi)
    ((6, 10, 4, 12), 0.9, 1.0, 2, 1, 20),    # extents that are not powers of two, 2 Metropolis passes
    ((2, 2, 2, 2), 1.0, 1.0, 1, 3, 20),      # smallest lattice: every neighbour wraps
])
def test_fp64_parity_geometries(u1, orc, dims, beta, aniso, metrop, ovrn, nsw):
    g, og, lat, ref = make_pair(u1, orc, dims, beta, aniso, True, seed=77)
    say(f"fp64 parity {dims} beta={beta} xi={aniso}")
    m = u1.BENCH_METRO
    om = orc.metro(1, 4, 0.5)
    for s in range(nsw):
        assert lat.update(metrop, ovrn, m, 77, s, count_accepts=True) == \
            orc.update(og, ref, metrop, ovrn, om, 77, s, count_accepts=True)
    assert np.array_equal(bits(lat.to_host()), bits(ref))
    assert np.max(np.abs(lat.plaquette() - orc.plaquette(og, ref))) <= 1e-12


@pytest.mark.parametrize("dims", [(24, 24, 24, 24), (48, 8, 6, 10)])
def test_fp64_parity_multichunk_planes(u1, orc, dims):
    """Planes of P = 288 and 192 even/odd sites, split into several workgroup
    chunks (a ragged tail for most tuned shapes; workgroup counts that need not
    divide by the 8 XCDs): 8 sweeps bit-identical to the oracle with identical
    accept counts, then a fused-plaquette sweep."""
    g, og, lat, ref = make_pair(u1, orc, dims, 1.0, 1.0, True, seed=31)
    say(f"fp64 parity multi-chunk {dims}")
    m = u1.BENCH_METRO
    om = orc.metro(1, 4, 0.5)
    for s in range(8):
        assert lat.update(1, 3, m, 31, s, count_accepts=True) == \
            orc.update(og, ref, 1, 3, om, 31, s, count_accepts=True), s
    assert np.array_equal(bits(lat.to_host()), bits(ref))
    assert np.max(np.abs(lat.update_plaquette(1, 3, m, 31, 8) - orc.plaquette(og, orc_step(orc, og, ref, om, 31, 8)))) <= 1e-12
    assert np.array_equal(bits(lat.to_host()), bits(ref))


def orc_step(orc, og, ref, om, seed, sweep):
    orc.update(og, ref, 1, 3, om, seed, sweep)
    return ref


@pytest.mark.parametrize("dims", [
    (2, 2, 2, 2),      # P = 2: one mostly idle workgroup per plane
    (10, 6, 4, 6),     # P = 30, ragged extents
    (16, 16, 4, 4),    # P = 128: exact chunks
    (24, 20, 4, 4),    # P = 240: tail chunk with idle threads
    (48, 48, 4, 4),    # P = 1152: two sites per thread
    (64, 32, 2, 2),    # P = 1024: four sites per thread (binary32)
])
def test_plaquette_decompositions(u1, orc, dims):
    """Plaquette over both parities with the plane-chunk decomposition: FP64
    within 1e-12 of the oracle, FP32 links within 1e-6 relative."""
    g, og, lat, ref = make_pair(u1, orc, dims, 1.0, 1.0, True, seed=5)
    say(f"plaquette decomposition {dims}")
    p64 = orc.plaquette(og, ref)
    assert np.max(np.abs(lat.plaquette() - p64)) <= 1e-12
    p32 = lat.astype(u1.F32).plaquette()
    for k in (0, 2):
        assert abs(p32[k] - p64[k]) <= 1e-6 * abs(p64[k])
    for k in (1, 3):  # imaginary parts: O(1/sqrt(V)) for a hot start
        assert abs(p32[k] - p64[k]) <= 1e-6


@pytest.mark.parametrize("dims,aniso,metrop,ovrn", [
    ((8, 8, 8, 8), 1.0, 1, 3),
    ((6, 10, 4, 12), 1.0, 2, 0),   # Metropolis-only sweep: the last Metropolis pass measures
    ((4, 4, 4, 8), 2.0, 1, 2),     # anisotropic: plaquettes are unweighted
    ((2, 2, 2, 2), 1.0, 1, 1),
])
def test_fused_plaquette_fp64(u1, orc, dims, aniso, metrop, ovrn):
    """update_plaquette: the same links as update(), and the plaquette of the
    result within 1e-13 of the oracle's."""
    g, og, lat, ref = make_pair(u1, orc, dims, 1.0, aniso, True, seed=9)
    say(f"fused plaquette fp64 {dims} xi={aniso} metrop={metrop} ovrn={ovrn}")
    m = u1.BENCH_METRO
    for s in range(3):
        p = lat.update_plaquette(metrop, ovrn, m, 9, s)
        orc.update(og, ref, metrop, ovrn, orc.metro(1, 4, 0.5), 9, s)
        assert np.array_equal(bits(lat.to_host()), bits(ref))
        assert np.max(np.abs(p - orc.plaquette(og, ref))) <= 1e-13, (p, orc.plaquette(og, ref))


def test_fused_plaquette_fp32():
    """FP32: the fused plaquette (hardware sin/cos pieces) against the
    separate pass on the same updated links, 48x48x4x4 and 16^4."""
    import latticeu1_amd as u1

    for dims in ((48, 48, 4, 4), (16, 16, 16, 16)):
        g = u1.geom(dims, 1.0)
        lat = u1.Lattice.cold(g, u1.F32)
        lat.hotstart(21)
        for s in range(3):
            p = lat.update_plaquette(1, 3, u1.BENCH_METRO, 21, s)
            q = lat.plaquette()
            for k in (0, 2):
                assert abs(p[k] - q[k]) <= 2e-6 * abs(q[k]), (dims, s, p, q)
            for k in (1, 3):
                assert abs(p[k] - q[k]) <= 2e-6, (dims, s, p, q)


@pytest.mark.parametrize("dims,nsw", [((16, 16, 16, 16), 20), ((48, 48, 48, 48), 10)])
def test_fused_plaquette_fp32_vs_fp64(u1, orc, dims, nsw):
    """The quantity the benchmark times: the FP32 sweep's in-sweep plaquette
    (plaquette.cu:93-118 summed from the update's staple pieces) against the
    FP64 plaquette of the same updated links widened to binary64, from the
    GPU FP64 pass and from the oracle: <= 1e-6 relative for SS, ST and <P>."""
    g = u1.geom(dims, 1.0)
    lat = u1.Lattice.cold(g, u1.F32)
    lat.hotstart(21)
    say(f"fused fp32 plaquette vs fp64 {dims}")
    worst = 0.0
    for s in range(nsw):
        p = lat.update_plaquette(1, 3, u1.BENCH_METRO, 21, s)
        if s < nsw - 3:
            continue
        w = lat.astype(u1.F64)
        q = w.plaquette()
        if s == nsw - 1:
            qo = orc.plaquette(orc.geom(dims, 1.0), w.to_host())
            assert np.max(np.abs(q - qo)) <= 1e-12
        for k in (0, 2):
            worst = max(worst, abs(p[k] - q[k]) / abs(q[k]))
        worst = max(worst, abs(u1.mean_plaquette(p) - u1.mean_plaquette(q)) / u1.mean_plaquette(q))
        for k in (1, 3):
            assert abs(p[k] - q[k]) <= 1e-6, (s, p, q)
    print(f"[fused fp32 plaquette] {dims}: max relative error {worst:.2e}", flush=True)
    assert worst <= 1e-6


def test_stream_reference(u1):
    """u1_selftest_stream moves 36 n bytes at a plausible HBM rate and checks its arguments."""
    n = 1 << 21
    for pattern in (0, 1, 2):
        r = u1.stream_rate(pattern, n, 5)
        say(f"stream pattern {pattern}: {r:.0f} GB/s")
        assert 500.0 < r < 9000.0
    with pytest.raises(u1.U1Error):
        u1.stream_rate(2, n + 4, 1)
    with pytest.raises(u1.U1Error):
        u1.stream_rate(3, n, 1)


def test_phase_write_set_gpu(u1, orc):
    """Each (parity, mu) launch modifies only its own half-plane."""
    dims = (8, 4, 6, 4)
    g, og, lat, ref = make_pair(u1, orc, dims, seed=5)
    V = int(np.prod(dims))
    for kind in (0, 1):
        for parity in (0, 1):
            for mu in range(4):
                before = lat.to_host()
                lat.phase(kind, parity, mu, u1.BENCH_METRO, 5, sweep=kind)
                after = lat.to_host()
                ch = np.nonzero(bits(after) != bits(before))[0]
                lo = mu * V + parity * (V // 2)
                assert ch.size > 0 and ch.min() >= lo and ch.max() < lo + V // 2


def test_masked_multilevel_update_parity(u1, orc):
    dims = (4, 4, 4, 8)
    g, og, lat, ref = make_pair(u1, orc, dims, seed=9)
    for ml, stream in ((4, 3), (2, 4)):
        for s in range(3):
            lat.update(1, 3, u1.BENCH_METRO, 9, s, stream=stream, c3=5, mask_ml=ml)
            orc.update(og, ref, 1, 3, orc.metro(1, 4, 0.5), 9, s, stream=stream, c3=5, mask_ml=ml)
    assert np.array_equal(bits(lat.to_host()), bits(ref))


# ------------------------------------------------------------ FP64 measurements
@pytest.fixture(scope="module")
def thermal(u1, orc):
    """A thermalised 8^3 x 8 FP64 configuration (GPU path, bit-equal to the oracle's)."""
    dims = (8, 8, 8, 8)
    g, og, lat, ref = make_pair(u1, orc, dims, 1.0, 1.0, True, seed=4321)
    for s in range(40):
        lat.update(1, 3, u1.BENCH_METRO, 4321, s)
    for s in range(40):
        orc.update(og, ref, 1, 3, orc.metro(1, 4, 0.5), 4321, s)
    assert np.array_equal(bits(lat.to_host()), bits(ref))
    return g, og, lat, ref


def test_polyakov_and_correlators_fp64(u1, orc, thermal):
    g, og, lat, ref = thermal
    say("fp64 measurements")
    L_gpu, L_orc = lat.polyakov(), orc.polyakov(og, ref)
    assert abs(L_gpu - L_orc) <= 1e-12 * max(1.0, abs(L_orc))
    for mode in (0, 1):
        P_gpu, P_orc = lat.poly_volume(mode), orc.poly_volume(og, ref, mode)
        assert np.allclose(P_gpu, P_orc, rtol=1e-12, atol=1e-15)
        c_gpu, c_orc = lat.poly2(mode), orc.poly2(og, ref, mode)
        assert np.allclose(c_gpu, c_orc, rtol=1e-12, atol=1e-15)
    mh = u1.metro(1, 100, 0.5)
    c_gpu = lat.poly2(2, mh, seed=11, c1=3, c3=1)
    c_orc = orc.poly2(og, ref, 2, orc.metro(1, 100, 0.5), seed=11, c1=3, c3=1)
    assert np.allclose(c_gpu, c_orc, rtol=1e-11, atol=1e-15)
    for mode in (0, 1, 2):
        t_gpu = lat.temporal_links(mode, mh, seed=3, c1=2)
        t_orc = orc.temporal_links(og, ref, mode, orc.metro(1, 100, 0.5), seed=3, c1=2)
        assert np.allclose(t_gpu, t_orc, rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("mh_mode,n4,k4,n2,k2", [
    (1, 2, 1, 3, 1),
    (2, 2, 1, 3, 1),
    (0, 3, 0, 4, 2),   # no level-4 updates, two level-2 sweeps per step (graph replays 12 times)
])
def test_multilevel_fp64(u1, orc, mh_mode, n4, k4, n2, k2):
    """The level-2 step runs as one captured graph replayed n4 * n2 times with
    device-side Philox counters; every replay must match the oracle's eager loop."""
    dims = (4, 4, 4, 8)
    g, og, lat, ref = make_pair(u1, orc, dims, 1.0, 1.0, True, seed=19)
    for s in range(5):
        lat.update(1, 3, u1.BENCH_METRO, 19, s)
        orc.update(og, ref, 1, 3, orc.metro(1, 4, 0.5), 19, s)
    say(f"multilevel fp64 mh_mode={mh_mode} n4={n4} k4={k4} n2={n2} k2={k2}")
    mh = u1.metro(1, 20, 0.5)
    r_gpu, f_gpu = lat.multilevel(n4, k4, n2, k2, 1, 2, u1.BENCH_METRO, 19, call_id=2, mh_mode=mh_mode, mh=mh,
                                  field=True)
    r_orc, f_orc = orc.multilevel(og, ref, n4, k4, n2, k2, 1, 2, orc.metro(1, 4, 0.5), 19, call_id=2,
                                  mh_mode=mh_mode, mh=orc.metro(1, 20, 0.5), field=True)
    assert np.allclose(r_gpu, r_orc, rtol=1e-11, atol=1e-16)
    assert np.allclose(f_gpu, f_orc, rtol=1e-11, atol=1e-16)
    # lattice passed in is not modified (multilevel works on a copy)
    assert np.array_equal(bits(lat.to_host()), bits(ref))


def test_multilevel_eager_matches_graph(u1, orc):
    """u1_set_graph_capture(0): the level-2 steps enqueued eagerly with the
    same device counter give the graph replays' result bit for bit."""
    dims = (4, 4, 4, 8)
    g, og, lat, ref = make_pair(u1, orc, dims, 1.0, 1.0, True, seed=23)
    for s in range(3):
        lat.update(1, 3, u1.BENCH_METRO, 23, s)
    say("multilevel eager vs graph")
    mh = u1.metro(1, 20, 0.5)
    out = {}
    try:
        for graphs in (True, False):
            u1.set_graph_capture(graphs)
            out[graphs] = lat.multilevel(2, 1, 3, 2, 1, 2, u1.BENCH_METRO, 23, call_id=4, mh_mode=2, mh=mh,
                                         field=True)
    finally:
        u1.set_graph_capture(True)
    assert np.array_equal(bits(out[True][0].view(np.float64)), bits(out[False][0].view(np.float64)))
    assert np.array_equal(bits(out[True][1].view(np.float64)), bits(out[False][1].view(np.float64)))
    r_orc = orc.multilevel(og, lat.to_host(), 2, 1, 3, 2, 1, 2, orc.metro(1, 4, 0.5), 23, call_id=4, mh_mode=2,
                           mh=orc.metro(1, 20, 0.5))
    assert np.allclose(out[False][0], r_orc, rtol=1e-11, atol=1e-16)


@pytest.mark.parametrize("n4,n2", [(3, 4), (5, 2)])
def test_multilevel_level_averages_fp64(u1, orc, n4, n2):
    """k4 = k2 = 0: the captured level-2 graph replays n4 * n2 times over one
    frozen field, so the level averages (1/n2, 1/n4, T2 cleared per level-4
    step; multilevel.cu:665-700) must give back Poly2 with the multihit links."""
    dims = (4, 4, 4, 8)
    g, og, lat, ref = make_pair(u1, orc, dims, 1.0, 1.0, True, seed=44)
    for s in range(3):
        lat.update(1, 3, u1.BENCH_METRO, 44, s)
        orc.update(og, ref, 1, 3, orc.metro(1, 4, 0.5), 44, s)
    say(f"multilevel level averages n4={n4} n2={n2}")
    r_gpu = lat.multilevel(n4, 0, n2, 0, 1, 3, u1.BENCH_METRO, 44, mh_mode=1)
    assert np.allclose(r_gpu, orc.poly2(og, ref, 1), rtol=1e-11, atol=1e-16)
    assert np.allclose(r_gpu, lat.poly2(u1.POLY_MULTIHIT), rtol=1e-11, atol=1e-16)


def test_multilevel_errors(u1):
    g = u1.geom((4, 4, 4, 6))
    lat = u1.Lattice.cold(g)
    with pytest.raises(u1.U1Error):
        lat.multilevel(1, 1, 1, 1, 1, 1, u1.BENCH_METRO, 1)
    with pytest.raises(u1.U1Error):
        u1.Lattice.cold(u1.geom((4, 4, 3, 4)))  # allocation is fine...
        bad = u1.Lattice(u1.geom((4, 4, 3, 4)), 0, u1.DeviceArray(4 * 48))
        bad.plaquette()  # ...but odd extents are rejected (parameters.cu:163-168)


def test_geometry_and_argument_limits(u1):
    """Geometry and argument checks run before any device work: odd or tiny
    extents (parameters.cu:163-168), 4V >= 2^31 (32-bit byte offsets within a
    direction plane), bad proposals, nhit and pass counts all fail cleanly."""
    import ctypes as C

    out = np.zeros(4)
    for dims in [(4, 4, 3, 4), (0, 4, 4, 4), (152, 152, 152, 152), (2, 2, 2, 1)]:
        with pytest.raises(u1.U1Error) as e:
            u1.call("u1_plaquette", C.byref(u1.geom(dims)), u1.F32, None, u1._dp(out))
        assert e.value.code == 1, dims
    g = u1.geom((150, 150, 150, 150))  # 4V = 2.0e9 < 2^31: geometry accepted, null lattice rejected
    with pytest.raises(u1.U1Error):
        u1.call("u1_plaquette", C.byref(g), u1.F32, None, u1._dp(out))
    ok = u1.geom((4, 4, 4, 4))
    lat = u1.Lattice.cold(ok, u1.F32)
    for m in [u1.metro(7, 4, 0.5), u1.metro(1, 0, 0.5), u1.metro(1, 4, 4.0), u1.metro(1, 9000, 0.5)]:
        with pytest.raises(u1.U1Error):
            lat.update(1, 3, m, 1, 0)
    with pytest.raises(u1.U1Error):
        lat.update(-1, 3, u1.BENCH_METRO, 1, 0)
    with pytest.raises(u1.U1Error):
        lat.update_plaquette(0, 0, u1.BENCH_METRO, 1, 0)
    assert np.all(lat.to_host() == 0.0)  # nothing ran
    e = u1.Ensemble([ok], [5], [0], u1.F32, 1, 3, u1.BENCH_METRO)
    for bad in (lambda: e.sweep(-1), lambda: u1.call("u1_ensemble_sweep_plaquette", e.ptr, -2, u1._dp(out))):
        with pytest.raises(u1.U1Error):
            bad()
    assert e.info()[0] == 0 and np.all(e.read(0) == 0.0)
    e.destroy()


def test_layout_and_fields_fp64(u1, orc, thermal):
    g, og, lat, ref = thermal
    no = lat.convert_layout(True)
    assert np.array_equal(bits(no.to_host()), bits(orc.convert(og, ref, True)))
    back = no.convert_layout(False)
    assert np.array_equal(bits(back.to_host()), bits(ref))
    for st in (True, False):
        F, m = lat.plaquette_fields(st, True)
        Fo, mo = orc.plaquette_fields(og, ref, st, True)
        assert np.allclose(F, Fo, atol=1e-13) and np.allclose(m, mo, atol=1e-13)
        Fn, mn = no.plaquette_fields(st, False)
        assert np.allclose(Fn, Fo, atol=1e-13) and np.allclose(mn, mo, atol=1e-13)


# ------------------------------------------------------------ FP32 bulk
def test_fp32_fixed_configuration(u1, orc, thermal):
    """FP32 angles (FP64 reductions) on a fixed FP64 configuration: relative 1e-6."""
    g, og, lat, ref = thermal
    f32 = lat.astype(u1.F32)
    say("fp32 fixed configuration")
    p32, p64 = f32.plaquette(), orc.plaquette(og, ref)
    for k in (0, 2):
        assert abs(p32[k] - p64[k]) <= 1e-6 * abs(p64[k])
    SV = 8 ** 3
    L32, L64 = f32.polyakov(), orc.polyakov(og, ref)
    assert abs(L32 - L64) <= 1e-6 * max(abs(L64), SV ** -0.5)
    for mode in (0, 1):
        c32, c64 = f32.poly2(mode), orc.poly2(og, ref, mode)
        scale = max(np.max(np.abs(c64)), 1e-300)
        assert np.max(np.abs(c32 - c64)) <= 1e-6 * scale, (mode, np.max(np.abs(c32 - c64)) / scale)


@pytest.mark.parametrize("aniso", [1.0, 1.7])
def test_fp32_overrelaxation_tracks_fp64(u1, orc, aniso):
    """From a configuration exactly representable in FP32, one OR sweep in
    FP32 stays within a few FP32 ulps of the FP64 sweep (the binary32 staple
    weights space-space terms by 1/xi and space-time terms by xi with no test
    of xi; at xi = 1 both weights are exactly 1)."""
    dims = (8, 8, 8, 8)
    g = u1.geom(dims, 1.0, aniso)
    og = orc.geom(dims, 1.0, aniso)
    start = orc.hotstart(og, 3).astype(np.float32).astype(np.float64)
    a64 = u1.Lattice.from_host(g, start, u1.F64)
    a32 = u1.Lattice.from_host(g, start, u1.F32)
    a64.update(0, 1, u1.BENCH_METRO, 3, 0)
    a32.update(0, 1, u1.BENCH_METRO, 3, 0)
    d = np.angle(np.exp(1j * (a64.to_host() - a32.to_host().astype(np.float64))))
    assert np.median(np.abs(d)) < 1e-5


@pytest.mark.parametrize("proposal,nhit,step,beta,aniso", [(0, 1, 0.0, 1.0, 1.0), (0, 2, 0.0, 0.7, 1.0),
                                                           (1, 4, 0.5, 1.0, 1.0), (1, 3, 2.5, 1.3, 1.0),
                                                           (1, 4, 0.5, 1.0, 2.0)])
def test_fp32_metropolis_tracks_fp64(u1, orc, proposal, nhit, step, beta, aniso):
    """The binary32 hit (proposal FMA, 2^(beta log2(e) dRe + 32) > raw word,
    binary64 wrap after every second pair of hits and after the last) reads
    the same Philox words as the binary64 hit, which matches the oracle bit
    for bit.  From a configuration
    exactly representable in FP32, one Metropolis pass in each precision: the
    accept decisions agree except for the rare draws within rounding of the
    threshold (and the neighbours they perturb later in the pass), and the
    angles agree to FP32 rounding."""
    dims = (8, 8, 8, 8)
    g = u1.geom(dims, beta, aniso)
    og = orc.geom(dims, beta, aniso)
    ref = orc.hotstart(og, 17)
    for s in range(5):
        orc.update(og, ref, 1, 3, orc.metro(1, 4, 0.5), 17, s)
    start = ref.astype(np.float32).astype(np.float64)
    m = u1.metro(proposal, nhit, step)
    a64 = u1.Lattice.from_host(g, start, u1.F64)
    a32 = u1.Lattice.from_host(g, start, u1.F32)
    n64 = a64.update(1, 0, m, 23, 7, count_accepts=True)
    n32 = a32.update(1, 0, m, 23, 7, count_accepts=True)
    d = np.abs(np.angle(np.exp(1j * (a64.to_host() - a32.to_host().astype(np.float64)))))
    links = d.size
    moved = np.mean(np.abs(np.angle(np.exp(1j * (a64.to_host() - start)))) > 1e-6)
    say(f"fp32 Metropolis vs fp64: proposal {proposal} nhit {nhit} step {step} beta {beta} xi {aniso}: accepts "
        f"{n32} vs {n64} of {links * nhit} hits, links moved {moved:.3f}, links off by > 1e-4: {np.sum(d > 1e-4)}, "
        f"median |d| {np.median(d):.1e}")
    assert moved > 0.2
    assert abs(n32 - n64) <= 2e-3 * links * nhit
    assert np.sum(d > 1e-4) <= 2e-3 * links
    assert np.median(d) < 1e-6


def test_fp32_chain_statistics_match_fp64(u1):
    """Ensemble means: FP32 and FP64 chains (same Philox streams) agree within errors."""
    dims = (8, 8, 8, 8)
    g = u1.geom(dims, 1.0)
    vals = {}
    for prec in (u1.F64, u1.F32):
        lat = u1.Lattice.cold(g, prec)
        lat.hotstart(99)
        v = []
        for s in range(400):
            lat.update(1, 3, u1.BENCH_METRO, 99, s)
            if s >= 100:
                v.append(u1.mean_plaquette(lat.plaquette()))
        vals[prec] = np.array(v)
    a, b = vals[u1.F64], vals[u1.F32]
    err = math.sqrt(a.var(ddof=1) / len(a) * 10 + b.var(ddof=1) / len(b) * 10)  # tau_int <~ 5
    say(f"fp64 <P> = {a.mean():.6f}, fp32 <P> = {b.mean():.6f}, combined err ~ {err:.2e}")
    assert abs(a.mean() - b.mean()) < 4 * err


# ------------------------------------------------------------ batched replicas
@pytest.mark.parametrize("graph", [True, False])
def test_ensemble_fp64_matches_oracle(u1, orc, graph):
    specs = [((4, 4, 4, 4), 0.9, 1.0, 1), ((6, 4, 4, 8), 1.05, 1.0, 0), ((8, 8, 8, 8), 1.1, 1.0, 1),
             ((4, 4, 4, 8), 1.0, 2.0, 1)]
    geoms = [u1.geom(d, b, x) for d, b, x, _ in specs]
    seeds = [101 + i for i in range(len(specs))]
    hot = [h for *_, h in specs]
    e = u1.Ensemble(geoms, seeds, hot, u1.F64, 1, 3, u1.BENCH_METRO)
    say(f"ensemble fp64 graph={graph}")
    e.sweep(6, graph=graph)
    p = e.plaquette()
    for r, (d, b, x, h) in enumerate(specs):
        og = orc.geom(d, b, x)
        ref = orc.hotstart(og, seeds[r], replica=r) if h else orc.cold(og)
        for s in range(6):
            orc.update(og, ref, 1, 3, orc.metro(1, 4, 0.5), seeds[r], s, c3=r)
        assert np.array_equal(bits(e.read(r)), bits(ref)), f"replica {r}"
        assert np.max(np.abs(p[r] - orc.plaquette(og, ref))) <= 1e-12
    e.destroy()


@pytest.mark.parametrize("prec", [0, 1])
def test_ensemble_sweep_plaquette(u1, prec):
    """Ensemble sweeps with the plaquette summed inside the sweep (one graph,
    per-replica fold, series chunked on the device) against a twin ensemble
    swept plainly and measured by the separate batched plaquette: identical
    links, plaquettes within 1e-13 (FP64) or 2e-6 relative (FP32)."""
    dims = [(4, 4, 4, 4), (6, 4, 4, 8), (8, 8, 8, 8), (4, 4, 4, 8), (12, 6, 4, 4)]
    geoms = [u1.geom(d, 0.9 + 0.05 * i, 1.5 if i == 3 else 1.0) for i, d in enumerate(dims)]
    seeds = [501 + i for i in range(len(dims))]
    hot = [1, 0, 1, 1, 1]
    a = u1.Ensemble(geoms, seeds, hot, prec, 1, 3, u1.BENCH_METRO)
    b = u1.Ensemble(geoms, seeds, hot, prec, 1, 3, u1.BENCH_METRO)
    say(f"ensemble sweep_plaquette prec={prec}")
    n = 300  # more than one device-side series chunk
    series = a.sweep_plaquette(n)
    assert series.shape == (n, len(dims), 4)
    tol_abs, tol_rel = (1e-13, 0.0) if prec == 0 else (2e-6, 2e-6)
    for s in range(n):
        b.sweep(1, graph=True)
        if s % 37 == 0 or s == n - 1:
            q = b.plaquette()
            assert np.all(np.abs(series[s] - q) <= tol_abs + tol_rel * np.abs(q)), (s, series[s], q)
    for r in range(len(dims)):
        assert np.array_equal(bits(a.read(r)), bits(b.read(r))), r
    assert a.info()[0] == b.info()[0] == n
    a.destroy()
    b.destroy()


def test_ensemble_fp32_matches_single_lattices(u1):
    """FP32 has no oracle, so the batched path is pinned to the single-lattice
    path: replica i of an FP32 ensemble (graph sweeps) equals, bit for bit, an
    FP32 lattice hot-started and updated alone with the same seed and replica
    index (Philox counter word 3), anisotropy included."""
    specs = [((8, 8, 8, 8), 0.95, 1.0, 1), ((12, 12, 12, 16), 1.05, 1.3, 1), ((6, 4, 4, 8), 1.0, 1.0, 0)]
    geoms = [u1.geom(d, b, x) for d, b, x, _ in specs]
    seeds = [11, 22, 33]
    e = u1.Ensemble(geoms, seeds, [h for *_, h in specs], u1.F32, 1, 3, u1.BENCH_METRO)
    e.sweep(20, graph=True)
    for i, (g, seed, spec) in enumerate(zip(geoms, seeds, specs)):
        lat = u1.Lattice.cold(g, u1.F32)
        if spec[3]:
            lat.hotstart(seed, replica=i)
        for s in range(20):
            lat.update(1, 3, u1.BENCH_METRO, seed, s, c3=i)
        assert np.array_equal(e.read(i), lat.to_host().astype(np.float64)), i
    p = e.plaquette()
    assert 0.2 < u1.mean_plaquette(p[0]) < 0.8 and 0.2 < u1.mean_plaquette(p[1]) < 0.8
    s, n = e.info()
    assert s == 20 and n == 4 * (8 ** 4 + 12 ** 3 * 16 + 6 * 4 * 4 * 8)
    e.destroy()


def test_tuner_records_entries(u1, tmp_path):
    p = tmp_path / "tune.csv"
    u1.tune_save(str(p))
    txt = p.read_text()
    assert "Metropolis_f64" in txt and "OverRelaxation_f64" in txt


@pytest.mark.parametrize("prec", [0, 1])
def test_checkpoint_resume_is_exact(u1, tmp_path, prec):
    """Counter-based RNG: (configuration, seed, sweep) resumes the chain exactly."""
    dims = (8, 6, 4, 8)
    g = u1.geom(dims, 1.0)
    a = u1.Lattice.cold(g, prec)
    a.hotstart(5)
    for s in range(4):
        a.update(1, 3, u1.BENCH_METRO, 5, s)
    path = str(tmp_path / "cfg.u1")
    a.save(path, 5, 4)
    for s in range(4, 8):
        a.update(1, 3, u1.BENCH_METRO, 5, s)
    b = u1.Lattice.cold(g, prec)
    seed, sweep = b.load(path)
    assert (seed, sweep) == (5, 4)
    for s in range(sweep, 8):
        b.update(1, 3, u1.BENCH_METRO, seed, s)
    assert np.array_equal(a.to_host().view(np.uint8), b.to_host().view(np.uint8))
    # the file holds FP64 angles in lexicographic order
    raw = np.fromfile(path, dtype=np.float64, offset=64)
    assert raw.size == 4 * int(np.prod(dims))
