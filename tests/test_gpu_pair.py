"""Fused (parity, 0) + (parity, 1) update launches (k_pair, u1_set_pair_fusion).

One workgroup owns a whole (z, t) plane and runs both phases with a workgroup
barrier between them.  The bar is the unfused path's: FP64 bit-identical to
the oracle after every sweep with identical accept counts; FP32 bit-identical
to the FP32 two-launch path.  Fusion is forced on (mode 2) here, with the
tuner on and off, on planes that fit one step, need several steps per phase
(P > workgroup x sites per thread), or leave most of the workgroup idle.
"""
import numpy as np
import pytest

from tests.test_gpu_parity import bits, make_pair, say

pytestmark = pytest.mark.gpu


@pytest.fixture
def forced(u1):
    u1.set_pair_fusion(2)
    u1.set_pass_fusion(0)  # the pairs, not the fused parity passes that would replace them
    yield
    u1.set_pair_fusion(1)
    u1.set_pass_fusion(1)
    u1.tune_enable(True)


@pytest.mark.parametrize("dims,beta,aniso,metrop,ovrn,tune", [
    ((8, 8, 8, 8), 1.0, 1.0, 1, 3, False),    # P = 32: one mostly idle workgroup per plane
    ((6, 10, 4, 12), 0.9, 1.0, 2, 1, True),    # ragged extents, tuned pair shape
    ((4, 4, 4, 8), 1.5, 2.0, 1, 3, False),     # anisotropic
    ((48, 24, 4, 4), 1.0, 1.0, 1, 2, False),   # P = 576 > 512: two steps per phase in binary64
    ((2, 2, 2, 2), 1.0, 1.0, 1, 3, False),     # every neighbour wraps
])
def test_fp64_pair_parity(u1, orc, forced, dims, beta, aniso, metrop, ovrn, tune):
    u1.tune_enable(tune)
    g, og, lat, ref = make_pair(u1, orc, dims, beta, aniso, True, seed=41)
    say(f"fp64 pair parity {dims} beta={beta} xi={aniso} metrop={metrop} ovrn={ovrn} tune={tune}")
    m = u1.BENCH_METRO
    om = orc.metro(1, 4, 0.5)
    for s in range(10):
        assert lat.update(metrop, ovrn, m, 41, s, count_accepts=True) == \
            orc.update(og, ref, metrop, ovrn, om, 41, s, count_accepts=True), s
        assert np.array_equal(bits(lat.to_host()), bits(ref)), s
    # the measuring phase (1, 1) runs inside the last pass's second pair
    p = lat.update_plaquette(metrop, ovrn, m, 41, 10)
    orc.update(og, ref, metrop, ovrn, om, 41, 10)
    assert np.array_equal(bits(lat.to_host()), bits(ref))
    assert np.max(np.abs(p - orc.plaquette(og, ref))) <= 1e-13


def test_fp64_pair_metropolis_only_measurement(u1, orc, forced):
    """Metropolis-only sweeps: the last Metropolis pass measures, through its pair."""
    u1.tune_enable(False)
    g, og, lat, ref = make_pair(u1, orc, (6, 10, 4, 12), 1.0, 1.0, True, seed=8)
    for s in range(3):
        p = lat.update_plaquette(2, 0, u1.BENCH_METRO, 8, s)
        orc.update(og, ref, 2, 0, orc.metro(1, 4, 0.5), 8, s)
        assert np.array_equal(bits(lat.to_host()), bits(ref))
        assert np.max(np.abs(p - orc.plaquette(og, ref))) <= 1e-13


def test_fp64_pair_masked_and_multilevel(u1, orc, forced):
    """Multilevel masks (t % ml == 0 spatial links frozen, uniform over a
    plane) through the pair kernel, eagerly and inside the captured graph."""
    u1.tune_enable(False)
    dims = (4, 4, 4, 8)
    g, og, lat, ref = make_pair(u1, orc, dims, seed=9)
    for ml, stream in ((4, 3), (2, 4)):
        for s in range(3):
            lat.update(1, 3, u1.BENCH_METRO, 9, s, stream=stream, c3=5, mask_ml=ml)
            orc.update(og, ref, 1, 3, orc.metro(1, 4, 0.5), 9, s, stream=stream, c3=5, mask_ml=ml)
    assert np.array_equal(bits(lat.to_host()), bits(ref))
    mh = u1.metro(1, 20, 0.5)
    r_gpu = lat.multilevel(2, 1, 3, 1, 1, 2, u1.BENCH_METRO, 19, call_id=2, mh_mode=1, mh=mh)
    r_orc = orc.multilevel(og, ref, 2, 1, 3, 1, 1, 2, orc.metro(1, 4, 0.5), 19, call_id=2, mh_mode=1,
                           mh=orc.metro(1, 20, 0.5))
    assert np.allclose(r_gpu, r_orc, rtol=1e-11, atol=1e-16)


@pytest.mark.parametrize("dims,tune", [((16, 16, 16, 16), False), ((48, 48, 4, 4), False),
                                       ((48, 48, 4, 4), True), ((24, 24, 8, 8), True), ((48, 48, 48, 48), True)])
def test_fp32_pair_matches_two_launches(u1, dims, tune):
    """FP32 has no oracle: the fused path is pinned to the two-launch FP32
    path, bit for bit, over sweeps with and without the fused plaquette."""
    g = u1.geom(dims, 1.0)
    out = {}
    try:
        u1.tune_enable(tune)
        u1.set_pass_fusion(0)
        for mode in (0, 2):
            u1.set_pair_fusion(mode)
            lat = u1.Lattice.cold(g, u1.F32)
            lat.hotstart(13)
            acc = [lat.update(1, 3, u1.BENCH_METRO, 13, s, count_accepts=True) for s in range(4)]
            p = lat.update_plaquette(1, 3, u1.BENCH_METRO, 13, 4)
            out[mode] = (lat.to_host(), acc, p)
    finally:
        u1.set_pair_fusion(1)
        u1.set_pass_fusion(1)
        u1.tune_enable(True)
    say(f"fp32 pair vs two launches {dims} tune={tune}: accepts {out[2][1]}")
    assert np.array_equal(out[0][0], out[2][0])
    assert out[0][1] == out[2][1]
    assert np.max(np.abs(out[0][2] - out[2][2])) <= 1e-12


@pytest.mark.parametrize("block,spt,bz", [(384, 3, 2), (576, 2, 2), (384, 3, 1), (576, 2, 3), (320, 4, 2),
                                          (576, 1, 2), (768, 1, 4), (896, 1, 2), (1024, 1, 4), (192, 4, 2),
                                          (128, 3, 2)])
def test_fp32_pair_shapes_match_two_launches(u1, tmp_path, block, spt, bz):
    """Every fused-pair shape the tuner can choose on a 48 x 48 plane
    (P = 1152 sites per parity: one exact step of 384 x 3 or 576 x 2; one
    step with idle slots; two steps, the second full or shorter; more
    steps), in lexicographic, XCD-ordered and tiled plane orders (block.z 1
    to 4), forced through the tune cache, bit for bit against two launches in
    binary32."""
    dims = (48, 48, 4, 6)
    vol = "x".join(map(str, dims))
    cache = tmp_path / "shapes.csv"
    rows = [f"{vol}\t{name}_f32\tthreads=halfvolume\t{b}\t{s}\t{z}\t0\t1\t1\t0\t# forced"
            for name, b, s, z in (("Metropolis", 256, 1, 2), ("OverRelaxation", 256, 1, 2),
                                  ("MetropolisPair", block, spt, bz), ("OverRelaxationPair", block, spt, bz))]
    cache.write_text("tunecache\tlatticeu1_amd-1\tgpu_arch=gfx950\nvolume\tname\taux\n" + "\n".join(rows) + "\n")
    g = u1.geom(dims, 1.0)
    out = {}
    try:
        u1.tune_enable(False)
        u1.set_pass_fusion(0)
        u1.tune_load(str(cache))
        for mode in (0, 2):
            u1.set_pair_fusion(mode)
            lat = u1.Lattice.cold(g, u1.F32)
            lat.hotstart(17)
            acc = [lat.update(1, 3, u1.BENCH_METRO, 17, s, count_accepts=True) for s in range(3)]
            p = lat.update_plaquette(1, 3, u1.BENCH_METRO, 17, 3)
            out[mode] = (lat.to_host(), acc, p)
    finally:
        u1.set_pair_fusion(1)
        u1.set_pass_fusion(1)
        u1.tune_enable(True)
    say(f"fp32 pair shape {block} x {spt} (block.z {bz}) vs two launches: accepts {out[2][1]}")
    assert np.array_equal(out[0][0], out[2][0])
    assert out[0][1] == out[2][1]
    assert np.max(np.abs(out[0][2] - out[2][2])) <= 1e-12


@pytest.mark.parametrize("block,form", [(512, "two steps"), (192, "more steps"), (64, "more steps, one wave")])
def test_fp64_pair_forms_with_multilevel_mask(u1, orc, tmp_path, block, form):
    """Multilevel masks (frozen planes leave before the barrier) through the
    pair forms with steps: a 24 x 48 plane of 576 sites per parity in two
    steps of 512 threads (phase 1 keeping the angles it shares with phase 0)
    or in more steps, binary64 bit for bit against the oracle, eager sweeps
    with two mask widths and a measuring sweep without a mask."""
    dims = (24, 48, 2, 8)
    vol = "x".join(map(str, dims))
    cache = tmp_path / "forms.csv"
    rows = [f"{vol}\t{name}_f64\tthreads=halfvolume\t{b}\t1\t2\t0\t1\t1\t0\t# forced"
            for name, b in (("Metropolis", 64), ("OverRelaxation", 64), ("MetropolisPair", block),
                            ("OverRelaxationPair", block))]
    cache.write_text("tunecache\tlatticeu1_amd-1\tgpu_arch=gfx950\nvolume\tname\taux\n" + "\n".join(rows) + "\n")
    try:
        u1.tune_enable(False)
        u1.set_pass_fusion(0)
        u1.set_pair_fusion(1)
        u1.tune_load(str(cache))
        g, og, lat, ref = make_pair(u1, orc, dims, 1.0, 1.0, True, seed=31)
        say(f"fp64 pair {form} ({block} threads) with multilevel masks {dims}")
        om = orc.metro(1, 4, 0.5)
        for ml, stream in ((4, 3), (2, 4)):
            for s in range(2):
                lat.update(1, 3, u1.BENCH_METRO, 31, s, stream=stream, c3=6, mask_ml=ml)
                orc.update(og, ref, 1, 3, om, 31, s, stream=stream, c3=6, mask_ml=ml)
            assert np.array_equal(bits(lat.to_host()), bits(ref)), ml
        p = lat.update_plaquette(1, 3, u1.BENCH_METRO, 31, 2)
        orc.update(og, ref, 1, 3, om, 31, 2)
        assert np.array_equal(bits(lat.to_host()), bits(ref))
        assert np.max(np.abs(p - orc.plaquette(og, ref))) <= 1e-13
    finally:
        u1.set_pass_fusion(1)
        u1.tune_enable(True)


def test_pair_fusion_mode_checked(u1):
    with pytest.raises(u1.U1Error):
        u1.set_pair_fusion(3)
    with pytest.raises(u1.U1Error):
        u1.set_pair_fusion(-1)
    u1.set_pair_fusion(1)


def test_kernel_timing_classes(u1):
    """u1_bench_updates' kernel timing (back-to-back launches per class on a
    scratch copy): launches per sweep by class, with and without fused pairs
    (1 Metropolis + 3 OR passes, plaquette fused into the last pass), and the
    same plaquette series either way."""
    g = u1.geom((8, 8, 8, 8), 1.0)
    try:
        u1.tune_enable(False)
        got = {}
        for mode, pair, chain in (("plain", 0, 0), ("pairs", 2, 0), ("passes", 0, 2)):
            u1.set_pair_fusion(pair)
            u1.set_pass_fusion(chain)
            lat = u1.Lattice.cold(g, u1.F32)
            lat.hotstart(3)
            ms, km, series = lat.bench(1, 3, u1.BENCH_METRO, 3, 0, 12, 1, kernel_timing=True)
            got[mode] = (ms, km, series)
            assert km[19] == 32 and ms > 0 and abs(km[18] - ms / 12) <= 1e-6 * ms
            assert all(km[c] > 0 for c in range(9) if km[9 + c] > 0)
    finally:
        u1.set_pair_fusion(1)
        u1.set_pass_fusion(1)
        u1.tune_enable(True)
    # launches per sweep: metropolis, or, or+plaq, metropolis pair, or pair,
    # or pair+plaq, metropolis pass, or pass, or pass+plaq
    assert list(got["plain"][1][9:18]) == [8, 21, 3, 0, 0, 0, 0, 0, 0]
    assert list(got["pairs"][1][9:18]) == [4, 10, 2, 2, 5, 1, 0, 0, 0]
    assert list(got["passes"][1][9:18]) == [0, 0, 0, 0, 0, 0, 2, 5, 1]
    for mode in ("pairs", "passes"):
        assert np.array_equal(got["plain"][2], got[mode][2]) or \
            np.max(np.abs(got["plain"][2] - got[mode][2])) <= 1e-12


@pytest.mark.parametrize("tune,graph", [(False, True), (True, True), (False, False)])
def test_ensemble_pairs_fp64_matches_oracle(u1, orc, tune, graph):
    """Batched replicas with fused pairs forced (per-replica plane tables, one
    workgroup per plane of every replica, planes from 2 to 96 sites covered in
    one or more steps): every replica bit-identical to the oracle, and the
    measuring sweep (its last parity-1 pair runs unfused) too."""
    specs = [((4, 4, 4, 4), 0.9, 1.0, 1), ((6, 4, 4, 8), 1.05, 1.0, 0), ((8, 8, 8, 8), 1.1, 1.0, 1),
             ((4, 4, 4, 8), 1.0, 2.0, 1), ((16, 12, 4, 6), 1.0, 1.0, 1), ((2, 2, 2, 2), 1.0, 1.0, 1)]
    geoms = [u1.geom(d, b, x) for d, b, x, _ in specs]
    seeds = [301 + i for i in range(len(specs))]
    hot = [h for *_, h in specs]
    try:
        u1.tune_enable(tune)
        u1.set_pair_fusion(2)
        e = u1.Ensemble(geoms, seeds, hot, u1.F64, 1, 3, u1.BENCH_METRO)
        say(f"ensemble fp64 fused pairs tune={tune} graph={graph}")
        e.sweep(4, graph=graph)
        series = e.sweep_plaquette(2)
    finally:
        u1.set_pair_fusion(1)
        u1.tune_enable(True)
    for r, (d, b, x, h) in enumerate(specs):
        og = orc.geom(d, b, x)
        ref = orc.hotstart(og, seeds[r], replica=r) if h else orc.cold(og)
        for s in range(4):
            orc.update(og, ref, 1, 3, orc.metro(1, 4, 0.5), seeds[r], s, c3=r)
        for s in range(4, 6):
            orc.update(og, ref, 1, 3, orc.metro(1, 4, 0.5), seeds[r], s, c3=r)
            assert np.max(np.abs(series[s - 4, r] - orc.plaquette(og, ref))) <= 1e-13, (r, s)
        assert np.array_equal(bits(e.read(r)), bits(ref)), f"replica {r}"
    e.destroy()


def test_ensemble_pairs_fp32_match_unfused(u1):
    """FP32 batched replicas: fused pairs forced against no fusion, bit for bit."""
    dims = [(8, 8, 8, 8), (12, 12, 12, 16), (6, 4, 4, 8), (24, 24, 8, 8)]
    geoms = [u1.geom(d, 1.0) for d in dims]
    seeds = [71, 72, 73, 74]
    out = {}
    try:
        u1.tune_enable(False)
        for mode in (0, 2):
            u1.set_pair_fusion(mode)
            e = u1.Ensemble(geoms, seeds, [1, 1, 1, 1], u1.F32, 1, 3, u1.BENCH_METRO)
            e.sweep(5, graph=True)
            out[mode] = [e.read(r) for r in range(len(dims))]
            e.destroy()
    finally:
        u1.set_pair_fusion(1)
        u1.tune_enable(True)
    for r in range(len(dims)):
        assert np.array_equal(out[0][r], out[2][r]), r


def test_fp64_parity_at_48x4(u1, orc):
    """The production-size path in binary64 at 48^4: tuned launch shapes
    (XCD order, plane tiles), then fused pairs forced, two sweeps each, bit
    for bit against the oracle, with the fused plaquette of the last sweep."""
    dims = (48, 48, 48, 48)
    g, og, lat, ref = make_pair(u1, orc, dims, 1.0, 1.0, True, seed=4848)
    m, om = u1.BENCH_METRO, orc.metro(1, 4, 0.5)
    try:
        for mode, sweeps in ((1, (0, 1)), (2, (2, 3))):
            u1.set_pair_fusion(mode)
            for s in sweeps:
                say(f"fp64 48^4 parity, pair fusion mode {mode}, sweep {s}")
                assert lat.update(1, 3, m, 4848, s, count_accepts=True) == \
                    orc.update(og, ref, 1, 3, om, 4848, s, count_accepts=True), s
            assert np.array_equal(bits(lat.to_host()), bits(ref)), mode
        p = lat.update_plaquette(1, 3, m, 4848, 4)
        orc.update(og, ref, 1, 3, om, 4848, 4)
    finally:
        u1.set_pair_fusion(1)
    assert np.array_equal(bits(lat.to_host()), bits(ref))
    assert np.max(np.abs(p - orc.plaquette(og, ref))) <= 1e-12
