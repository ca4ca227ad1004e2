"""Fused parity passes (k_chain, u1_set_pass_fusion).

The four phases (parity, 0..3) of a pass run in one launch; a workgroup of
phase mu waits on device-side completion counters of the phase-(mu - 1)
planes it depends on (latticeu1_amd/csrc/u1_chain.h; the schedule and the
dependency set are checked on the CPU by tests/cpu_checks/check_chain.cpp).
The bar is the unfused path's: FP64 bit-identical to the oracle after every
sweep with identical accept counts, the in-sweep plaquette within 1e-13 of the
oracle's; FP32 bit-identical to the unfused FP32 path.  No wait may time out
(u1_sync raises if one did).
"""
import numpy as np
import pytest

from tests.test_gpu_parity import bits, make_pair, say

pytestmark = pytest.mark.gpu


@pytest.fixture
def forced(u1):
    u1.set_pass_fusion(2)
    yield
    u1.set_pass_fusion(1)
    u1.tune_enable(True)


@pytest.mark.parametrize("dims,beta,aniso,metrop,ovrn,proposal,tune", [
    ((8, 8, 8, 8), 1.0, 1.0, 1, 3, 1, False),     # lag 8 over 8 planes per tile: nearly serial
    ((4, 4, 2, 4), 0.9, 1.0, 1, 3, 0, False),     # Nz = 2: both z neighbours are one plane
    ((6, 10, 4, 12), 1.1, 1.0, 2, 1, 1, True),    # ragged extents, tuned shape
    ((4, 4, 4, 8), 1.5, 2.0, 1, 3, 1, False),     # anisotropic
    ((16, 12, 6, 8), 1.0, 1.0, 1, 2, 0, False),   # several chunks per plane
])
def test_fp64_pass_parity(u1, orc, forced, dims, beta, aniso, metrop, ovrn, proposal, tune):
    u1.tune_enable(tune)
    g, og, lat, ref = make_pair(u1, orc, dims, beta, aniso, True, seed=53)
    say(f"fp64 fused-pass parity {dims} beta={beta} xi={aniso} metrop={metrop} ovrn={ovrn} "
        f"proposal={proposal} tune={tune}")
    m = u1.metro(proposal, 4, 0.5)
    om = orc.metro(proposal, 4, 0.5)
    for s in range(6):
        assert lat.update(metrop, ovrn, m, 53, s, count_accepts=True) == \
            orc.update(og, ref, metrop, ovrn, om, 53, s, count_accepts=True), s
        assert np.array_equal(bits(lat.to_host()), bits(ref)), s
    # the measuring phases (1, 1..3) run inside the last pass's parity-1 launch
    p = lat.update_plaquette(metrop, ovrn, m, 53, 6)
    orc.update(og, ref, metrop, ovrn, om, 53, 6)
    assert np.array_equal(bits(lat.to_host()), bits(ref))
    assert np.max(np.abs(p - orc.plaquette(og, ref))) <= 1e-13
    u1.sync()


def test_fp64_pass_metropolis_only_measurement(u1, orc, forced):
    """Metropolis-only sweeps: the last Metropolis pass measures, fused."""
    u1.tune_enable(False)
    g, og, lat, ref = make_pair(u1, orc, (6, 10, 4, 12), 1.0, 1.0, True, seed=8)
    for s in range(3):
        p = lat.update_plaquette(2, 0, u1.BENCH_METRO, 8, s)
        orc.update(og, ref, 2, 0, orc.metro(1, 4, 0.5), 8, s)
        assert np.array_equal(bits(lat.to_host()), bits(ref))
        assert np.max(np.abs(p - orc.plaquette(og, ref))) <= 1e-13


def test_fp64_pass_at_48x4(u1, orc, forced):
    """The production size: three sweeps, bit for bit against the oracle."""
    u1.tune_enable(False)
    g, og, lat, ref = make_pair(u1, orc, (48, 48, 48, 48), 1.0, 1.0, True, seed=61)
    m, om = u1.BENCH_METRO, orc.metro(1, 4, 0.5)
    for s in range(3):
        assert lat.update(1, 3, m, 61, s, count_accepts=True) == orc.update(og, ref, 1, 3, om, 61, s,
                                                                             count_accepts=True)
    assert np.array_equal(bits(lat.to_host()), bits(ref))
    u1.sync()


def test_pass_not_used_with_masks(u1, orc, forced):
    """Multilevel masks keep the unfused phases: masked sweeps and a multilevel
    call with pass fusion forced still match the oracle."""
    u1.tune_enable(False)
    dims = (4, 4, 4, 8)
    g, og, lat, ref = make_pair(u1, orc, dims, seed=9)
    for ml, stream in ((4, 3), (2, 4)):
        for s in range(2):
            lat.update(1, 3, u1.BENCH_METRO, 9, s, stream=stream, c3=5, mask_ml=ml)
            orc.update(og, ref, 1, 3, orc.metro(1, 4, 0.5), 9, s, stream=stream, c3=5, mask_ml=ml)
    assert np.array_equal(bits(lat.to_host()), bits(ref))
    r_gpu = lat.multilevel(2, 1, 3, 1, 1, 2, u1.BENCH_METRO, 19, call_id=2, mh_mode=1)
    r_orc = orc.multilevel(og, ref, 2, 1, 3, 1, 1, 2, orc.metro(1, 4, 0.5), 19, call_id=2, mh_mode=1)
    assert np.allclose(r_gpu, r_orc, rtol=1e-11, atol=1e-16)


def _fp32_run(u1, g, seed, sweeps):
    lat = u1.Lattice.cold(g, u1.F32)
    lat.hotstart(seed)
    acc = [lat.update(1, 3, u1.BENCH_METRO, seed, s, count_accepts=True) for s in range(sweeps)]
    p = lat.update_plaquette(1, 3, u1.BENCH_METRO, seed, sweeps)
    return lat.to_host(), acc, p


@pytest.mark.parametrize("dims,tune", [((16, 16, 16, 16), False), ((48, 48, 48, 48), False),
                                       ((48, 48, 48, 48), True), ((24, 24, 8, 32), True), ((64, 32, 16, 16), False)])
def test_fp32_pass_matches_unfused(u1, dims, tune):
    """FP32 has no oracle: the fused passes are pinned to the unfused FP32
    path (no pairs, no passes), bit for bit, over sweeps with and without the
    fused plaquette."""
    g = u1.geom(dims, 1.0)
    out = {}
    try:
        u1.tune_enable(tune)
        for mode in (0, 2):
            u1.set_pair_fusion(0)
            u1.set_pass_fusion(mode)
            out[mode] = _fp32_run(u1, g, 13, 4)
        u1.sync()
    finally:
        u1.set_pair_fusion(1)
        u1.set_pass_fusion(1)
        u1.tune_enable(True)
    say(f"fp32 fused pass vs unfused {dims} tune={tune}: accepts {out[2][1]}")
    assert np.array_equal(out[0][0], out[2][0])
    assert out[0][1] == out[2][1]
    assert np.max(np.abs(out[0][2] - out[2][2])) <= 1e-12


@pytest.mark.parametrize("block,spt,tz,margin", [(64, 1, 2, 0), (128, 1, 8, 16), (192, 1, 4, 8), (256, 1, 2, 64),
                                                 (384, 1, 8, 0), (512, 1, 4, 32), (64, 2, 8, 8), (128, 2, 2, 16),
                                                 (256, 2, 4, 0), (384, 2, 8, 64), (512, 2, 2, 8), (256, 1, 1, 16)])
def test_fp32_pass_shapes_match_unfused(u1, tmp_path, block, spt, tz, margin):
    """Fused-pass shapes the tuner can choose, forced through the tune cache on
    48 x 48 planes (P = 1152 sites per parity: from 18 chunks of 64 down to
    one chunk of 1024 slots) in 1 x 8 to 8 x 1 tilings of the (z, t) planes
    and with schedule margins from 0 to 64 steps, bit for bit against the
    unfused path."""
    dims = (48, 48, 8, 8)
    vol = "x".join(map(str, dims))
    cache = tmp_path / "chain_shapes.csv"
    rows = [f"{vol}\t{name}_f32\tthreads=halfvolume\t{b}\t{s}\t{z}\t{gx}\t1\t1\t0\t# forced"
            for name, b, s, z, gx in (("Metropolis", 256, 1, 2, 0), ("OverRelaxation", 256, 1, 2, 0),
                                      ("MetropolisPair", 0, 1, 1, 0), ("OverRelaxationPair", 0, 1, 1, 0),
                                      ("MetropolisChain", block, spt, tz, margin),
                                      ("OverRelaxationChain", block, spt, tz, margin))]
    cache.write_text("tunecache\tlatticeu1_amd-1\tgpu_arch=gfx950\nvolume\tname\taux\n" + "\n".join(rows) + "\n")
    g = u1.geom(dims, 1.0)
    out = {}
    try:
        u1.tune_enable(False)
        u1.tune_load(str(cache))
        for mode in (0, 1):  # 1: the cached decision, i.e. the forced shape
            u1.set_pass_fusion(mode)
            out[mode] = _fp32_run(u1, g, 17, 3)
        u1.sync()
    finally:
        u1.set_pass_fusion(1)
        u1.tune_enable(True)
    say(f"fp32 fused pass shape {block} x {spt}, {tz} z tiles, margin {margin}, vs unfused: accepts {out[1][1]}")
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1] == out[1][1]
    assert np.max(np.abs(out[0][2] - out[1][2])) <= 1e-12


def test_fp64_pass_shapes_match_oracle(u1, orc, tmp_path):
    """Binary64 fused passes in one and two sites per thread, forced through
    the tune cache: bit for bit against the oracle."""
    dims = (16, 8, 4, 8)
    vol = "x".join(map(str, dims))
    for block, spt, tz in ((64, 2, 4), (128, 1, 2), (192, 1, 1)):
        cache = tmp_path / f"c{block}_{spt}.csv"
        rows = [f"{vol}\t{name}_f64\tthreads=halfvolume\t{block}\t{spt}\t{tz}\t0\t1\t1\t0\t# forced"
                for name in ("MetropolisChain", "OverRelaxationChain")]
        cache.write_text("tunecache\tlatticeu1_amd-1\tgpu_arch=gfx950\nvolume\tname\taux\n" + "\n".join(rows) + "\n")
        try:
            u1.tune_enable(False)
            u1.tune_load(str(cache))
            g, og, lat, ref = make_pair(u1, orc, dims, 1.0, 1.0, True, seed=71)
            for s in range(3):
                lat.update(1, 3, u1.BENCH_METRO, 71, s)
                orc.update(og, ref, 1, 3, orc.metro(1, 4, 0.5), 71, s)
            assert np.array_equal(bits(lat.to_host()), bits(ref)), (block, spt)
        finally:
            u1.tune_enable(True)


def test_pass_fusion_mode_checked(u1):
    with pytest.raises(u1.U1Error):
        u1.set_pass_fusion(3)
    with pytest.raises(u1.U1Error):
        u1.set_pass_fusion(-1)
    u1.set_pass_fusion(1)
