"""The C++ drop-in layer end to end: a reference-style driver (tests/dropin/
u1_mini.cpp, built by build() with g++ against include/u1 and libu1.so) runs on
the GPU; its lattice, observables and data files are checked against the oracle.
"""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "latticeu1_amd", "u1_mini")


def test_dropin_driver_matches_oracle(orc, tmp_path):
    assert os.path.exists(EXE), "build() has not produced latticeu1_amd/u1_mini"
    print("[gpu] drop-in driver", flush=True)
    out = subprocess.run([EXE, "20"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    vals = {}
    for line in out.stdout.splitlines():
        parts = line.split()
        if parts and parts[0].isupper() and len(parts) > 1:
            try:
                vals[parts[0]] = [float(v) for v in parts[1:]]
            except ValueError:
                pass
    dims = (8, 8, 8, 8)
    og = orc.geom(dims, 1.0, 1.0)
    m = orc.metro(1, 4, 0.5)
    ref = orc.hotstart(og, 1234)
    for s in range(20):
        orc.update(og, ref, 1, 3, m, 1234, s)
    lat = np.fromfile(tmp_path / "lattice.bin", dtype=np.float64)
    assert np.array_equal(lat.view(np.uint64), ref.view(np.uint64))
    p = orc.plaquette(og, ref)
    assert np.allclose(vals["PLAQ"], p, atol=1e-12)
    L = orc.polyakov(og, ref)
    assert abs(complex(*vals["POLY"]) - L) < 1e-12
    assert abs(complex(*vals["POLY2"]) - orc.poly2(og, ref, 0)[0]) < 1e-12
    assert abs(complex(*vals["POLY2MH"]) - orc.poly2(og, ref, 1)[0]) < 1e-12
    ml = orc.multilevel(og, ref, 2, 1, 2, 1, 1, 3, m, 1234, call_id=0, mh_mode=1)
    assert abs(complex(*vals["ML"]) - ml[0]) < 1e-11 * max(1.0, abs(ml[0]))
    assert vals["CONVERT_ROUNDTRIP"] == [1.0]
    F, means = orc.plaquette_fields(og, ref, True, True)
    assert abs(complex(*vals["FIELD0"]) - means[0]) < 1e-12
    hp = orc.plaquette_host(og, ref)
    assert np.allclose(vals["HOSTPLAQ"], hp, atol=1e-12)
    # host-array metropolis() / overrelaxation() (update.cu:51-98) from the
    # final lattice, host generator (seed 1234, sweeps 0, 1, 2, 3)
    host = ref.copy()
    acc = 0
    for s in range(2):
        acc += orc.update(og, host, 1, 0, m, 1234, 2 * s, count_accepts=True)
        orc.update(og, host, 0, 1, m, 1234, 2 * s + 1)
    got = np.fromfile(tmp_path / "host_passes.bin", dtype=np.float64)
    assert np.array_equal(got.view(np.uint64), host.view(np.uint64))
    assert vals["HOSTACCEPT"] == [float(acc)]
    # FP32 overloads: the same kernels through the C ABI from Python (hot
    # start, 5 sweeps of seed 1234) give the driver's FP32 plaquette
    import latticeu1_amd as u1

    u1.init(0)
    l32 = u1.Lattice.cold(u1.geom(dims, 1.0, 1.0), u1.F32)
    l32.hotstart(1234)
    for s in range(5):
        l32.update(1, 3, u1.BENCH_METRO, 1234, s)
    p32 = l32.plaquette()
    assert abs(vals["PLAQ32"][0] - p32[0]) <= 1e-14 and abs(vals["PLAQ32"][1] - p32[2]) <= 1e-14
    name = "8_8_8_8_beta_1_xi_1_mtr_1_ovr_3_iter_20"
    for fn in (f"Pot_{name}.dat", f"Pot_mhit_{name}.dat", f"Pot_mlevel_{name}_2_1_2_1_1_3.dat"):
        rows = (tmp_path / fn).read_text().strip().splitlines()
        assert len(rows) == 4 and rows[0].split("\t")[0] == "1"
    assert (tmp_path / "tunecache.csv").read_text().startswith("tunecache")


RUN = os.path.join(ROOT, "latticeu1_amd", "u1_run")


def test_run_driver_files_match_oracle(orc, tmp_path):
    """latticeu1_amd/u1_run, the reference's u1.cpp loop on the drop-in layer:
    plaq_/poly_ files every printiter sweeps (line formats of u1.cpp:84-128),
    and Poly2 / multihit / MultiLevel files from iteration 1000 on.  The FP64
    chain is the oracle's (default proposal: independent, 1 hit)."""
    assert os.path.exists(RUN), "build() has not produced latticeu1_amd/u1_run"
    print("[gpu] run driver 8^4, 1000 sweeps", flush=True)
    out = subprocess.run([RUN, "8", "8", "8", "8", "1.0", "1.0", "1", "3", "1000", "250", "1"], cwd=tmp_path,
                         capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    name = "8_8_8_8_beta_1_xi_1_mtr_1_ovr_3"
    plaq = [line.split("\t") for line in (tmp_path / f"plaq_{name}.dat").read_text().strip().splitlines()]
    poly = [line.split("\t") for line in (tmp_path / f"poly_{name}.dat").read_text().strip().splitlines()]
    assert [int(r[0]) for r in plaq] == [0, 250, 500, 750, 1000]
    assert [int(r[0]) for r in poly] == [0, 250, 500, 750, 1000]
    og = orc.geom((8, 8, 8, 8), 1.0, 1.0)
    m = orc.metro(0, 1, 0.5)
    ref = orc.hotstart(og, 1234)
    k = 0
    for it in range(1001):
        if it > 0:
            orc.update(og, ref, 1, 3, m, 1234, it - 1)
        if it % 250 == 0:
            p = orc.plaquette(og, ref)
            assert np.allclose([float(v) for v in plaq[k][1:5]], p, rtol=1e-10, atol=1e-12), (it, plaq[k], p)
            L = orc.polyakov(og, ref)
            assert len(poly[k]) == 4 and poly[k][3].startswith("|L|: ")
            assert abs(complex(float(poly[k][1]), float(poly[k][2])) - L) < 1e-11, (it, poly[k], L)
            assert abs(float(poly[k][3][5:]) - abs(L)) < 1e-11
            k += 1
    c0 = orc.poly2(og, ref, 0)
    rows = (tmp_path / f"Pot_{name}_iter_1000.dat").read_text().strip().splitlines()
    assert len(rows) == 4
    for r, row in enumerate(rows):
        f = row.split("\t")
        assert int(f[0]) == r + 1
        assert abs(complex(float(f[1]), float(f[2])) - c0[r]) < 1e-11 * max(1.0, abs(c0[r]))
    # Pot_mhit_ (analytic multihit, polyakov.cu:634-650) and Pot_mlevel_
    # (MultiLevel(10, 1, 20, 1, 2, 5), the run's first call: call id 0,
    # multilevel.cu:712-730) against the oracle on the same configuration
    want = {f"Pot_mhit_{name}_iter_1000.dat": orc.poly2(og, ref, 1),
            f"Pot_mlevel_{name}_iter_1000_10_1_20_1_2_5.dat":
                orc.multilevel(og, ref, 10, 1, 20, 1, 2, 5, m, 1234, call_id=0, mh_mode=1)}
    for fn, c in want.items():
        rows = (tmp_path / fn).read_text().strip().splitlines()
        assert len(rows) == 4, fn
        for r, row in enumerate(rows):
            f = row.split("\t")
            assert int(f[0]) == r + 1
            assert abs(complex(float(f[1]), float(f[2])) - c[r]) <= 1e-10 * max(abs(c[r]), 1e-12), (fn, r, row, c[r])
