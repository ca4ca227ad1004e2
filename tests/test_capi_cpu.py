"""Host-side checks of the C ABI library (no GPU needed, no compute calls).

* libu1amd.so loads and exports every function include/latticeu1_amd.h declares;
* calls before u1_init fail with U1_ERR_STATE and an error message;
* the tune cache file round-trips and malformed files are rejected
  (tune.cpp:195-209 treats a bad cache as an error).
"""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    with open(os.path.join(ROOT, "include", "latticeu1_amd.h")) as f:
        txt = f.read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(u1_\w+)\s*\(", txt, flags=re.M)))


def test_header_declares_expected_entry_points():
    names = declared_functions()
    for must in ["u1_init", "u1_update", "u1_hotstart", "u1_plaquette", "u1_polyakov", "u1_poly2", "u1_multilevel",
                 "u1_convert_layout", "u1_plaquette_fields", "u1_ensemble_create", "u1_tune_save"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    import latticeu1_amd as u1

    lib = u1.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python mirror binds every declared entry point
    bound = set(u1._SIGS) | {"u1_last_error"}
    assert set(declared_functions()) <= bound


def test_calls_before_init_fail_cleanly():
    import latticeu1_amd as u1

    lib = u1.load()
    g = u1.geom((4, 4, 4, 4))
    out = (C.c_double * 4)()
    assert lib.u1_plaquette(C.byref(g), 0, None, out) == 3
    assert b"u1_init" in lib.u1_last_error()
    with pytest.raises(u1.U1Error):
        u1.call("u1_update", C.byref(g), 0, None, 1, 3, C.byref(u1.BENCH_METRO), 1, 0, 0, 0, 0, None)


def test_tune_cache_roundtrip(tmp_path):
    import latticeu1_amd as u1

    lib = u1.load()
    p = tmp_path / "tunecache.csv"
    p.write_text("tunecache\tlatticeu1_amd-1\tgpu_arch=gfx950\nvolume\tname\taux\n"
                 "8x8x8x8\tMetropolis_f64\tthreads=halfvolume\t128\t1\t1\t0\t1\t1\t0\t# x\n"
                 "8x8x8x8\tOverRelaxation_f32\tthreads=halfvolume\t192\t2\t4\t0\t1\t1\t0\t# x\n"
                 "8x8x8x8\tOverRelaxationPair_f32\tthreads=halfvolume\t576\t2\t2\t0\t1\t1\t0\t# x\n"
                 "8x8x8x8\tMetropolisPair_f64\tthreads=halfvolume\t0\t1\t1\t0\t1\t1\t0\t# x\n"
                 "8x8x8x8\tMetropolisPair_f32\tthreads=halfvolume\t384\t3\t2\t0\t1\t1\t0\t# x\n"
                 "8x8x8x8\tOverRelaxationChain_f64\tthreads=halfvolume\t64\t2\t8\t128\t1\t1\t0\t# x\n"
                 "8x8x8x8\tMetropolisChain_f32\tthreads=halfvolume\t0\t1\t0\t0\t1\t1\t0\t# x\n")
    u1.tune_load(str(p))
    n = C.c_int32(0)
    assert lib.u1_tune_entries(C.byref(n)) == 0 and n.value >= 4
    q = tmp_path / "out.csv"
    u1.tune_save(str(q))
    txt = q.read_text()
    assert txt.startswith("tunecache\tlatticeu1_amd-1") and "8x8x8x8\tMetropolis_f64" in txt
    # block.z = 1 + XCD order + 2 x tiled planes survives the round trip
    assert "8x8x8x8\tMetropolis_f64\tthreads=halfvolume\t128\t1\t1\t" in txt
    assert "8x8x8x8\tOverRelaxation_f32\tthreads=halfvolume\t192\t2\t4\t" in txt
    # fused pair shapes (workgroup = whole plane, up to 1024 threads) and "two launches" (block 0)
    assert "8x8x8x8\tOverRelaxationPair_f32\tthreads=halfvolume\t576\t2\t2\t" in txt
    assert "8x8x8x8\tMetropolisPair_f64\tthreads=halfvolume\t0\t1\t1\t" in txt
    # three sites per thread (binary32 pairs, up to 512 threads)
    assert "8x8x8x8\tMetropolisPair_f32\tthreads=halfvolume\t384\t3\t2\t" in txt
    # fused parity passes: block.z = tiles along z, grid.x = schedule margin; block 0 = unfused
    assert "8x8x8x8\tOverRelaxationChain_f64\tthreads=halfvolume\t64\t2\t8\t128\t" in txt
    assert "8x8x8x8\tMetropolisChain_f32\tthreads=halfvolume\t0\t1\t" in txt
    bad = tmp_path / "bad.csv"
    bad.write_text("not a cache\n")
    with pytest.raises(u1.U1Error):
        u1.tune_load(str(bad))
    stale = tmp_path / "stale.csv"
    stale.write_text("tunecache\tother-version\tgpu_arch=gfx950\n")
    with pytest.raises(u1.U1Error):
        u1.tune_load(str(stale))


@pytest.mark.parametrize("name,shape", [("Metropolis_f64", s) for s in ("100\t1\t1", "0\t1\t1", "1024\t1\t1",
                                                                       "128\t3\t1", "128\t1\t5", "128\t1\t0",
                                                                       "64\t4\t2")]
                         + [("OverRelaxationPair_f32", "1088\t2\t1"), ("OverRelaxationPair_f32", "600\t2\t1"),
                            ("MetropolisPair_f64", "1024\t1\t1"), ("MetropolisPair_f64", "256\t4\t1"),
                            ("MetropolisPair_f64", "0\t2\t1")])
def test_tune_cache_rejects_unknown_launch_shapes(tmp_path, name, shape):
    """A cache entry must be one of the tuner's own shapes (whole wavefronts,
    <= 512 threads, 1/2/4 sites per thread, 4 only for binary32; block.z 1-4 =
    1 + XCD order + 2 x tiled planes; fused pairs: whole wavefronts up to 1024
    threads in binary32 and 512 in binary64, or block 0 with one site per
    thread for two launches), and a file with one bad line changes nothing."""
    import latticeu1_amd as u1

    lib = u1.load()
    head = "tunecache\tlatticeu1_amd-1\tgpu_arch=gfx950\nvolume\tname\taux\n"
    good = "6x6x6x6\tOverRelaxation_f32\tthreads=halfvolume\t256\t4\t2\t0\t1\t1\t0\t# x\n"
    n0 = C.c_int32(0)
    assert lib.u1_tune_entries(C.byref(n0)) == 0
    p = tmp_path / "bad_shape.csv"
    p.write_text(head + good + f"6x6x6x6\t{name}\tthreads=halfvolume\t{shape}\t0\t1\t1\t0\t# x\n")
    with pytest.raises(u1.U1Error) as e:
        u1.tune_load(str(p))
    assert e.value.code == 5 and "launch shape" in str(e.value)
    n1 = C.c_int32(0)
    assert lib.u1_tune_entries(C.byref(n1)) == 0 and n1.value == n0.value


@pytest.mark.parametrize("shape", ["128\t3\t2\t16", "100\t1\t2\t16", "1024\t1\t2\t16", "256\t1\t3\t16",
                                   "256\t1\t16\t16", "256\t1\t2\t-1", "256\t1\t2\t5000", "0\t2\t2\t0"])
def test_tune_cache_rejects_bad_fused_pass_entries(tmp_path, shape):
    """Fused-pass entries: 64-512 threads in whole wavefronts, 1 or 2 sites per
    thread, 1, 2, 4 or 8 tiles along z, a margin in [0, 1024]; or block 0
    (unfused) with one site per thread.  A bad line changes nothing."""
    import latticeu1_amd as u1

    lib = u1.load()
    head = "tunecache\tlatticeu1_amd-1\tgpu_arch=gfx950\nvolume\tname\taux\n"
    n0 = C.c_int32(0)
    assert lib.u1_tune_entries(C.byref(n0)) == 0
    p = tmp_path / "bad_chain.csv"
    p.write_text(head + f"6x6x6x6\tOverRelaxationChain_f32\tthreads=halfvolume\t{shape}\t1\t1\t0\t# x\n")
    with pytest.raises(u1.U1Error) as e:
        u1.tune_load(str(p))
    assert e.value.code == 5 and "launch shape" in str(e.value)
    n1 = C.c_int32(0)
    assert lib.u1_tune_entries(C.byref(n1)) == 0 and n1.value == n0.value


def test_pass_fusion_mode_checked_on_host():
    import latticeu1_amd as u1

    u1.load()
    for bad in (-1, 3):
        with pytest.raises(u1.U1Error):
            u1.set_pass_fusion(bad)
    u1.set_pass_fusion(1)


def test_config_file_validation_on_host(tmp_path):
    """u1_config_load checks the header before touching the device."""
    import struct

    import latticeu1_amd as u1

    g = u1.geom((4, 4, 4, 4))
    bad = tmp_path / "bad.cfg"
    bad.write_bytes(b"garbage" * 20)
    with pytest.raises(u1.U1Error) as e:
        u1.call("u1_config_load", str(bad).encode(), C.byref(g), 0, None, None, None)
    assert e.value.code == 5
    hdr = struct.pack("<8sI4iIddQQ", b"U1AMDCFG", 1, 4, 4, 4, 6, 0, 1.0, 1.0, 7, 3)
    assert len(hdr) == 64
    other = tmp_path / "other.cfg"
    other.write_bytes(hdr)
    with pytest.raises(u1.U1Error) as e:
        u1.call("u1_config_load", str(other).encode(), C.byref(g), 0, None, None, None)
    assert e.value.code == 1  # geometry mismatch
