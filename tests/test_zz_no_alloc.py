"""u1_plaquette_noalloc: the plaquette through the library's fixed device
scratch (no allocation), against u1_plaquette through the C ABI.

Two calls in a row, the second on a smaller lattice (the scratch keeps the
first call's rows, which the second must not read), both bit-identical to
u1_plaquette (the same kernels and fold order); then a lattice too large for
the scratch is rejected before any launch.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(u1, lat):
    out = u1.DeviceArray(4, np.float64)
    u1.call("u1_plaquette_noalloc", C.byref(lat.g), lat.prec, lat.buf.ptr, out.ptr)
    return out.to_host()


@pytest.mark.parametrize("prec", [0, 1])
def test_plaquette_noalloc_matches_plaquette(u1, prec):
    for dims, seed in (((16, 16, 16, 32), 5), ((8, 8, 8, 8), 6)):  # the second call is smaller
        print(f"[gpu] plaquette_noalloc prec={prec} {dims}", flush=True)
        g = u1.geom(dims, 1.0)
        lat = u1.Lattice.cold(g, prec)
        lat.hotstart(seed)
        for s in range(2):
            lat.update(1, 3, u1.BENCH_METRO, seed, s)
        a = _noalloc(u1, lat)
        b = lat.plaquette()
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), (dims, a, b)


def test_plaquette_noalloc_rejects_oversized(u1):
    small = u1.DeviceArray(4, np.float64)
    out = u1.DeviceArray(4, np.float64)
    g = u1.geom((128, 128, 128, 200), 1.0)  # needs more rows than the scratch holds
    with pytest.raises(u1.U1Error) as ei:
        u1.call("u1_plaquette_noalloc", C.byref(g), u1.F64, small.ptr, out.ptr)
    assert ei.value.code == 1 and "too large" in str(ei.value)
