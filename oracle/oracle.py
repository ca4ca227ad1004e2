"""ctypes wrapper of the FP64 host oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU baseline,
never as the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "_build", "liboracle.so")
LIB_LIBM = os.path.join(_HERE, "_build", "liboracle_libm.so")


class OGeom(C.Structure):
    _fields_ = [("dims", C.c_int32 * 4), ("beta", C.c_double), ("aniso", C.c_double)]


class OMetro(C.Structure):
    _fields_ = [("proposal", C.c_int32), ("nhit", C.c_int32), ("step", C.c_double)]


_D = C.POINTER(C.c_double)
_U32 = C.POINTER(C.c_uint32)
_SIGS = {
    "orc_sincos": ([_D, _D, _D, C.c_int64], None),
    "orc_exp": ([_D, _D, C.c_int64], None),
    "orc_atan2": ([_D, _D, _D, C.c_int64], None),
    "orc_bessel_i0i1": ([_D, _D, _D, C.c_int64], None),
    "orc_philox": ([_U32, _U32, _U32, C.c_int64], None),
    "orc_threads": ([], C.c_int),
    "orc_set_threads": ([C.c_int], None),
    "orc_eo_coords": ([C.POINTER(OGeom), C.c_int32, C.c_int32, C.POINTER(C.c_int32)], None),
    "orc_eo_neighbor": ([C.POINTER(OGeom), C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32],
                        C.c_int32),
    "orc_staple": ([C.POINTER(OGeom), _D, C.c_int32, C.c_int32, C.c_int32, _D], None),
    "orc_hotstart": ([C.POINTER(OGeom), _D, C.c_uint64, C.c_uint32], None),
    "orc_update": ([C.POINTER(OGeom), _D, C.c_int32, C.c_int32, C.POINTER(OMetro), C.c_uint64, C.c_uint32, C.c_uint32,
                    C.c_uint32, C.c_int32, C.POINTER(C.c_uint64)], None),
    "orc_metropolis_phase": ([C.POINTER(OGeom), _D, C.c_int32, C.c_int32, C.POINTER(OMetro), C.c_uint64, C.c_uint32,
                              C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint64)], None),
    "orc_overrelax_phase": ([C.POINTER(OGeom), _D, C.c_int32, C.c_int32], None),
    "orc_plaquette": ([C.POINTER(OGeom), _D, _D], None),
    "orc_plaquette_host": ([C.POINTER(OGeom), _D, _D], None),
    "orc_polyakov": ([C.POINTER(OGeom), _D, _D], None),
    "orc_poly_volume": ([C.POINTER(OGeom), _D, C.c_int32, C.POINTER(OMetro), C.c_uint64, C.c_uint32, C.c_uint32,
                         C.c_uint32, _D], None),
    "orc_pp": ([C.POINTER(OGeom), _D, C.c_int32, _D], None),
    "orc_temporal_links": ([C.POINTER(OGeom), _D, C.c_int32, C.POINTER(OMetro), C.c_uint64, C.c_uint32, C.c_uint32,
                            C.c_uint32, _D], None),
    "orc_multilevel": ([C.POINTER(OGeom), _D, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                        C.POINTER(OMetro), C.c_int32, C.POINTER(OMetro), C.c_uint64, C.c_uint32, _D, _D], None),
    "orc_convert": ([C.POINTER(OGeom), _D, _D, C.c_int32], None),
    "orc_plaquette_fields": ([C.POINTER(OGeom), _D, C.c_int32, C.c_int32, _D, _D], None),
}


def build() -> None:
    subprocess.run(["make", "-C", _HERE, "-s"], check=True)


class Oracle:
    def __init__(self, libm: bool = False):
        path = LIB_LIBM if libm else LIB
        if not os.path.exists(path):
            build()
        self.lib = C.CDLL(path)
        for name, (args, res) in _SIGS.items():
            fn = getattr(self.lib, name)
            fn.argtypes = args
            fn.restype = res

    # ---- helpers
    @staticmethod
    def geom(dims, beta=1.0, aniso=1.0) -> OGeom:
        g = OGeom()
        for i in range(4):
            g.dims[i] = int(dims[i])
        g.beta = float(beta)
        g.aniso = float(aniso)
        return g

    @staticmethod
    def metro(proposal=0, nhit=1, step=0.5) -> OMetro:
        return OMetro(int(proposal), int(nhit), float(step))

    @staticmethod
    def _d(a: np.ndarray):
        assert a.dtype == np.float64 and a.flags.c_contiguous
        return a.ctypes.data_as(_D)

    # ---- elementary functions
    def sincos(self, x: np.ndarray):
        x = np.ascontiguousarray(x, dtype=np.float64)
        s = np.empty_like(x)
        c = np.empty_like(x)
        self.lib.orc_sincos(self._d(x), self._d(s), self._d(c), x.size)
        return s, c

    def exp(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        self.lib.orc_exp(self._d(x), self._d(y), x.size)
        return y

    def atan2(self, y, x):
        y = np.ascontiguousarray(y, dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64)
        o = np.empty_like(x)
        self.lib.orc_atan2(self._d(y), self._d(x), self._d(o), x.size)
        return o

    def bessel_i0i1(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        a = np.empty_like(x)
        b = np.empty_like(x)
        self.lib.orc_bessel_i0i1(self._d(x), self._d(a), self._d(b), x.size)
        return a, b

    def philox(self, ctr: np.ndarray, key: np.ndarray) -> np.ndarray:
        ctr = np.ascontiguousarray(ctr, dtype=np.uint32).reshape(-1, 4)
        key = np.ascontiguousarray(key, dtype=np.uint32).reshape(-1, 2)
        out = np.empty_like(ctr)
        self.lib.orc_philox(ctr.ctypes.data_as(_U32), key.ctypes.data_as(_U32), out.ctypes.data_as(_U32), ctr.shape[0])
        return out

    def threads(self) -> int:
        return self.lib.orc_threads()

    def set_threads(self, n: int) -> None:
        self.lib.orc_set_threads(n)

    # ---- geometry
    def eo_coords(self, g, id_, parity):
        x = (C.c_int32 * 4)()
        self.lib.orc_eo_coords(C.byref(g), id_, parity, x)
        return list(x)

    def eo_neighbor(self, g, id_, parity, mu, lmu, nu=0, lnu=0):
        return self.lib.orc_eo_neighbor(C.byref(g), id_, parity, mu, lmu, nu, lnu)

    def staple(self, g, lat, id_, parity, mu) -> complex:
        out = np.zeros(2)
        self.lib.orc_staple(C.byref(g), self._d(lat), id_, parity, mu, self._d(out))
        return complex(out[0], out[1])

    # ---- updates
    def cold(self, g) -> np.ndarray:
        v = g.dims[0] * g.dims[1] * g.dims[2] * g.dims[3]
        return np.zeros(4 * v)

    def hotstart(self, g, seed, replica=0) -> np.ndarray:
        lat = self.cold(g)
        self.lib.orc_hotstart(C.byref(g), self._d(lat), seed, replica)
        return lat

    def update(self, g, lat, metrop, ovrn, m, seed, sweep, stream=0, c3=0, mask_ml=0, count_accepts=False) -> int:
        acc = C.c_uint64(0)
        self.lib.orc_update(C.byref(g), self._d(lat), metrop, ovrn, C.byref(m), seed, stream, sweep, c3, mask_ml,
                            C.byref(acc) if count_accepts else None)
        return int(acc.value)

    def metropolis_phase(self, g, lat, parity, mu, m, seed, sweep=0, pass_=0, stream=0, c3=0) -> int:
        acc = C.c_uint64(0)
        self.lib.orc_metropolis_phase(C.byref(g), self._d(lat), parity, mu, C.byref(m), seed, stream, sweep, pass_, c3,
                                      C.byref(acc))
        return int(acc.value)

    def overrelax_phase(self, g, lat, parity, mu) -> None:
        self.lib.orc_overrelax_phase(C.byref(g), self._d(lat), parity, mu)

    # ---- measurements
    def plaquette(self, g, lat) -> np.ndarray:
        out = np.zeros(4)
        self.lib.orc_plaquette(C.byref(g), self._d(lat), self._d(out))
        return out

    def plaquette_host(self, g, lat) -> np.ndarray:
        out = np.zeros(2)
        self.lib.orc_plaquette_host(C.byref(g), self._d(lat), self._d(out))
        return out

    def polyakov(self, g, lat) -> complex:
        out = np.zeros(2)
        self.lib.orc_polyakov(C.byref(g), self._d(lat), self._d(out))
        return complex(out[0], out[1])

    def poly_volume(self, g, lat, mode, mh: Optional[OMetro] = None, seed=0, stream=2, c1=0, c3=0) -> np.ndarray:
        sv = g.dims[0] * g.dims[1] * g.dims[2]
        P = np.zeros(2 * sv)
        self.lib.orc_poly_volume(C.byref(g), self._d(lat), mode, C.byref(mh) if mh else None, seed, stream, c1, c3,
                                 self._d(P))
        return P[0::2] + 1j * P[1::2]

    def pp(self, g, P: np.ndarray, rmax: int) -> np.ndarray:
        Pr = np.empty(2 * P.size)
        Pr[0::2] = P.real
        Pr[1::2] = P.imag
        out = np.zeros(2 * rmax)
        self.lib.orc_pp(C.byref(g), self._d(Pr), rmax, self._d(out))
        return out[0::2] + 1j * out[1::2]

    def poly2(self, g, lat, mode, mh=None, seed=0, c1=0, c3=0) -> np.ndarray:
        return self.pp(g, self.poly_volume(g, lat, mode, mh, seed, 2, c1, c3), g.dims[0] // 2)

    def temporal_links(self, g, lat, mode, mh=None, seed=0, stream=2, c1=0, c3=0) -> np.ndarray:
        v = g.dims[0] * g.dims[1] * g.dims[2] * g.dims[3]
        out = np.zeros(2 * v)
        self.lib.orc_temporal_links(C.byref(g), self._d(lat), mode, C.byref(mh) if mh else None, seed, stream, c1, c3,
                                    self._d(out))
        return out[0::2] + 1j * out[1::2]

    def multilevel(self, g, lat, n4, k4, n2, k2, metrop, ovrn, m, seed, call_id=0, mh_mode=1, mh=None, field=False):
        r = g.dims[0] // 2
        out = np.zeros(2 * r)
        sv = g.dims[0] * g.dims[1] * g.dims[2]
        pps = np.zeros(2 * sv * r * 3) if field else None
        self.lib.orc_multilevel(C.byref(g), self._d(lat), n4, k4, n2, k2, metrop, ovrn, C.byref(m), mh_mode,
                                C.byref(mh) if mh else None, seed, call_id, self._d(out),
                                self._d(pps) if field else None)
        res = out[0::2] + 1j * out[1::2]
        if field:
            return res, pps[0::2] + 1j * pps[1::2]
        return res

    def convert(self, g, lat, eo_to_no: bool) -> np.ndarray:
        out = np.zeros_like(lat)
        self.lib.orc_convert(C.byref(g), self._d(lat), self._d(out), 1 if eo_to_no else 0)
        return out

    def plaquette_fields(self, g, lat, spacetime: bool, eo_order: bool = True):
        size = g.dims[0] * g.dims[1] * g.dims[2] * (g.dims[3] if spacetime else 1)
        F = np.zeros(12 * size)
        means = np.zeros(12)
        self.lib.orc_plaquette_fields(C.byref(g), self._d(lat), 1 if spacetime else 0, 1 if eo_order else 0,
                                      self._d(F), self._d(means))
        return (F[0::2] + 1j * F[1::2]).reshape(6, size), means[0::2] + 1j * means[1::2]
