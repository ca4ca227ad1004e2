"""latticeu1_amd -- Python host mirror of the gfx950 LatticeU1 hot path.

A thin ctypes layer over the C ABI in ``include/latticeu1_amd.h``
(``libu1amd.so``).  It mirrors the reference's operator interface for the hot
path -- ``HotStart``/``UpdateLattice`` (update.h:10-12), ``Plaquette``
(plaquette.h:12), ``Polyakov``/``Poly2`` (polyakov.h:13-16), ``MultiLevel``
(multilevel.h:9), ``LatticeConvert`` (array.h:186), ``PlaquetteFields``
(plaquette.h:34) -- with explicit geometry instead of process globals.  Tests
and ``bench.py`` use it; the C++ drop-in layer (include/u1/*.h) calls the same
C ABI.

Nothing here computes on the CPU: every operation is a HIP kernel launched by
the library, and loading fails loudly if the library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libu1amd.so")

F64, F32 = 0, 1
PROPOSAL_INDEPENDENT, PROPOSAL_DELTA = 0, 1
POLY_PLAIN, POLY_MULTIHIT, POLY_MULTIHIT_STOCHASTIC = 0, 1, 2
STREAM_UPDATE, STREAM_HOTSTART, STREAM_MULTIHIT, STREAM_ML_L4, STREAM_ML_L2, STREAM_ML_MULTIHIT = range(6)
H2H, H2D, D2H, D2D = range(4)
ERRORS = {1: "U1_ERR_ARG", 2: "U1_ERR_HIP", 3: "U1_ERR_STATE", 4: "U1_ERR_UNSUPPORTED", 5: "U1_ERR_IO"}


class U1Error(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Geom(C.Structure):
    _fields_ = [("dims", C.c_int32 * 4), ("beta", C.c_double), ("aniso", C.c_double)]


class Metro(C.Structure):
    _fields_ = [("proposal", C.c_int32), ("nhit", C.c_int32), ("step", C.c_double)]


def geom(dims: Sequence[int], beta: float = 1.0, aniso: float = 1.0) -> Geom:
    g = Geom()
    for i in range(4):
        g.dims[i] = int(dims[i])
    g.beta = float(beta)
    g.aniso = float(aniso)
    return g


def metro(proposal: int = PROPOSAL_INDEPENDENT, nhit: int = 1, step: float = 0.5) -> Metro:
    return Metro(int(proposal), int(nhit), float(step))


REFERENCE_METRO = metro(PROPOSAL_INDEPENDENT, 1, 0.0)  # update.cu:181-182
BENCH_METRO = metro(PROPOSAL_DELTA, 4, 0.5)  # theta + U[-0.5, 0.5), 4 hits

_P = C.c_void_p
_SIGS = {
    "u1_init": (C.c_int,),
    "u1_finalize": (),
    "u1_sync": (),
    "u1_device_name": (C.c_char_p, C.c_size_t),
    "u1_malloc": (C.POINTER(C.c_void_p), C.c_size_t),
    "u1_free": (_P,),
    "u1_memcpy": (_P, _P, C.c_size_t, C.c_int),
    "u1_zero": (_P, C.c_size_t),
    "u1_convert_precision": (_P, C.c_int, _P, C.c_int, C.c_int64),
    "u1_hotstart": (C.POINTER(Geom), C.c_int, _P, C.c_uint64, C.c_uint32),
    "u1_update": (C.POINTER(Geom), C.c_int, _P, C.c_int32, C.c_int32, C.POINTER(Metro), C.c_uint64, C.c_uint32,
                  C.c_uint32, C.c_uint32, C.c_int32, C.POINTER(C.c_uint64)),
    "u1_update_plaquette": (C.POINTER(Geom), C.c_int, _P, C.c_int32, C.c_int32, C.POINTER(Metro), C.c_uint64,
                            C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(C.c_double)),
    "u1_phase": (C.POINTER(Geom), C.c_int, _P, C.c_int32, C.c_int32, C.c_int32, C.POINTER(Metro), C.c_uint64,
                 C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int32),
    "u1_selftest_math": (C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int64),
    "u1_selftest_philox": (C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.c_int64),
    "u1_selftest_stream": (C.c_int32, C.c_int64, C.c_int32, C.POINTER(C.c_double)),
    "u1_plaquette": (C.POINTER(Geom), C.c_int, _P, C.POINTER(C.c_double)),
    "u1_plaquette_noalloc": (C.POINTER(Geom), C.c_int, _P, _P),
    "u1_polyakov": (C.POINTER(Geom), C.c_int, _P, C.POINTER(C.c_double)),
    "u1_poly_volume": (C.POINTER(Geom), C.c_int, _P, C.c_int32, C.POINTER(Metro), C.c_uint64, C.c_uint32, C.c_uint32,
                       C.c_uint32, _P),
    "u1_pp": (C.POINTER(Geom), _P, C.c_int32, C.POINTER(C.c_double)),
    "u1_poly2": (C.POINTER(Geom), C.c_int, _P, C.c_int32, C.POINTER(Metro), C.c_uint64, C.c_uint32, C.c_uint32,
                 C.POINTER(C.c_double)),
    "u1_temporal_links": (C.POINTER(Geom), C.c_int, _P, C.c_int32, C.POINTER(Metro), C.c_uint64, C.c_uint32,
                          C.c_uint32, C.c_uint32, _P),
    "u1_multilevel": (C.POINTER(Geom), C.c_int, _P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                      C.POINTER(Metro), C.c_int32, C.POINTER(Metro), C.c_uint64, C.c_uint32, C.POINTER(C.c_double),
                      _P),
    "u1_convert_layout": (C.POINTER(Geom), C.c_int, _P, _P, C.c_int32),
    "u1_plaquette_fields": (C.POINTER(Geom), C.c_int, _P, C.c_int32, C.c_int32, _P, C.POINTER(C.c_double)),
    "u1_config_save": (C.c_char_p, C.POINTER(Geom), C.c_int, _P, C.c_uint64, C.c_uint64),
    "u1_config_load": (C.c_char_p, C.POINTER(Geom), C.c_int, _P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)),
    "u1_ensemble_create": (C.POINTER(_P), C.c_int32, C.POINTER(Geom), C.POINTER(C.c_uint64), C.POINTER(C.c_int32),
                           C.c_int, C.c_int32, C.c_int32, C.POINTER(Metro)),
    "u1_ensemble_sweep": (_P, C.c_int32, C.c_int32),
    "u1_ensemble_sweep_plaquette": (_P, C.c_int32, C.POINTER(C.c_double)),
    "u1_ensemble_bench": (_P, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)),
    "u1_ensemble_plaquette": (_P, C.POINTER(C.c_double)),
    "u1_ensemble_read": (_P, C.c_int32, C.POINTER(C.c_double)),
    "u1_ensemble_write": (_P, C.c_int32, C.POINTER(C.c_double)),
    "u1_ensemble_info": (_P, C.POINTER(C.c_uint64), C.POINTER(C.c_int64)),
    "u1_ensemble_destroy": (_P,),
    "u1_tune_enable": (C.c_int32,),
    "u1_set_pair_fusion": (C.c_int32,),
    "u1_set_pass_fusion": (C.c_int32,),
    "u1_set_graph_capture": (C.c_int32,),
    "u1_set_verbosity": (C.c_int32,),
    "u1_tune_load": (C.c_char_p,),
    "u1_tune_save": (C.c_char_p,),
    "u1_tune_entries": (C.POINTER(C.c_int32),),
    "u1_bench_updates": (C.POINTER(Geom), C.c_int, _P, C.c_int32, C.c_int32, C.POINTER(Metro), C.c_uint64, C.c_uint32,
                         C.c_int32, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)),
}

_lib: Optional[C.CDLL] = None


def load() -> C.CDLL:
    """Load libu1amd.so; raises if it has not been built (no CPU fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(LIB_PATH)
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = list(args)
        fn.restype = C.c_int
    lib.u1_last_error.argtypes = []
    lib.u1_last_error.restype = C.c_char_p
    _lib = lib
    return lib


def _check(rc: int) -> None:
    if rc != 0:
        raise U1Error(rc, load().u1_last_error().decode())


def call(name: str, *args) -> None:
    _check(getattr(load(), name)(*args))


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


_initialised = False


def init(device: int = 0) -> None:
    """Start(gpuid) (cuda_error_check.h:45): select the device, create the stream."""
    global _initialised
    call("u1_init", device)
    _initialised = True


def finalize() -> None:
    global _initialised
    call("u1_finalize")
    _initialised = False


def sync() -> None:
    call("u1_sync")


def device_name() -> str:
    buf = C.create_string_buffer(256)
    call("u1_device_name", buf, 256)
    return buf.value.decode()


class DeviceArray:
    """Owning device buffer (the reference's Array<Real>(Device, n), array.h:19-182)."""

    def __init__(self, n: int, dtype=np.float64):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        p = C.c_void_p()
        call("u1_malloc", C.byref(p), self.nbytes)
        self.ptr = p

    @classmethod
    def from_host(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        call("u1_memcpy", d.ptr, a.ctypes.data_as(C.c_void_p), d.nbytes, H2D)
        return d

    def to_host(self) -> np.ndarray:
        out = np.empty(self.n, self.dtype)
        call("u1_memcpy", out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes, D2H)
        return out

    def copy_from(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        call("u1_memcpy", self.ptr, a.ctypes.data_as(C.c_void_p), self.nbytes, H2D)

    def free(self) -> None:
        if self.ptr is not None and self.ptr.value:
            call("u1_free", self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            if _initialised and self.ptr is not None and self.ptr.value:
                load().u1_free(self.ptr)
        except Exception:
            pass


@dataclass
class Lattice:
    """A gauge field on the device: 4V angles in even/odd SoA layout."""

    g: Geom
    prec: int
    buf: DeviceArray

    @property
    def volume(self) -> int:
        d = self.g.dims
        return d[0] * d[1] * d[2] * d[3]

    @classmethod
    def cold(cls, g: Geom, prec: int = F64) -> "Lattice":
        d = g.dims
        v = d[0] * d[1] * d[2] * d[3]
        return cls(g, prec, DeviceArray(4 * v, np.float64 if prec == F64 else np.float32))

    @classmethod
    def from_host(cls, g: Geom, angles: np.ndarray, prec: int = F64) -> "Lattice":
        dt = np.float64 if prec == F64 else np.float32
        return cls(g, prec, DeviceArray.from_host(np.asarray(angles, dtype=dt)))

    def to_host(self) -> np.ndarray:
        return self.buf.to_host()

    def astype(self, prec: int) -> "Lattice":
        out = Lattice.cold(self.g, prec)
        call("u1_convert_precision", self.buf.ptr, self.prec, out.buf.ptr, prec, 4 * self.volume)
        return out

    # -- update.h
    def hotstart(self, seed: int, replica: int = 0) -> None:
        call("u1_hotstart", C.byref(self.g), self.prec, self.buf.ptr, seed, replica)

    def update(self, metrop: int, ovrn: int, m: Metro, seed: int, sweep: int, stream: int = STREAM_UPDATE,
               c3: int = 0, mask_ml: int = 0, count_accepts: bool = False) -> int:
        acc = C.c_uint64(0)
        call("u1_update", C.byref(self.g), self.prec, self.buf.ptr, metrop, ovrn, C.byref(m), seed, stream, sweep, c3,
             mask_ml, C.byref(acc) if count_accepts else None)
        return int(acc.value)

    def update_plaquette(self, metrop: int, ovrn: int, m: Metro, seed: int, sweep: int, stream: int = STREAM_UPDATE,
                         c3: int = 0) -> np.ndarray:
        """update() fused with the plaquette of the result (u1_plaquette form)."""
        out = np.zeros(4)
        call("u1_update_plaquette", C.byref(self.g), self.prec, self.buf.ptr, metrop, ovrn, C.byref(m), seed, stream,
             sweep, c3, _dp(out))
        return out

    def phase(self, kind: int, parity: int, mu: int, m: Metro, seed: int, sweep: int = 0, pass_: int = 0,
              stream: int = STREAM_UPDATE, c3: int = 0, mask_ml: int = 0) -> None:
        call("u1_phase", C.byref(self.g), self.prec, self.buf.ptr, kind, parity, mu, C.byref(m), seed, stream, sweep,
             pass_, c3, mask_ml)

    # -- plaquette.h / polyakov.h
    def plaquette(self) -> np.ndarray:
        out = np.zeros(4)
        call("u1_plaquette", C.byref(self.g), self.prec, self.buf.ptr, _dp(out))
        return out

    def polyakov(self) -> complex:
        out = np.zeros(2)
        call("u1_polyakov", C.byref(self.g), self.prec, self.buf.ptr, _dp(out))
        return complex(out[0], out[1])

    def poly_volume(self, mode: int, mh: Optional[Metro] = None, seed: int = 0, stream: int = STREAM_MULTIHIT,
                    c1: int = 0, c3: int = 0) -> np.ndarray:
        d = self.g.dims
        sv = d[0] * d[1] * d[2]
        P = DeviceArray(2 * sv)
        call("u1_poly_volume", C.byref(self.g), self.prec, self.buf.ptr, mode, C.byref(mh) if mh else None, seed,
             stream, c1, c3, P.ptr)
        h = P.to_host()
        return h[0::2] + 1j * h[1::2]

    def poly2(self, mode: int, mh: Optional[Metro] = None, seed: int = 0, c1: int = 0, c3: int = 0) -> np.ndarray:
        r = self.g.dims[0] // 2
        out = np.zeros(2 * r)
        call("u1_poly2", C.byref(self.g), self.prec, self.buf.ptr, mode, C.byref(mh) if mh else None, seed, c1, c3,
             _dp(out))
        return out[0::2] + 1j * out[1::2]

    def temporal_links(self, mode: int, mh: Optional[Metro] = None, seed: int = 0, stream: int = STREAM_MULTIHIT,
                       c1: int = 0, c3: int = 0) -> np.ndarray:
        P = DeviceArray(2 * self.volume)
        call("u1_temporal_links", C.byref(self.g), self.prec, self.buf.ptr, mode, C.byref(mh) if mh else None, seed,
             stream, c1, c3, P.ptr)
        h = P.to_host()
        return h[0::2] + 1j * h[1::2]

    # -- multilevel.h
    def multilevel(self, n4: int, k4: int, n2: int, k2: int, metrop: int, ovrn: int, m: Metro, seed: int,
                   call_id: int = 0, mh_mode: int = POLY_MULTIHIT, mh: Optional[Metro] = None,
                   field: bool = False):
        r = self.g.dims[0] // 2
        out = np.zeros(2 * r)
        d = self.g.dims
        pps = DeviceArray(2 * d[0] * d[1] * d[2] * r * 3) if field else None
        call("u1_multilevel", C.byref(self.g), self.prec, self.buf.ptr, n4, k4, n2, k2, metrop, ovrn, C.byref(m),
             mh_mode, C.byref(mh) if mh else None, seed, call_id, _dp(out), pps.ptr if pps else None)
        res = out[0::2] + 1j * out[1::2]
        if field:
            h = pps.to_host()
            return res, h[0::2] + 1j * h[1::2]
        return res

    # -- configuration files
    def save(self, path: str, seed: int, sweep: int) -> None:
        call("u1_config_save", path.encode(), C.byref(self.g), self.prec, self.buf.ptr, seed, sweep)

    def load(self, path: str):
        seed, sweep = C.c_uint64(0), C.c_uint64(0)
        call("u1_config_load", path.encode(), C.byref(self.g), self.prec, self.buf.ptr, C.byref(seed), C.byref(sweep))
        return int(seed.value), int(sweep.value)

    # -- array.h LatticeConvert
    def convert_layout(self, eo_to_no: bool) -> "Lattice":
        out = Lattice.cold(self.g, self.prec)
        call("u1_convert_layout", C.byref(self.g), self.prec, self.buf.ptr, out.buf.ptr, 1 if eo_to_no else 0)
        return out

    def plaquette_fields(self, spacetime: bool, eo_order: bool = True):
        d = self.g.dims
        size = self.volume if spacetime else d[0] * d[1] * d[2]
        F = DeviceArray(12 * size)
        means = np.zeros(12)
        call("u1_plaquette_fields", C.byref(self.g), self.prec, self.buf.ptr, 1 if spacetime else 0,
             1 if eo_order else 0, F.ptr, _dp(means))
        h = F.to_host()
        return (h[0::2] + 1j * h[1::2]).reshape(6, size), means[0::2] + 1j * means[1::2]

    def bench(self, metrop: int, ovrn: int, m: Metro, seed: int, sweep0: int, nsweeps: int, measure_plaq: int,
              kernel_timing: bool = False):
        """measure_plaq: 0 none, 1 fused into the sweep, 2 separate plaquette pass.
        kernel_timing: afterwards, time every launch class of the sweep on a
        scratch copy (the chain is not advanced); returns the 20 values of
        u1_bench_updates' kernel_ms (else None)."""
        measure_plaq = int(measure_plaq)
        series = np.zeros(4 * nsweeps) if measure_plaq else None
        el = C.c_double(0)
        km = np.zeros(20) if kernel_timing else None
        call("u1_bench_updates", C.byref(self.g), self.prec, self.buf.ptr, metrop, ovrn, C.byref(m), seed, sweep0,
             nsweeps, measure_plaq, _dp(series) if measure_plaq else None, C.byref(el),
             _dp(km) if kernel_timing else None)
        return el.value, km, (series.reshape(nsweeps, 4) if measure_plaq else None)


class Ensemble:
    """Batched independent replicas of uneven geometry and coupling (no
    counterpart in the reference, which holds one lattice per process)."""

    def __init__(self, geoms: Sequence[Geom], seeds: Sequence[int], hot: Sequence[int], prec: int, metrop: int,
                 ovrn: int, m: Metro):
        n = len(geoms)
        self.geoms = list(geoms)
        garr = (Geom * n)(*geoms)
        sarr = (C.c_uint64 * n)(*[int(s) for s in seeds])
        harr = (C.c_int32 * n)(*[int(h) for h in hot])
        self._m = m
        p = C.c_void_p()
        call("u1_ensemble_create", C.byref(p), n, garr, sarr, harr, prec, metrop, ovrn, C.byref(m))
        self.ptr = p
        self.n = n

    def sweep(self, nsweeps: int = 1, graph: bool = True) -> None:
        call("u1_ensemble_sweep", self.ptr, nsweeps, 1 if graph else 0)

    def sweep_plaquette(self, nsweeps: int) -> np.ndarray:
        """nsweeps graph sweeps with each replica's plaquette summed inside the
        sweep; returns (nsweeps, nrep, 4) in plaquette() form."""
        out = np.zeros(4 * self.n * nsweeps)
        call("u1_ensemble_sweep_plaquette", self.ptr, nsweeps, _dp(out))
        return out.reshape(nsweeps, self.n, 4)

    def bench(self, nsweeps: int, kernel_timing: bool = False):
        """sweep_plaquette timed with HIP events on the library stream: returns
        (ms, kernel_ms or None, series); kernel_ms as Lattice.bench."""
        out = np.zeros(4 * self.n * nsweeps)
        ms = C.c_double(0.0)
        kms = np.zeros(20) if kernel_timing else None
        call("u1_ensemble_bench", self.ptr, nsweeps, _dp(out), C.byref(ms), _dp(kms) if kernel_timing else None)
        return ms.value, kms, out.reshape(nsweeps, self.n, 4)

    def plaquette(self) -> np.ndarray:
        out = np.zeros(4 * self.n)
        call("u1_ensemble_plaquette", self.ptr, _dp(out))
        return out.reshape(self.n, 4)

    def read(self, r: int) -> np.ndarray:
        d = self.geoms[r].dims
        out = np.zeros(4 * d[0] * d[1] * d[2] * d[3])
        call("u1_ensemble_read", self.ptr, r, _dp(out))
        return out

    def write(self, r: int, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=np.float64)
        call("u1_ensemble_write", self.ptr, r, _dp(a))

    def info(self):
        s = C.c_uint64(0)
        t = C.c_int64(0)
        call("u1_ensemble_info", self.ptr, C.byref(s), C.byref(t))
        return int(s.value), int(t.value)

    def destroy(self) -> None:
        if self.ptr is not None:
            call("u1_ensemble_destroy", self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            if _initialised and self.ptr is not None:
                load().u1_ensemble_destroy(self.ptr)
        except Exception:
            pass


def selftest_math(fn: int, x: np.ndarray, y: Optional[np.ndarray] = None) -> np.ndarray:
    """Evaluate a shared elementary function on the device (see u1_selftest_math)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y if y is not None else x, dtype=np.float64)
    nout = 2 if fn in (0, 3, 5, 7) else 1
    out = np.zeros(nout * x.size)
    call("u1_selftest_math", fn, _dp(x), _dp(y), _dp(out), x.size)
    return out.reshape(x.size, nout) if nout == 2 else out


def selftest_philox(ctr: np.ndarray, key: np.ndarray) -> np.ndarray:
    ctr = np.ascontiguousarray(ctr, dtype=np.uint32).reshape(-1, 4)
    key = np.ascontiguousarray(key, dtype=np.uint32).reshape(-1, 2)
    out = np.zeros_like(ctr)
    u32 = C.POINTER(C.c_uint32)
    call("u1_selftest_philox", ctr.ctypes.data_as(u32), key.ctypes.data_as(u32), out.ctypes.data_as(u32),
         ctr.shape[0])
    return out


def set_verbosity(level: int) -> None:
    """0 silent, 1 the tuner's choices, 2 every tuning candidate (stderr)."""
    call("u1_set_verbosity", level)


def tune_enable(on: bool) -> None:
    call("u1_tune_enable", 1 if on else 0)


def set_graph_capture(on: bool) -> None:
    """Multilevel level-2 steps from one captured hipGraph (True, default) or
    enqueued eagerly (False); identical results."""
    call("u1_set_graph_capture", 1 if on else 0)


def set_pair_fusion(mode: int) -> None:
    """Fused (parity, 0) + (parity, 1) update launches on single lattices:
    0 never, 1 where the tuner measures them faster (default), 2 always."""
    call("u1_set_pair_fusion", mode)


def set_pass_fusion(mode: int) -> None:
    """Fused parity passes (the four phases of a parity in one launch, ordered
    by device-side completion counters) on single lattices: 0 never, 1 where
    the tuner measures them faster (default), 2 always."""
    call("u1_set_pass_fusion", mode)


def tune_save(path: str) -> None:
    call("u1_tune_save", path.encode())


def tune_load(path: str) -> None:
    call("u1_tune_load", path.encode())


def mean_plaquette(p4: np.ndarray) -> float:
    """<P> = 1 - Re(plaq_SS + plaq_ST)/2 (the drop-in return convention is 1 - <e^{iP}>)."""
    return 1.0 - 0.5 * (p4[0] + p4[2])


def stream_rate(pattern: int, n: int, reps: int = 20) -> float:
    """GB/s of the streaming reference moving 36 n bytes (u1_selftest_stream)."""
    out = C.c_double(0)
    call("u1_selftest_stream", pattern, n, reps, C.byref(out))
    return out.value
