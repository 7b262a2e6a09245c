"""ctypes binding of the CPU oracle (lbm_oracle.h) -- test infrastructure only."""
from __future__ import annotations

import ctypes
import subprocess
from ctypes import POINTER, byref, c_double, c_int, c_int64, c_uint8, c_void_p
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liblbm_oracle.so"

NOSLIP, MOVINGWALL, FREESLIP, OUTFLOW, INFLOW, PRESSURE = 1, 2, 3, 4, 5, 6
BC_KINDS = {"noslip": NOSLIP, "movingwall": MOVINGWALL, "freeslip": FREESLIP,
            "outflow": OUTFLOW, "inflow": INFLOW, "pressure": PRESSURE}

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def load() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        lib = ctypes.CDLL(str(LIB_PATH))
        dp = POINTER(c_double)
        lib.ora_create.restype = c_void_p
        lib.ora_create.argtypes = [c_int, c_int64, c_int64, c_int64, c_double, POINTER(c_int)]
        lib.ora_destroy.argtypes = [c_void_p]
        lib.ora_lattice.argtypes = [c_int, POINTER(c_int), dp]
        lib.ora_cs.restype = c_double
        lib.ora_set_boundary_box.restype = c_int
        lib.ora_set_boundary_box.argtypes = [c_void_p, c_int, dp] + [c_int64] * 6
        lib.ora_set_solid_mask.restype = c_int
        lib.ora_set_solid_mask.argtypes = [c_void_p, POINTER(c_uint8)]
        lib.ora_finalize.restype = c_int
        lib.ora_finalize.argtypes = [c_void_p]
        lib.ora_set_brinkman.restype = c_int
        lib.ora_set_brinkman.argtypes = [c_void_p, dp, dp]
        for n in ("ora_stream", "ora_swap", "ora_collide"):
            getattr(lib, n).argtypes = [c_void_p]
        lib.ora_step.argtypes = [c_void_p, c_int64]
        lib.ora_get_pdf.argtypes = [c_void_p, dp]
        lib.ora_set_pdf.argtypes = [c_void_p, dp]
        lib.ora_moments.argtypes = [c_void_p, dp, dp]
        lib.ora_conservation.argtypes = [c_void_p, dp]
        lib.ora_get_kinds.argtypes = [c_void_p, POINTER(c_uint8)]
        lib.ora_init_equilibrium.argtypes = [c_void_p, dp, dp]
        lib.ora_num_threads.restype = c_int
        _lib = lib
    return _lib


def lattice(q: int):
    """(c[Q,3] int, w[Q] double) as the oracle holds them."""
    c = np.zeros((q, 3), dtype=np.int32)
    w = np.zeros(q)
    load().ora_lattice(q, c.ctypes.data_as(POINTER(c_int)), w.ctypes.data_as(POINTER(c_double)))
    return c.astype(np.int64), w


def cs() -> float:
    return load().ora_cs()


def _dp(a):
    return a.ctypes.data_as(POINTER(c_double))


class OracleDomain:
    """Reference-structured CPU domain (two AoS fields, handlers, two-pass collide)."""

    def __init__(self, xl, yl, zl, tau, q=19, periodic=(False, False, False)):
        lib = load()
        per = (c_int * 3)(*[1 if p else 0 for p in periodic])
        self._h = lib.ora_create(q, xl, yl, zl, tau, per)
        if not self._h:
            raise ValueError("ora_create failed")
        self.q, self.xl, self.yl, self.zl, self.tau = q, xl, yl, zl, tau
        self.periodic = tuple(bool(p) for p in periodic)

    def close(self):
        if getattr(self, "_h", None):
            load().ora_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def ghost_shape(self):
        return (self.zl + 2, self.yl + 2, self.xl + 2)

    def set_boundary_box(self, kind, params, x0, xe, y0, ye, z0, ze):
        if isinstance(kind, str):
            kind = BC_KINDS[kind]
        p = np.ascontiguousarray(params if params is not None else [0, 0, 0, 1], dtype=np.float64)
        rc = load().ora_set_boundary_box(self._h, kind, _dp(p), x0, xe, y0, ye, z0, ze)
        if rc:
            raise ValueError(f"ora_set_boundary_box: {rc}")

    def set_solid_mask(self, fluid_mask):
        m = np.ascontiguousarray(fluid_mask, dtype=np.uint8).reshape(-1)
        rc = load().ora_set_solid_mask(self._h, m.ctypes.data_as(POINTER(c_uint8)))
        if rc:
            raise ValueError(f"ora_set_solid_mask: {rc}")

    def set_brinkman(self, alpha, g=(0.0, 0.0, 0.0)):
        gv = np.ascontiguousarray(g, dtype=np.float64)
        if alpha is None:
            load().ora_set_brinkman(self._h, None, _dp(gv))
            return
        a = np.ascontiguousarray(alpha, dtype=np.float64).reshape(-1)
        rc = load().ora_set_brinkman(self._h, _dp(a), _dp(gv))
        if rc:
            raise ValueError(f"ora_set_brinkman: {rc}")

    def finalize(self):
        rc = load().ora_finalize(self._h)
        if rc:
            raise ValueError(f"ora_finalize: {rc}")

    def step(self, n=1):
        load().ora_step(self._h, n)

    def stream(self):
        load().ora_stream(self._h)

    def swap(self):
        load().ora_swap(self._h)

    def collide(self):
        load().ora_collide(self._h)

    def get_pdf(self):
        out = np.empty(self.ghost_shape + (self.q,))
        load().ora_get_pdf(self._h, _dp(out))
        return out

    def set_pdf(self, f):
        a = np.ascontiguousarray(f, dtype=np.float64)
        assert a.size == np.prod(self.ghost_shape) * self.q
        load().ora_set_pdf(self._h, _dp(a))

    def init_equilibrium(self, rho, u):
        r = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1)
        v = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
        assert r.size == self.xl * self.yl * self.zl and v.size == 3 * r.size
        load().ora_init_equilibrium(self._h, _dp(r), _dp(v))

    def moments(self):
        rho = np.empty((self.zl, self.yl, self.xl))
        u = np.empty((self.zl, self.yl, self.xl, 3))
        load().ora_moments(self._h, _dp(rho), _dp(u))
        return rho, u

    def conservation(self):
        out = np.zeros(4)
        load().ora_conservation(self._h, _dp(out))
        return out

    def kinds(self):
        out = np.empty(self.ghost_shape, dtype=np.uint8)
        load().ora_get_kinds(self._h, out.ctypes.data_as(POINTER(c_uint8)))
        return out


def num_threads() -> int:
    return load().ora_num_threads()
