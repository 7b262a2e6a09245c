"""ctypes binding of the C ABI in include/lbm_hip.h (liblbm_hip.so).

This is the Python face of the drop-in boundary: the same entry points a C or
C++ caller uses, wrapped in a small `Domain` class that mirrors the reference's
`Domain` lifecycle (construct, setBoundaryCondition..., nullcollide/finalize,
stream/swap/collide steps, cell/moment reads). The library must be built
(`python -m lbm_amd.build`); there is no fallback implementation.
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, byref, c_char_p, c_double, c_int, c_uint8, c_uint64, c_void_p
from pathlib import Path

import numpy as np

import os

LIB_PATH = Path(os.environ.get("LBM_HIP_LIB", Path(__file__).resolve().parent / "liblbm_hip.so"))

STORAGE = {"f64": 0, "f32": 1, "f16": 2, "bf16": 3, "f64s": 4}
STORAGE_BYTES = {"f64": 8, "f32": 4, "f16": 2, "bf16": 2, "f64s": 8}

NOSLIP, MOVINGWALL, FREESLIP, OUTFLOW, INFLOW, PRESSURE = 1, 2, 3, 4, 5, 6
BC_KINDS = {"noslip": NOSLIP, "movingwall": MOVINGWALL, "freeslip": FREESLIP,
            "outflow": OUTFLOW, "inflow": INFLOW, "pressure": PRESSURE}

ERRORS = {-1: "LBM_ERR_ARG", -2: "LBM_ERR_STATE", -3: "LBM_ERR_HIP", -4: "LBM_ERR_ALLOC",
          -5: "LBM_ERR_UNSUPPORTED", -6: "LBM_ERR_GEOMETRY"}


class LbmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Desc(ctypes.Structure):
    _fields_ = [("q", c_int), ("storage", c_int), ("device", c_int), ("periodic", c_int * 3),
                ("xl", c_uint64), ("yl", c_uint64), ("zl", c_uint64), ("tau", c_double),
                ("origin", c_double * 3), ("spacing", c_double * 3), ("layout", c_int)]


class Info(ctypes.Structure):
    _fields_ = [("ncells_interior", c_uint64), ("ncells_fluid", c_uint64),
                ("ncells_near", c_uint64), ("bytes_per_value", c_uint64),
                ("field_bytes", c_uint64), ("device_bytes", c_uint64), ("vec", c_int),
                ("block_x", c_int), ("block_y", c_int), ("nconds", c_int), ("cell_tiles", c_int),
                ("layout", c_int), ("steps_per_pass", c_int), ("z_chunk", c_int)]


# (name, restype, argtypes) of every exported entry point in include/lbm_hip.h
SIGNATURES = [
    ("lbm_last_error", c_char_p, []),
    ("lbm_abi_version", c_int, []),
    ("lbm_lattice", c_int, [c_int, POINTER(c_int), POINTER(c_double), POINTER(c_double)]),
    ("lbm_desc_init", c_int, [POINTER(Desc)]),
    ("lbm_domain_create", c_int, [POINTER(Desc), POINTER(c_void_p)]),
    ("lbm_domain_destroy", c_int, [c_void_p]),
    ("lbm_set_boundary_box", c_int, [c_void_p, c_int, POINTER(c_double)] + [c_uint64] * 6),
    ("lbm_set_solid_mask", c_int, [c_void_p, POINTER(c_uint8)]),
    ("lbm_set_brinkman", c_int, [c_void_p, POINTER(c_double), POINTER(c_double)]),
    ("lbm_finalize", c_int, [c_void_p]),
    ("lbm_step", c_int, [c_void_p, c_uint64]),
    ("lbm_sync", c_int, [c_void_p]),
    ("lbm_get_pdf", c_int, [c_void_p, POINTER(c_double)]),
    ("lbm_set_pdf", c_int, [c_void_p, POINTER(c_double)]),
    ("lbm_init_equilibrium", c_int, [c_void_p, POINTER(c_double), POINTER(c_double)]),
    ("lbm_moments", c_int, [c_void_p, POINTER(c_double), POINTER(c_double)]),
    ("lbm_moments_noalloc", c_int, [c_void_p, POINTER(c_double), POINTER(c_double)]),
    ("lbm_conservation", c_int, [c_void_p, POINTER(c_double)]),
    ("lbm_time_steps", c_int, [c_void_p, c_uint64, POINTER(c_double)]),
    ("lbm_copy_bandwidth", c_int, [c_int, c_uint64, c_int, POINTER(c_double)]),
    ("lbm_copy_planes_bandwidth", c_int, [c_int, c_int, c_uint64, c_int, POINTER(c_double)]),
    ("lbm_adj_create", c_int, [c_void_p, c_void_p, POINTER(c_void_p)]),
    ("lbm_adj_destroy", c_int, [c_void_p]),
    ("lbm_adj_forward", c_int, [c_void_p, POINTER(c_double), POINTER(c_double)]),
    ("lbm_adj_backward", c_int, [c_void_p, c_double, POINTER(c_double), POINTER(c_double)]),
    ("lbm_jvp", c_int, [c_void_p, POINTER(c_double), POINTER(c_double), POINTER(c_double),
                        POINTER(c_double)]),
    ("lbm_adj_timing", c_int, [c_void_p, POINTER(c_double)]),
    ("lbm_batch_create", c_int, [POINTER(c_void_p), c_int, POINTER(c_void_p)]),
    ("lbm_batch_destroy", c_int, [c_void_p]),
    ("lbm_batch_tune", c_int, [c_void_p, c_int, POINTER(c_double)]),
    ("lbm_batch_step", c_int, [c_void_p, c_uint64]),
    ("lbm_batch_sync", c_int, [c_void_p]),
    ("lbm_batch_time_steps", c_int, [c_void_p, c_uint64, POINTER(c_double)]),
    ("lbm_batch_get_info", c_int, [c_void_p, c_void_p]),
    ("lbm_batch_set_config", c_int, [c_void_p, c_int, c_int, c_int, c_int]),
    ("lbm_batch_status", c_int, [c_void_p, POINTER(c_int), c_int]),
    ("lbm_adj_forward_device", c_int, [c_void_p, c_void_p, POINTER(c_double)]),
    ("lbm_adj_backward_device", c_int, [c_void_p, c_double, c_void_p, POINTER(c_double)]),
    ("lbm_batch_adj_create", c_int, [c_void_p, c_uint64, c_uint64, POINTER(c_void_p)]),
    ("lbm_batch_adj_create2", c_int, [c_void_p, c_uint64, c_uint64, c_uint64, POINTER(c_void_p)]),
    ("lbm_batch_adj_levels", c_int, [c_void_p, POINTER(c_uint64)]),
    ("lbm_batch_adj_destroy", c_int, [c_void_p]),
    ("lbm_batch_adj_forward", c_int, [c_void_p, POINTER(c_double)]),
    ("lbm_batch_adj_backward", c_int, [c_void_p, POINTER(c_double)]),
    ("lbm_batch_adj_timing", c_int, [c_void_p, POINTER(c_double)]),
    ("lbm_get_info", c_int, [c_void_p, POINTER(Info)]),
    ("lbm_set_tuning", c_int, [c_void_p, c_int, c_int]),
    ("lbm_set_time_blocking", c_int, [c_void_p, c_int, c_int]),
    ("lbm_set_index_width", c_int, [c_void_p, c_int]),
]

_lib = None


def _bind_torch_runtime() -> None:
    """Loads torch's HIP runtime first, where torch is installed (without
    importing torch). torch ships its own libamdhip64 (soname
    libamdhip64.so.7) and asks for it by the name libamdhip64.so. Loaded
    after this library, which links /opt/rocm's, it would be a second HIP
    runtime in the process, under which torch finds no GPU. Loaded first,
    this library binds to it, torch reuses it when imported, and device
    memory passes between the two (lbm_adj_*_device with torch tensors).
    LBM_HIP_NO_TORCH=1 keeps /opt/rocm's runtime."""
    if os.environ.get("LBM_HIP_NO_TORCH"):
        return
    import importlib.util
    spec = importlib.util.find_spec("torch")
    if spec is None or not spec.origin:
        return
    hip = Path(spec.origin).parent / "lib" / "libamdhip64.so"
    if hip.exists():
        ctypes.CDLL(str(hip), mode=ctypes.RTLD_GLOBAL)


def load() -> ctypes.CDLL:
    """Loads liblbm_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m lbm_amd.build`")
        _bind_torch_runtime()
        lib = ctypes.CDLL(str(LIB_PATH))
        for name, res, args in SIGNATURES:
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def _check(rc: int) -> None:
    if rc != 0:
        raise LbmError(rc, load().lbm_last_error().decode())


def lattice(q: int):
    """(c[Q,3], w[Q], cs) as the library holds them; needs no GPU."""
    c = np.zeros((q, 3), dtype=np.int32)
    w = np.zeros(q)
    cs = c_double()
    _check(load().lbm_lattice(q, c.ctypes.data_as(POINTER(c_int)),
                              w.ctypes.data_as(POINTER(c_double)), byref(cs)))
    return c.astype(np.int64), w, cs.value


def copy_bandwidth(nbytes: int, reps: int = 10, device: int = 0) -> dict:
    """Plain device copy of nbytes (lbm_copy_bandwidth): ms per copy and GB/s."""
    out = np.zeros(2)
    _check(load().lbm_copy_bandwidth(device, nbytes, reps, out.ctypes.data_as(POINTER(c_double))))
    return {"ms": float(out[0]), "gbs": float(out[1]), "bytes": int(nbytes // 16 * 16)}


def copy_planes_bandwidth(q: int, plane_elems: int, reps: int = 10, device: int = 0) -> dict:
    """q-plane fp32 copy (lbm_copy_planes_bandwidth): ms per copy and GB/s."""
    out = np.zeros(2)
    _check(load().lbm_copy_planes_bandwidth(device, q, plane_elems, reps, out.ctypes.data_as(POINTER(c_double))))
    return {"ms": float(out[0]), "gbs": float(out[1]), "q": q, "plane_elems": int(plane_elems)}


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(POINTER(c_double))


def _f64(a, size: int, what: str) -> np.ndarray:
    """Contiguous flat fp64 copy of a, checked to hold exactly `size` values:
    the C ABI reads that many from the pointer."""
    v = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
    if v.size != size:
        raise ValueError(f"{what}: expected {size} values, got {v.size}")
    return v


class Domain:
    """One single-device domain (include/lbm_hip.h: lbm_domain)."""

    def __init__(self, xl: int, yl: int, zl: int, tau: float, q: int = 19, storage: str = "f64",
                 periodic=(False, False, False), device: int = 0,
                 origin=(0.0, 0.0, 0.0), spacing=(1.0, 1.0, 1.0), layout: str = "ab"):
        lib = load()
        d = Desc()
        _check(lib.lbm_desc_init(byref(d)))
        d.q, d.storage, d.device = q, STORAGE[storage], device
        d.layout = {"ab": 0, "aa": 1}[layout]
        for k in range(3):
            d.periodic[k] = 1 if periodic[k] else 0
            d.origin[k] = origin[k]
            d.spacing[k] = spacing[k]
        d.xl, d.yl, d.zl, d.tau = xl, yl, zl, tau
        h = c_void_p()
        _check(lib.lbm_domain_create(byref(d), byref(h)))
        self._h = h
        self.q, self.xl, self.yl, self.zl, self.tau = q, xl, yl, zl, tau
        self.storage = storage
        self.device = device
        self.periodic = tuple(bool(p) for p in periodic)

    # -- lifecycle -----------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_h", None):
            load().lbm_domain_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- geometry ------------------------------------------------------------
    def set_boundary_box(self, kind, params, x0, xe, y0, ye, z0, ze) -> None:
        if isinstance(kind, str):
            kind = BC_KINDS[kind]
        p = np.ascontiguousarray(params if params is not None else [0, 0, 0, 1], dtype=np.float64)
        _check(load().lbm_set_boundary_box(self._h, kind, _dptr(p), x0, xe, y0, ye, z0, ze))

    def set_solid_mask(self, fluid_mask: np.ndarray) -> None:
        m = np.ascontiguousarray(fluid_mask, dtype=np.uint8).reshape(-1)
        if m.size != self.xl * self.yl * self.zl:
            raise ValueError(f"fluid mask: expected {self.xl * self.yl * self.zl} values, got {m.size}")
        _check(load().lbm_set_solid_mask(self._h, m.ctypes.data_as(POINTER(c_uint8))))

    def set_brinkman(self, alpha, g=(0.0, 0.0, 0.0)) -> None:
        gv = _f64(g, 3, "body acceleration")
        if alpha is None:
            _check(load().lbm_set_brinkman(self._h, None, _dptr(gv)))
            return
        a = _f64(alpha, self.xl * self.yl * self.zl, "porosity")
        _check(load().lbm_set_brinkman(self._h, _dptr(a), _dptr(gv)))

    def finalize(self) -> None:
        _check(load().lbm_finalize(self._h))

    # -- time stepping -------------------------------------------------------
    def step(self, n: int = 1) -> None:
        _check(load().lbm_step(self._h, n))

    def sync(self) -> None:
        _check(load().lbm_sync(self._h))

    def time_steps(self, n: int) -> dict:
        out = np.zeros(4)
        _check(load().lbm_time_steps(self._h, n, _dptr(out)))
        return {"total_ms": out[0], "bulk_ms": out[1], "near_ms": out[2], "steps": int(out[3])}

    # -- fields --------------------------------------------------------------
    @property
    def ghost_shape(self):
        return (self.zl + 2, self.yl + 2, self.xl + 2)

    def get_pdf(self) -> np.ndarray:
        """Ghosted collide field as (zl+2, yl+2, xl+2, Q) doubles."""
        out = np.empty(self.ghost_shape + (self.q,), dtype=np.float64)
        _check(load().lbm_get_pdf(self._h, _dptr(out)))
        return out

    def set_pdf(self, f: np.ndarray) -> None:
        a = _f64(f, int(np.prod(self.ghost_shape)) * self.q, "pdf field")
        _check(load().lbm_set_pdf(self._h, _dptr(a)))

    def init_equilibrium(self, rho: np.ndarray, u: np.ndarray) -> None:
        n = self.xl * self.yl * self.zl
        r = _f64(rho, n, "density")
        v = _f64(u, 3 * n, "velocity")
        _check(load().lbm_init_equilibrium(self._h, _dptr(r), _dptr(v)))

    def moments(self):
        """(rho, u): interior (zl, yl, xl) and (zl, yl, xl, 3) doubles."""
        rho = np.empty((self.zl, self.yl, self.xl), dtype=np.float64)
        u = np.empty((self.zl, self.yl, self.xl, 3), dtype=np.float64)
        _check(load().lbm_moments(self._h, _dptr(rho), _dptr(u)))
        return rho, u

    def moments_noalloc(self):
        """moments() through lbm_moments_noalloc (fixed per-GPU staging, no
        device allocation; at most 2^20 interior cells)."""
        rho = np.empty((self.zl, self.yl, self.xl), dtype=np.float64)
        u = np.empty((self.zl, self.yl, self.xl, 3), dtype=np.float64)
        _check(load().lbm_moments_noalloc(self._h, _dptr(rho), _dptr(u)))
        return rho, u

    def conservation(self) -> np.ndarray:
        out = np.zeros(4)
        _check(load().lbm_conservation(self._h, _dptr(out)))
        return out

    def info(self) -> dict:
        i = Info()
        _check(load().lbm_get_info(self._h, byref(i)))
        return {k: getattr(i, k) for k, _ in Info._fields_}

    def set_tuning(self, vec: int = 0, block_x: int = 0) -> None:
        _check(load().lbm_set_tuning(self._h, vec, block_x))

    def set_index_width(self, bits: int = 32) -> None:
        """32: 32-bit byte offsets where a plane is under 4 GiB (default);
        64: 64-bit element indices in the cell-tile and in-place kernels."""
        _check(load().lbm_set_index_width(self._h, bits))

    def set_time_blocking(self, steps_per_pass: int = 0, z_chunk: int = 0) -> None:
        """0 (default): two steps per pass where that measured faster
        (bulk-only domains in fp32/fp16/bf16; fp32 D3Q15/D3Q19 domains with
        boundary cells); 2: two steps per pass wherever it applies (also fp32
        D3Q27 domains with boundary cells); 1: single steps."""
        _check(load().lbm_set_time_blocking(self._h, steps_per_pass, z_chunk))


class AdjDesc(ctypes.Structure):
    _fields_ = [("objective", c_int), ("plane_x", c_int), ("target_u", POINTER(c_double)),
                ("nsteps", c_uint64), ("checkpoint_every", c_uint64)]


OBJ_PLANE_VELOCITY_MISFIT, OBJ_KINETIC_ENERGY = 1, 2


class Tape:
    """Differentiation tape over a finalized Domain (lbm_adj_* / lbm_jvp).

    forward(alpha) -> J, backward(dJ) -> (grad_alpha, grad_wall), jvp(alpha,
    dalpha) -> (J, dJ). alpha / grads are interior arrays (zl, yl, xl)."""

    def __init__(self, dom: Domain, nsteps: int, objective: int = OBJ_KINETIC_ENERGY,
                 plane_x: int = 0, target_u=None, checkpoint_every: int = 0):
        self.dom = dom
        self._target = None
        d = AdjDesc()
        d.objective, d.plane_x, d.nsteps, d.checkpoint_every = objective, plane_x, nsteps, checkpoint_every
        if target_u is not None:
            self._target = _f64(target_u, dom.yl * dom.zl * 3, "target velocity plane")
            d.target_u = _dptr(self._target)
        h = c_void_p()
        _check(load().lbm_adj_create(dom._h, byref(d), byref(h)))
        self._h = h
        self.nsteps = nsteps

    def close(self):
        if getattr(self, "_h", None):
            load().lbm_adj_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _alpha(self, alpha):
        if alpha is None:
            return None, None
        a = _f64(alpha, self.dom.xl * self.dom.yl * self.dom.zl, "porosity")
        return a, _dptr(a)

    def forward(self, alpha=None) -> float:
        a, p = self._alpha(alpha)
        j = c_double()
        _check(load().lbm_adj_forward(self._h, p, byref(j)))
        return j.value

    def backward(self, dJ: float = 1.0, want_alpha: bool = True):
        shape = (self.dom.zl, self.dom.yl, self.dom.xl)
        g = np.zeros(shape) if want_alpha else None
        gw = np.zeros(3)
        _check(load().lbm_adj_backward(self._h, dJ, _dptr(g) if want_alpha else None, _dptr(gw)))
        return g, gw

    def forward_device(self, alpha) -> float:
        """forward() with a porosity torch tensor on the domain's GPU (float64,
        (zl, yl, xl), contiguous); no host copy."""
        import torch
        self._check_device_tensor(alpha, "porosity")
        torch.cuda.current_stream(alpha.device).synchronize()  # the library reads it on its own stream
        j = c_double()
        _check(load().lbm_adj_forward_device(self._h, c_void_p(alpha.data_ptr()), byref(j)))
        return j.value

    def backward_device(self, dJ: float = 1.0):
        """backward() with the porosity gradient returned as a torch tensor on
        the domain's GPU; (grad_alpha, grad_wall)."""
        import torch
        g = torch.empty((self.dom.zl, self.dom.yl, self.dom.xl), dtype=torch.float64,
                        device=f"cuda:{self.dom.device}")
        gw = np.zeros(3)
        _check(load().lbm_adj_backward_device(self._h, dJ, c_void_p(g.data_ptr()), _dptr(gw)))
        return g, gw

    def _check_device_tensor(self, t, what):
        n = self.dom.xl * self.dom.yl * self.dom.zl
        if not (getattr(t, "is_cuda", False) and str(t.dtype) == "torch.float64" and t.is_contiguous()
                and t.numel() == n and t.device.index == self.dom.device):
            raise ValueError(f"{what}: expected a contiguous float64 tensor of {n} values on cuda:{self.dom.device}")

    def jvp(self, alpha, dalpha):
        a, p = self._alpha(alpha)
        da = _f64(dalpha, self.dom.xl * self.dom.yl * self.dom.zl, "porosity direction")
        j, dj = c_double(), c_double()
        _check(load().lbm_jvp(self._h, p, _dptr(da), byref(j), byref(dj)))
        return j.value, dj.value

    def timing(self) -> dict:
        out = np.zeros(4)
        _check(load().lbm_adj_timing(self._h, _dptr(out)))
        return {"recompute_ms": out[0], "adjoint_ms": out[1], "forward_ms": out[2], "steps": int(out[3])}


BATCH_ROWS_SEPARATE, BATCH_ROWS_FUSED, BATCH_CELLS = 0, 1, 2


class BatchInfo(ctypes.Structure):
    _fields_ = [("ndomains", c_uint64), ("ncells", c_uint64), ("ntiles", c_uint64),
                ("nnear", c_uint64), ("vec", c_int), ("blocks_per_cu", c_int),
                ("bx_policy", c_int), ("ncu", c_int), ("kernel", c_int)]


class Batch:
    """Batched stepping of finalized domains (lbm_batch_*). The domains stay
    owned by the caller (kept alive here)."""

    def __init__(self, domains):
        self.domains = list(domains)
        arr = (c_void_p * len(self.domains))(*[d._h.value for d in self.domains])
        h = c_void_p()
        _check(load().lbm_batch_create(arr, len(self.domains), byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            load().lbm_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def tune(self, trial_steps: int = 5) -> float:
        ms = c_double()
        _check(load().lbm_batch_tune(self._h, trial_steps, byref(ms)))
        return ms.value

    def step(self, n: int) -> None:
        _check(load().lbm_batch_step(self._h, n))

    def sync(self) -> None:
        _check(load().lbm_batch_sync(self._h))

    def time_steps(self, n: int) -> dict:
        out = np.zeros(2)
        _check(load().lbm_batch_time_steps(self._h, n, _dptr(out)))
        return {"ms": float(out[0]), "cells": int(out[1]), "mlups": out[1] * n / (out[0] * 1e3)}

    def set_config(self, kernel: int = 2, vec: int = 1, bx_policy: int = 0, blocks_per_cu: int = 8) -> None:
        """kernel: BATCH_ROWS_SEPARATE, BATCH_ROWS_FUSED or BATCH_CELLS."""
        _check(load().lbm_batch_set_config(self._h, kernel, vec, bx_policy, blocks_per_cu))

    def status(self, reset: bool = False) -> np.ndarray:
        """Per-domain divergence flags (bool array in domain order): True if a
        batched step read a NaN or infinite value in that domain's fluid
        cells since creation or the last reset."""
        out = np.zeros(len(self.domains), dtype=np.int32)
        _check(load().lbm_batch_status(self._h, out.ctypes.data_as(POINTER(c_int)), 1 if reset else 0))
        return out != 0

    def info(self) -> dict:
        i = BatchInfo()
        _check(load().lbm_batch_get_info(self._h, ctypes.cast(byref(i), c_void_p)))
        return {k: getattr(i, k) for k, _ in BatchInfo._fields_}


class BatchTape:
    """Batched adjoint over a Batch (lbm_batch_adj_*): forward() -> J per
    domain, backward() -> (ndomains, 3) lid-velocity gradients."""

    def __init__(self, batch: Batch, nsteps: int, checkpoint_every: int = 0, sub_every: int = 0):
        """checkpoint_every = 0: automatic (sqrt(N), or two levels where
        those slots do not fit); sub_every > 0: two levels, checkpoints every
        checkpoint_every steps and, inside each segment, every sub_every."""
        self.batch = batch
        h = c_void_p()
        if sub_every:
            _check(load().lbm_batch_adj_create2(batch._h, nsteps, checkpoint_every, sub_every, byref(h)))
        else:
            _check(load().lbm_batch_adj_create(batch._h, nsteps, checkpoint_every, byref(h)))
        self._h = h
        self.n = len(batch.domains)

    def levels(self) -> dict:
        out = (ctypes.c_uint64 * 3)()
        _check(load().lbm_batch_adj_levels(self._h, out))
        return {"checkpoint_every": int(out[0]), "sub_every": int(out[1]), "slots_per_domain": int(out[2])}

    def close(self):
        if getattr(self, "_h", None):
            load().lbm_batch_adj_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def forward(self) -> np.ndarray:
        j = np.zeros(self.n)
        _check(load().lbm_batch_adj_forward(self._h, _dptr(j)))
        return j

    def backward(self) -> np.ndarray:
        g = np.zeros((self.n, 3))
        _check(load().lbm_batch_adj_backward(self._h, _dptr(g)))
        return g

    def timing(self) -> dict:
        out = np.zeros(4)
        _check(load().lbm_batch_adj_timing(self._h, _dptr(out)))
        return {"recompute_ms": out[0], "adjoint_ms": out[1], "forward_ms": out[2], "steps": int(out[3])}
