"""Closed forms used by the tests: textbook second-order equilibrium with
c_s^2 = 1/3, moments, and seeded synthetic inputs. Independent of the oracle
and of the library (which use the reference's C_S literal); tolerances in the
tests allow for the 1.3e-12 relative difference that literal makes."""
from __future__ import annotations

import numpy as np


def feq_textbook(c: np.ndarray, w: np.ndarray, rho, u) -> np.ndarray:
    """w_q rho (1 + 3 c.u + 4.5 (c.u)^2 - 1.5 u.u); rho (...), u (..., 3) -> (..., Q)."""
    rho = np.asarray(rho, dtype=np.float64)
    u = np.asarray(u, dtype=np.float64)
    cu = np.einsum("qd,...d->...q", c.astype(np.float64), u)
    uu = np.einsum("...d,...d->...", u, u)[..., None]
    return w * rho[..., None] * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * uu)


def moments(c: np.ndarray, f: np.ndarray):
    """(rho, u) of pdfs f (..., Q)."""
    rho = f.sum(-1)
    j = np.einsum("...q,qd->...d", f, c.astype(np.float64))
    return rho, j / rho[..., None]


def second_moment(c: np.ndarray, f: np.ndarray) -> np.ndarray:
    cc = np.einsum("qa,qb->qab", c.astype(np.float64), c.astype(np.float64))
    return np.einsum("...q,qab->...ab", f, cc)


def tgv_fields(xl, yl, zl, u0=0.02, k=None, rho_noise=1e-4, seed=12345):
    """2-D Taylor-Green vortex in x-y, uniform in z, plus a seeded density
    perturbation rho = 1 + rho_noise * N(0,1). Cell centres at x = 0..xl-1."""
    if k is None:
        k = 2 * np.pi / xl
    x = np.arange(xl)[None, None, :]
    y = np.arange(yl)[None, :, None]
    ux = u0 * np.sin(k * x) * np.cos(k * y) * np.ones((zl, 1, 1))
    uy = -u0 * np.cos(k * x) * np.sin(k * y) * np.ones((zl, 1, 1))
    u = np.stack([ux, uy, np.zeros_like(ux)], axis=-1)
    rng = np.random.default_rng(seed)
    rho = 1.0 + rho_noise * rng.standard_normal((zl, yl, xl))
    return rho, u
