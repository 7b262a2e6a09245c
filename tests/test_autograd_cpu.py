"""Coarse-to-fine porosity map of the inverse problem (CPU)."""
import numpy as np

from lbm_amd import autograd


def test_trilinear_rows_partition_unity():
    m = autograd.trilinear_matrix(8, 32)
    assert np.allclose(m.sum(1), 1.0)
    assert (m >= 0).all()


def test_upsample_reproduces_linear_periodic_profile():
    # a coarse cosine is reproduced at fine cell centres up to interpolation error
    nc, nf = 16, 64
    xc = (np.arange(nc) + 0.5) / nc
    xf = (np.arange(nf) + 0.5) / nf
    c = np.cos(2 * np.pi * xc)
    m = autograd.trilinear_matrix(nc, nf)
    assert np.abs(m @ c - np.cos(2 * np.pi * xf)).max() < (2 * np.pi / nc) ** 2


def test_upsample_adjoint_is_transpose():
    rng = np.random.default_rng(0)
    c = rng.standard_normal((4, 5, 6))
    f = rng.standard_normal((8, 10, 12))
    lhs = (autograd.upsample_np(c, f.shape) * f).sum()
    rhs = (c * autograd.upsample_np_adjoint(f, c.shape)).sum()
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
