"""Lattice tables and C_S: pinned by closed forms (exact rationals, isotropy,
inversion symmetry), the reference's index formulas and the C literal.

References: include/model.h:13-139, include/lbmdefinitions.h:47.
"""
from fractions import Fraction

import numpy as np
import pytest

from lbm_amd import capi
from oracle import ora

# Rational weights of the three lattices (numerator over denominator), as the
# reference's k.0/n literals define them (include/model.h:29-34, 69-75, 114-122).
RATIONAL = {15: (72, {0: 16, 1: 8, 3: 1}), 19: (36, {0: 12, 1: 2, 2: 1}), 27: (216, {0: 64, 1: 16, 2: 4, 3: 1})}


def _tables(q):
    c, w, cs = capi.lattice(q)
    return c, w, cs


@pytest.mark.parametrize("q", [15, 19, 27])
def test_weights_are_correctly_rounded_rationals(q):
    c, w, _ = _tables(q)
    den, by_norm = RATIONAL[q]
    for i in range(q):
        k = by_norm[int(np.abs(c[i]).sum())]
        assert w[i] == float(Fraction(k, den)), (q, i)


@pytest.mark.parametrize("q", [15, 19, 27])
def test_isotropy_to_fourth_order(q):
    """Exact rational moments: sum w = 1, sum w c_a c_b = d_ab/3,
    sum w c_a c_b c_c c_d = (d_ab d_cd + d_ac d_bd + d_ad d_bc)/9, odd = 0."""
    c, _, _ = _tables(q)
    den, by_norm = RATIONAL[q]
    w = [Fraction(by_norm[int(np.abs(ci).sum())], den) for ci in c]
    d = lambda a, b: 1 if a == b else 0
    assert sum(w) == 1
    for a in range(3):
        assert sum(wi * ci[a] for wi, ci in zip(w, c)) == 0
        for b in range(3):
            assert sum(wi * ci[a] * ci[b] for wi, ci in zip(w, c)) == Fraction(d(a, b), 3)
            for e in range(3):
                assert sum(wi * ci[a] * ci[b] * ci[e] for wi, ci in zip(w, c)) == 0
                for f in range(3):
                    m4 = sum(wi * ci[a] * ci[b] * ci[e] * ci[f] for wi, ci in zip(w, c))
                    assert m4 == Fraction(d(a, b) * d(e, f) + d(a, e) * d(b, f) + d(a, f) * d(b, e), 9)


@pytest.mark.parametrize("q", [15, 19, 27])
def test_directions_distinct_and_inverse_pairs(q):
    c, _, _ = _tables(q)
    assert len({tuple(ci) for ci in c}) == q
    assert np.abs(c).max() == 1
    for i in range(q):
        assert tuple(c[q - 1 - i]) == tuple(-c[i])  # inv(q) = Q-1-q
    assert tuple(c[q // 2]) == (0, 0, 0)


def _velocity_index(q, u, v, w):
    """The reference's index formulas (include/model.h:41-44, 82-85, 130-133),
    C integer semantics (division truncates toward zero)."""
    if q == 15:
        t = (u + v)
        return 5 * w + 2 * v + u + 7 - ((w + 2) % 2) * int(t / 2)
    if q == 19:
        return (6 + (v + 1) * 3 + u) if w == 0 else ((w + 1) * 7 + (v + 1) * 2 + u)
    return 9 * w + 3 * v + u + 13


@pytest.mark.parametrize("q", [15, 19, 27])
def test_velocity_index_matches_table_order(q):
    c, _, _ = _tables(q)
    for i in range(q):
        assert _velocity_index(q, *c[i]) == i


@pytest.mark.parametrize("q", [15, 19, 27])
def test_oracle_tables_identical(q):
    c, w, _ = _tables(q)
    oc, ow = ora.lattice(q)
    assert np.array_equal(c, oc)
    assert np.array_equal(w.view(np.uint64), ow.view(np.uint64))


def test_cs_literal():
    """C_S = 0.57735026919L narrowed to double (include/lbmdefinitions.h:47).
    The oracle gets it from the C compiler's long double literal; the library
    spells it in hex. Products as measured for the reference's expressions."""
    _, _, cs = _tables(19)
    assert cs == ora.cs() == float.fromhex("0x1.279a745904047p-1")
    assert abs(cs / np.sqrt(1 / 3) - 1 - 6.48e-13) < 1e-14
    assert cs * cs == 0.33333333333376547
    assert 2 * cs * cs == 0.66666666666753094
    assert 2 * cs * cs * cs * cs == 0.22222222222279839


@pytest.mark.parametrize("q,expected", [(15, 1.0), (19, 1.0000000000000002), (27, 0.9999999999999999)])
def test_in_order_weight_sum(q, expected):
    """The initial density of every cell is the in-order fp64 sum of the weights."""
    _, w, _ = _tables(q)
    s = 0.0
    for x in w:
        s += x
    assert s == expected
