"""Closed-form checks of the velocity sets (LatticeModel.cpp:51-151, 208-334)."""
import numpy as np
import pytest


@pytest.mark.parametrize("q,rest", [(19, 18), (27, 0)])
def test_lattice_moments(oracle, q, rest):
    c, w, opp = oracle.lattice(q)
    assert abs(w.sum() - 1.0) < 1e-15
    assert np.allclose((w[:, None] * c).sum(0), 0.0, atol=1e-16)
    m2 = np.einsum("d,da,db->ab", w, c, c)
    assert np.allclose(m2, np.eye(3) / 3.0, atol=1e-15)
    # fourth-order isotropy (needed for Navier-Stokes): sum w c_a^2 c_b^2 = 1/9 (a!=b)
    assert abs((w * c[:, 0] ** 2 * c[:, 1] ** 2).sum() - 1.0 / 9.0) < 1e-15
    for d in range(q):
        assert np.array_equal(c[opp[d]], -c[d])
    assert np.array_equal(c[rest], [0, 0, 0])
    assert len({tuple(v) for v in c}) == q
