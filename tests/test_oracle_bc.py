"""Zou-He, bounce-back, collision and the lattice tables of the oracle checked
against the reference's own Boundary::ApplyBc and LatticeModel (built from its
sources) and moment identities."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref", "ref_driver")


def near_eq(rng, q, n, amp=0.01):
    c = None
    import pyoracle
    c, w, _ = pyoracle.lattice(q)
    rho = 1.0 + amp * rng.standard_normal(n)
    u = 0.03 * rng.standard_normal((n, 3))
    cu = u @ c.T
    usq = (u * u).sum(1)[:, None]
    feq = w * rho[:, None] * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * usq)
    return feq * (1 + amp * rng.standard_normal(feq.shape))


@pytest.mark.skipif(not os.path.exists(REF), reason="reference Zou-He not built (needs /root/reference)")
def test_inlet_matches_reference_build(oracle, tmp_path):
    rng = np.random.default_rng(7)
    f = near_eq(rng, 19, 500)
    src = tmp_path / "in.bin"
    dst = tmp_path / "out.bin"
    with open(src, "wb") as fh:
        fh.write(np.int64(len(f)).tobytes())
        fh.write(f.tobytes())
    subprocess.check_call([REF, "zouhe", str(src), str(dst)])
    ref = np.fromfile(dst, np.float64).reshape(-1, 19)
    bc = oracle.make_bc(0, 1, -1, u=(0.0, -0.1, 0.0))
    mine = np.array([oracle.zouhe_node(19, bc, (0.0, -0.1, 0.0), row) for row in f])
    assert np.array_equal(mine.view(np.uint64), ref.view(np.uint64))


@pytest.mark.skipif(not os.path.exists(REF), reason="reference build not present (needs /root/reference)")
def test_lattice_matches_reference_build(oracle, tmp_path):
    """The oracle's D3Q19 tables against the reference's own LatticeModel
    (LatticeModel.cpp:51-151): weights bit for bit, directions and the
    opposite-direction map exactly, and c_s^2 = 1/3."""
    dst = tmp_path / "lattice.bin"
    subprocess.check_call([REF, "lattice", str(dst)])
    raw = dst.read_bytes()
    q = int(np.frombuffer(raw[:4], np.int32)[0])
    rec = np.frombuffer(raw[4:4 + q * 24], dtype=[("w", "<f8"), ("c", "<i4", 3), ("opp", "<i4")])
    cs2 = np.frombuffer(raw[4 + q * 24:], np.float64)[0]
    c, w, opp = oracle.lattice(19)
    assert q == 19
    assert np.array_equal(rec["w"].view(np.uint64), np.asarray(w, np.float64).view(np.uint64))
    assert np.array_equal(rec["c"], c) and np.array_equal(rec["opp"], opp)
    assert cs2 == 1.0 / 3.0


def moments(q, f):
    import pyoracle
    c, _, _ = pyoracle.lattice(q)
    rho = f.sum()
    return rho, (c.T @ f) / rho


FACES = [(a, s) for a in range(3) for s in (1, -1)]


@pytest.mark.parametrize("q", [19, 27])
@pytest.mark.parametrize("axis,sigma", FACES)
def test_velocity_face_imposes_u(oracle, q, axis, sigma):
    rng = np.random.default_rng(axis * 7 + sigma + q)
    c, _, _ = oracle.lattice(q)
    for f in near_eq(rng, q, 20):
        u = 0.04 * rng.standard_normal(3)
        out = oracle.zouhe_node(q, oracle.make_bc(0, axis, sigma, u=u), u, f)
        rho, uu = moments(q, out)
        assert np.allclose(uu, u, atol=2e-15, rtol=0)
        # only the incoming populations change
        changed = np.nonzero(out != f)[0]
        assert all(c[d][axis] == sigma for d in changed)


@pytest.mark.parametrize("q", [19, 27])
@pytest.mark.parametrize("axis,sigma", FACES)
def test_pressure_face_imposes_rho(oracle, q, axis, sigma):
    rng = np.random.default_rng(11 + axis * 3 + sigma + q)
    for f in near_eq(rng, q, 20):
        out = oracle.zouhe_node(q, oracle.make_bc(1, axis, sigma, rho=1.01), (0, 0, 0), f)
        rho, uu = moments(q, out)
        assert abs(rho - 1.01) < 4e-15
        t = [a for a in range(3) if a != axis]
        assert np.allclose(uu[t], 0.0, atol=2e-15)


def test_x_outlet_typo_residual(oracle):
    """Simulation.cpp:1124-1129 uses f10 where f16 belongs in N_xz; the
    x-min outlet then leaves rho u_z = f16 - f10 instead of 0."""
    rng = np.random.default_rng(3)
    for f in near_eq(rng, 19, 50):
        out = oracle.zouhe_node(19, oracle.make_bc(1, 0, 1, rho=1.01, typo=1), (0, 0, 0), f)
        rho, uu = moments(19, out)
        assert abs(rho - 1.01) < 4e-15
        assert abs(uu[1]) < 2e-15
        assert abs(rho * uu[2] - (f[16] - f[10])) < 4e-15


@pytest.mark.parametrize("q", [19, 27])
def test_collision_identities(oracle, q):
    c, w, _ = oracle.lattice(q)
    fnew, rho, u = oracle.collide_node(q, 1.0 / 0.8, w)
    assert np.array_equal(u, [0, 0, 0])
    assert np.allclose(fnew, w, rtol=1e-15, atol=0)
    rng = np.random.default_rng(5)
    for f in near_eq(rng, q, 10):
        rho, uu = moments(q, f)
        # omega = 1 gives f_eq, whose moments are (rho, rho u)
        feq, r1, u1 = oracle.collide_node(q, 1.0, f)
        assert abs(r1 - rho) < 1e-15 and np.allclose(u1, uu, atol=1e-15)
        r2, u2 = moments(q, feq)
        assert abs(r2 - rho) < 1e-14 and np.allclose(u2, uu, atol=1e-14)
        # BGK: f_new - f_eq = (1 - omega)(f - f_eq)
        om = 1.0 / 0.66
        fn, _, _ = oracle.collide_node(q, om, f)
        assert np.allclose(fn - feq, (1 - om) * (f - feq), atol=1e-15)
