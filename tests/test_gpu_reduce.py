"""lbm_reduce: the FP64 mass and per-condition flux sums equal, bit for bit,
the same sums formed in numpy in the device's documented fixed order
(tests/fixed_order.py) from the moments the library reads back, and repeated
calls give the same bits.  Two pipes: 11k nodes (one thread per node) and
332k nodes (more than 1024 x 256, so every thread adds several)."""
import numpy as np
import pytest

from fixed_order import block_sum, node_sum
from lbm_cases import pipe_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import pv2013_amd as P
    c = P.Context(0)
    yield c
    c.close()


def expected(flags, bcs, rho, u):
    fluid = flags > 0
    mass = node_sum(rho[fluid])
    flux = [block_sum(rho[flags == 3 + b] * u[bc["axis"]][flags == 3 + b]) for b, bc in enumerate(bcs)]
    return mass, flux


@pytest.mark.parametrize("dims, radius", [((24, 24, 40), 9.5), ((50, 50, 200), 23.0)])
@pytest.mark.parametrize("storage", [0, 1])
def test_reduce_matches_fixed_order(ctx, oracle, dims, radius, storage):
    import pv2013_amd as P
    cl, bcs = pipe_case(oracle, dims, radius, 0.04)
    dom = P.Domain(ctx, dims, cl, 0.8, [P.make_bc(**b) for b in bcs], storage=storage, scheme=P.SCHEME_AA,
                   fast_axis=0)  # node order = site order
    print("reduce: pipe", dims, "storage", storage, "nodes", dom.stats()["nodes"], flush=True)
    dom.step(30)
    rho, u = dom.macros()
    m, flux = dom.reduce()
    m2, flux2 = dom.reduce()
    em, eflux = expected(cl, bcs, rho, u)
    assert np.float64(m).tobytes() == np.float64(em).tobytes()
    assert np.array_equal(np.asarray(flux).view(np.uint64), np.asarray(eflux).view(np.uint64))
    assert m == m2 and np.array_equal(flux, flux2)
    assert flux[0] > 0  # the inlet drives mass in along +z
    dom.close()
