"""The OpenMP CPU-baseline port (oracle/mpm_oracle_omp.c) runs the same step
as the oracle: equal to rounding (its node sums go colour by colour), the
same bits for any thread count (each node takes one cell per colour), and it
refuses what it does not cover. CPU only."""
import numpy as np
import pytest

import pyoracle
from mpm_mi355x import capi, scenes


def runs(s, steps, threads):
    o = pyoracle.Oracle(s)
    assert o.run(1, steps) == 0
    p = pyoracle.Oracle(s, fast="omp")
    assert p.run_omp(1, steps, threads) == 0
    return o, p


@pytest.mark.parametrize("scheme", [capi.SCHEME_PIC, capi.SCHEME_FLIP])
@pytest.mark.parametrize("idimc", [(1, 1, 1), (2, 1, 2)])
def test_omp_port_matches_the_oracle(scheme, idimc):
    s = scenes.tiny(n_cells=8, scheme=scheme, idimc=idimc)
    o, p = runs(s, 40, 4)
    np.testing.assert_allclose(p.parts.coord, o.parts.coord, rtol=0, atol=1e-13)
    np.testing.assert_allclose(p.parts.vel, o.parts.vel, rtol=0, atol=1e-10)
    np.testing.assert_allclose(p.parts.stress, o.parts.stress, rtol=0, atol=1e-8 * np.abs(o.parts.stress).max())


def test_omp_port_is_independent_of_the_thread_count():
    s = scenes.tiny(n_cells=8)
    a = pyoracle.Oracle(s, fast="omp")
    b = pyoracle.Oracle(s, fast="omp")
    assert a.run_omp(1, 20, 1) == 0 and b.run_omp(1, 20, 5) == 0
    for f in ("coord", "vel", "stress", "def_grad"):
        assert np.array_equal(getattr(a.parts, f), getattr(b.parts, f)), f


def test_omp_port_refuses_what_it_does_not_cover():
    s = scenes.tiny_j2(n_cells=6)
    assert pyoracle.Oracle(s, fast="omp").run_omp(1, 2, 2) == -1000
    s = scenes.tiny(n_cells=6, scheme=capi.SCHEME_APIC)
    assert pyoracle.Oracle(s, fast="omp").run_omp(1, 2, 2) == -1000
