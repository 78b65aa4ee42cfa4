"""Pins the oracle's control-point coordinates and its TPIC / APIC transfers
(module/mesh.h:32-47, module/map_and_interpolate.h:57-176) by what the
mathematics fixes:

* control points sit where both shape functions reproduce linear fields,
  sum_a N_a(x) x_a = x and sum_a x_a (x) grad N_a = I, in every cell, the
  quadratic B-spline boundary cells included;
* G2P of a nodal field v = A x + b gives the TPIC velocity gradient A and the
  APIC product B D^-1 = A, for a non-symmetric A (a transposed index or a
  sign fails);
* P2G of particles carrying v_p = A x_p + b with C_p = A gives nodal
  velocities A x_a + b at every node that has mass;
* the affine term carries no net momentum: sum_a p_a = sum_p m_p v_p.
CPU only."""
import ctypes as C

import numpy as np
import pytest

import pyoracle
from mpm_mi355x import capi, scenes

A = np.array([[0.3, -1.2, 0.5], [2.0, 0.1, -0.7], [-0.4, 0.9, -0.2]])
B = np.array([0.25, -0.5, 1.5])
IDIMC = [(1, 1, 1), (2, 2, 2), (2, 1, 2)]


def filled_scene(idimc, scheme, n=3000, seed=5):
    xmin, xmax, nelem = (0.0, -0.5, 1.0), (1.4, 0.5, 2.2), (7, 5, 6)
    rng = np.random.default_rng(seed)
    lo, hi = np.array(xmin), np.array(xmax)
    x = lo + rng.random((n, 3)) * (hi - lo) * 0.999999
    x[:40] = lo + (hi - lo) * rng.integers(0, 2, (40, 3)) * 0.999999  # corners: boundary cells
    parts = capi.Particles(n, coord=x, mass=0.5 + rng.random(n), vol0=np.full(n, 1e-3))
    return scenes.Scene("affine", xmin, xmax, nelem, [(capi.MODEL_LINELAST, [1e5, 0.3])], 1000.0, 1e-5, 1,
                        (0.0, 0.0, 0.0), scheme=scheme, parts=parts, idimc=idimc)


def node_xyz(o):
    st = o.st
    nx, ny, nz = st.nxc[0], st.nxc[1], st.nxc[2]
    ids = np.arange(o.nodec)
    i, j, k = ids % nx, (ids // nx) % ny, ids // (nx * ny)
    return np.stack([np.array([o.node_coord(0, a) for a in i]), np.array([o.node_coord(1, a) for a in j]),
                     np.array([o.node_coord(2, a) for a in k])], 1)


@pytest.mark.parametrize("idimc", IDIMC)
def test_control_points_reproduce_linear_fields(idimc):
    o = pyoracle.Oracle(filled_scene(idimc, pyoracle.PIC))
    st = o.st
    x = o.parts.coord
    n = len(x)
    cells = np.array([o.cell_of(x[p]) for p in range(n)], np.int32)
    nen = (idimc[0] + 1) * (idimc[1] + 1) * (idimc[2] + 1)
    nc = np.zeros((n, nen), np.int32)
    sf = np.zeros((n, nen))
    dsf = np.zeros((n, nen, 3))
    o.lib.orc_maksf_batch(C.byref(st), n, pyoracle.P(cells, C.c_int32), pyoracle.P(np.ascontiguousarray(x)),
                          pyoracle.P(nc, C.c_int32), pyoracle.P(sf), pyoracle.P(dsf))
    xc = node_xyz(o)[nc]                                   # [n, nen, 3]
    np.testing.assert_allclose(sf.sum(1), 1.0, atol=1e-14)
    np.testing.assert_allclose(np.einsum("pa,pai->pi", sf, xc), x, rtol=0, atol=1e-14)
    np.testing.assert_allclose(np.einsum("pai,paj->pij", xc, dsf), np.broadcast_to(np.eye(3), (n, 3, 3)), atol=1e-12)
    # the quadratic axes' end points are the grid ends
    for a in range(3):
        if idimc[a] == 2:
            assert o.node_coord(a, 0) == st.xmin[a] and o.node_coord(a, st.nxc[a] - 1) == st.xmax[a]


def linear_nodal_field(o):
    return (node_xyz(o) @ A.T + B)


@pytest.mark.parametrize("idimc", IDIMC)
@pytest.mark.parametrize("scheme", [pyoracle.TPIC, pyoracle.APIC])
def test_g2p_recovers_the_gradient_of_a_linear_field(idimc, scheme):
    o = pyoracle.Oracle(filled_scene(idimc, scheme))
    assert o.bin() == 0
    v = linear_nodal_field(o)
    o.nvel[:] = v.T.ravel()
    o.naccel[:] = 0.0
    x0 = o.parts.coord.copy()
    o.g2p()
    np.testing.assert_allclose(o.parts.vel, x0 @ A.T + B, rtol=0, atol=1e-12)
    if scheme == pyoracle.TPIC:
        got = o.cmat
    else:
        # D is singular (or nearly so) where a particle sits on an
        # interpolating node: for linear tents D_ii = (x_p - x_0)(x_1 - x_p),
        # and the open B-spline interpolates at the grid ends. The first 40
        # particles sit on grid corners (the reference has the same property).
        assert np.isfinite(o.dinv[40:]).all()
        got = o.cmat[40:] @ o.dinv[40:]
    np.testing.assert_allclose(got, np.broadcast_to(A, got.shape), rtol=0, atol=1e-9)


@pytest.mark.parametrize("idimc", IDIMC)
@pytest.mark.parametrize("scheme", [pyoracle.TPIC, pyoracle.APIC])
def test_p2g_of_affine_particles_gives_the_linear_field(idimc, scheme):
    o = pyoracle.Oracle(filled_scene(idimc, scheme))
    assert o.bin() == 0
    o.parts.vel[:] = o.parts.coord @ A.T + B
    o.cmat[:] = A
    if scheme == pyoracle.APIC:
        o.dinv[:] = np.eye(3)
    o.p2g()
    has = o.nmass > 1e-12
    got = o.nvel.reshape(3, -1).T[has]
    np.testing.assert_allclose(got, linear_nodal_field(o)[has], rtol=0, atol=1e-11)


@pytest.mark.parametrize("scheme", [pyoracle.TPIC, pyoracle.APIC])
def test_affine_term_carries_no_net_momentum(scheme):
    o = pyoracle.Oracle(filled_scene((2, 1, 2), scheme))
    assert o.bin() == 0
    rng = np.random.default_rng(9)
    o.parts.vel[:] = rng.normal(size=o.parts.vel.shape)
    o.cmat[:] = rng.normal(size=o.cmat.shape) * 10
    o.dinv[:] = rng.normal(size=o.dinv.shape)
    o.p2g()
    tot = o.nmome.reshape(3, -1).sum(1)
    ref = (o.parts.mass[:, None] * o.parts.vel).sum(0)
    np.testing.assert_allclose(tot, ref, rtol=1e-12, atol=1e-10)


@pytest.mark.parametrize("scheme", [pyoracle.TPIC, pyoracle.APIC])
def test_rigid_translation_stays_exact(scheme):
    s = scenes.tiny(n_cells=6, model=capi.MODEL_LINELAST, props=(1e5, 0.3), scheme=scheme)
    s.bb = (0.0, 0.0, 0.0)
    s.bcs = {}
    s.parts.vel[:] = (0.5, -0.25, 0.125)
    o = pyoracle.Oracle(s)
    assert o.run(1, 5) == 0
    np.testing.assert_allclose(o.parts.vel, np.broadcast_to([0.5, -0.25, 0.125], o.parts.vel.shape), rtol=0,
                               atol=1e-14)
    # a uniform field has zero gradient (up to the rounding of the sums)
    assert np.abs(o.cmat).max() < 1e-12
