"""Pins the oracle's implicit-path transfers (Particle2Node / Node2Particle of
module/solid/implicit/var_trans_solid_implicit.cpp:20-167 with the Newmark
helpers of module/nonlinear_time_integration.cpp) by what the mathematics
fixes, not by restating them:

* the node update from a displacement increment satisfies the Newmark-beta
  relations u = dt v_n + dt^2/2 ((1 - 2 beta) a_n + 2 beta a_n+1) and
  v_n+1 = v_n + dt ((1 - gamma) a_n + gamma a_n+1) for random fields (a wrong
  coefficient, sign or dropped term fails);
* P2G conserves mass, momentum and sum m a, and without BCs a uniform
  particle velocity and acceleration come back at every node with mass;
* G2P of a uniform displacement translates every particle by it exactly,
  leaves F unchanged and gives the uniform Newmark acceleration;
* G2P of a linear displacement u = A x (non-symmetric A) gives dF = I + A and
  particle displacement A x_p, for linear and quadratic B-spline shape
  functions (a transposed index fails);
* rigid particles are neither moved nor accelerated; FLIP velocities follow
  the Newmark velocity update.
CPU only."""
import numpy as np
import pytest

import pyoracle
from mpm_mi355x import capi, scenes

A = np.array([[0.3, -1.2, 0.5], [2.0, 0.1, -0.7], [-0.4, 0.9, -0.2]]) * 1e-3
GAMMA, BETA = 0.6, 0.3025  # generalised-alpha values for spectral radius 0.818


def filled_scene(idimc=(1, 1, 1), scheme=capi.SCHEME_PIC, n=2500, seed=8, materials=None):
    xmin, xmax, nelem = (0.0, -0.5, 1.0), (1.4, 0.5, 2.2), (7, 5, 6)
    rng = np.random.default_rng(seed)
    lo, hi = np.array(xmin), np.array(xmax)
    x = lo + rng.random((n, 3)) * (hi - lo) * 0.999999
    parts = capi.Particles(n, coord=x, vel=rng.standard_normal((n, 3)), mass=0.5 + rng.random(n),
                           vol0=np.full(n, 1e-3))
    mats = materials or [(capi.MODEL_LINELAST, [1e5, 0.3])]
    return scenes.Scene("implicit", xmin, xmax, nelem, mats, 1000.0, 1e-3, 1, (0.0, 0.0, 0.0), scheme=scheme,
                        parts=parts, idimc=idimc)


def node_xyz(o):
    st = o.st
    nx, ny = st.nxc[0], st.nxc[1]
    ids = np.arange(o.nodec)
    i, j, k = ids % nx, (ids // nx) % ny, ids // (nx * ny)
    return np.stack([np.array([o.node_coord(0, a) for a in i]), np.array([o.node_coord(1, a) for a in j]),
                     np.array([o.node_coord(2, a) for a in k])], 1)


def blocks(v):
    """(nodec, 3) -> [u|v|w] blocks"""
    return np.ascontiguousarray(v.T).reshape(-1)


def test_newmark_node_update_satisfies_newmark_relations():
    s = filled_scene()
    o = pyoracle.Oracle(s)
    assert o.bin() == 0
    dt = s.dt
    rng = np.random.default_rng(1)
    nn = o.nodec
    v0, a0, u = rng.standard_normal(3 * nn), rng.standard_normal(3 * nn), 1e-3 * rng.standard_normal(3 * nn)
    o.nvel[:] = v0
    o.naccel[:] = a0
    nb = pyoracle.newmark_para(GAMMA, BETA, dt)
    accel = np.zeros((s.n, 3))
    assert o.implicit_g2p(accel, u, nb, GAMMA) == 0
    v1, a1 = o.nvel.copy(), o.naccel.copy()
    np.testing.assert_allclose(dt * v0 + 0.5 * dt * dt * ((1 - 2 * BETA) * a0 + 2 * BETA * a1), u, rtol=0, atol=1e-12)
    np.testing.assert_allclose(v0 + dt * ((1 - GAMMA) * a0 + GAMMA * a1), v1, rtol=0, atol=1e-9)


def test_implicit_p2g_conserves_and_reproduces_uniform_fields():
    s = filled_scene()
    s.bcs = {}
    o = pyoracle.Oracle(s)
    assert o.bin() == 0
    m = o.parts.mass
    rng = np.random.default_rng(2)
    accel = rng.standard_normal((s.n, 3))
    o.implicit_p2g(accel)
    nn = o.nodec
    np.testing.assert_allclose(o.nmass.sum(), m.sum(), rtol=1e-13)
    np.testing.assert_allclose(o.nvof.sum(), o.parts.vol.sum(), rtol=1e-13)
    np.testing.assert_allclose(o.nmome.reshape(3, nn).sum(1), (m[:, None] * o.parts.vel).sum(0), rtol=1e-12,
                               atol=1e-12)
    np.testing.assert_allclose(o.nforce.reshape(3, nn).sum(1), (m[:, None] * accel).sum(0), rtol=1e-12, atol=1e-12)
    # uniform particle fields come back at every node with mass
    vu, au = np.array([0.3, -2.0, 1.1]), np.array([-9.81, 0.5, 0.25])
    o.parts.vel[:] = vu
    o.implicit_p2g(np.tile(au, (s.n, 1)))
    on = o.nmass > s.mtol
    np.testing.assert_allclose(o.nvel.reshape(3, nn)[:, on].T, np.broadcast_to(vu, (on.sum(), 3)), rtol=1e-13)
    np.testing.assert_allclose(o.naccel.reshape(3, nn)[:, on].T, np.broadcast_to(au, (on.sum(), 3)), rtol=1e-13)
    assert np.all(o.nvel.reshape(3, nn)[:, ~on] == 0.0)


def test_implicit_p2g_applies_bcs():
    s = filled_scene()
    o = pyoracle.Oracle(s)
    nn = o.nodec
    floor = np.arange(0, nn, 7, dtype=np.int32)
    s.bcs = {1: (floor, np.full(len(floor), 0.25)), 3: (floor[:5], np.zeros(5))}
    o = pyoracle.Oracle(s)
    assert o.bin() == 0
    o.implicit_p2g(np.ones((s.n, 3)), facl=0.5)
    np.testing.assert_array_equal(o.nvel[nn + floor[5:]], 0.125)     # v list: fbc * facl
    np.testing.assert_array_equal(o.nvel[nn + floor[:5]], 0.0)        # rigid list wins, all components
    np.testing.assert_array_equal(o.nvel[floor[:5]], 0.0)
    np.testing.assert_array_equal(o.naccel[nn + floor], 0.0)          # acceleration zeroed
    np.testing.assert_array_equal(o.naccel[floor[:5]], 0.0)


@pytest.mark.parametrize("scheme", [capi.SCHEME_PIC, capi.SCHEME_FLIP])
def test_implicit_g2p_uniform_displacement(scheme):
    s = filled_scene(scheme=scheme)
    o = pyoracle.Oracle(s)
    assert o.bin() == 0
    x0, F0, v0 = o.parts.coord.copy(), o.parts.def_grad.copy(), o.parts.vel.copy()
    u0 = np.array([2e-4, -1e-4, 3e-4])
    nn = o.nodec
    dt = s.dt
    nb = pyoracle.newmark_para(GAMMA, BETA, dt)
    accel = np.zeros((s.n, 3))
    assert o.implicit_g2p(accel, np.repeat(u0, nn), nb, GAMMA) == 0
    np.testing.assert_allclose(o.parts.coord - x0, np.broadcast_to(u0, x0.shape), rtol=0, atol=1e-15)
    np.testing.assert_allclose(o.parts.def_grad, F0, rtol=0, atol=1e-13)
    np.testing.assert_allclose(o.parts.stress, 0.0, atol=1e-6)
    a_expect = nb[3] * u0                          # node v = a = 0 before the step
    np.testing.assert_allclose(accel, np.broadcast_to(a_expect, accel.shape), rtol=1e-13)
    if scheme == capi.SCHEME_FLIP:
        np.testing.assert_allclose(o.parts.vel, v0 + dt * GAMMA * a_expect, rtol=1e-13, atol=1e-14)
    else:
        np.testing.assert_allclose(o.parts.vel, np.broadcast_to(nb[0] * u0, v0.shape), rtol=1e-13)


@pytest.mark.parametrize("idimc", [(1, 1, 1), (2, 2, 2), (2, 1, 2)])
def test_implicit_g2p_linear_displacement(idimc):
    s = filled_scene(idimc=idimc)
    o = pyoracle.Oracle(s)
    assert o.bin() == 0
    x0 = o.parts.coord.copy()
    F0 = o.parts.def_grad.copy()
    u = node_xyz(o) @ A.T                          # u_a = A x_a
    nb = pyoracle.newmark_para(GAMMA, BETA, s.dt)
    accel = np.zeros((s.n, 3))
    assert o.implicit_g2p(accel, blocks(u), nb, GAMMA) == 0
    dF = np.eye(3) + A
    np.testing.assert_allclose(o.parts.def_grad, dF @ F0, rtol=0, atol=1e-14)
    np.testing.assert_allclose(o.parts.det_def_grad, np.linalg.det(dF), rtol=1e-14)
    np.testing.assert_allclose(o.parts.vol, np.linalg.det(dF) * o.parts.vol0, rtol=1e-14)
    np.testing.assert_allclose(o.parts.coord - x0, x0 @ A.T, rtol=0, atol=1e-15)


def test_implicit_g2p_rigid_particles_stay():
    s = filled_scene(materials=[(capi.MODEL_LINELAST, [1e5, 0.3]), (capi.MODEL_RIGID, [0.0, 0.0])])
    s.parts.matid = (np.arange(s.n) % 3 == 0).astype(np.int32)
    o = pyoracle.Oracle(s)
    assert o.bin() == 0
    x0 = o.parts.coord.copy()
    nb = pyoracle.newmark_para(GAMMA, BETA, s.dt)
    accel = np.ones((s.n, 3))
    assert o.implicit_g2p(accel, np.full(3 * o.nodec, 1e-4), nb, GAMMA) == 0
    rigid = s.parts.matid == 1
    np.testing.assert_array_equal(o.parts.coord[rigid], x0[rigid])
    np.testing.assert_array_equal(accel[rigid], 0.0)
    np.testing.assert_array_equal(o.parts.vel[rigid], 0.0)          # PIC: velocity reset, nothing gathered
    np.testing.assert_allclose(o.parts.coord[~rigid] - x0[~rigid], 1e-4, rtol=1e-11)


def test_implicit_g2p_rejects_affine_schemes():
    s = filled_scene(scheme=capi.SCHEME_APIC)
    o = pyoracle.Oracle(s)
    assert o.bin() == 0
    nb = pyoracle.newmark_para(GAMMA, BETA, s.dt)
    assert o.implicit_g2p(np.zeros((s.n, 3)), np.zeros(3 * o.nodec), nb, GAMMA) == -1
