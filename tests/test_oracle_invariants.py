"""Pins of the FP64 oracle where the reference cannot be compiled
(shape functions, binning, P2G/G2P need MPI headers in the reference tree):
closed forms, invariants, brute force, and an independent vectorised
restatement of the explicit step written in a different style."""
import ctypes as C
import itertools

import numpy as np
import pytest

import pyoracle
from pyoracle import P
from mpm_mi355x import capi, scenes

orc = pyoracle.load()


def state(nelem=(4, 3, 5), xmin=(0.0, -1.0, 0.5), xmax=(2.0, 0.5, 3.0), idimc=(1, 1, 1)):
    st = pyoracle.OrcState()
    for i in range(3):
        st.xmin[i], st.xmax[i], st.nex[i], st.idimc[i] = xmin[i], xmax[i], nelem[i], idimc[i]
    orc.orc_grid_init(C.byref(st))
    return st


def maksf(st, m, x):
    nen = (st.idimc[0] + 1) * (st.idimc[1] + 1) * (st.idimc[2] + 1)
    nc = np.zeros(nen, np.int32)
    sf = np.zeros(nen)
    dsf = np.zeros(3 * nen)
    x = np.ascontiguousarray(x, np.float64)
    k = orc.orc_maksf(C.byref(st), m, P(x), P(nc, C.c_int32), P(sf), P(dsf))
    assert k == nen
    return nc, sf, dsf.reshape(nen, 3)


def node_xyz(st, nid):
    nx, ny = st.nxc[0], st.nxc[1]
    i, j, k = nid % nx, (nid // nx) % ny, nid // (nx * ny)
    return np.array([st.xmin[0] + st.dx[0] * i, st.xmin[1] + st.dx[1] * j, st.xmin[2] + st.dx[2] * k])


def test_maksf_linear_closed_form_and_invariants():
    st = state()
    rng = np.random.default_rng(0)
    for _ in range(300):
        c = [rng.integers(0, st.nex[a]) for a in range(3)]
        m = c[0] + c[1] * st.nex[0] + c[2] * st.nex[0] * st.nex[1]
        x = np.array([st.xmin[a] + st.dx[a] * (c[a] + rng.random()) for a in range(3)])
        nc, sf, dsf = maksf(st, m, x)
        # stencil order k, j, i over the cell's corners
        want_ids = [c[0] + i + st.nxc[0] * (c[1] + j) + st.nxc[0] * st.nxc[1] * (c[2] + k)
                    for k in range(2) for j in range(2) for i in range(2)]
        assert list(nc) == want_ids
        xa = np.array([node_xyz(st, n) for n in nc])
        want = np.prod(1.0 - np.abs(x - xa) / np.array(st.dx[:]), axis=1)
        np.testing.assert_allclose(sf, want, rtol=0, atol=1e-14)
        assert abs(sf.sum() - 1.0) < 1e-14                           # partition of unity
        np.testing.assert_allclose(dsf.sum(axis=0), 0.0, atol=1e-12)  # sum of gradients
        np.testing.assert_allclose((sf[:, None] * xa).sum(axis=0), x, atol=1e-14)  # linear reproduction
        np.testing.assert_allclose((dsf[:, :, None] * xa[:, None, :]).sum(axis=0), np.eye(3), atol=1e-12)


def test_maksf_gradient_sign_rule_at_node():
    """shape_function.cpp:45-50: dN = +1/h for x < x_a, -1/h otherwise (x == x_a
    included). At a particle exactly on its cell's lower corner the lower node
    takes -1/h and the upper +1/h."""
    st = state(nelem=(4, 4, 4), xmin=(0, 0, 0), xmax=(1, 1, 1))
    m = 1 + 4 * 1 + 16 * 1
    x = np.array([0.25, 0.25 + 0.1, 0.25 + 0.2])
    nc, sf, dsf = maksf(st, m, x)
    h = 0.25
    # x component of the gradient on the 1D factor: corners i=0 have x == x_a
    wy = [1 - 0.1 / h, 0.1 / h]
    wz = [1 - 0.2 / h, 0.2 / h]
    for idx, (k, j, i) in enumerate(itertools.product(range(2), range(2), range(2))):
        sx = -1.0 / h if i == 0 else 1.0 / h
        assert dsf[idx, 0] == pytest.approx(sx * wy[j] * wz[k], rel=1e-14)


def test_maksf_quadratic_interior_invariants():
    st = state(nelem=(8, 8, 8), xmin=(0, 0, 0), xmax=(1, 1, 1), idimc=(2, 2, 2))
    rng = np.random.default_rng(3)
    for _ in range(100):
        c = [rng.integers(2, 6) for a in range(3)]  # away from the boundary branches
        m = c[0] + c[1] * 8 + c[2] * 64
        x = np.array([(c[a] + rng.random()) / 8.0 for a in range(3)])
        nc, sf, dsf = maksf(st, m, x)
        assert len(nc) == 27
        assert abs(sf.sum() - 1.0) < 1e-13
        np.testing.assert_allclose(dsf.sum(axis=0), 0.0, atol=1e-11)
        # control points of the open B-spline sit at cell centres: x_a = xmin + h (i - 0.5)
        nx, ny = st.nxc[0], st.nxc[1]
        ids = np.asarray(nc)
        xa = np.stack([(ids % nx - 0.5) / 8.0, ((ids // nx) % ny - 0.5) / 8.0, (ids // (nx * ny) - 0.5) / 8.0], 1)
        np.testing.assert_allclose((sf[:, None] * xa).sum(axis=0), x, atol=1e-13)


def test_maksf_quadratic_boundary_partition_of_unity():
    st = state(nelem=(6, 6, 6), xmin=(0, 0, 0), xmax=(1, 1, 1), idimc=(2, 2, 2))
    rng = np.random.default_rng(4)
    for _ in range(200):
        c = [rng.integers(0, 6) for a in range(3)]
        m = c[0] + c[1] * 6 + c[2] * 36
        x = np.array([(c[a] + rng.random()) / 6.0 for a in range(3)])
        _, sf, dsf = maksf(st, m, x)
        assert abs(sf.sum() - 1.0) < 1e-13
        np.testing.assert_allclose(dsf.sum(axis=0), 0.0, atol=1e-10)


def test_cell_of_brute_force_and_faces():
    st = state()
    rng = np.random.default_rng(1)
    for _ in range(2000):
        x = np.array([st.xmin[a] + (st.xmax[a] - st.xmin[a]) * rng.random() for a in range(3)])
        c = orc.orc_cell_of(C.byref(st), P(x))
        ijk = [int(np.floor((x[a] - st.xmin[a]) / st.dx[a])) for a in range(3)]
        assert c == ijk[0] + ijk[1] * st.nex[0] + ijk[2] * st.nex[0] * st.nex[1]
    # a position on a cell face belongs to the higher cell; the last cell exists
    x = np.array([st.xmin[0] + st.dx[0] * 2, st.xmin[1], st.xmin[2] + st.dx[2] * 4])
    assert orc.orc_cell_of(C.byref(st), P(x)) == 2 + 0 + 4 * st.nex[0] * st.nex[1]
    x = np.array([st.xmax[0], st.xmin[1], st.xmin[2]])
    assert orc.orc_cell_of(C.byref(st), P(x)) == -1  # outside (the reference has no check)


def test_bin_is_stable_sort_by_cell():
    s = scenes.tiny(n_cells=6)
    o = pyoracle.Oracle(s)
    rng = np.random.default_rng(2)
    o.parts.coord[:] = rng.random(o.parts.coord.shape) * 0.999
    assert o.bin() == 0
    cells = np.array([o.cell_of(x) for x in o.parts.coord])
    order = np.argsort(cells, kind="stable")
    # walk the linked lists in cell order
    walk = []
    for m in range(o.nelem):
        p = o.idepf[m]
        cnt = 0
        while p != -1:
            walk.append(p)
            p = o.idp2p[p]
            cnt += 1
        assert cnt == o.numep[m]
    assert np.array_equal(np.array(walk), order)


# ---------------------------------------------------------------------------
# Independent vectorised restatement of one explicit step (PIC/FLIP, MUSL)
# ---------------------------------------------------------------------------
def np_weights(s, x):
    dx = np.array(s.dx)
    xmin = np.array(s.xmin)
    c = np.floor((x - xmin) / dx).astype(int)
    nx = np.array(s.nelem) + 1
    out_ids, out_w, out_g = [], [], []
    for k, j, i in itertools.product(range(2), range(2), range(2)):
        o = np.array([i, j, k])
        idx = c + o
        xa = xmin + dx * idx
        r = x - xa
        w1 = 1.0 - np.abs(r) / dx
        g1 = np.where(r < 0, 1.0, -1.0) / dx
        N = w1[:, 0] * w1[:, 1] * w1[:, 2]
        G = np.stack([g1[:, 0] * w1[:, 1] * w1[:, 2], w1[:, 0] * g1[:, 1] * w1[:, 2], w1[:, 0] * w1[:, 1] * g1[:, 2]], 1)
        out_ids.append(idx[:, 0] + nx[0] * idx[:, 1] + nx[0] * nx[1] * idx[:, 2])
        out_w.append(N)
        out_g.append(G)
    return np.stack(out_ids, 1), np.stack(out_w, 1), np.stack(out_g, 1)


def np_step(s, parts, facl=1.0):
    n, nodec = parts.n, s.nodec
    x = parts.coord.copy()
    ids, N, G = np_weights(s, x)
    m = parts.mass
    nm = np.bincount(ids.ravel(), (N * m[:, None]).ravel(), nodec)
    p = np.stack([np.bincount(ids.ravel(), (N * (m * parts.vel[:, c])[:, None]).ravel(), nodec) for c in range(3)])
    sig = parts.stress
    S = np.zeros((n, 3, 3))
    vo = [(0, 0, 0), (1, 1, 1), (2, 2, 2), (1, 2, 3), (2, 1, 3), (0, 2, 4), (2, 0, 4), (0, 1, 5), (1, 0, 5)]
    for a, b, q in vo:
        S[:, a, b] = sig[:, q]
    fint = -np.einsum("p,pab,pkb->pka", parts.vol, S, G)
    fext = N[:, :, None] * (m[:, None] * np.asarray(s.bb)[None, :] * facl)[:, None, :]
    f = np.stack([np.bincount(ids.ravel(), (fint + fext)[:, :, c].ravel(), nodec) for c in range(3)])
    act = nm > s.mtol
    v = np.where(act, p / np.where(act, nm, 1.0), 0.0)
    a = np.where(act, f / np.where(act, nm, 1.0), 0.0)
    for comp, (nbc, fbc) in s.bcs.items():
        comps = [comp] if comp < 3 else [0, 1, 2]
        for c in comps:
            v[c, nbc] = np.asarray(fbc) * facl
            a[c, nbc] = 0.0
    if s.scheme == capi.SCHEME_FLIP:
        vp = parts.vel + s.dt * np.einsum("pk,cpk->pc", N, a[:, ids])
    else:
        v = v + s.dt * a
        vp = np.einsum("pk,cpk->pc", N, v[:, ids])
    p2 = np.stack([np.bincount(ids.ravel(), (N * (m * vp[:, c])[:, None]).ravel(), nodec) for c in range(3)])
    v2 = np.where(act, p2 / np.where(act, nm, 1.0), 0.0)
    for comp, (nbc, fbc) in s.bcs.items():
        comps = [comp] if comp < 3 else [0, 1, 2]
        for c in comps:
            v2[c, nbc] = np.asarray(fbc) * facl
    L = np.einsum("cpk,pkb->pcb", v2[:, ids], G)
    dF = np.eye(3)[None] + s.dt * L
    F = np.einsum("pab,pbc->pac", dF, parts.def_grad)
    J = np.linalg.det(F)
    xn = x + s.dt * np.einsum("pk,cpk->pc", N, v2[:, ids])
    return dict(nmass=nm, nmome2=p2, nforce=f, vp=vp, x=xn, F=F, J=J, dF=dF, nvel2=v2)


@pytest.mark.parametrize("scheme", [capi.SCHEME_PIC, capi.SCHEME_FLIP])
def test_step_against_vectorised_restatement(scheme):
    s = scenes.tiny(n_cells=6, scheme=scheme, model=capi.MODEL_NEOHOOKEAN)
    rng = np.random.default_rng(11)
    # give particles a non-trivial stress and deformation to exercise every term
    s.parts.stress[:] = rng.standard_normal(s.parts.stress.shape) * 50.0
    s.parts.def_grad[:] = np.eye(3) + 1e-3 * rng.standard_normal((s.n, 3, 3))
    want = np_step(s, s.parts)
    o = pyoracle.Oracle(s)
    assert o.bin() == 0
    o.p2g(1.0)
    np.testing.assert_allclose(o.nmass, want["nmass"], rtol=1e-12, atol=1e-18)
    np.testing.assert_allclose(o.nforce.reshape(3, -1), want["nforce"], rtol=1e-10, atol=1e-12)
    o.solve_grid()
    o.g2p(1.0)
    np.testing.assert_allclose(o.parts.vel, want["vp"], rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(o.nmome.reshape(3, -1), want["nmome2"], rtol=1e-12, atol=1e-16)
    np.testing.assert_allclose(o.parts.coord, want["x"], rtol=0, atol=1e-15)
    np.testing.assert_allclose(o.parts.def_grad, want["F"], rtol=0, atol=1e-14)
    np.testing.assert_allclose(o.parts.det_def_grad, want["J"], rtol=1e-14)
    np.testing.assert_allclose(o.parts.vol, want["J"] * s.parts.vol0, rtol=1e-14)


def test_conservation_in_p2g():
    s = scenes.tiny(n_cells=8, gravity=(0.0, 0.0, 0.0))
    s.bcs = {}
    rng = np.random.default_rng(8)
    s.parts.stress[:] = rng.standard_normal(s.parts.stress.shape) * 100.0
    o = pyoracle.Oracle(s)
    o.bin()
    o.p2g(1.0)
    M = s.parts.mass.sum()
    assert o.nmass.sum() == pytest.approx(M, rel=1e-14)
    mom = (s.parts.mass[:, None] * s.parts.vel).sum(0)
    np.testing.assert_allclose(o.nmome.reshape(3, -1).sum(1), mom, rtol=1e-12, atol=1e-14)
    # internal forces sum to zero (partition of unity of the gradients)
    np.testing.assert_allclose(o.nforce.reshape(3, -1).sum(1), 0.0, atol=1e-9)
    # gravity alone: total external force = M g
    s2 = scenes.tiny(n_cells=8, gravity=(0.0, -9.81, 0.0))
    s2.bcs = {}
    o2 = pyoracle.Oracle(s2)
    o2.bin()
    o2.p2g(1.0)
    np.testing.assert_allclose(o2.nforce.reshape(3, -1).sum(1), [0.0, -9.81 * M, 0.0], rtol=1e-12, atol=1e-9)


def test_free_fall_is_exact_under_pic():
    """No BCs, no stress, uniform initial velocity: v = v0 + g t and the block
    translates rigidly (F stays I to round-off)."""
    s = scenes.tiny(n_cells=8, omega=0.0, gravity=(0.0, 0.0, -9.81), model=capi.MODEL_LINELAST, props=(1e6, 0.3))
    s.bcs = {}
    s.parts.vel[:] = [0.1, 0.0, 0.0]
    s.parts.coord[:, 2] += 0.1
    o = pyoracle.Oracle(s)
    nst = 20
    assert o.run(1, nst) == 0
    t = nst * s.dt
    np.testing.assert_allclose(o.parts.vel[:, 2], -9.81 * t, rtol=1e-10)
    np.testing.assert_allclose(o.parts.vel[:, 0], 0.1, rtol=1e-10)
    np.testing.assert_allclose(o.parts.def_grad, np.tile(np.eye(3), (s.n, 1, 1)), atol=1e-12)
    # positions: x = x0 + sum_k dt v(t_k) with v(t_k) = v0 + g k dt
    dz = sum(s.dt * (-9.81 * k * s.dt) for k in range(1, nst + 1))
    np.testing.assert_allclose(o.parts.coord[:, 2], s.parts.coord[:, 2] + dz, atol=1e-12)


def test_neohookean_uniaxial_closed_form():
    E, nu = 5e5, 0.35
    lam = nu * E / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    for st in (0.9, 1.0, 1.05, 1.3):
        F = np.diag([st, 1.0, 1.0])
        s = np.zeros(6)
        pr = np.zeros(8)
        pr[:2] = (E, nu)
        orc.orc_update_stress(1, P(s), P(np.ascontiguousarray(F)), P(np.eye(3)), P(pr), st, st)
        want_xx = (mu * (st * st - 1.0) + lam * np.log(st)) / st
        want_yy = lam * np.log(st) / st
        np.testing.assert_allclose(s, [want_xx, want_yy, want_yy, 0, 0, 0], rtol=1e-13, atol=1e-9)


def test_linelast_hooke_increment():
    E, nu = 1e6, 0.3
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    eps = 1e-4
    dF = np.eye(3)
    dF[0, 0] += eps
    dF[0, 1] = 0.5 * eps
    dF[1, 0] = 0.5 * eps
    s = np.zeros(6)
    pr = np.zeros(8)
    pr[:2] = (E, nu)
    orc.orc_update_stress(0, P(s), P(np.eye(3)), P(np.ascontiguousarray(dF)), P(pr), 1.0, 1.0)
    np.testing.assert_allclose(s, [(lam + 2 * mu) * eps, lam * eps, lam * eps, 0, 0, mu * eps], rtol=1e-9)


def test_det3_permutation_expansion():
    rng = np.random.default_rng(12)
    for _ in range(100):
        A = np.ascontiguousarray(rng.standard_normal((3, 3)))
        perm_det = 0.0
        for p in itertools.permutations(range(3)):
            sgn = np.linalg.det(np.eye(3)[list(p)])
            perm_det += sgn * A[0, p[0]] * A[1, p[1]] * A[2, p[2]]
        assert orc.orc_det3(P(A)) == pytest.approx(perm_det, rel=1e-12, abs=1e-14)


def j2(props, s, ep, de):
    s = np.ascontiguousarray(s, np.float64).copy()
    ep_ = np.array([ep])
    dg = np.zeros(1)
    r = np.zeros(1)
    pl = orc.orc_j2_return(P(np.asarray(props + [0] * (8 - len(props)), np.float64)), P(s), P(ep_),
                           P(np.ascontiguousarray(de, np.float64)), P(dg), P(r))
    return pl, s, ep_[0], dg[0], r[0]


def test_j2_return_mapping_properties():
    props = [2e6, 0.3, 2e3, 1e4]
    E, nu, sy, H = props
    mu = E / (2 * (1 + nu))
    rng = np.random.default_rng(21)
    for _ in range(300):
        s0 = rng.standard_normal(6) * 500.0
        de = rng.standard_normal(6) * 1e-3
        ep0 = abs(rng.standard_normal()) * 1e-3
        pl, s, ep, dg, r = j2(props, s0, ep0, de)
        pm = s[:3].mean()
        d = s[:3] - pm
        q = np.sqrt(3 * (0.5 * (d @ d) + s[3:] @ s[3:]))
        assert dg >= 0.0
        if pl:
            assert abs(q - (sy + H * ep)) <= 1e-9 * sy     # on the yield surface
            assert ep == pytest.approx(ep0 + dg)
        else:
            assert q <= sy + H * ep0 + 1e-9                  # elastic: unchanged trial
            assert dg == 0.0 and ep == ep0


def test_j2_uniaxial_hardening_closed_form():
    """Pure shear strain increments far past yield: the equivalent stress
    follows sigma_y + H eps_p exactly, and the elastic-plastic tangent in
    q-space is 3 mu H / (3 mu + H)."""
    props = [2e6, 0.3, 2e3, 1e4]
    E, nu, sy, H = props
    mu = E / (2 * (1 + nu))
    s, ep = np.zeros(6), 0.0
    gam = 1e-3
    qs = []
    for _ in range(10):
        pl, s, ep, dg, r = j2(props, s, ep, [0, 0, 0, 0, 0, gam])
        qs.append(np.sqrt(3.0) * abs(s[5]))
    dq = np.diff(qs[2:])
    tangent = 3 * mu * H / (3 * mu + H)
    # in pure shear q = sqrt(3)|tau| and dq = tangent * d(eps_eq) with d(eps_eq) = gam / sqrt(3)
    np.testing.assert_allclose(dq, tangent * gam / np.sqrt(3.0), rtol=1e-9)


# ------------------------------------------------------------------ Tait fluid (no reference counterpart)
def _tait(prop, jac, dF, dt):
    lib = pyoracle.load()
    s = np.zeros(6)
    lib.orc_tait_stress(pyoracle.P(np.asarray(prop, np.float64)), jac, pyoracle.P(np.asarray(dF, np.float64)), dt,
                        pyoracle.P(s))
    return s


@pytest.mark.parametrize("jac", [0.97, 0.999, 1.0, 1.004])
def test_tait_pressure_closed_form(jac):
    rho0, c, gam, mu = 1000.0, 30.0, 7.0, 1e-3
    s = _tait([rho0, c, gam, mu], jac, np.eye(3).ravel(), 1e-5)
    # p = rho0 c^2 / gamma ((rho / rho0)^gamma - 1), rho = rho0 / J; at rest no viscous part
    p = rho0 * c * c / gam * ((1.0 / jac) ** gam - 1.0)
    assert np.allclose(s[:3], -p, rtol=1e-12, atol=1e-9)
    assert np.all(s[3:] == 0.0)
    if jac < 1:
        assert p > 0  # compression pushes back


def test_tait_viscous_shear_and_dilatation():
    rho0, c, gam, mu, dt = 1000.0, 30.0, 7.0, 0.5, 1e-4
    g = 3.0  # shear rate dv_x/dy
    dF = np.eye(3)
    dF[0, 1] = g * dt  # L_xy = g
    s = _tait([rho0, c, gam, mu], 1.0, dF.ravel(), dt)
    assert np.allclose(s, [0, 0, 0, 0, 0, mu * g], rtol=1e-12, atol=1e-12)
    # a pure dilatation rate has no deviatoric part: isotropic stress only
    dF = np.eye(3) * (1 + 2e-3 * dt)
    s = _tait([rho0, c, gam, mu], 1.0, dF.ravel(), dt)
    assert np.allclose(s, 0.0, atol=1e-9)
    # a deviatoric stretch rate: sigma_xx - sigma_yy = 2 mu (L_xx - L_yy)
    dF = np.diag([1 + 1e-3 * dt, 1 - 1e-3 * dt, 1.0])
    s = _tait([rho0, c, gam, mu], 1.0, dF.ravel(), dt)
    assert np.isclose(s[0] - s[1], 2 * mu * 2e-3, rtol=1e-9)
    assert np.isclose(s[0] + s[1] + s[2], 0.0, atol=1e-15)
