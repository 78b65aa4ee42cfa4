"""Pins the oracle's particle shifting (DeltaCorrectionParticleShifting,
module/material_point.cpp:174-227) and nodal volume (MakNodalVol,
module/mesh.cpp:160-202), which need MPI to build in the reference: closed
forms of the nodal volume, no shift for the undisturbed particle lattice,
descent of the excess-volume norm, mirror symmetry, and the correction
itself from the mathematics it implements.

The correction is one Gauss-Newton step on the excess-volume norm
E(x) = sum_a max(0, nvof_a - nvol_a)^2 with nvof_a = sum_p N_a(x_p) V_p:
geup_p = 2 V_p sum_a grad N_a(x_p) nei_a is dE/dx_p at fixed V_p, and
delta_p = -E geup_p / sum_q |geup_q|^2, so that sum_p delta_p . geup_p = -E.
The finite-difference tests below predict delta from E alone; the oracle
builds with a planted mistake (no factor 2 in geup, vol0 for vol) fail
them."""
import numpy as np
import pytest

import pyoracle
from mpm_mi355x import capi, scenes


def lattice_scene(n_cells=6, ppc=(2, 2, 2), lo=(0, 0, 0), hi=None, jitter=0.0, seed=3):
    hi = hi or (n_cells,) * 3
    h = 1.0 / n_cells
    nelem = (n_cells,) * 3
    coord = scenes.seed_block((0.0, 0.0, 0.0), (h, h, h), nelem, lo, hi, ppc)
    if jitter:
        rng = np.random.default_rng(seed)
        coord = coord + rng.uniform(-jitter, jitter, coord.shape) * h
    n = len(coord)
    volp = h ** 3 / np.prod(ppc)
    parts = capi.Particles(n, coord=coord, mass=np.full(n, 1000 * volp), vol0=np.full(n, volp))
    return scenes.Scene("lattice", (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), nelem, [(capi.MODEL_LINELAST, [1e6, 0.3])], 1000.0,
                        1e-5, 1, (0.0, 0.0, 0.0), parts=parts, shifting=True, npxye=ppc)


@pytest.mark.parametrize("ppc", [(2, 2, 2), (3, 1, 2), (4, 4, 4)])
def test_nodal_volume_closed_form(ppc):
    # two or more Gauss-type points per axis integrate the tent functions
    # exactly: interior nodes h^3, face nodes h^3/2, edges h^3/4, corners h^3/8
    # (one point per axis at the cell centre integrates them exactly too)
    s = lattice_scene(4, ppc)
    o = pyoracle.Oracle(s)
    nx = 5
    h3 = (1.0 / 4) ** 3
    i, j, k = np.meshgrid(np.arange(nx), np.arange(nx), np.arange(nx), indexing="ij")
    faces = sum(m.astype(int) for m in (i == 0, i == nx - 1, j == 0, j == nx - 1, k == 0, k == nx - 1))
    expect = h3 / 2.0 ** faces
    got = o.nvol.reshape(nx, nx, nx).transpose(2, 1, 0)  # node id i + nx j + nx^2 k
    assert np.allclose(got, expect, rtol=1e-13, atol=0)
    assert np.isclose(o.nvol.sum(), 1.0, rtol=1e-13)


def test_nodal_volume_quadratic_partition_of_unity():
    s = lattice_scene(5, (2, 2, 2))
    s.idimc = (2, 2, 2)
    o = pyoracle.Oracle(s)
    assert np.isclose(o.nvol.sum(), 1.0, rtol=1e-13)
    assert np.all(o.nvol >= 0.0)


def delta_of(s):
    o = pyoracle.Oracle(s)
    assert o.bin() == 0
    o.p2g(1.0)
    return o, o.shift_delta()


def test_no_shift_for_the_reference_lattice():
    # particles exactly at the MakNodalVol points with their volume: the
    # nodal volumes match, nothing is in excess, nothing moves
    s = lattice_scene(5, (2, 2, 2))
    o, d = delta_of(s)
    assert np.max(np.abs(o.nvof - o.nvol)) < 1e-15
    assert np.max(np.abs(d)) < 1e-14 / 5


def excess(o, coord):
    p = o.parts
    p.coord[:] = coord
    assert o.bin() == 0
    o.p2g(1.0)
    return float(np.sum(np.maximum(0.0, o.nvof - o.nvol) ** 2))


def test_shift_descends_the_excess_volume_norm():
    # a block squeezed towards its centre has excess nodal volume inside;
    # one shifting correction lowers the excess-volume norm
    s = lattice_scene(8, (2, 2, 2), lo=(2, 2, 2), hi=(6, 6, 6))
    c = np.array([0.5, 0.5, 0.5])
    s.parts.coord[:] = c + 0.9 * (s.parts.coord - c)
    o, d = delta_of(s)
    x0 = o.parts.coord.copy()
    e0 = excess(o, x0)
    assert e0 > 0 and np.max(np.abs(d)) > 0
    e1 = excess(o, x0 + d)
    assert e1 < e0


def test_shift_is_mirror_symmetric():
    s = lattice_scene(8, (2, 2, 2), lo=(2, 2, 2), hi=(6, 6, 6))
    c = np.array([0.5, 0.5, 0.5])
    s.parts.coord[:] = c + 0.9 * (s.parts.coord - c)
    _, d = delta_of(s)
    x = s.parts.coord
    # partner of each particle under x -> 1 - x
    key = {tuple(np.round(p, 12)): i for i, p in enumerate(x)}
    tol = 1e-9 * np.max(np.abs(d))
    for i, p in enumerate(x):
        q = (round(1.0 - p[0], 12), round(p[1], 12), round(p[2], 12))
        j = key[q]
        assert abs(d[j, 0] + d[i, 0]) <= tol
        assert abs(d[j, 1] - d[i, 1]) <= tol and abs(d[j, 2] - d[i, 2]) <= tol


def squeezed_scene(idimc, seed=11):
    """A block squeezed towards its centre, jittered, with particle volumes
    V != V0, so that every term of the correction is exercised."""
    s = lattice_scene(8, (2, 2, 2), lo=(2, 2, 2), hi=(6, 6, 6), jitter=0.12, seed=seed)
    c = np.array([0.5, 0.5, 0.5])
    s.parts.coord[:] = c + 0.9 * (s.parts.coord - c)
    rng = np.random.default_rng(seed + 1)
    s.parts.vol[:] = s.parts.vol0 * rng.uniform(0.7, 1.3, s.n)
    s.idimc = idimc
    return s


def fd_gradient(o, x0, eps):
    """Central differences of E with respect to every particle coordinate."""
    g = np.zeros_like(x0)
    for p in range(len(x0)):
        for a in range(3):
            x = x0.copy()
            x[p, a] += eps
            ep = excess(o, x)
            x[p, a] -= 2 * eps
            em = excess(o, x)
            g[p, a] = (ep - em) / (2 * eps)
    return g


def gauss_newton_mismatch(idimc, fast=False):
    s = squeezed_scene(idimc)
    o = pyoracle.Oracle(s, fast=fast)
    x0 = o.parts.coord.copy()
    e0 = excess(o, x0)                # bins and nvof at x0
    d = o.shift_delta()
    h = 1.0 / 8
    g = fd_gradient(o, x0, 1e-7 * h)
    pred = -e0 * g / np.sum(g * g)
    ident = float(np.sum(d * g))      # sum_p delta_p . dE/dx_p
    return e0, d, pred, ident


@pytest.mark.parametrize("idimc", [(1, 1, 1), (2, 2, 2)])
def test_shift_delta_is_the_gauss_newton_step_of_the_excess_norm(idimc):
    e0, d, pred, ident = gauss_newton_mismatch(idimc)
    assert e0 > 0 and np.max(np.abs(d)) > 0
    # the step length: sum_p delta_p . dE/dx_p = -E
    assert abs(ident + e0) <= 1e-6 * e0, (ident, e0)
    # the step itself, particle by particle
    assert np.max(np.abs(d - pred)) <= 1e-6 * np.max(np.abs(pred))


@pytest.mark.parametrize("mutant", [1, 2])
def test_planted_mistakes_in_the_correction_are_caught(mutant):
    """The same check on oracle builds with a planted mistake: 1 drops the
    factor 2 of the gradient (the step doubles), 2 takes vol0 for vol."""
    e0, d, pred, ident = gauss_newton_mismatch((1, 1, 1), fast=("mutant", mutant))
    err = np.max(np.abs(d - pred)) / np.max(np.abs(pred))
    assert err > 1e-2, err
    if mutant == 1:
        assert abs(ident + 2 * e0) <= 1e-6 * e0


@pytest.mark.parametrize("idimc", [(1, 1, 1), (2, 2, 2), (2, 1, 2)])
def test_maksf_gradient_is_the_derivative_of_the_weights(idimc):
    """dsf = d sf / dx by central differences inside the cells, for the tent
    functions and the quadratic B-spline, boundary cells included."""
    import ctypes as C
    nelem = (6, 5, 7)
    st = pyoracle.OrcState()
    lo, hi = (0.0, -0.5, 1.0), (1.2, 0.5, 2.4)
    for i in range(3):
        st.xmin[i], st.xmax[i], st.nex[i], st.idimc[i] = lo[i], hi[i], nelem[i], idimc[i]
    orc = pyoracle.load()
    orc.orc_grid_init(C.byref(st))
    rng = np.random.default_rng(5)
    n = 3000
    ijk = np.stack([rng.integers(0, nelem[a], n) for a in range(3)], 1)
    cells = (ijk[:, 0] + ijk[:, 1] * nelem[0] + ijk[:, 2] * nelem[0] * nelem[1]).astype(np.int32)
    h = np.array([st.dx[a] for a in range(3)])
    frac = rng.uniform(0.01, 0.99, (n, 3))
    x = np.array(lo) + (ijk + frac) * h
    nen = int(np.prod(np.array(idimc) + 1))

    def maksf(xx):
        nc = np.zeros((n, nen), np.int32)
        sf = np.zeros((n, nen))
        dsf = np.zeros((n, nen, 3))
        orc.orc_maksf_batch(C.byref(st), n, pyoracle.P(cells, C.c_int32), pyoracle.P(np.ascontiguousarray(xx)),
                            pyoracle.P(nc, C.c_int32), pyoracle.P(sf), pyoracle.P(dsf))
        return nc, sf, dsf

    nc, sf, dsf = maksf(x)
    for a in range(3):
        eps = 1e-6 * h[a]
        xp, xm = x.copy(), x.copy()
        xp[:, a] += eps
        xm[:, a] -= eps
        ncp, sfp, _ = maksf(xp)
        ncm, sfm, _ = maksf(xm)
        assert np.array_equal(ncp, nc) and np.array_equal(ncm, nc)
        fd = (sfp - sfm) / (2 * eps)
        assert np.max(np.abs(fd - dsf[:, :, a])) <= 1e-7 / h[a], a
