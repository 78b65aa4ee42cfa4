"""Pins the oracle's particle shifting (DeltaCorrectionParticleShifting,
module/material_point.cpp:174-227) and nodal volume (MakNodalVol,
module/mesh.cpp:160-202), which need MPI to build in the reference: closed
forms of the nodal volume, no shift for the undisturbed particle lattice,
descent of the excess-volume norm, mirror symmetry."""
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
