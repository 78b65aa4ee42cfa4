"""Synthetic scenes for the explicit solid MPM step.

Particles are seeded like the reference's generators
(data/generate/solid/solid_generator.cpp:92-188, module/material_point.cpp:16-50):
cells in (k, j, i) order, then sub-cell points in (kp, jp, ip) order at the
offsets of GaussianDistribution (module/mesh.cpp:12-54), each with mass
rho * dx*dy*dz / npe. Node-list BCs are built in node-id order like the
generator's lists. Random velocity noise comes from a SplitMix64 stream
(Steele, Lea & Flood 2014) with a fixed seed per scene.
"""
from dataclasses import dataclass, field

import numpy as np

from . import capi


def splitmix64(seed, count):
    """count uint64 draws of SplitMix64 starting from state `seed`."""
    M = np.uint64(0xFFFFFFFFFFFFFFFF)
    with np.errstate(over="ignore"):
        state = (np.uint64(seed) + np.arange(1, count + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)) & M
        z = state
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed, count, lo, hi):
    u = (splitmix64(seed, count) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return lo + (hi - lo) * u


def gauss_offsets(npx, dx):
    """GaussianDistribution (module/mesh.cpp:12-54) for one axis."""
    table = {
        1: [0.0],
        2: [-0.25, 0.25],
        3: [-1.0 / 3.0, 0.0, 1.0 / 3.0],
        4: [-0.375, -0.125, 0.125, 0.375],
        5: [-0.4, -0.2, 0.0, 0.2, 0.4],
        6: [-5.0 / 12.0, -3.0 / 12.0, -1.0 / 12.0, 1.0 / 12.0, 3.0 / 12.0, 5.0 / 12.0],
    }
    return [c * dx for c in table[npx]]


@dataclass
class Scene:
    name: str
    xmin: tuple
    xmax: tuple
    nelem: tuple
    materials: list           # [(model, [props])]
    rho: float
    dt: float
    steps: int
    bb: tuple
    mtol: float = 1e-12
    scheme: int = capi.SCHEME_PIC
    nlstep: int = 1
    bcs: dict = field(default_factory=dict)   # comp -> (nbc, fbc)
    parts: object = None
    idimc: tuple = (1, 1, 1)                 # 1 = linear tent, 2 = quadratic B-spline (open, per axis)
    shifting: bool = False                   # reference particle shifting
    npxye: tuple = (2, 2, 2)                 # points per cell of the nodal-volume reference (MakNodalVol)
    fbar: bool = False                       # F-bar projection
    mat_dt: list = None                      # per-material time steps (FP32 engine), or None: dt for all

    @property
    def dx(self):
        return tuple((self.xmax[i] - self.xmin[i]) / self.nelem[i] for i in range(3))

    @property
    def nodec(self):
        return int(np.prod(np.asarray(self.nelem) + np.asarray(self.idimc)))

    @property
    def nelem_total(self):
        return int(np.prod(self.nelem))

    @property
    def n(self):
        return self.parts.n


def seed_block(xmin, dx, nelem, lo_cells, hi_cells, npxye):
    """Particle coordinates of the cells [lo, hi) per axis, reference order."""
    ox = gauss_offsets(npxye[0], dx[0])
    oy = gauss_offsets(npxye[1], dx[1])
    oz = gauss_offsets(npxye[2], dx[2])
    ks = np.arange(lo_cells[2], hi_cells[2])
    js = np.arange(lo_cells[1], hi_cells[1])
    is_ = np.arange(lo_cells[0], hi_cells[0])
    K, J, I = np.meshgrid(ks, js, is_, indexing="ij")
    ecx = dx[0] * (I.ravel() + 0.5) + xmin[0]
    ecy = dx[1] * (J.ravel() + 0.5) + xmin[1]
    ecz = dx[2] * (K.ravel() + 0.5) + xmin[2]
    pts = []
    for kp in range(npxye[2]):
        for jp in range(npxye[1]):
            for ip in range(npxye[0]):
                pts.append((ox[ip], oy[jp], oz[kp]))
    pts = np.asarray(pts)  # (npe, 3)
    x = ecx[:, None] + pts[None, :, 0]
    y = ecy[:, None] + pts[None, :, 1]
    z = ecz[:, None] + pts[None, :, 2]
    return np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)


def node_lists(nelem, select, idimc=(1, 1, 1)):
    """Node ids (reference numbering) in id order where select(i, j, k) holds."""
    nx, ny, nz = nelem[0] + idimc[0], nelem[1] + idimc[1], nelem[2] + idimc[2]
    K, J, I = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    ids = (I + nx * J + nx * ny * K).ravel()
    mask = select(I.ravel(), J.ravel(), K.ravel(), nx, ny, nz)
    return ids[mask].astype(np.int32)


def _bc(ids):
    return (ids, np.zeros(len(ids)))


def cfg1(steps=1000, seed=0x5EED0001):
    """2D elastic block (reference slab form: one cell through the thickness,
    vertical = z): 64x64 cells on the unit square, block [0.25,0.75]x[0,0.5]
    at 2x2 points per cell = 4,096 particles, linear elastic E=1e6, nu=0.3,
    rho=1000, dt=1e-5. Bottom nodes fixed, side nodes roller, v = 0
    everywhere (plane strain)."""
    h = 1.0 / 64.0
    xmin, xmax, nelem = (0.0, 0.0, 0.0), (1.0, h, 1.0), (64, 1, 64)
    dx = (h, h, h)
    coord = seed_block(xmin, dx, nelem, (16, 0, 0), (48, 1, 32), (2, 1, 2))
    n = len(coord)
    rho = 1000.0
    volp = dx[0] * dx[1] * dx[2] / 4.0
    vel = uniform(seed, 3 * n, -1e-3, 1e-3).reshape(n, 3)
    vel[:, 1] = 0.0
    parts = capi.Particles(n, coord=coord, vel=vel, mass=np.full(n, rho * volp), vol0=np.full(n, volp))
    bottom = node_lists(nelem, lambda i, j, k, nx, ny, nz: k == 0)
    sides = node_lists(nelem, lambda i, j, k, nx, ny, nz: (i == 0) | (i == nx - 1) | (k == 0))
    allv = node_lists(nelem, lambda i, j, k, nx, ny, nz: np.ones_like(i, dtype=bool))
    return Scene("cfg1-2d-linelast", xmin, xmax, nelem, [(capi.MODEL_LINELAST, [1e6, 0.3])], rho, 1e-5, steps,
                 (0.0, 0.0, -9.81), bcs={0: _bc(sides), 1: _bc(allv), 2: _bc(bottom)}, parts=parts)


def cfg2(steps=1000, ppc=(2, 2, 2), n_cells=128, seed=0x5EED0002, idimc=(1, 1, 1)):
    """3D Neo-Hookean cube: 128^3 cells on the unit cube, cube
    [0.25,0.75]x[0,0.5]x[0.25,0.75]; ppc (2,2,2) gives 2,097,152 particles
    (geometry as stated), (2,4,2) gives 4,194,304. E=5e5, nu=0.35, rho=1000,
    angular velocity 2 rad/s about the y axis through (0.5, y, 0.5) plus
    +-1e-3 m/s noise, gravity -y, floor nodes (y = 0) fixed, dt = 5e-6."""
    h = 1.0 / n_cells
    xmin, xmax, nelem = (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (n_cells,) * 3
    dx = (h, h, h)
    q = n_cells // 4
    coord = seed_block(xmin, dx, nelem, (q, 0, q), (3 * q, 2 * q, 3 * q), ppc)
    n = len(coord)
    rho = 1000.0
    volp = h ** 3 / float(np.prod(ppc))
    omega = 2.0
    vel = np.zeros((n, 3))
    vel[:, 0] = omega * (coord[:, 2] - 0.5)
    vel[:, 2] = -omega * (coord[:, 0] - 0.5)
    vel += uniform(seed, 3 * n, -1e-3, 1e-3).reshape(n, 3)
    parts = capi.Particles(n, coord=coord, vel=vel, mass=np.full(n, rho * volp), vol0=np.full(n, volp))
    floor = node_lists(nelem, lambda i, j, k, nx, ny, nz: j == 0, idimc)
    bcs = {0: _bc(floor), 1: _bc(floor), 2: _bc(floor)}
    name = "cfg2-3d-neohookean-%dppc" % int(np.prod(ppc))
    if tuple(idimc) != (1, 1, 1):
        name += "-bspline%d%d%d" % tuple(idimc)
    return Scene(name, xmin, xmax, nelem, [(capi.MODEL_NEOHOOKEAN, [5e5, 0.35])], rho, 5e-6, steps,
                 (0.0, -9.81, 0.0), bcs=bcs, parts=parts, idimc=tuple(idimc))


def tiny(n_cells=8, ppc=(2, 2, 2), model=capi.MODEL_NEOHOOKEAN, props=(5e5, 0.35), scheme=capi.SCHEME_PIC,
         seed=0x5EED0009, dt=1e-5, omega=2.0, gravity=(0.0, -9.81, 0.0), nlstep=1, idimc=(1, 1, 1)):
    """Small 3D scene for parity tests: the cfg2 geometry on an n_cells^3 grid."""
    h = 1.0 / n_cells
    xmin, xmax, nelem = (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (n_cells,) * 3
    dx = (h, h, h)
    q = max(1, n_cells // 4)
    coord = seed_block(xmin, dx, nelem, (q, 0, q), (3 * q, 2 * q, 3 * q), ppc)
    n = len(coord)
    rho = 1000.0
    volp = h ** 3 / float(np.prod(ppc))
    vel = np.zeros((n, 3))
    vel[:, 0] = omega * (coord[:, 2] - 0.5)
    vel[:, 2] = -omega * (coord[:, 0] - 0.5)
    vel += uniform(seed, 3 * n, -1e-3, 1e-3).reshape(n, 3)
    parts = capi.Particles(n, coord=coord, vel=vel, mass=np.full(n, rho * volp), vol0=np.full(n, volp))
    floor = node_lists(nelem, lambda i, j, k, nx, ny, nz: j == 0, idimc)
    bcs = {0: _bc(floor), 1: _bc(floor), 2: _bc(floor)}
    return Scene("tiny", xmin, xmax, nelem, [(model, list(props))], rho, dt, 10, gravity, scheme=scheme,
                 nlstep=nlstep, bcs=bcs, parts=parts, idimc=tuple(idimc))


def cfg3(steps=2000, h=1.0 / 150.0, rho=2000.0, dt=4e-5, seed=0x5EED0003):
    """Column collapse with J2 plasticity (linear hardening E=2e6, nu=0.3,
    sigma_y=2e3, H=1e4): 192x96x96 cells. The spacing is not given for this
    case; h = 1/150 makes the domain [0,1.28]x[0,0.64]x[0,0.64] and the column
    [0,0.3]x[0,0.6]x[0,0.5] exactly 45x90x75 cells, at 2x2x2 points per cell
    2,430,000 particles. rho = 2000 and dt = 4e-5 (about 0.22 of the elastic
    CFL limit h/c) are likewise choices of this build. Gravity -y; floor nodes
    (y = 0) fixed; x = 0 and z = 0 planes are symmetry rollers (u = 0, w = 0);
    the far walls are rollers too."""
    nelem = (192, 96, 96)
    xmin = (0.0, 0.0, 0.0)
    xmax = (192 * h, 96 * h, 96 * h)
    dx = (h, h, h)
    hi = (int(round(0.3 / h)), int(round(0.6 / h)), int(round(0.5 / h)))
    coord = seed_block(xmin, dx, nelem, (0, 0, 0), hi, (2, 2, 2))
    n = len(coord)
    volp = h ** 3 / 8.0
    vel = uniform(seed, 3 * n, -1e-3, 1e-3).reshape(n, 3)
    parts = capi.Particles(n, coord=coord, vel=vel, mass=np.full(n, rho * volp), vol0=np.full(n, volp))
    floor = node_lists(nelem, lambda i, j, k, nx, ny, nz: j == 0)
    xwalls = node_lists(nelem, lambda i, j, k, nx, ny, nz: (i == 0) | (i == nx - 1) | (j == 0))
    zwalls = node_lists(nelem, lambda i, j, k, nx, ny, nz: (k == 0) | (k == nz - 1) | (j == 0))
    bcs = {0: _bc(xwalls), 1: _bc(floor), 2: _bc(zwalls)}
    return Scene("cfg3-j2-column-%dk" % (n // 1000), xmin, xmax, nelem,
                 [(capi.MODEL_J2, [2e6, 0.3, 2e3, 1e4])], rho, dt, steps, (0.0, -9.81, 0.0), bcs=bcs, parts=parts)


def tiny_j2(n_cells=12, dt=4e-5, sigma_y=2e3, shear=0.0):
    """Small J2 column for parity tests: the cfg3 column on an n_cells^3 grid
    of spacing 1/n_cells; `shear` adds an initial velocity v_x = shear * y."""
    h = 1.0 / n_cells
    nelem = (n_cells,) * 3
    xmin, xmax = (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)
    hi = (max(1, int(0.3 * n_cells)), max(1, int(0.6 * n_cells)), max(1, int(0.5 * n_cells)))
    coord = seed_block(xmin, (h, h, h), nelem, (1, 0, 1), (1 + hi[0], hi[1], 1 + hi[2]), (2, 2, 2))
    n = len(coord)
    rho = 2000.0
    volp = h ** 3 / 8.0
    vel = uniform(0x5EED0013, 3 * n, -1e-3, 1e-3).reshape(n, 3)
    vel[:, 0] += shear * coord[:, 1]
    parts = capi.Particles(n, coord=coord, vel=vel, mass=np.full(n, rho * volp), vol0=np.full(n, volp))
    floor = node_lists(nelem, lambda i, j, k, nx, ny, nz: j == 0)
    bcs = {0: _bc(floor), 1: _bc(floor), 2: _bc(floor)}
    return Scene("tiny-j2", xmin, xmax, nelem, [(capi.MODEL_J2, [2e6, 0.3, sigma_y, 1e4])], rho, dt, 10,
                 (0.0, -9.81, 0.0), bcs=bcs, parts=parts)


CONFIGS = {"cfg1": cfg1, "cfg2": cfg2, "cfg3": cfg3}


TAIT_PROPS = [1000.0, 30.0, 7.0, 1e-3]  # rho0, c, gamma, mu


def _walls(nelem, idimc=(1, 1, 1)):
    """Free-slip box: u = 0 on x walls, v = 0 on floor and ceiling, w = 0 on z walls."""
    xw = node_lists(nelem, lambda i, j, k, nx, ny, nz: (i == 0) | (i == nx - 1), idimc)
    yw = node_lists(nelem, lambda i, j, k, nx, ny, nz: (j == 0) | (j == ny - 1), idimc)
    zw = node_lists(nelem, lambda i, j, k, nx, ny, nz: (k == 0) | (k == nz - 1), idimc)
    return {0: _bc(xw), 1: _bc(yw), 2: _bc(zw)}


def cfg5(steps=2000, seed=0x5EED0005, ppc=(2, 2, 2)):
    """Dam break, weakly compressible Tait fluid (rho0 = 1000, c = 30,
    gamma = 7, mu = 1e-3): 256x128x128 cells of h = 1/128 on [0,2]x[0,1]x[0,1];
    water column [0,0.5]x[0,0.5]x[0,1] = 64x64x128 cells at 2x2x2 points per
    cell = 4,194,304 particles (ppc (4,2,4): 32 per cell, 16,777,216, the
    stated count); gravity -y; free-slip walls; dt = 2e-5."""
    h = 1.0 / 128.0
    nelem = (256, 128, 128)
    xmin, xmax = (0.0, 0.0, 0.0), (2.0, 1.0, 1.0)
    coord = seed_block(xmin, (h, h, h), nelem, (0, 0, 0), (64, 64, 128), ppc)
    n = len(coord)
    volp = h ** 3 / float(np.prod(ppc))
    rho = TAIT_PROPS[0]
    vel = uniform(seed, 3 * n, -1e-3, 1e-3).reshape(n, 3)
    parts = capi.Particles(n, coord=coord, vel=vel, mass=np.full(n, rho * volp), vol0=np.full(n, volp))
    return Scene("cfg5-tait-dam-break-%dk" % (n // 1000), xmin, xmax, nelem, [(capi.MODEL_TAIT, list(TAIT_PROPS))],
                 rho, 2e-5, steps, (0.0, -9.81, 0.0), bcs=_walls(nelem), parts=parts)


def tiny_tait(n_cells=12, dt=2e-5, idimc=(1, 1, 1)):
    """Small dam break for parity tests: water [0,1/4]x[0,1/2]x[0,1) in a unit
    box of n_cells^3 cells."""
    h = 1.0 / n_cells
    nelem = (n_cells,) * 3
    xmin, xmax = (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)
    coord = seed_block(xmin, (h, h, h), nelem, (0, 0, 0), (max(1, n_cells // 4), n_cells // 2, n_cells), (2, 2, 2))
    n = len(coord)
    volp = h ** 3 / 8.0
    rho = TAIT_PROPS[0]
    vel = uniform(0x5EED0015, 3 * n, -1e-3, 1e-3).reshape(n, 3)
    parts = capi.Particles(n, coord=coord, vel=vel, mass=np.full(n, rho * volp), vol0=np.full(n, volp))
    return Scene("tiny-tait", xmin, xmax, nelem, [(capi.MODEL_TAIT, list(TAIT_PROPS))], rho, dt, 10,
                 (0.0, -9.81, 0.0), bcs=_walls(nelem, idimc), parts=parts, idimc=tuple(idimc))


CONFIGS["cfg5"] = cfg5


def cfg4_scene(k, seed=0x5EED0004):
    """Scene k of the ensemble: a 2D elastic block dropped onto a fixed floor
    (reference slab form, vertical = z, v = 0 everywhere). Seeded per scene:
    grid G^2 with G in {64, 128, 256}; particle count N ~ U[2k, min(64k, G^2)]
    at 2x2 points per cell (a w_x x w_z block of cells, w_x*w_z*4 ~ N); Young's
    modulus log-uniform in [1e5, 1e7], nu ~ U[0.2, 0.45], linear elastic or
    Neo-Hookean (alternating), drop height (block bottom) ~ U[0.2, 0.6],
    rho = 1000; dt = 0.3 h / c with c the P-wave speed."""
    r = uniform(seed ^ (k * 0x9E3779B1), 8, 0.0, 1.0)
    G = (64, 128, 256)[min(2, int(r[0] * 3))]
    nmax = min(64000, G * G)
    N = 2000 + r[1] * (nmax - 2000)
    cells = max(1, int(N / 4))
    aspect = 0.5 + r[2]
    wx = max(1, min(G // 2, int(round(np.sqrt(cells * aspect)))))
    wz = max(1, min(G // 2, cells // wx))
    E = 10 ** (5 + 2 * r[3])
    nu = 0.2 + 0.25 * r[4]
    model = capi.MODEL_LINELAST if k % 2 == 0 else capi.MODEL_NEOHOOKEAN
    h = 1.0 / G
    z0 = int(round((0.2 + 0.4 * r[5]) * G))
    z0 = min(z0, G - wz - 2)
    x0 = (G - wx) // 2
    xmin, xmax, nelem = (0.0, 0.0, 0.0), (1.0, h, 1.0), (G, 1, G)
    coord = seed_block(xmin, (h, h, h), nelem, (x0, 0, z0), (x0 + wx, 1, z0 + wz), (2, 1, 2))
    n = len(coord)
    rho = 1000.0
    volp = h ** 3 / 4.0
    vel = uniform(seed + k, 3 * n, -1e-3, 1e-3).reshape(n, 3)
    vel[:, 1] = 0.0
    parts = capi.Particles(n, coord=coord, vel=vel, mass=np.full(n, rho * volp), vol0=np.full(n, volp))
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    c = np.sqrt((lam + 2 * mu) / rho)
    bottom = node_lists(nelem, lambda i, j, k_, nx, ny, nz: k_ == 0)
    allv = node_lists(nelem, lambda i, j, k_, nx, ny, nz: np.ones_like(i, dtype=bool))
    return Scene("cfg4-scene%03d-%d" % (k, G), xmin, xmax, nelem, [(model, [E, nu])], rho, 0.3 * h / c, 500,
                 (0.0, 0.0, -9.81), bcs={0: _bc(bottom), 1: _bc(allv), 2: _bc(bottom)}, parts=parts)


def cfg4(n_scenes=256, seed=0x5EED0004):
    """The ensemble: n_scenes seeded scenes (cfg4_scene); returns (scenes, bucket ids by grid size)."""
    sc = [cfg4_scene(k, seed) for k in range(n_scenes)]
    bucket = [{64: 0, 128: 1, 256: 2}[s.nelem[0]] for s in sc]
    return sc, bucket


def atlas(scene_list, name="atlas", per_scene_dt=True):
    """Packs 2D slab scenes of one grid (nelem (G, 1, G), same spacing, density
    and gravity) into one grid so that a single engine advances them all:
    scene k sits in tile (k mod T, k div T) of (G+1) x (G+1) cells, one empty
    gap cell between tiles so that no node is shared; each scene keeps its
    material (matid = k), its floor (the tile's bottom node row: u = w = 0)
    and v = 0 everywhere. Each scene keeps its own CFL time step as its
    material's time step (per_scene_dt, FP32 engine), or all take the
    smallest. Returns (scene, layout) with layout[k] = (first particle,
    count, origin)."""
    G = scene_list[0].nelem[0]
    h = scene_list[0].dx[0]
    for sc in scene_list:
        assert sc.nelem == (G, 1, G) and sc.dx[0] == h and sc.rho == scene_list[0].rho
    ns = len(scene_list)
    T = int(np.ceil(np.sqrt(ns)))
    rows = (ns + T - 1) // T
    stride = G + 1
    nelem = (T * stride, 1, rows * stride)
    xmin = (0.0, 0.0, 0.0)
    xmax = (nelem[0] * h, h, nelem[2] * h)
    coords, vels, masses, vols, mat, layout = [], [], [], [], [], []
    first = 0
    for k, sc in enumerate(scene_list):
        ox, oz = (k % T) * stride * h, (k // T) * stride * h
        c = sc.parts.coord.copy()
        c[:, 0] += ox
        c[:, 2] += oz
        coords.append(c)
        vels.append(sc.parts.vel)
        masses.append(sc.parts.mass)
        vols.append(sc.parts.vol0)
        mat.append(np.full(sc.n, k, np.int32))
        layout.append((first, sc.n, (ox, 0.0, oz)))
        first += sc.n
    n = first
    parts = capi.Particles(n, coord=np.concatenate(coords), vel=np.concatenate(vels), mass=np.concatenate(masses),
                           vol0=np.concatenate(vols), matid=np.concatenate(mat))
    floor = node_lists(nelem, lambda i, j, k_, nx, ny, nz: (k_ % stride) == 0)
    allv = node_lists(nelem, lambda i, j, k_, nx, ny, nz: np.ones_like(i, dtype=bool))
    mats = [sc.materials[0] for sc in scene_list]
    dt = min(sc.dt for sc in scene_list)
    mat_dt = [sc.dt for sc in scene_list] if per_scene_dt else None
    return (Scene(name, xmin, xmax, nelem, mats, scene_list[0].rho, dt, scene_list[0].steps, scene_list[0].bb,
                  bcs={0: _bc(floor), 1: _bc(allv), 2: _bc(floor)}, parts=parts, mat_dt=mat_dt), layout)


def cfg4_atlases(n_scenes=256, seed=0x5EED0004, per_scene_dt=True):
    """The ensemble packed per grid-size bucket: [(atlas scene, layout, member scene indices)]."""
    sc, bucket = cfg4(n_scenes, seed)
    out = []
    for b in range(3):
        idx = [i for i, x in enumerate(bucket) if x == b]
        if idx:
            a, lay = atlas([sc[i] for i in idx], "cfg4-atlas-%d" % (64 << b), per_scene_dt)
            out.append((a, lay, idx))
    return out
