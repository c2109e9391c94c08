"""Seeded synthetic RVE generator writing the reference's input formats.

The reference ships one hand-made example and no generator; the workloads here are defined by
their parameters (sphere count, radius range, mesh density, materials, load), and this module turns
a parameter set plus an integer seed into a deterministic `Problem`:

* spheres: random sequential addition (RSA) in the unit cube with a minimum surface gap and a wall
  clearance, radii uniform in [rmin, rmax];
* boundary: the cube surface meshed with m x m four-node quads per face, shared nodes
  (NN = 6 m^2 + 2), node order per element (-1,-1), (1,-1), (1,1), (-1,1) so that d/dxi x d/deta
  points outward -- the convention of the reference example (BEMElasticHelper3D.cpp:218-262);
* boundary conditions: symmetric supports (u_x = 0 on x = 0, u_y = 0 on y = 0, u_z = 0 on z = 0,
  tangential tractions zero there), traction t_x = 1 on x = 1, other faces traction free; or a
  prescribed macro strain u = eps . x on every face;
* stress points: uniform in [0.05, 0.95]^3.

Random numbers come from splitmix64 implemented here, so every machine produces the same files.
"""
from __future__ import annotations

import math

import numpy as np

from .formats import Problem

MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / 9007199254740992.0))

    def randint(self, lo: int, hi: int) -> int:
        """Uniform integer in [lo, hi]."""
        return lo + self.next_u64() % (hi - lo + 1)


def cube_mesh(m: int):
    """Cube [0,1]^3 surface, m x m bilinear quads per face, shared nodes, outward normals.

    Returns (nodes (NN,3), conn (NE,4) one-based, face (NE,) in 0..5 = x0,x1,y0,y1,z0,z1)."""
    ids = {}
    nodes = []

    def nid(i, j, k):
        key = (i, j, k)
        if key not in ids:
            ids[key] = len(nodes)
            nodes.append((i / m, j / m, k / m))
        return ids[key]

    conn = []
    face = []
    # each face: fixed axis, fixed value, and two in-plane axes (u, v) with u x v = outward normal
    faces = [
        (0, 0, 2, 1),  # x = 0, normal -x: e_z x e_y = -e_x
        (0, m, 1, 2),  # x = 1, normal +x: e_y x e_z = +e_x
        (1, 0, 0, 2),  # y = 0, normal -y: e_x x e_z = -e_y
        (1, m, 2, 0),  # y = 1, normal +y: e_z x e_x = +e_y
        (2, 0, 1, 0),  # z = 0, normal -z: e_y x e_x = -e_z
        (2, m, 0, 1),  # z = 1, normal +z: e_x x e_y = +e_z
    ]
    for f, (ax, val, ua, va) in enumerate(faces):
        for a in range(m):
            for b in range(m):
                quad = []
                for (du, dv) in ((0, 0), (1, 0), (1, 1), (0, 1)):
                    c = [0, 0, 0]
                    c[ax] = val
                    c[ua] = a + du
                    c[va] = b + dv
                    quad.append(nid(*c) + 1)
                conn.append(quad)
                face.append(f)
    return (np.array(nodes, dtype=np.float64), np.array(conn, dtype=np.int32),
            np.array(face, dtype=np.int32))


def symmetric_support_bc(nodes, conn, face, load=1.0):
    """u_x=0 on x=0, u_y=0 on y=0, u_z=0 on z=0 (rollers), t_x=load on x=1, other tractions 0."""
    NN = nodes.shape[0]
    NE = conn.shape[0]
    bc = np.zeros((3 * NN, 2))
    bc[:, 0] = 1.0
    for i in range(NN):
        for d in range(3):
            if nodes[i, d] == 0.0:
                bc[3 * i + d, 0] = 0.0
    tbc = np.zeros((NE, 24))
    tbc[:, :12] = 1.0
    for e in range(NE):
        f = face[e]
        if f in (0, 2, 4):            # symmetry plane: normal traction unknown
            d = f // 2
            for k in range(4):
                tbc[e, 3 * k + d] = 0.0
        elif f == 1:                   # loaded face
            for k in range(4):
                tbc[e, 12 + 3 * k + 0] = load
    return bc, tbc


def macro_strain_bc(nodes, conn, eps):
    """Prescribed displacement u = eps . x on every node; all tractions unknown."""
    NN = nodes.shape[0]
    NE = conn.shape[0]
    bc = np.zeros((3 * NN, 2))
    u = nodes @ np.asarray(eps, dtype=np.float64).T
    bc[:, 1] = u.reshape(-1)
    tbc = np.zeros((NE, 24))
    return bc, tbc


def rsa_spheres(rng: SplitMix64, num: int, rmin: float, rmax: float, gap: float, wall: float,
                vf_cap: float | None = None, max_tries: int = 2000000):
    centres = np.zeros((num, 3))
    radii = np.zeros(num)
    placed = 0
    tries = 0
    vol = 0.0
    while placed < num:
        r = rng.uniform(rmin, rmax)
        if vf_cap is not None and vol + 4.0 / 3.0 * math.pi * r ** 3 > vf_cap:
            break
        lo = r + wall
        hi = 1.0 - r - wall
        ok = False
        while not ok:
            tries += 1
            if tries > max_tries:
                raise RuntimeError(f"RSA could not place sphere {placed} of {num}")
            c = np.array([rng.uniform(lo, hi) for _ in range(3)])
            if placed == 0:
                ok = True
                break
            dist = np.sqrt(((centres[:placed] - c) ** 2).sum(1))
            ok = bool(np.all(dist >= radii[:placed] + r + gap))
        centres[placed] = c
        radii[placed] = r
        vol += 4.0 / 3.0 * math.pi * r ** 3
        placed += 1
    return centres[:placed].copy(), radii[:placed].copy()


def make_rve(seed: int, num: int, rmin: float, rmax: float, m: int, *, gap: float = 0.01,
             wall: float = 0.02, E0: float = 1.0, nu0: float = 0.3, E1: float = 10.0,
             nu1: float = 0.25, nump: int = 1000, load: float = 1.0,
             title: str = "rve_symmetric_support") -> Problem:
    rng = SplitMix64(seed)
    centres, radii = rsa_spheres(rng, num, rmin, rmax, gap, wall)
    nodes, conn, face = cube_mesh(m)
    bc, tbc = symmetric_support_bc(nodes, conn, face, load)
    prng = SplitMix64(seed ^ 0x5EED5EED5EED5EED)
    points = np.array([[prng.uniform(0.05, 0.95) for _ in range(3)] for _ in range(nump)])
    points = points.reshape(nump, 3)
    return Problem(title, E0, nu0, E1, nu1, nodes, conn, bc, tbc, centres, radii, points)


# Workload definitions (sizes and materials of the benchmark configurations).
WORKLOADS = {
    # 64 spheres, 8x8 quads/face: n = 3*386 + 60*64 = 4998
    "rve64_m8": dict(num=64, rmin=0.04, rmax=0.06, m=8),
    # 1000 spheres, 16x16 quads/face: n = 3*1538 + 60*1000 = 64614
    "rve1000_m16": dict(num=1000, rmin=0.02, rmax=0.03, m=16),
    # 2500 spheres, 24x24 quads/face: n = 3*3458 + 60*2500 = 160374
    "rve2500_m24": dict(num=2500, rmin=0.015, rmax=0.022, m=24),
}


def make_workload(name: str, seed: int = 1, **over) -> Problem:
    kw = dict(WORKLOADS[name])
    kw.update(over)
    return make_rve(seed, **kw)


def unit_strain_tensor(v6) -> np.ndarray:
    """The symmetric tensor of six components in pair order 00 01 02 11 12 22."""
    e = np.zeros((3, 3))
    for (i, j), v in zip(((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)), v6):
        e[i, j] = e[j, i] = v
    return e


def rsa_spheres_np(rng: np.random.Generator, num: int, rmin: float, rmax: float, gap: float,
                   wall: float, vf_cap: float | None = None, tries: int = 20000, chunk: int = 256):
    """Random sequential addition with candidates drawn in chunks (numpy): each sphere takes the
    first non-overlapping candidate of its chunks; placement stops early when the volume-fraction cap
    would be exceeded or a sphere finds no place in `tries` candidates (a jammed packing)."""
    centres = np.zeros((num, 3))
    radii = np.zeros(num)
    placed, vol = 0, 0.0
    while placed < num:
        r = rng.uniform(rmin, rmax)
        if vf_cap is not None and vol + 4.0 / 3.0 * math.pi * r ** 3 > vf_cap:
            break
        lo, hi = r + wall, 1.0 - r - wall
        spot = None
        drawn, size = 0, 8  # chunks double while the candidates keep colliding
        while drawn < tries:
            cand = rng.uniform(lo, hi, size=(size, 3))
            drawn += size
            if placed == 0:
                spot = cand[0]
                break
            d2 = ((cand[:, None, :] - centres[None, :placed, :]) ** 2).sum(-1)
            ok = np.all(d2 >= (radii[None, :placed] + r + gap) ** 2, axis=1)
            if ok.any():
                spot = cand[int(np.argmax(ok))]
                break
            size = min(2 * size, chunk)
        if spot is None:
            break
        centres[placed] = spot
        radii[placed] = r
        vol += 4.0 / 3.0 * math.pi * r ** 3
        placed += 1
    return centres[:placed].copy(), radii[:placed].copy()


def make_batch_rve(seed: int, index: int, *, nmin: int = 8, nmax: int = 256, rmin: float = 0.03,
                   rmax: float = 0.08, vf_cap: float = 0.3, m: int = 8, gap: float = 0.005,
                   wall: float = 0.02, E0: float = 1.0, nu0: float = 0.3):
    """One RVE of the batched workload (seed, index): sphere count drawn uniform in [nmin, nmax]
    (fewer when the volume-fraction cap or a jammed packing stops the placement), radii uniform in
    [rmin, rmax], particle modulus log-uniform in [2, 100] x E0 and Poisson ratio uniform in
    [0.2, 0.35], an m x m mesh per face with the displacement u = eps . x prescribed on every node for
    a random unit macro strain eps (uniform direction in the six-component space). numpy's PCG64
    seeded with (seed, index) draws everything, so every machine builds the same RVEs. Returns
    (Problem, eps as 6 components)."""
    rng = np.random.default_rng([seed, index])
    num = int(rng.integers(nmin, nmax + 1))
    E1 = E0 * math.exp(rng.uniform(math.log(2.0), math.log(100.0)))
    nu1 = float(rng.uniform(0.2, 0.35))
    g = rng.standard_normal(6)
    eps6 = g / np.linalg.norm(g)
    centres, radii = rsa_spheres_np(rng, num, rmin, rmax, gap, wall, vf_cap=vf_cap)
    nodes, conn, face = cube_mesh(m)
    bc, tbc = macro_strain_bc(nodes, conn, unit_strain_tensor(eps6))
    prob = Problem(f"batch_rve_{seed}_{index}", E0, nu0, E1, nu1, nodes, conn, bc, tbc, centres, radii,
                   np.zeros((0, 3)))
    return prob, eps6


def _batch_item(args):
    seed, i, kw = args
    return make_batch_rve(seed, i, **kw)


def make_batch(seed: int, count: int, workers: int = 1, **kw):
    """count RVEs of the batched workload: [(Problem, eps6)]; with workers > 1 built in that many
    processes (each RVE depends only on (seed, index), so the result is the same)."""
    items = [(seed, i, kw) for i in range(count)]
    if workers <= 1:
        return [_batch_item(a) for a in items]
    from concurrent.futures import ProcessPoolExecutor
    with ProcessPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(_batch_item, items, chunksize=8))
