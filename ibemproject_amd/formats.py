"""Readers and writers for the reference's elastic-3D sphere input/output text formats.

Formats (configElastic3D.h:76-129, configElastic3D.cpp:32-197, postProcessorElastic3D.cpp:578-585):

* ``position.txt``  -- ``num``, ``E1``, ``nu1``, then ``x y z r`` per sphere.
* ``BEM.txt``       -- title line, ``E0``, ``nu0``, ``NN``, ``NE`` (one per line, read with getline),
  then NN x 3 node coordinates, NE x 4 one-based connectivity, 3NN x 2 displacement BC rows
  ``(flag, value)`` (flag 1 = displacement unknown, 0 = prescribed), NE x 24 traction rows
  (12 flags, 1 = traction prescribed, 0 = unknown; then 12 values), dofs node-major (node, xyz).
* ``postprocess.txt`` -- ``nump`` then ``x y z`` per point.
* ``stress.txt``    -- one row per point ``s00 s01 s02 s11 s12 s22`` in ``std::scientific``
  (6 significant digits).

The reference files mix ``\\r``, ``\\r\\r\\n`` and ``\\n`` line ends; the readers treat any of them as
whitespace, which is what the reference's ``operator>>`` does for the numeric body.
"""
from __future__ import annotations

import dataclasses
import os

import numpy as np


@dataclasses.dataclass
class Problem:
    """One RVE in the reference's input formats (all arrays float64 unless noted)."""

    title: str
    E0: float
    nu0: float
    E1: float
    nu1: float
    nodes: np.ndarray        # (NN, 3)
    conn: np.ndarray         # (NE, 4) int32, one-based like the reference
    bc: np.ndarray           # (3NN, 2)
    tbc: np.ndarray          # (NE, 24)
    centres: np.ndarray      # (num, 3)
    radii: np.ndarray        # (num,)
    points: np.ndarray       # (nump, 3)

    @property
    def NN(self) -> int:
        return int(self.nodes.shape[0])

    @property
    def NE(self) -> int:
        return int(self.conn.shape[0])

    @property
    def num(self) -> int:
        return int(self.centres.shape[0])

    @property
    def nump(self) -> int:
        return int(self.points.shape[0])

    @property
    def n78(self) -> int:
        return 3 * self.NN + 78 * self.num

    @property
    def n60(self) -> int:
        return 3 * self.NN + 60 * self.num


def _tokens(data: bytes) -> list[bytes]:
    return data.replace(b"\r", b" ").split()


def read_position(path: str):
    tok = _tokens(open(path, "rb").read())
    num = int(tok[0])
    E1 = float(tok[1])
    nu1 = float(tok[2])
    vals = np.array([float(t) for t in tok[3:3 + 4 * num]], dtype=np.float64)
    if vals.size != 4 * num:
        raise ValueError(f"{path}: expected {4 * num} sphere values, found {vals.size}")
    vals = vals.reshape(num, 4)
    return E1, nu1, vals[:, :3].copy(), vals[:, 3].copy()


def read_bem(path: str):
    raw = open(path, "rb").read()
    # header: five getline() lines (configElastic3D.cpp:189-209); split on '\n' and strip '\r'
    lines = raw.split(b"\n")
    head = []
    pos = 0
    consumed = 0
    while len(head) < 5:
        if pos >= len(lines):
            raise ValueError(f"{path}: truncated header")
        ln = lines[pos]
        consumed += len(ln) + 1
        pos += 1
        s = ln.replace(b"\r", b"").strip()
        if len(head) == 0:
            head.append(s.decode(errors="replace"))
        elif s:
            head.append(s.decode())
        else:
            # the reference's getline would return an empty line here and stod would throw;
            # tolerate blank lines between header fields
            continue
    title = head[0]
    E0 = float(head[1])
    nu0 = float(head[2])
    NN = int(head[3])
    NE = int(head[4])
    tok = _tokens(raw[consumed:])
    need = 3 * NN + 4 * NE + 6 * NN + 24 * NE
    if len(tok) < need:
        raise ValueError(f"{path}: expected {need} numbers after the header, found {len(tok)}")
    v = np.array([float(t) for t in tok[:need]], dtype=np.float64)
    p = 0
    nodes = v[p:p + 3 * NN].reshape(NN, 3).copy()
    p += 3 * NN
    conn = v[p:p + 4 * NE].reshape(NE, 4).astype(np.int32)
    p += 4 * NE
    bc = v[p:p + 6 * NN].reshape(3 * NN, 2).copy()
    p += 6 * NN
    tbc = v[p:p + 24 * NE].reshape(NE, 24).copy()
    return title, E0, nu0, nodes, conn, bc, tbc


def read_postprocess(path: str) -> np.ndarray:
    tok = _tokens(open(path, "rb").read())
    nump = int(tok[0])
    v = np.array([float(t) for t in tok[1:1 + 3 * nump]], dtype=np.float64)
    if v.size != 3 * nump:
        raise ValueError(f"{path}: expected {3 * nump} point values, found {v.size}")
    return v.reshape(nump, 3)


def read_problem(directory: str) -> Problem:
    E1, nu1, centres, radii = read_position(os.path.join(directory, "position.txt"))
    title, E0, nu0, nodes, conn, bc, tbc = read_bem(os.path.join(directory, "BEM.txt"))
    pp = os.path.join(directory, "postprocess.txt")
    points = read_postprocess(pp) if os.path.exists(pp) else np.zeros((0, 3))
    return Problem(title, E0, nu0, E1, nu1, nodes, conn, bc, tbc, centres, radii, points)


def _g(x: float) -> str:
    return "%.17g" % x


def write_problem(prob: Problem, directory: str) -> None:
    """Write position.txt, BEM.txt, postprocess.txt and calculationType.txt (type 1)."""
    os.makedirs(directory, exist_ok=True)
    with open(os.path.join(directory, "calculationType.txt"), "w") as f:
        f.write("1\n")
    with open(os.path.join(directory, "position.txt"), "w") as f:
        f.write(f"{prob.num}\n{_g(prob.E1)}\n{_g(prob.nu1)}\n")
        for c, r in zip(prob.centres, prob.radii):
            f.write(f"{_g(c[0])} {_g(c[1])} {_g(c[2])} {_g(r)}\n")
    with open(os.path.join(directory, "BEM.txt"), "w") as f:
        f.write(f"{prob.title}\n{_g(prob.E0)}\n{_g(prob.nu0)}\n{prob.NN}\n{prob.NE}\n")
        for x in prob.nodes:
            f.write(f"{_g(x[0])} {_g(x[1])} {_g(x[2])}\n")
        for c in prob.conn:
            f.write(f"{int(c[0])} {int(c[1])} {int(c[2])} {int(c[3])}\n")
        for row in prob.bc:
            f.write(f"{int(row[0])} {_g(row[1])}\n")
        for row in prob.tbc:
            f.write(" ".join([str(int(v)) for v in row[:12]] + [_g(v) for v in row[12:]]) + "\n")
    with open(os.path.join(directory, "postprocess.txt"), "w") as f:
        f.write(f"{prob.nump}\n")
        for x in prob.points:
            f.write(f"{_g(x[0])} {_g(x[1])} {_g(x[2])}\n")


def write_stress(path: str, s: np.ndarray) -> None:
    """stress.txt as the reference writes it: std::scientific, 6 digits, space separated."""
    with open(path, "w") as f:
        for row in s:
            f.write(" ".join("%.6e" % v for v in row) + "\n")


def read_stress(path: str) -> np.ndarray:
    rows = []
    for ln in open(path, "rb").read().replace(b"\r", b"\n").split(b"\n"):
        t = ln.split()
        if len(t) == 6:
            rows.append([float(v) for v in t])
    return np.array(rows, dtype=np.float64)
