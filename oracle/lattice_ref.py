"""TEST INFRASTRUCTURE ONLY -- restatement of the reference neighbour build.

Follows pyhea/lattice/lattice.py step by step, including its float16
coordinates: supercell coordinates in basis order (lattice.py:146-194), the
periodic wrap ``c - floor(c / L) * L`` (lattice.py:53-60), the site index
from integer part and fractional basis offset (lattice.py:77-144) and the
ascending sort of every shell (lattice.py:196-323).  Pure Python loops: use
on small cells only.  Used by tests/test_lattice.py to pin the product's
closed-form builder (pyhea_amd/lattice.py).
"""
from __future__ import annotations

import numpy as np

_BCC_SHELLS = [
    [(s1 * 0.5, s2 * 0.5, s3 * 0.5) for s1 in (1, -1) for s2 in (1, -1) for s3 in (1, -1)],
    [(1.0, 0.0, 0.0), (-1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, -1.0, 0.0), (0.0, 0.0, 1.0),
     (0.0, 0.0, -1.0)],
    None,  # filled below: <110>
]
_XY = [(1.0, 1.0), (-1.0, 1.0), (1.0, -1.0), (-1.0, -1.0)]
_110 = ([(a, b, 0.0) for a, b in _XY] + [(a, 0.0, b) for a, b in _XY]
        + [(0.0, a, b) for a, b in _XY])
_BCC_SHELLS[2] = _110
_FCC_SHELLS = [
    [(a * 0.5, b * 0.5, 0.0) for a, b in _XY] + [(a * 0.5, 0.0, b * 0.5) for a, b in _XY]
    + [(0.0, a * 0.5, b * 0.5) for a, b in _XY],
    _BCC_SHELLS[1],
    _110,
]
_BASIS = {
    "BCC": [(0.0, 0.0, 0.0), (0.5, 0.5, 0.5)],
    "FCC": [(0.0, 0.0, 0.0), (0.5, 0.5, 0.0), (0.5, 0.0, 0.5), (0.0, 0.5, 0.5)],
}


def coords(cell_dim: int, latt_typ: str) -> np.ndarray:
    basis = _BASIS[latt_typ]
    out = []
    for ix in range(cell_dim):
        for iy in range(cell_dim):
            for iz in range(cell_dim):
                for bx, by, bz in basis:
                    out.append([ix + bx, iy + by, iz + bz])
    return np.array(out, dtype=np.float16)


def _pbc(c: np.ndarray, L: np.ndarray) -> np.ndarray:
    return c - np.floor(c / L) * L


def _site_id(c, cell_dim: int, latt_typ: str) -> int:
    basis = _BASIS[latt_typ]
    nb = len(basis)
    frac = tuple(float(np.mod(v, 1)) for v in c)
    loc = None
    for i, b in enumerate(basis):
        if frac == b:
            loc = i
    if loc is None:
        raise ValueError(f"Invalid coordinate for {latt_typ} lattice: {c}")
    idx = nb * (int(c[0]) * cell_dim ** 2 + int(c[1]) * cell_dim + int(c[2])) + loc
    return idx % (nb * cell_dim ** 3)


def neighbor_list(cell_dim: int, latt_typ: str):
    """Reference nested form: list[N] of [shell1, shell2, shell3] arrays."""
    cs = coords(cell_dim, latt_typ)
    L = np.array([cell_dim] * 3, dtype=np.float16)
    shells = _BCC_SHELLS if latt_typ == "BCC" else _FCC_SHELLS
    out = []
    for i in range(len(cs)):
        per = []
        for vecs in shells:
            ids = [_site_id(_pbc(cs[i] + np.array(d, dtype=np.float16), L), cell_dim, latt_typ)
                   for d in vecs]
            per.append(np.sort(ids))
        out.append(per)
    return out
