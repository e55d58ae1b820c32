"""Periodic FCC/BCC supercells and their neighbour shells.

Mirrors ``pyhea/lattice/lattice.py`` (class ``lattice``): same site
numbering, ``id = b*(x*L^2 + y*L + z) + loc`` with the basis orders of
lattice.py:146-194 (index formulas lattice.py:105, 142), same neighbour
vectors per shell (lattice.py:207-221, 272-287) and the same ascending sort
of every shell list (lattice.py:236-320).  The reference builds the lists
with per-site Python loops over float16 coordinates; here the whole table is
computed in integer half-lattice units with numpy, so a 500,000-site lattice
takes about a second.

Shell sets:
  * ``"pyhea"`` -- exactly the reference's shells: BCC 8/6/12, FCC 12/6/12
    (its FCC "third shell" is the <110> shell, lattice.py:283-287).
  * ``"true"``  -- the first three coordination shells: BCC 8/6/12 (same as
    "pyhea"), FCC 12/6/24 with <1 1/2 1/2> as the third shell.
"""
from __future__ import annotations

import itertools

import numpy as np

# basis sites in half-lattice units, in the reference's order
_BASIS = {
    "BCC": np.array([[0, 0, 0], [1, 1, 1]], dtype=np.int64),
    "FCC": np.array([[0, 0, 0], [1, 1, 0], [1, 0, 1], [0, 1, 1]], dtype=np.int64),
}


def _signed_perms(base: tuple[int, int, int]) -> np.ndarray:
    out = set()
    for perm in itertools.permutations(base):
        for signs in itertools.product((1, -1), repeat=3):
            out.add(tuple(p * s for p, s in zip(perm, signs)))
    return np.array(sorted(out), dtype=np.int64)


# neighbour vectors in half-lattice units
_SHELLS = {
    ("BCC", "pyhea"): [(1, 1, 1), (2, 0, 0), (2, 2, 0)],
    ("BCC", "true"): [(1, 1, 1), (2, 0, 0), (2, 2, 0)],
    ("FCC", "pyhea"): [(1, 1, 0), (2, 0, 0), (2, 2, 0)],
    ("FCC", "true"): [(1, 1, 0), (2, 0, 0), (2, 1, 1)],
}


def shell_vectors(latt_typ: str, shells: str = "pyhea") -> list[np.ndarray]:
    """Neighbour vectors (in units of the lattice constant) of the 3 shells."""
    return [_signed_perms(v) / 2.0 for v in _SHELLS[(latt_typ, shells)]]


def site_positions(cell_dim: int, latt_typ: str) -> np.ndarray:
    """Half-unit integer positions [N, 3] in site-id order."""
    basis = _BASIS[latt_typ]
    L = int(cell_dim)
    cells = np.stack(np.meshgrid(np.arange(L), np.arange(L), np.arange(L), indexing="ij"), -1)
    cells = cells.reshape(-1, 3) * 2
    return (cells[:, None, :] + basis[None, :, :]).reshape(-1, 3)


def neighbor_table(cell_dim: int, latt_typ: str, shells: str = "pyhea"):
    """(nbr [N, sum Z] int32, shell_sizes) with each shell's columns ascending."""
    if latt_typ not in _BASIS:
        raise ValueError(f"Invalid lattice type: {latt_typ}")
    if shells not in ("pyhea", "true"):
        raise ValueError(f"Invalid shell set: {shells}")
    L = int(cell_dim)
    if L < 1:
        raise ValueError("cell_dim must be >= 1")
    basis = _BASIS[latt_typ]
    nb = len(basis)
    loc_of = np.full(8, -1, dtype=np.int64)  # fractional parity bits -> basis index
    for i, b in enumerate(basis):
        loc_of[b[0] * 4 + b[1] * 2 + b[2]] = i
    pos = site_positions(L, latt_typ)
    cols, sizes = [], []
    for v in _SHELLS[(latt_typ, shells)]:
        d = _signed_perms(v)
        q = np.mod(pos[:, None, :] + d[None, :, :], 2 * L)
        cell, frac = q // 2, q % 2
        loc = loc_of[frac[..., 0] * 4 + frac[..., 1] * 2 + frac[..., 2]]
        if (loc < 0).any():
            raise AssertionError("neighbour vector leaves the lattice")
        ids = nb * ((cell[..., 0] * L + cell[..., 1]) * L + cell[..., 2]) + loc
        cols.append(np.sort(ids, axis=1))
        sizes.append(d.shape[0])
    return np.ascontiguousarray(np.concatenate(cols, axis=1).astype(np.int32)), sizes


class lattice:  # noqa: N801  (reference class name)
    """Drop-in for ``pyhea.lattice.lattice`` (lattice.py:15-51).

    ``nbor`` is the flat int32 table; ``nbor_list`` gives the reference's
    nested form (list[N] of list[3] of arrays), built on first use.
    """

    def __init__(self, cell_dim, latt_typ, latt_con=1.0, latt_vec=None, valid=False,
                 shells: str = "pyhea"):
        self.latt_typ = latt_typ
        self.cell_dim = int(cell_dim[0] if np.ndim(cell_dim) else cell_dim)
        self.latt_con = float(latt_con)
        self.latt_vec = np.diag([self.cell_dim] * 3).astype(np.float64)
        self.shell_set = shells
        self.nbor, self.shell_sizes = neighbor_table(self.cell_dim, latt_typ, shells)
        self.coords = site_positions(self.cell_dim, latt_typ) / 2.0
        self._nested = None
        if valid:
            self.valid_nbors()

    @property
    def nbor_list(self):
        if self._nested is None:
            offs = np.cumsum([0] + list(self.shell_sizes))
            self._nested = [[row[offs[s]:offs[s + 1]] for s in range(3)] for row in self.nbor]
        return self._nested

    def calc_pbc_dist(self, a, b):
        """Minimum-image distance (lattice.py:63-74)."""
        L = self.cell_dim
        d = np.asarray(a, dtype=np.float64) - np.asarray(b, dtype=np.float64)
        d = d - L * np.round(d / L)
        return np.sqrt((d * d).sum(-1))

    def valid_nbors(self) -> bool:
        """Checks every listed neighbour lies at its shell's distance
        (lattice.py:326-433)."""
        offs = np.cumsum([0] + list(self.shell_sizes))
        vecs = shell_vectors(self.latt_typ, self.shell_set)
        ok = True
        for s in range(3):
            want = np.linalg.norm(vecs[s][0])
            nb = self.nbor[:, offs[s]:offs[s + 1]]
            dist = self.calc_pbc_dist(self.coords[:, None, :], self.coords[nb])
            ok &= bool(np.all(np.abs(dist - want) < 1e-2)) or self.cell_dim < 2
        return ok

    @property
    def shell1(self):
        return self.shell_sizes[0]

    @property
    def shell2(self):
        return self.shell_sizes[1]

    @property
    def shell3(self):
        return self.shell_sizes[2]

    @property
    def nbor_size(self):
        return int(sum(self.shell_sizes))

    @property
    def shells(self):
        return 3
