"""Output formats and an independent Warren-Cowley analyser.

* ``write_poscar`` -- VASP POSCAR as written by ``pyhea/io/output.py:64-107``
  (header "PyHEA v...", scale, cell vectors, species relabelled A, B, C...,
  counts, Cartesian coordinates grouped by species).
* ``write_lammps`` -- LAMMPS "atomic" data file in the layout dpdata writes
  for the reference (``output.py:14-62``; compare pyhea's examples/conf.lmp),
  without the dpdata dependency.
* ``read_lammps`` -- positions and types back from such a file.
* ``warren_cowley`` -- alpha_sab = 1 - n_sab / (Z_s N c_a c_b) from integer
  pair counts; replaces the OVITO WarrenCowleyParameters post-analysis
  (``pyhea/utils/analyze.py:12-45``) for validating outputs.
* ``species_from_positions`` -- maps Cartesian positions (in units of a/2)
  to this package's site numbering, so external structures can be analysed.
"""
from __future__ import annotations

import numpy as np

from .lattice import site_positions

__version__ = "1.0.0"


def warren_cowley(species, nbr, shell_sizes, n_species: int) -> np.ndarray:
    """alpha [S][K][K] of one lattice (float64, from integer counts)."""
    sp = np.asarray(species, dtype=np.int64)
    N = sp.shape[0]
    K = int(n_species)
    conc = np.bincount(sp, minlength=K) / N
    offs = np.cumsum([0] + list(shell_sizes))
    out = np.zeros((len(shell_sizes), K, K))
    for s, z in enumerate(shell_sizes):
        nb = np.asarray(nbr)[:, offs[s]:offs[s + 1]]
        pair = sp[:, None] * K + sp[nb]
        n = np.bincount(pair.ravel(), minlength=K * K).reshape(K, K).astype(np.float64)
        out[s] = 1.0 - n / (z * N * conc[:, None] * conc[None, :])
    return out


def fitness_of_alpha(alpha, target, weight) -> float:
    """sqrt(sum_s w_s sum_ij (alpha - target)^2) (the search objective)."""
    a = np.asarray(alpha)
    t = np.asarray(target).reshape(a.shape[0], a.shape[1], a.shape[2])
    return float(np.sqrt(sum(w * ((a[s] - t[s]) ** 2).sum() for s, w in enumerate(weight))))


def species_from_positions(half_pos, types, cell_dim: int, latt_typ: str) -> np.ndarray:
    """Species per site (this package's numbering) from integer positions in
    half-lattice units and per-atom types."""
    L = int(cell_dim)
    ref = site_positions(L, latt_typ)
    key = lambda p: (p[:, 0] % (2 * L)) * (4 * L * L) + (p[:, 1] % (2 * L)) * (2 * L) + (
        p[:, 2] % (2 * L))
    order = {int(k): i for i, k in enumerate(key(ref))}
    out = np.full(len(ref), -1, dtype=np.int64)
    for k, t in zip(key(np.asarray(half_pos, dtype=np.int64)), types):
        out[order[int(k)]] = int(t)
    if (out < 0).any():
        raise ValueError("positions do not cover every lattice site")
    return out.astype(np.uint8)


def write_poscar(species, latt, n_species: int, path: str) -> None:
    sp = np.asarray(species)
    counts = np.bincount(sp, minlength=n_species)
    labels = [chr(65 + i) for i in range(n_species)]
    with open(path, "w") as f:
        f.write(f"PyHEA v{__version__}\n")
        f.write(f"{latt.latt_con}\n")
        for row in np.asarray(latt.latt_vec, dtype=np.float64):
            f.write("\t".join(str(v) for v in row) + "\n")
        f.write("   ".join(labels) + "\n")
        f.write("   ".join(str(int(c)) for c in counts) + "\n")
        f.write("Cartesian\n")
        for t in range(n_species):
            for i in np.flatnonzero(sp == t):
                c = latt.coords[i]
                f.write(f"{c[0]}\t{c[1]}\t{c[2]}\n")


def write_lammps(species, latt, n_species: int, path: str) -> None:
    sp = np.asarray(species)
    a = float(latt.latt_con)
    box = latt.cell_dim * a
    order = np.argsort(sp, kind="stable")
    with open(path, "w") as f:
        f.write(f"{len(sp)} atoms\n{n_species} atom types\n")
        for ax in ("x", "y", "z"):
            f.write(f"{0.0:15.10f}{box:15.10f} {ax}lo {ax}hi\n")
        f.write(f"{0.0:15.10f}{0.0:15.10f}{0.0:15.10f} xy xz yz\n\nAtoms # atomic\n\n")
        for n, i in enumerate(order, start=1):
            c = np.asarray(latt.coords[i], dtype=np.float64) * a
            f.write(f"{n:6d} {int(sp[i]) + 1:6d} {c[0]:15.10f} {c[1]:15.10f} {c[2]:15.10f}\n")


def read_lammps(path: str):
    """(positions [n][3] float64, types [n] uint8 zero-based, box length)."""
    pos, typ, box = [], [], None
    with open(path) as f:
        for line in f:
            p = line.split()
            if len(p) == 4 and p[2] == "xlo":
                box = float(p[1]) - float(p[0])
            if len(p) == 5 and p[0].isdigit():
                typ.append(int(p[1]) - 1)
                pos.append([float(v) for v in p[2:]])
    return np.array(pos), np.array(typ, np.uint8), box
