"""Search problems: lattice + species counts + shell weights + target SRO.

The synthetic workloads used by the tests and the benchmark are generated
here, with every random choice seeded:

* counts: fractions -> integer counts by the largest-remainder rule (each
  species keeps at least one atom);
* target alpha: per shell a symmetric K x K matrix whose off-diagonal entries
  are uniform in [-0.15, 0.05] and whose diagonal is fixed by the
  Warren-Cowley sum rule sum_j c_j alpha_ij = 0.

``target_from_lower_triangle`` expands the reference's TARGET_SRO file
layout (lower triangle, row by row; pyhea/io/input.py:423-458) to the full
row-major K x K form for any K (the reference stops at K = 4, input.py:457).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .lattice import neighbor_table


def largest_remainder(fractions, n: int) -> list[int]:
    f = np.asarray(fractions, dtype=np.float64)
    f = f / f.sum()
    raw = f * n
    base = np.floor(raw).astype(np.int64)
    base = np.maximum(base, 1)
    rem = n - int(base.sum())
    order = np.argsort(-(raw - np.floor(raw)), kind="stable")
    i = 0
    while rem > 0:
        base[order[i % len(base)]] += 1
        rem -= 1
        i += 1
    while rem < 0:  # only when the >= 1 floor pushed the sum over n
        j = int(np.argmax(base))
        base[j] -= 1
        rem += 1
    return [int(x) for x in base]


def random_target(counts, n_shells: int, seed: int, lo: float = -0.15, hi: float = 0.05):
    """[S][K*K] symmetric targets obeying sum_j c_j alpha_ij = 0."""
    c = np.asarray(counts, dtype=np.float64)
    c = c / c.sum()
    K = len(c)
    rng = np.random.default_rng(seed)
    out = np.zeros((n_shells, K * K))
    for s in range(n_shells):
        a = np.zeros((K, K))
        iu = np.triu_indices(K, 1)
        a[iu] = rng.uniform(lo, hi, size=len(iu[0]))
        a = a + a.T
        for i in range(K):
            a[i, i] = -sum(c[j] * a[i, j] for j in range(K) if j != i) / c[i]
        out[s] = a.reshape(-1)
    return out


def target_from_lower_triangle(shells, K: int) -> np.ndarray:
    """Reference TARGET_SRO layout -> [3][K*K] (missing shells are zero)."""
    full = np.zeros((3, K * K))
    for s, tri in enumerate(list(shells)[:3]):
        tri = np.asarray(tri, dtype=np.float64)
        if tri.size != K * (K + 1) // 2:
            raise ValueError(f"Shell {s + 1} has incorrect number of values. "
                             f"Expected {K * (K + 1) // 2}, got {tri.size}")
        for i in range(K):
            for j in range(K):
                hi_, lo_ = max(i, j), min(i, j)
                full[s, i * K + j] = tri[hi_ * (hi_ + 1) // 2 + lo_]
    return full


@dataclass
class ProblemSpec:
    latt_typ: str
    cell_dim: int
    counts: list
    weight: list
    target: np.ndarray          # [S][K*K]
    shells: str = "pyhea"
    nbr: np.ndarray = field(default=None, repr=False)
    shell_sizes: list = None

    def __post_init__(self):
        if self.nbr is None:
            nbr, sizes = neighbor_table(self.cell_dim, self.latt_typ, self.shells)
            S = len(self.weight)
            used = int(sum(sizes[:S]))
            self.nbr = np.ascontiguousarray(nbr[:, :used])
            self.shell_sizes = list(sizes[:S])
        self.target = np.ascontiguousarray(np.asarray(self.target, dtype=np.float64)
                                           [: len(self.weight)])

    @property
    def n_sites(self) -> int:
        return int(self.nbr.shape[0])

    @property
    def K(self) -> int:
        return len(self.counts)

    @property
    def S(self) -> int:
        return len(self.weight)

    @property
    def sum_z(self) -> int:
        return int(sum(self.shell_sizes))


def make_problem(latt_typ: str, cell_dim: int, fractions, weight=(4.0, 1.0, 0.5),
                 shells: str = "true", target_seed: int = 2, counts=None) -> ProblemSpec:
    n_basis = 2 if latt_typ == "BCC" else 4
    n = n_basis * cell_dim ** 3
    if counts is None:
        counts = largest_remainder(fractions, n)
    tgt = random_target(counts, len(weight), target_seed)
    return ProblemSpec(latt_typ, cell_dim, list(counts), list(weight), tgt, shells)


# The five workloads, each as (problem, run parameters).  Seeds: target seed
# 1000*c + 2, stream seed 1000*c + 1 for workload c.
def workload(c: int, shells: str | None = None):
    if c == 0:
        p = make_problem("FCC", 10, [0.25] * 4, shells=shells or "true", target_seed=2,
                         counts=[1000] * 4)
        run = dict(solutions=128, parallel_task=256, total_iter=10, converge_depth=10, seed=1)
    elif c == 1:
        p = make_problem("BCC", 16, [0.30, 0.25, 0.20, 0.15, 0.10], shells=shells or "pyhea",
                         target_seed=1002)
        run = dict(solutions=512, parallel_task=256, total_iter=50, converge_depth=10, seed=1001)
    elif c == 3:
        p = make_problem("FCC", 30, [0.2] * 5, shells=shells or "true", target_seed=3002)
        run = dict(solutions=256, parallel_task=512, total_iter=1, converge_depth=10, seed=3001)
    elif c == 4:
        p = make_problem("FCC", 50, [0.25, 0.20, 0.20, 0.15, 0.10, 0.10],
                         shells=shells or "true", target_seed=4002)
        run = dict(solutions=64, parallel_task=256, total_iter=1, converge_depth=10, seed=4001)
    else:
        raise ValueError(f"no single-problem workload {c}")
    return p, run


def batched_workload(seed: int = 2002, n_lattices: int = 64):
    """64 lattices of uneven size, species count and type (batched workload)."""
    rng = np.random.default_rng(seed)
    probs = []
    for i in range(n_lattices):
        typ = "FCC" if i < n_lattices // 2 else "BCC"
        edge = int(rng.integers(6, 21))
        K = int(rng.integers(3, 7))
        frac = rng.dirichlet([2.0] * K)
        probs.append(make_problem(typ, edge, frac, shells="true", target_seed=seed * 100 + i))
    return probs
