"""Neighbour lists: the product's closed-form builder against the
restatement of pyhea/lattice/lattice.py (float16 coordinates, per-site
loops) and against invariants and known answers of the reference's own
tests (tests/lattice/test_lattice.py)."""
import itertools

import numpy as np
import pytest

import lattice_ref
from pyhea_amd.lattice import lattice, neighbor_table, shell_vectors, site_positions


@pytest.mark.parametrize("typ", ["FCC", "BCC"])
@pytest.mark.parametrize("L", [1, 2, 3, 4, 5])
def test_matches_reference_restatement(typ, L):
    ref = lattice_ref.neighbor_list(L, typ)
    nbr, sizes = neighbor_table(L, typ, "pyhea")
    offs = np.cumsum([0] + sizes)
    for i, per in enumerate(ref):
        for s in range(3):
            assert np.array_equal(nbr[i, offs[s]:offs[s + 1]], per[s]), (i, s)


@pytest.mark.parametrize("typ,sizes", [("BCC", [8, 6, 12]), ("FCC", [12, 6, 12])])
def test_reference_shell_sizes(typ, sizes):
    # test_lattice.py:78-116: 8/6/12 (BCC), 12/6/12 (FCC), in range, no duplicates
    lat = lattice([4, 4, 4], typ, 3.6, None)
    assert lat.shell_sizes == sizes
    n = {"BCC": 2, "FCC": 4}[typ] * 64
    assert lat.nbor.shape == (n, sum(sizes))
    offs = np.cumsum([0] + sizes)
    for s in range(3):
        blk = lat.nbor[:, offs[s]:offs[s + 1]]
        assert blk.min() >= 0 and blk.max() < n
        assert all(len(set(r)) == sizes[s] for r in blk)


def test_atom_counts():
    # test_lattice.py:12-43
    for L in (1, 2, 3, 4):
        assert lattice([L] * 3, "BCC", 1.0, None).coords.shape == (2 * L ** 3, 3)
        assert lattice([L] * 3, "FCC", 1.0, None).coords.shape == (4 * L ** 3, 3)


def test_invalid_type():
    with pytest.raises(ValueError):
        lattice([2, 2, 2], "HCP", 1.0, None)


@pytest.mark.parametrize("typ,shells", [("FCC", "pyhea"), ("FCC", "true"), ("BCC", "true")])
@pytest.mark.parametrize("L", [3, 4])
def test_brute_force_distance_binning(typ, shells, L):
    """Every site at exactly the shell distance (minimum image) is listed."""
    pos = site_positions(L, typ) / 2.0
    nbr, sizes = neighbor_table(L, typ, shells)
    offs = np.cumsum([0] + sizes)
    radii = [np.linalg.norm(v[0]) for v in shell_vectors(typ, shells)]
    d = pos[:, None, :] - pos[None, :, :]
    d -= L * np.round(d / L)
    dist = np.sqrt((d * d).sum(-1))
    for s, r in enumerate(radii):
        want = [np.flatnonzero(np.abs(dist[i] - r) < 1e-9) for i in range(len(pos))]
        got = nbr[:, offs[s]:offs[s + 1]]
        for i in range(len(pos)):
            assert np.array_equal(np.sort(want[i]), got[i])


@pytest.mark.parametrize("typ,shells", [("FCC", "true"), ("BCC", "pyhea"), ("FCC", "pyhea")])
def test_symmetric_relation(typ, shells):
    nbr, sizes = neighbor_table(5, typ, shells)
    offs = np.cumsum([0] + sizes)
    for s in range(3):
        blk = nbr[:, offs[s]:offs[s + 1]]
        pairs = {(i, int(j)) for i in range(len(blk)) for j in blk[i]}
        assert all((j, i) in pairs for i, j in pairs)


def test_known_distances():
    # test_lattice.py:128-150: sqrt(3)/2, 1, sqrt(2) for BCC shells
    lat = lattice([4, 4, 4], "BCC", 1.0, None)
    offs = np.cumsum([0] + lat.shell_sizes)
    for s, want in enumerate([np.sqrt(3) / 2, 1.0, np.sqrt(2)]):
        j = lat.nbor[0, offs[s]]
        assert abs(lat.calc_pbc_dist(lat.coords[0], lat.coords[j]) - want) < 1e-9
    assert lat.valid_nbors()


def test_true_fcc_third_shell():
    nbr, sizes = neighbor_table(6, "FCC", "true")
    assert sizes == [12, 6, 24]
    lat = lattice([6, 6, 6], "FCC", 1.0, None, shells="true")
    assert lat.valid_nbors()
    assert lat.nbor_size == 42


def test_nested_form_matches_flat():
    lat = lattice([3, 3, 3], "FCC", 1.0, None)
    nested = lat.nbor_list
    offs = np.cumsum([0] + lat.shell_sizes)
    for i in (0, 17, len(nested) - 1):
        for s in range(3):
            assert list(nested[i][s]) == list(lat.nbor[i, offs[s]:offs[s + 1]])
