"""Output formats and the Warren-Cowley analyser, pinned by the reference's
own converged example structure (examples/conf.lmp, via the fixture
tests/golden/conf_fcc10.npz): its first-shell alpha equals the example
TARGET_SRO to two decimals."""
import os

import numpy as np
import pytest

from oracle import Oracle
from pyhea_amd.lattice import lattice, neighbor_table
from pyhea_amd.output import (fitness_of_alpha, read_lammps, species_from_positions,
                              warren_cowley, write_lammps, write_poscar)

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def conf():
    d = np.load(os.path.join(GOLD, "conf_fcc10.npz"))
    sp = species_from_positions(d["positions"], d["types"], 10, "FCC")
    return sp, d["target1"]


def test_conf_lmp_first_shell_matches_target(conf):
    sp, target = conf
    nbr, sizes = neighbor_table(10, "FCC", "pyhea")
    alpha = warren_cowley(sp, nbr, sizes, 4)
    assert np.abs(alpha[0] - target).max() < 0.01
    # Warren-Cowley sum rule for every shell
    c = np.bincount(sp, minlength=4) / len(sp)
    assert np.allclose(alpha @ c, 0.0, atol=1e-12)


def test_analyser_agrees_with_oracle_fitness(conf):
    sp, target = conf
    nbr, sizes = neighbor_table(10, "FCC", "true")
    tgt = np.zeros((3, 16))
    tgt[0] = target.reshape(-1)
    w = [4.0, 1.0, 0.5]
    alpha = warren_cowley(sp, nbr, sizes, 4)
    o = Oracle(nbr, sizes, np.bincount(sp, minlength=4), w, tgt)
    assert abs(o.fitness(sp)[0] - fitness_of_alpha(alpha, tgt, w)) < 1e-12


def test_lammps_round_trip(tmp_path, conf):
    sp, _ = conf
    latt = lattice([10, 10, 10], "FCC", 3.615234375, None)
    f = str(tmp_path / "out.lmp")
    write_lammps(sp, latt, 4, f)
    pos, typ, box = read_lammps(f)
    assert abs(box - 36.15234375) < 1e-9
    back = species_from_positions(np.rint(pos / (3.615234375 / 2)).astype(int), typ, 10, "FCC")
    assert np.array_equal(back, sp)
    head = open(f).read().splitlines()[:8]
    assert head[0] == "4000 atoms" and head[1] == "4 atom types" and head[7] == "Atoms # atomic"


def test_poscar_layout(tmp_path):
    latt = lattice([2, 2, 2], "BCC", 2.87, None)
    sp = np.array([0, 1] * 8, np.uint8)
    f = str(tmp_path / "POSCAR")
    write_poscar(sp, latt, 2, f)
    lines = open(f).read().splitlines()
    assert lines[0].startswith("PyHEA v") and lines[1] == "2.87"
    assert lines[5].split() == ["A", "B"] and lines[6].split() == ["8", "8"]
    assert lines[7] == "Cartesian" and len(lines) == 8 + 16
