"""GPU parity: the HIP path against the oracle (and the reference engine
itself where oracle/_ref/ref_hcs was built), bit for bit."""
import numpy as np
import pytest

from oracle import Oracle, RefHarness
from pyhea_amd.hcs import Problem
from pyhea_amd.problem import make_problem

pytestmark = pytest.mark.gpu

CASES = [
    # (type, cell, fractions, shells, n_weights)
    ("FCC", 3, [0.4, 0.3, 0.3], "pyhea", 3),
    ("BCC", 4, [0.5, 0.25, 0.25], "true", 3),
    ("FCC", 2, [0.5, 0.5], "true", 3),
    ("FCC", 1, [0.5, 0.5], "pyhea", 3),
    ("BCC", 3, [0.2, 0.2, 0.2, 0.2, 0.2], "pyhea", 2),
    ("FCC", 4, [0.3, 0.2, 0.2, 0.1, 0.1, 0.1], "true", 3),
    ("BCC", 5, [0.25, 0.25, 0.25, 0.25], "pyhea", 1),
    ("FCC", 3, [0.15] * 6 + [0.1], "true", 3),
    ("BCC", 4, [0.125] * 8, "true", 2),
]


def _spec(case, seed=7):
    typ, cell, fr, shells, S = case
    return make_problem(typ, cell, fr, weight=[4.0, 1.0, 0.5][:S], shells=shells,
                        target_seed=seed)


def _bits(a):
    return np.asarray(a, dtype=np.float64).view(np.uint64)


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}{c[1]}K{len(c[2])}S{c[4]}")
def test_generate_and_fitness(gpu, case):
    spec = _spec(case)
    o, p = Oracle.from_spec(spec), Problem.from_spec(spec)
    lat, fit = p.generate(6, seed=11, iter=3, phase=1)
    ref = np.stack([o.permute(11, 3, 1, l) for l in range(6)])
    assert np.array_equal(lat, ref)
    assert np.array_equal(_bits(fit), _bits(o.fitness(ref)))
    assert np.array_equal(p.pair_counts(lat), np.stack([o.pair_counts(x) for x in lat]))


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}{c[1]}K{len(c[2])}S{c[4]}")
@pytest.mark.parametrize("task,depth", [(1, 1), (4, 3), (8, 2), (64, 3), (300, 2)])
def test_local_mc_bit_exact(gpu, case, task, depth):
    spec = _spec(case)
    o, p = Oracle.from_spec(spec), Problem.from_spec(spec)
    lat0 = np.stack([o.permute(5, 0, 0, l) for l in range(5)])
    fit0 = o.fitness(lat0)
    lo, fo, so = o.local_mc(lat0, fit0, task, depth, seed=21, iter=2, phase=2)
    lg, fg, sg = p.local_mc(lat0, fit0, task, depth, seed=21, iter=2, phase=2)
    assert np.array_equal(_bits(fg), _bits(fo))
    assert np.array_equal(lg, lo)
    for k in ("evals", "rounds", "accepts", "draws"):
        assert sg[k] == so[k], k


def test_local_mc_matches_reference_engine(gpu):
    if not RefHarness.available():
        pytest.skip("reference harness not built")
    spec = _spec(CASES[0])
    o, p = Oracle.from_spec(spec), Problem.from_spec(spec)
    lat0 = np.stack([o.permute(5, 0, 0, l) for l in range(4)])
    fit0 = o.fitness(lat0)
    lr, fr, words, _ = RefHarness(o).local_mc(lat0, fit0, 16, 3, 9, 0, 3)
    lg, fg, sg = p.local_mc(lat0, fit0, 16, 3, seed=9, iter=0, phase=3)
    assert np.array_equal(lg, lr) and np.array_equal(_bits(fg), _bits(fr))
    assert 2 * sg["draws"] == int(words.sum())


@pytest.mark.parametrize("case", CASES[:6], ids=lambda c: f"{c[0]}{c[1]}K{len(c[2])}S{c[4]}")
def test_run_hcs_bit_exact(gpu, case):
    spec = _spec(case)
    o, p = Oracle.from_spec(spec), Problem.from_spec(spec)
    lo, fo, _, so = o.run_hcs(6, 3, 16, 2, seed=77)
    lg, fg, sg = p.run_hcs(6, 3, 16, 2, seed=77)
    assert np.array_equal(_bits(fg), _bits(fo))
    assert np.array_equal(lg, lo)
    assert sg["evals"] == so["evals"] and sg["accepts"] == so["accepts"]


# Equimolar two- and three-species lattices: many proposals tie exactly and,
# near convergence, the best ones differ in the last bits, which is where the
# screen's window must still hold the reference's choice.
TIE_CASES = [
    ("BCC", 6, [0.5, 0.5], "true", 3),
    ("FCC", 4, [0.5, 0.5], "pyhea", 3),
    ("FCC", 4, [1 / 3] * 3, "true", 3),
    ("BCC", 5, [0.5, 0.5], "pyhea", 1),
    ("FCC", 6, [0.25] * 4, "true", 3),
]


@pytest.mark.parametrize("case", TIE_CASES, ids=lambda c: f"{c[0]}{c[1]}K{len(c[2])}S{c[4]}")
def test_local_mc_to_convergence_bit_exact(gpu, case):
    spec = _spec(case, seed=13)
    o, p = Oracle.from_spec(spec), Problem.from_spec(spec)
    lat0 = np.stack([o.permute(31, 0, 0, l) for l in range(4)])
    fit0 = o.fitness(lat0)
    lo, fo, so = o.local_mc(lat0, fit0, 256, 8, seed=31, iter=0, phase=2)
    lg, fg, sg = p.local_mc(lat0, fit0, 256, 8, seed=31, iter=0, phase=2)
    assert np.array_equal(_bits(fg), _bits(fo))
    assert np.array_equal(lg, lo)
    for k in ("evals", "rounds", "accepts", "draws"):
        assert sg[k] == so[k], k
    assert so["rounds"] > 8 * 4  # every lattice ran past its first rejections


@pytest.mark.parametrize("case,task", [
    # near the largest round whose per-proposal records fit 160 KiB of LDS
    # (about 80 bytes per proposal), seven passes of 256
    (("BCC", 4, [0.5, 0.3, 0.2], "true", 1), 1792),
    (("FCC", 3, [0.4, 0.3, 0.3], "true", 3), 1200),  # five passes per round, three shells
])
def test_local_mc_largest_rounds_bit_exact(gpu, case, task):
    spec = _spec(case, seed=17)
    o, p = Oracle.from_spec(spec), Problem.from_spec(spec)
    lat0 = np.stack([o.permute(41, 0, 0, l) for l in range(3)])
    fit0 = o.fitness(lat0)
    lo, fo, so = o.local_mc(lat0, fit0, task, 2, seed=41, iter=0, phase=2)
    lg, fg, sg = p.local_mc(lat0, fit0, task, 2, seed=41, iter=0, phase=2)
    assert np.array_equal(_bits(fg), _bits(fo))
    assert np.array_equal(lg, lo)
    for k in ("evals", "rounds", "accepts", "draws"):
        assert sg[k] == so[k], k


def test_round_too_large_for_lds_is_an_error(gpu):
    """Per-proposal walk records of 2048 proposals in three shells do not fit
    in 160 KiB of LDS: the call fails cleanly, it does not launch."""
    spec = _spec(("FCC", 3, [0.4, 0.3, 0.3], "true", 3))
    p = Problem.from_spec(spec)
    lat0, fit0 = p.generate(2, seed=3)
    with pytest.raises(ValueError, match="LDS"):
        p.local_mc(lat0, fit0, 2048, 2, seed=3, iter=0, phase=2)
