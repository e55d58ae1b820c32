"""Batched search: every problem of a batch equals the plain search of that
problem with seed + p, bit for bit, whatever the persistent grid size."""
import numpy as np
import pytest

from pyhea_amd.hcs import Batch, Problem
from pyhea_amd.problem import make_problem

pytestmark = pytest.mark.gpu


def _specs():
    rng = np.random.default_rng(4)
    out = []
    for i in range(6):
        typ = ["FCC", "BCC"][i % 2]
        K = int(rng.integers(2, 7))
        out.append(make_problem(typ, int(rng.integers(2, 5)), rng.dirichlet([2.0] * K) + 0.05,
                                shells="true", target_seed=50 + i))
    return out


def _bits(a):
    return np.asarray(a, np.float64).view(np.uint64)


@pytest.mark.parametrize("grid", [None, 1, 3])
def test_batch_equals_plain(gpu, grid):
    specs = _specs()
    probs = [Problem.from_spec(s) for s in specs]
    L, I, T, D, seed = 5, 2, 32, 2, 900
    b = Batch(probs, L, T, D, seed=seed)
    if grid:
        b.set_grid(grid)
    b.init()
    b.step(I)
    for q, p in enumerate(probs):
        lb, fb, sb = b.read(q)
        lp, fp, sp = p.run_hcs(L, I, T, D, seed=seed + q)
        assert np.array_equal(lb, lp), q
        assert np.array_equal(_bits(fb), _bits(fp)), q
        assert sb["evals"] == sp["evals"] and sb["accepts"] == sp["accepts"], q
    b.close()


def test_batch_rejects_mixed_shell_counts(gpu):
    a = make_problem("FCC", 2, [0.5, 0.5], weight=[1.0, 1.0], shells="true", target_seed=1)
    c = make_problem("FCC", 2, [0.5, 0.5], weight=[1.0], shells="true", target_seed=1)
    with pytest.raises(ValueError):
        Batch([Problem.from_spec(a), Problem.from_spec(c)], 2, 8, 1)


@pytest.mark.parametrize("tile", [64, 128, 256])
def test_launch_shape_does_not_change_results(gpu, tile):
    specs = _specs()[:3]
    probs = [Problem.from_spec(s) for s in specs]
    L, I, T, D, seed = 4, 2, 200, 2, 31
    b = Batch(probs, L, T, D, seed=seed)
    b.tile(tile)
    b.set_grid(7)
    b.init()
    b.step(I)
    for q, p in enumerate(probs):
        s = p.session(L, T, D, seed=seed + q)
        s.tile(tile)
        s.init()
        s.step(I)
        ls, fs, _ = s.read()
        lp, fp, _ = p.run_hcs(L, I, T, D, seed=seed + q)  # default tile
        lb, fb, _ = b.read(q)
        assert np.array_equal(ls, lp) and np.array_equal(_bits(fs), _bits(fp))
        assert np.array_equal(lb, lp) and np.array_equal(_bits(fb), _bits(fp))
    b.close()


def test_autotuner_picks_a_valid_shape(gpu, tmp_path):
    from pyhea_amd.autotune import tune_batch

    probs = [Problem.from_spec(s) for s in _specs()[:2]]
    best = tune_batch(probs, 4, 64, 2, cache_path=str(tmp_path / "t.json"), grids=[4, 8],
                      tiles=[64, 128])
    assert best["tile"] in (64, 128) and best["grid"] in (4, 8)
    again = tune_batch(probs, 4, 64, 2, cache_path=str(tmp_path / "t.json"))
    assert again == best


@pytest.mark.parametrize("K", [5, 6, 8])
def test_uniform_species_batch_equals_plain(gpu, K):
    """Batches whose problems all have K equal to the batch kernel's table
    width (5, 6, 8): each problem equals its plain run, which uses the
    exact-K single-problem kernel."""
    rng = np.random.default_rng(K)
    specs = [make_problem(typ, cell, rng.dirichlet([2.0] * K) + 0.05, shells="true",
                          target_seed=80 + i)
             for i, (typ, cell) in enumerate([("FCC", 3), ("BCC", 4), ("FCC", 4)])]
    probs = [Problem.from_spec(s) for s in specs]
    L, I, T, D, seed = 4, 2, 64, 2, 910
    b = Batch(probs, L, T, D, seed=seed)
    b.init()
    b.step(I)
    for q, p in enumerate(probs):
        lb, fb, sb = b.read(q)
        lp, fp, sp = p.run_hcs(L, I, T, D, seed=seed + q)
        assert np.array_equal(lb, lp), q
        assert np.array_equal(_bits(fb), _bits(fp)), q
        assert sb["evals"] == sp["evals"] and sb["accepts"] == sp["accepts"], q
    b.close()


@pytest.mark.parametrize("K", [3, 5, 7])
@pytest.mark.parametrize("guided", [False, True])
@pytest.mark.parametrize("hbm", [False, True])
def test_single_problem_batch_equals_session(gpu, monkeypatch, K, guided, hbm):
    """A batch of one problem launches the per-lattice kernel (its arguments
    in the kernarg segment, exact table width where K allows) instead of the
    persistent one; it must still equal the session of that problem bit for
    bit, plain and guided, with species in LDS and in HBM (7 species: the
    width-8 kernel, plain only)."""
    if guided and K == 7:
        pytest.skip("guided mode needs K <= 6")
    fr = np.random.default_rng(K).dirichlet([2.0] * K) + 0.05
    spec = make_problem("FCC", 3, fr, shells="true", target_seed=60 + K)
    monkeypatch.setenv("PYHEA_AMD_SPECIES_IN_HBM", "1" if hbm else "0")
    p = Problem.from_spec(spec)
    monkeypatch.delenv("PYHEA_AMD_SPECIES_IN_HBM")
    L, I, T, D, seed = 6, 2, 48, 3, 910
    b = Batch([p], L, T, D, seed=seed)
    s = p.session(L, T, D, seed=seed)
    if guided:
        b.guide(candidates=12, tau=0.1, accepts=4)
        s.guide(candidates=12, tau=0.1, accepts=4)
    b.init()
    b.step(I)
    s.init()
    s.step(I)
    lb, fb, sb = b.read(0)
    ls, fs, ss = s.read()
    assert np.array_equal(lb, ls)
    assert np.array_equal(_bits(fb), _bits(fs))
    assert sb["evals"] == ss["evals"] and sb["accepts"] == ss["accepts"] and sb["accepts"] > 0
    b.close()
    s.close()
