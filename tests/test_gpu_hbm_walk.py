"""Species read from HBM: lattices too large for LDS keep their species in
HBM, and the screen walk then takes each proposal's neighbour species
histograms from per-site neighbour records (one 64-bit word per site, kept
up to date by every accepted swap) instead of gathering every neighbour's
species.  The path is forced on small lattices with
PYHEA_AMD_SPECIES_IN_HBM=1 (read when a problem is created);
PYHEA_AMD_WALK_GATHER=1 keeps the gathering walk.  Both walks must
reproduce the oracle bit for bit (oracle/hcs_oracle.cpp, restating
local_parallel_monte_carlo, accelerate.cpp:601-694): the records carry the
same integer counts, so every decision, draw and stored fitness bit is the
same.  Cells whose sites list themselves and problems whose records do not
fit one word (K = 7, 8) keep the gathering walk; they are included as
controls."""
import numpy as np
import pytest

from oracle import Oracle
from pyhea_amd.hcs import Batch, Problem
from pyhea_amd.problem import make_problem

pytestmark = pytest.mark.gpu

CASES = [
    # (type, cell, fractions, shells, n_weights)
    ("FCC", 3, [0.4, 0.3, 0.3], "pyhea", 3),
    ("BCC", 4, [0.5, 0.25, 0.25], "true", 3),
    ("BCC", 3, [0.2, 0.2, 0.2, 0.2, 0.2], "pyhea", 2),
    ("FCC", 4, [0.3, 0.2, 0.2, 0.1, 0.1, 0.1], "true", 3),
    ("BCC", 5, [0.25, 0.25, 0.25, 0.25], "pyhea", 1),
    ("FCC", 1, [0.5, 0.5], "pyhea", 3),              # self neighbours: gathering walk
    ("FCC", 3, [0.15] * 6 + [0.1], "true", 3),       # K = 7: gathering walk
]
WALKS = ["records", "gather"]


def _id(c):
    return f"{c[0]}{c[1]}K{len(c[2])}S{c[4]}"


def _spec(case, seed=7):
    typ, cell, fr, shells, S = case
    return make_problem(typ, cell, fr, weight=[4.0, 1.0, 0.5][:S], shells=shells,
                        target_seed=seed)


def _bits(a):
    return np.asarray(a, dtype=np.float64).view(np.uint64)


def _problem(monkeypatch, spec, hbm=True, walk="records"):
    monkeypatch.setenv("PYHEA_AMD_SPECIES_IN_HBM", "1" if hbm else "0")
    monkeypatch.setenv("PYHEA_AMD_WALK_GATHER", "1" if walk == "gather" else "0")
    p = Problem.from_spec(spec)
    monkeypatch.delenv("PYHEA_AMD_SPECIES_IN_HBM")
    monkeypatch.delenv("PYHEA_AMD_WALK_GATHER")
    return p


@pytest.mark.parametrize("walk", WALKS)
@pytest.mark.parametrize("case", CASES, ids=_id)
@pytest.mark.parametrize("task,depth", [(8, 2), (64, 3), (300, 2)])
def test_local_mc_species_in_hbm_bit_exact(gpu, monkeypatch, case, walk, task, depth):
    spec = _spec(case)
    o, p = Oracle.from_spec(spec), _problem(monkeypatch, spec, walk=walk)
    lat0 = np.stack([o.permute(5, 0, 0, l) for l in range(5)])
    fit0 = o.fitness(lat0)
    lo, fo, so = o.local_mc(lat0, fit0, task, depth, seed=21, iter=2, phase=2)
    lg, fg, sg = p.local_mc(lat0, fit0, task, depth, seed=21, iter=2, phase=2)
    assert np.array_equal(_bits(fg), _bits(fo))
    assert np.array_equal(lg, lo)
    for k in ("evals", "rounds", "accepts", "draws"):
        assert sg[k] == so[k], k


@pytest.mark.parametrize("case", CASES[:5], ids=_id)
def test_run_hcs_species_in_hbm_bit_exact(gpu, monkeypatch, case):
    """Sessions (local searches I and II, merges) with the records walk."""
    spec = _spec(case)
    o, p = Oracle.from_spec(spec), _problem(monkeypatch, spec)
    lo, fo, _, so = o.run_hcs(6, 3, 16, 2, seed=77)
    lg, fg, sg = p.run_hcs(6, 3, 16, 2, seed=77)
    assert np.array_equal(lg, lo)
    assert np.array_equal(_bits(fg), _bits(fo))
    assert sg["evals"] == so["evals"] and sg["accepts"] == so["accepts"]


def test_batch_species_in_hbm_equals_plain(gpu, monkeypatch):
    """A batch whose problems read species from HBM (the records walk of each
    problem in its own slice of the batch's record buffer) equals the plain
    search of each problem with seed + p."""
    cases = [CASES[0], CASES[1], CASES[3], ("BCC", 4, [0.2] * 5, "true", 3),
             ("FCC", 3, [0.25] * 4, "true", 3)]  # one shell count per batch
    specs = [_spec(c, seed=40 + i) for i, c in enumerate(cases)]
    hbm = [_problem(monkeypatch, s) for s in specs]
    lds = [_problem(monkeypatch, s, hbm=False) for s in specs]
    L, I, T, D, seed = 5, 2, 48, 2, 600
    b = Batch(hbm, L, T, D, seed=seed)
    b.set_grid(3)
    b.init()
    b.step(I)
    for q, p in enumerate(lds):
        lb, fb, sb = b.read(q)
        lp, fp, sp = p.run_hcs(L, I, T, D, seed=seed + q)
        assert np.array_equal(lb, lp), q
        assert np.array_equal(_bits(fb), _bits(fp)), q
        assert sb["evals"] == sp["evals"] and sb["accepts"] == sp["accepts"], q
    b.close()


def _guided(p, accepts, L=6, I=2, T=64, D=3, seed=5):
    s = p.session(L, T, D, seed=seed)
    s.guide(candidates=8, tau=0.1, accepts=accepts)
    s.init()
    s.step(I)
    out = s.read()
    s.close()
    return out


@pytest.mark.parametrize("accepts", [1, 4])
@pytest.mark.parametrize("case", [CASES[1], CASES[3]], ids=_id)
def test_guided_species_in_hbm_matches_lds(gpu, monkeypatch, case, accepts):
    """Guided search (which always walks with its records): species in LDS
    and in HBM make the same decisions, with or without the switch that
    makes the plain walk gather, and the stored fitness is the oracle's
    recount."""
    spec = _spec(case)
    o = Oracle.from_spec(spec)
    runs = [_guided(_problem(monkeypatch, spec, hbm=h, walk=w), accepts)
            for h, w in ((False, "records"), (True, "records"), (True, "gather"))]
    l0, f0, s0 = runs[0]
    for lat, fit, st in runs[1:]:
        assert np.array_equal(lat, l0)
        assert np.array_equal(_bits(fit), _bits(f0))
        assert st["evals"] == s0["evals"] and st["accepts"] == s0["accepts"]
    assert np.allclose(o.fitness(l0), f0, rtol=1e-11, atol=0)
    assert s0["accepts"] > 0


def test_batch_mixed_walks_species_in_hbm(gpu, monkeypatch):
    """A batch holding a 7-species problem runs the width-8 kernel; its
    problems with at most 6 species still use records (8-bit packed walk
    histograms), the 7-species one gathers.  Each equals its plain search."""
    cases = [("FCC", 3, [0.15] * 6 + [0.1], "true", 3), ("BCC", 4, [0.5, 0.25, 0.25], "true", 3),
             ("FCC", 3, [0.3, 0.2, 0.2, 0.1, 0.1, 0.1], "true", 3)]
    specs = [_spec(c, seed=80 + i) for i, c in enumerate(cases)]
    hbm = [_problem(monkeypatch, s) for s in specs]
    lds = [_problem(monkeypatch, s, hbm=False) for s in specs]
    L, I, T, D, seed = 4, 2, 40, 2, 650
    b = Batch(hbm, L, T, D, seed=seed)
    b.init()
    b.step(I)
    for q, p in enumerate(lds):
        lb, fb, sb = b.read(q)
        lp, fp, sp = p.run_hcs(L, I, T, D, seed=seed + q)
        assert np.array_equal(lb, lp), q
        assert np.array_equal(_bits(fb), _bits(fp)), q
        assert sb["evals"] == sp["evals"] and sb["accepts"] == sp["accepts"], q
    b.close()
