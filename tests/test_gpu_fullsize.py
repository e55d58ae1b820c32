"""Parity at the benchmark workloads' full sizes.

Bit-exact against the oracle where the oracle finishes in seconds (local
search of lattices of the BCC 16^3 headline workload, and of one 108000-site
FCC 30^3 lattice with parallel_task 512, which runs the split-list screen);
elsewhere the size-independent properties the search guarantees: every
lattice keeps its composition, the population comes back sorted, each
stored fitness is the fitness of its lattice (the GPU recount, itself
bit-exact with the oracle on a sample, up to the gamma drift the reference
carries forward), the counters agree, and a run repeats bit for bit."""
import numpy as np
import pytest

from oracle import Oracle
from pyhea_amd.hcs import Batch, Problem
from pyhea_amd.problem import workload

pytestmark = pytest.mark.gpu


def _bits(f):
    return np.asarray(f, dtype=np.float64).view(np.uint64)


def _check_population(spec, prob, o, lat, fit, st, T, sample=4, guided=False):
    assert np.all(np.diff(fit) >= 0)
    for row in lat:
        assert np.array_equal(np.bincount(row, minlength=spec.K), spec.counts)
    rec = prob.fitness(lat)
    assert np.allclose(rec, fit, rtol=1e-11, atol=0)
    idx = np.linspace(0, lat.shape[0] - 1, sample).astype(int)
    assert np.array_equal(_bits(o.fitness(lat[idx])), _bits(rec[idx]))
    assert st["evals"] >= T * st["rounds"]
    if guided:  # plus at most 64 re-screens per further-accept step
        assert st["evals"] <= T * st["rounds"] + 64 * (st["accepts"] + st["rounds"])
    else:
        assert st["evals"] % T == 0


def _session(prob, run, L, tile=None, guide=None, iters=1):
    s = prob.session(L, run["parallel_task"], run["converge_depth"], seed=run["seed"])
    if tile:
        s.tile(tile)
    if guide:
        s.guide(**guide)
    s.init()
    s.step(iters)
    out = s.read()
    s.close()
    return out


def test_headline_local_search_bit_exact(gpu):
    """BCC 16^3, 5 species, parallel_task 256, converge_depth 10: local
    search I from eight permuted lattices, every bit against the oracle."""
    spec, run = workload(1)
    o, p = Oracle.from_spec(spec), Problem.from_spec(spec)
    lat0 = np.stack([o.permute(run["seed"], 0, 1, l) for l in range(8)])
    fit0 = o.fitness(lat0)
    T, D = run["parallel_task"], run["converge_depth"]
    lo, fo, so = o.local_mc(lat0, fit0, T, D, seed=run["seed"], iter=0, phase=2)
    lg, fg, sg = p.local_mc(lat0, fit0, T, D, seed=run["seed"], iter=0, phase=2)
    assert np.array_equal(_bits(fg), _bits(fo))
    assert np.array_equal(lg, lo)
    for k in ("evals", "rounds", "accepts", "draws"):
        assert sg[k] == so[k], k


def test_headline_session_properties(gpu):
    """The benchmark's own session: 512 solutions, tile 128, one global
    iteration replayed as a HIP graph."""
    spec, run = workload(1)
    o, p = Oracle.from_spec(spec), Problem.from_spec(spec)
    lat, fit, st = _session(p, run, run["solutions"], tile=128)
    _check_population(spec, p, o, lat, fit, st, run["parallel_task"])
    lat2, fit2, st2 = _session(p, run, run["solutions"], tile=128)
    assert np.array_equal(lat, lat2) and np.array_equal(_bits(fit), _bits(fit2))
    for k in ("evals", "rounds", "accepts", "draws"):
        assert st[k] == st2[k], k


def test_fcc30_local_search_bit_exact(gpu):
    """FCC 30^3 (108000 sites, shells 12/6/24), parallel_task 512: the third
    shell's list is walked in two parts; one lattice to convergence against
    the oracle (about 10 s of oracle time)."""
    spec, run = workload(3)
    o, p = Oracle.from_spec(spec), Problem.from_spec(spec)
    lat0 = o.permute(run["seed"], 0, 1, 0)[None]
    fit0 = o.fitness(lat0)
    T, D = run["parallel_task"], run["converge_depth"]
    lo, fo, so = o.local_mc(lat0, fit0, T, D, seed=run["seed"], iter=0, phase=2)
    lg, fg, sg = p.local_mc(lat0, fit0, T, D, seed=run["seed"], iter=0, phase=2)
    assert np.array_equal(_bits(fg), _bits(fo))
    assert np.array_equal(lg, lo)
    for k in ("evals", "rounds", "accepts", "draws"):
        assert sg[k] == so[k], k


def test_fcc30_guided_properties_and_batch_equality(gpu):
    """The time-to-fitness workload in guided mode (8 candidates, tau 0.1, 8
    accepts per round), 256 solutions: properties of one global iteration,
    and the batched persistent launch of the same problem gives the same
    bits as the session."""
    spec, run = workload(3)
    o, p = Oracle.from_spec(spec), Problem.from_spec(spec)
    g = dict(candidates=8, tau=0.1, accepts=8)
    lat, fit, st = _session(p, run, run["solutions"], guide=g)
    _check_population(spec, p, o, lat, fit, st, run["parallel_task"], guided=True)
    b = Batch([p], run["solutions"], run["parallel_task"], run["converge_depth"], seed=run["seed"])
    b.guide(**g)
    b.init()
    b.step(1)
    lb, fb, sb = b.read(0)
    b.close()
    assert np.array_equal(lb, lat) and np.array_equal(_bits(fb), _bits(fit))
    for k in ("evals", "rounds", "accepts", "draws"):
        assert sb[k] == st[k], k


@pytest.mark.parametrize("guided", [False, True])
def test_fcc50_properties(gpu, guided):
    """FCC 50^3 (500000 sites, 6 species): species do not fit in LDS and are
    read from HBM; eight solutions, one global iteration, repeated."""
    spec, run = workload(4)
    o, p = Oracle.from_spec(spec), Problem.from_spec(spec)
    g = dict(candidates=8, tau=0.1, accepts=8) if guided else None
    lat, fit, st = _session(p, run, 8, guide=g)
    _check_population(spec, p, o, lat, fit, st, run["parallel_task"], sample=2, guided=guided)
    lat2, fit2, _ = _session(p, run, 8, guide=g)
    assert np.array_equal(lat, lat2) and np.array_equal(_bits(fit), _bits(fit2))


def test_fcc50_configured_launches(gpu):
    """FCC 50^3 as configured: 64 solutions, the guided search in one
    persistent launch per phase (8 candidates, tau 0.1, up to 8 accepts per
    round) and the plain session, both with the neighbour-records walk; one
    global iteration each."""
    spec, run = workload(4)
    o, p = Oracle.from_spec(spec), Problem.from_spec(spec)
    L, T = run["solutions"], run["parallel_task"]
    b = Batch([p], L, T, run["converge_depth"], seed=run["seed"])
    b.guide(candidates=8, tau=0.1, accepts=8)
    b.init()
    b.step(1)
    lat, fit, st = b.read(0)
    b.close()
    assert lat.shape == (L, spec.n_sites)
    _check_population(spec, p, o, lat, fit, st, T, sample=2, guided=True)
    assert st["accepts"] > 8 * L
    lat, fit, st = _session(p, run, L)
    _check_population(spec, p, o, lat, fit, st, T, sample=2)
    assert st["accepts"] > L
