"""The two workloads that are neither the headline nor the time-to-fitness
problem, on the GPU:

* the reference's own example configuration (examples/config.yaml: FCC
  10x10x10 = 4000 sites, 4 x 1000 atoms, parallel_task 256, converge_depth
  10) in both FCC shell sets: bit for bit against the reference engine's
  golden vectors and the oracle, then the whole configured run (128
  solutions, 10 iterations) through the search's properties;
* global loops whose sorts meet equal fitness keys among more than 16
  lattices: the population order equals the reference's std::sort;
* the batched workload as configured: 64 problems (cell edges 6-20, half
  FCC and half BCC, K 3-6 with Dirichlet(2) fractions), 128 solutions each,
  one persistent launch per phase: properties of every problem, and the
  smallest, largest, a K = 3 and a K = 6 problem bit-identical to their
  plain sessions."""
import numpy as np
import pytest

from oracle import Oracle
from pyhea_amd.hcs import Batch, Problem
from pyhea_amd.problem import batched_workload
from test_fixtures_c0_ties import c0_case, tie_cases

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.asarray(a, np.float64).view(np.uint64)


def _properties(spec, prob, o, lat, fit, st, T, sample=2):
    assert np.all(np.diff(fit) >= 0)
    for row in lat:
        assert np.array_equal(np.bincount(row, minlength=spec.K), spec.counts)
    rec = prob.fitness(lat)
    assert np.allclose(rec, fit, rtol=1e-11, atol=0)
    idx = np.linspace(0, lat.shape[0] - 1, sample).astype(int)
    assert np.array_equal(_bits(o.fitness(lat[idx])), _bits(rec[idx]))
    assert st["evals"] % T == 0 and st["evals"] >= T * st["rounds"]


@pytest.mark.parametrize("shells", ["true", "pyhea"])
def test_c0_local_search_matches_reference(gpu, shells):
    spec, run, fx = c0_case(shells)
    p = Problem.from_spec(spec)
    lat, fit = p.generate(4, seed=run["seed"], iter=0, phase=1)
    assert np.array_equal(lat, fx["lat_in"])
    assert np.array_equal(_bits(fit), fx["fit_in_bits"])
    lg, fg, sg = p.local_mc(lat, fit, run["parallel_task"], run["converge_depth"],
                            seed=run["seed"], iter=0, phase=2)
    assert np.array_equal(lg, fx["ls1_lat"])
    assert np.array_equal(_bits(fg), fx["ls1_fit_bits"])
    assert 2 * sg["draws"] == int(fx["ls1_words"].sum())


@pytest.mark.parametrize("shells", ["true", "pyhea"])
def test_c0_global_loop_matches_reference(gpu, shells):
    spec, run, fx = c0_case(shells)
    p = Problem.from_spec(spec)
    lg, fg, _ = p.run_hcs(4, 2, run["parallel_task"], run["converge_depth"], seed=run["seed"])
    assert np.array_equal(lg, fx["hcs_lat"])
    assert np.array_equal(_bits(fg), fx["hcs_fit_bits"][-1])


@pytest.mark.parametrize("shells", ["true", "pyhea"])
def test_c0_configured_run(gpu, shells):
    """solutions 128, total_iter 10: properties of the final population, the
    best fitness of every iteration never rising, and a repeat bit-identical."""
    spec, run, _ = c0_case(shells)
    o, p = Oracle.from_spec(spec), Problem.from_spec(spec)
    L, T, D, I = run["solutions"], run["parallel_task"], run["converge_depth"], run["total_iter"]
    outs = []
    for _ in range(2):
        s = p.session(L, T, D, seed=run["seed"])
        s.init()
        best = []
        for _ in range(I):
            s.step(1)
            best.append(s.best())
        outs.append(s.read())
        s.close()
        assert all(b1 <= b0 for b0, b1 in zip(best, best[1:]))
    lat, fit, st = outs[0]
    _properties(spec, p, o, lat, fit, st, T, sample=4)
    assert np.array_equal(lat, outs[1][0]) and np.array_equal(_bits(fit), _bits(outs[1][1]))


def test_tie_order_matches_reference(gpu):
    for ent, spec, lat, fbits, _ in tie_cases():
        p = Problem.from_spec(spec)
        lg, fg, _ = p.run_hcs(ent["solutions"], ent["total_iter"], ent["task"], ent["depth"],
                              seed=ent["seed"])
        assert np.array_equal(lg, lat[-1]), ent["problem"]
        assert np.array_equal(_bits(fg), fbits[-1]), ent["problem"]
        p.close()


def test_batched_workload_as_configured(gpu):
    specs = batched_workload()
    assert len(specs) == 64
    probs = [Problem.from_spec(sp) for sp in specs]
    L, T, D, seed = 128, 256, 10, 2002
    b = Batch(probs, L, T, D, seed=seed)
    b.init()
    b.step(1)
    res = [b.read(q) for q in range(len(specs))]
    b.close()
    for q, (sp, p) in enumerate(zip(specs, probs)):
        lat, fit, st = res[q]
        _properties(sp, p, Oracle.from_spec(sp), lat, fit, st, T)
    sizes = [sp.n_sites for sp in specs]
    Ks = [sp.K for sp in specs]
    pick = {int(np.argmin(sizes)), int(np.argmax(sizes)), Ks.index(3), Ks.index(6)}
    for q in sorted(pick):
        s = probs[q].session(L, T, D, seed=seed + q)
        s.init()
        s.step(1)
        lp, fp, stp = s.read()
        s.close()
        lat, fit, st = res[q]
        assert np.array_equal(lat, lp), q
        assert np.array_equal(_bits(fit), _bits(fp)), q
        for k in ("evals", "rounds", "accepts", "draws"):
            assert st[k] == stp[k], (q, k)
