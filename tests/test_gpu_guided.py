"""Guided proposals: the acceptance rule stays exact, so every stored
fitness must agree with the oracle's recount of the returned lattice (to
gamma-drift precision), composition is preserved, the population is sorted,
and a run is reproducible from its seed.  With several accepts per round the
extra swaps are checked the same way (they are accepted only on an exact
fitness decrease, from proposals whose walks the round's swaps left valid)."""
import numpy as np
import pytest

from oracle import Oracle
from pyhea_amd.hcs import Batch, Problem
from pyhea_amd.problem import make_problem

pytestmark = pytest.mark.gpu


def _spec(typ="FCC", cell=4, K=4, seed=11):
    fr = np.random.default_rng(seed).dirichlet([3.0] * K) + 0.05
    return make_problem(typ, cell, fr, shells="true", target_seed=seed)


def _run(prob, guided, L=8, I=3, T=64, D=3, seed=5, accepts=1):
    s = prob.session(L, T, D, seed=seed)
    if guided:
        s.guide(candidates=8, tau=0.1, accepts=accepts)
    s.init()
    s.step(I)
    out = s.read()
    s.close()
    return out


@pytest.mark.parametrize("accepts", [1, 4])
@pytest.mark.parametrize("typ,cell,K", [("FCC", 4, 4), ("BCC", 5, 5), ("FCC", 3, 6)])
def test_guided_exact_acceptance(gpu, typ, cell, K, accepts):
    spec = _spec(typ, cell, K)
    prob, o = Problem.from_spec(spec), Oracle.from_spec(spec)
    lat, fit, st = _run(prob, True, accepts=accepts)
    assert np.all(np.diff(fit) >= 0)
    assert np.allclose(o.fitness(lat), fit, rtol=1e-11, atol=0)
    for row in lat:
        assert np.array_equal(np.bincount(row, minlength=K), spec.counts)
    # every round screens its 64 proposals; further accepts re-screen at most
    # the 64 shortlisted ones per step
    assert st["evals"] >= 64 * st["rounds"] and st["accepts"] > 0
    assert st["evals"] <= 64 * st["rounds"] + 64 * (st["accepts"] + st["rounds"])
    if accepts == 1:
        assert st["evals"] == 64 * st["rounds"]
    lat2, fit2, _ = _run(prob, True, accepts=accepts)
    assert np.array_equal(lat, lat2) and np.array_equal(fit, fit2)


def test_guided_uniform_candidates(gpu, monkeypatch):
    """PYHEA_AMD_GUIDE_BLOCK=0 draws both candidate sites over the whole
    lattice: the captured candidates of a slot then spread over several
    blocks, and the search it steers keeps the exact acceptance."""
    spec = _spec("BCC", 5, 5)
    monkeypatch.setenv("PYHEA_AMD_GUIDE_BLOCK", "0")
    prob = Problem.from_spec(spec)
    monkeypatch.delenv("PYHEA_AMD_GUIDE_BLOCK")
    o = Oracle.from_spec(spec)
    lat0, _ = prob.generate(2, seed=17)
    pairs, _ = prob.guided_scores(lat0, 64, 8, seed=23)
    blk = pairs[..., 0].reshape(2, 64, 8) // 16
    assert (blk != blk[..., :1]).any(axis=-1).mean() > 0.9
    lat, fit, st = _run(prob, True, accepts=4)
    assert np.allclose(o.fitness(lat), fit, rtol=1e-11, atol=0)
    assert st["accepts"] > 0
    lb, fb, _ = _run(Problem.from_spec(spec), True, accepts=4)  # default: blocks
    assert not np.array_equal(lb, lat)


def test_guided_batch_matches_guided_sessions(gpu):
    specs = [_spec("FCC", 3, 4, 21), _spec("BCC", 4, 5, 22)]
    probs = [Problem.from_spec(s) for s in specs]
    b = Batch(probs, 4, 32, 2, seed=70)
    b.guide(candidates=8, tau=0.1, accepts=3)
    b.init()
    b.step(2)
    for q, p in enumerate(probs):
        s = p.session(4, 32, 2, seed=70 + q)
        s.guide(candidates=8, tau=0.1, accepts=3)
        s.init()
        s.step(2)
        ls, fs, _ = s.read()
        lb, fb, _ = b.read(q)
        assert np.array_equal(lb, ls) and np.array_equal(fb, fs)


def test_guide_rejects_records_wider_than_64_bits(gpu):
    """A site's neighbour counts must fit one 64-bit record: each species but
    the last takes a field as wide as its shell needs, here (6-1) x (5+5+4)
    = 70 bits for shells of 30, 30 and 8 neighbours (a circulant graph on
    200 sites, offsets +-1..15, +-16..30, +-31..34: symmetric, no self)."""
    N, K = 200, 6
    offs = [list(range(1, 16)), list(range(16, 31)), list(range(31, 35))]
    cols = [d for sh in offs for d in (*sh, *[-x for x in sh])]
    nbr = (np.arange(N)[:, None] + np.array(cols)[None, :]) % N
    counts = np.array([40, 40, 30, 30, 30, 30])
    prob = Problem(nbr, [30, 30, 8], counts, [4.0, 1.0, 0.5], np.zeros((3, K, K)))
    s = prob.session(4, 16, 2, seed=1)
    with pytest.raises(ValueError, match="64"):
        s.guide(candidates=8, tau=0.1, accepts=1)
    s.close()


@pytest.mark.parametrize("K", [5, 6])
def test_guided_uniform_species_batch_matches_sessions(gpu, K):
    """Guided batch whose problems all have K equal to the table width
    against guided sessions of each problem (exact-K kernels)."""
    specs = [_spec("FCC", 3, K, 31), _spec("BCC", 4, K, 32)]
    probs = [Problem.from_spec(s) for s in specs]
    b = Batch(probs, 4, 32, 2, seed=71)
    b.guide(candidates=8, tau=0.1, accepts=3)
    b.init()
    b.step(2)
    for q, p in enumerate(probs):
        s = p.session(4, 32, 2, seed=71 + q)
        s.guide(candidates=8, tau=0.1, accepts=3)
        s.init()
        s.step(2)
        ls, fs, _ = s.read()
        lb, fb, _ = b.read(q)
        assert np.array_equal(lb, ls) and np.array_equal(fb, fs)


@pytest.mark.parametrize("typ,cell,K", [("FCC", 4, 4), ("BCC", 5, 5), ("FCC", 3, 6)])
def test_guided_scores_are_relaxation_site_gains(gpu, typ, cell, K):
    """The guided kernel's predicted change of F^2 for a candidate swap
    (i: a -> b, j: b -> a), from its neighbour-count records and gain table,
    equals u[i][b] + u[j][a] of the relaxation's site gains (hea_site_gain)
    at the same lattice with dn = dF^2/dn (the relaxation's gradient at the
    one-hot state): the coupling that ranks and samples the proposals."""
    import torch

    from pyhea_amd.relax import RelaxObjective, site_gain

    spec = _spec(typ, cell, K)
    prob = Problem.from_spec(spec)
    lat, _ = prob.generate(3, seed=17)
    pairs, pred = prob.guided_scores(lat, 64, 8, seed=23)
    n = torch.as_tensor(prob.pair_counts(lat), dtype=torch.float64, device="cuda")
    dn = RelaxObjective(prob).count_gradient(n)
    u = site_gain(prob, torch.as_tensor(lat, device="cuda"), dn).cpu().numpy().astype(np.float64)
    i, j = pairs[..., 0], pairs[..., 1]
    assert i.min() >= 0 and max(i.max(), j.max()) < spec.n_sites
    # a slot's 8 candidates take their first sites from one block of 16
    # consecutive sites and their second sites from another (the lattice's
    # last block may be shorter: 250 and 108 sites here)
    for x in (i, j):
        blk = x.reshape(x.shape[0], 64, 8) // 16
        assert (blk == blk[..., :1]).all()
    ll = np.arange(lat.shape[0])[:, None]
    a, b = lat[ll, i], lat[ll, j]
    live = a != b
    assert np.array_equal(np.isnan(pred), ~live) and live.mean() > 0.5
    want = u[ll, i, b] + u[ll, j, a]
    # float32 gains on both sides (the kernel's table is the FP64 gradient
    # rounded to float32, summed over 3 shells x (K-1) species)
    scale = np.abs(u).max()
    assert np.allclose(pred[live], want[live], rtol=1e-4, atol=1e-5 * scale)
    # the capture leaves the lattices alone and is reproducible
    pairs2, pred2 = prob.guided_scores(lat, 64, 8, seed=23)
    assert np.array_equal(pairs, pairs2) and np.array_equal(pred, pred2, equal_nan=True)
