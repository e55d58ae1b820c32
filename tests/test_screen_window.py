"""The local search's screen window on the CPU: the kernels' own templates
(pyhea_amd/csrc/hea_screen.h: safe step, entry bounds, beta, screen, exact
new entries), compiled for the host by tests/native/screen_check.cpp,
against the oracle's sequential FP64 updates (oracle/hcs_oracle.cpp,
restating calculate_fitness_incremental, accelerate.cpp:498-570).

For every unlike-species pair of random lattices with drifted gamma tables:
  * the new gamma entries of the fast path (one fma) or the replay equal the
    oracle's sequential +-delta updates bit for bit;
  * |err(P) - err(G) - X(P)| <= 2 beta (the bound the window-decided accepts
    and rejects rely on);
  * the oracle's first strict minimum of F lies inside the window
    X <= X_min + 2 beta.
The problems are chosen to be hard for the arithmetic: equimolar species
put the gamma entries next to powers of two (1/4 for two species, 1/16 for
four), BCC 4^3 with shells 8/6/12 makes 1/(Z N) exactly 2^-10 for the first
shell, a zero target makes every proposal's change tiny and nearly equal.
The measured deviation stays below 2e-4 of 2 beta (beta is a worst-case
bound), so the check also runs with beta shrunk 1e5-fold, where it must
fail.  The safe step is checked alone on random values, values at binade
edges and constructed rounding ties."""
import os
import subprocess

import numpy as np
import pytest

from pyhea_amd.problem import make_problem

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CASES = {
    "fcc3-k4-equimolar": (("FCC", 3, [0.25] * 4, "true", 1, [27] * 4), False),
    "bcc4-k2-equimolar": (("BCC", 4, [0.5, 0.5], "pyhea", 2, [64, 64]), False),
    "bcc4-k4-zero-target": (("BCC", 4, [0.25] * 4, "pyhea", 3, [32] * 4), True),
    "fcc4-k5": (("FCC", 4, [0.3, 0.25, 0.2, 0.15, 0.1], "true", 4, None), False),
    "fcc3-k6": (("FCC", 3, [0.3, 0.2, 0.2, 0.1, 0.1, 0.1], "pyhea", 5, None), False),
}


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("screen") / "screen_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off",
                    "-I", os.path.join(ROOT, "oracle"), "-I", os.path.join(ROOT, "pyhea_amd", "csrc"),
                    os.path.join(ROOT, "tests", "native", "screen_check.cpp"), "-o", exe],
                   check=True)
    return exe


def _problem_file(tmp_path, name):
    (typ, cell, fr, shells, seed, counts), zero_target = CASES[name]
    spec = make_problem(typ, cell, fr, shells=shells, target_seed=seed, counts=counts)
    path = str(tmp_path / f"{name}.bin")
    z = list(spec.shell_sizes[:spec.S]) + [0] * (3 - spec.S)
    target = np.zeros_like(spec.target) if zero_target else spec.target
    with open(path, "wb") as f:
        f.write(np.array([spec.n_sites, spec.K, spec.S] + z, np.int32).tobytes())
        f.write(np.ascontiguousarray(spec.nbr, np.int32).tobytes())
        f.write(np.array(spec.counts, np.int32).tobytes())
        f.write(np.array(spec.weight, np.float64).tobytes())
        f.write(np.ascontiguousarray(target, np.float64).tobytes())
    return path, spec


def _run(harness, path, seed, beta_scale):
    return subprocess.run([harness, path, "3", "200", str(seed), repr(beta_scale)],
                          capture_output=True, text=True)


@pytest.mark.parametrize("name", sorted(CASES))
def test_window_bound_and_exact_entries(harness, tmp_path, name):
    path, spec = _problem_file(tmp_path, name)
    r = _run(harness, path, 7, 1.0)
    assert r.returncode == 0, r.stdout + r.stderr
    ok, pairs, entries, replayed, ratio = r.stdout.split()
    assert ok == "ok"
    n_tab = 3 * spec.S * spec.K * spec.K  # table entries over the three lattices
    assert int(pairs) > 3 * spec.n_sites and int(entries) > 0
    assert 0 < int(replayed) <= n_tab  # the replay path is exercised
    assert 0.0 < float(ratio) < 1.0


def test_shrunk_beta_is_caught(harness, tmp_path):
    caught = []
    for name in sorted(CASES):
        path, _ = _problem_file(tmp_path, name)
        r = _run(harness, path, 7, 1e-5)
        assert r.returncode in (0, 1), r.stderr
        if r.returncode == 1:
            assert r.stdout.startswith(("bound:", "window:")), r.stdout
            caught.append(name)
    assert len(caught) >= 3, caught


def test_safe_step_fast_path_equals_sequential_updates(harness):
    r = subprocess.run([harness, "--steps", "300000", "11"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    ok, cases, fast, naive_miss = r.stdout.split()
    assert ok == "ok" and int(cases) > 250000
    assert int(fast) > 100000      # the fast path is taken and exact
    assert int(naive_miss) > 1000  # the unguarded step would fail on these inputs
