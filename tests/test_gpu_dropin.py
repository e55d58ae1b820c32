"""The drop-in surface end to end on the GPU: the pybind11 module called
exactly as pyhea/model/opt.py:67-68 and :89-90 call the reference
(positional args, nested neighbour lists), the YAML -> opt_model path, and
the reference engine as the judge of the result."""
import numpy as np
import pytest

from oracle import Oracle, RefHarness
from pyhea_amd.hcs import Problem
from pyhea_amd.lattice import lattice
from pyhea_amd.problem import make_problem

pytestmark = pytest.mark.gpu


def test_pybind_names_match_c_abi(gpu):
    from pyhea_amd.cpp import accelerate as acc

    assert acc.cuda_available()
    spec = make_problem("FCC", 3, [0.5, 0.3, 0.2], shells="pyhea", target_seed=3)
    latt = lattice([3, 3, 3], "FCC", 3.6, None)
    nested = latt.nbor_list  # list[N] of list[3] of arrays, as lattice.py builds it
    target = np.zeros((3, 9))
    target[:3] = spec.target
    for name in ("run_local_parallel_hcs", "run_local_parallel_hcs_cuda"):
        fn = getattr(acc, name)
        latts, fits = fn(6, 2, 16, 2, 0.001, nested, spec.counts, spec.weight, target, seed=42)
        assert len(latts) == 6 and len(latts[0]) == spec.n_sites
        assert fits == sorted(fits)
        lat, fit, _ = Problem.from_spec(spec).run_hcs(6, 2, 16, 2, seed=42)
        assert np.array_equal(np.array(latts, np.uint8), lat)
        assert np.array_equal(np.array(fits), fit)


def test_pybind_errors_raise(gpu):
    from pyhea_amd.cpp import accelerate as acc

    spec = make_problem("BCC", 2, [0.5, 0.5], shells="pyhea", target_seed=3)
    with pytest.raises(ValueError):
        acc.run_local_parallel_hcs(4, 1, 0, 2, 0.001, spec.nbr, spec.counts, spec.weight,
                                   np.zeros((3, 4)), shell_size=spec.shell_sizes)


def test_yaml_to_engine_matches_reference_engine(gpu, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)  # target_sro is relative to the working directory
    if not RefHarness.available():
        pytest.skip("reference harness not built")
    from pyhea_amd.config import input_config
    from pyhea_amd.opt import opt_model

    (tmp_path / "TARGET_SRO").write_text("First Shell\n-0.05\n0.02 0.03\n")
    cfg = tmp_path / "config.yaml"
    cfg.write_text("type: 2\nelement: [54, 54]\ncell_dim: [3, 3, 3]\ndevice: gpu\n"
                   "solutions: 4\ntotal_iter: 2\nparallel_task: 8\nconverge_depth: 2\n"
                   "weight: [4.0, 1.0, 0.5]\nstructure: POSCAR_FCC\ntarget_sro: TARGET_SRO\n"
                   "seed: 77\n")
    conf = input_config(str(cfg))
    latt = lattice(conf.cell_dim, conf.latt_type, 3.6, None)
    latts, fits = opt_model(latt, conf).run_optimization()
    o = Oracle(latt.nbor, latt.shell_sizes, conf.element, conf.weight, conf.target_sro)
    snaps, _, _ = RefHarness(o).run_hcs(4, 2, 8, 2, 77)
    assert np.array_equal(np.array(latts, np.uint8), snaps[-1][0])
    assert np.array_equal(np.array(fits), snaps[-1][1])


def test_pybind_default_seed_is_fresh_per_call(gpu):
    """Without `seed` every call draws a fresh stream key, as the reference
    seeds from std::random_device (accelerate.cpp:241-242): two calls differ."""
    from pyhea_amd.cpp import accelerate as acc

    spec = make_problem("FCC", 3, [0.5, 0.3, 0.2], shells="pyhea", target_seed=3)
    target = np.zeros((3, 9))
    target[:3] = spec.target
    runs = [acc.run_local_parallel_hcs(4, 1, 16, 2, 0.001, spec.nbr, spec.counts, spec.weight,
                                       target, shell_size=spec.shell_sizes) for _ in range(2)]
    assert not np.array_equal(np.array(runs[0][0]), np.array(runs[1][0]))
