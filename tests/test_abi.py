"""The C-ABI library loads, exports every function include/pyhea_hcs.h
declares, and rejects invalid problems on the host before touching a GPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from pyhea_amd import _native as N
from pyhea_amd.problem import make_problem

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    with open(os.path.join(ROOT, "include", "pyhea_hcs.h")) as f:
        src = f.read()
    return sorted(set(re.findall(r"^\s*(?:int|int32_t|const char\*)\s+(hea_\w+)\(", src, re.M)))


def test_header_and_binding_agree():
    assert _declared() == sorted(N.EXPORTS)


def test_library_exports_every_symbol():
    lib = N.lib()
    for name in _declared():
        assert hasattr(lib, name), name


def test_pybind_module_keeps_reference_names():
    from pyhea_amd.cpp import accelerate

    for name in ("run_local_parallel_hcs", "run_local_parallel_hcs_cuda", "cuda_available"):
        assert callable(getattr(accelerate, name))
    doc = accelerate.run_local_parallel_hcs.__doc__
    for kw in ("nnet", "step", "task", "depth", "thr", "nbor", "element", "weight",
               "target_sro"):
        assert kw in doc


def _desc(spec, **over):
    nbr = np.ascontiguousarray(over.pop("nbr", spec.nbr), np.int32)
    counts = np.ascontiguousarray(over.pop("counts", spec.counts), np.int32)
    weight = np.ascontiguousarray(over.pop("weight", spec.weight), np.float64)
    target = np.ascontiguousarray(over.pop("target", spec.target), np.float64)
    d = N.HeaProblemDesc()
    d.n_sites = over.pop("n_sites", nbr.shape[0])
    d.n_species = over.pop("n_species", len(counts))
    d.n_shells = over.pop("n_shells", len(weight))
    sizes = over.pop("shell_size", spec.shell_sizes)
    for s in range(3):
        d.shell_size[s] = sizes[s] if s < len(sizes) else 0
    d.nbr = nbr.ctypes.data_as(C.POINTER(C.c_int32))
    d.counts = counts.ctypes.data_as(C.POINTER(C.c_int32))
    d.weight = weight.ctypes.data_as(C.POINTER(C.c_double))
    d.target = target.ctypes.data_as(C.POINTER(C.c_double))
    d._keep = (nbr, counts, weight, target)
    return d


def _create(d):
    h = C.c_void_p()
    rc = N.lib().hea_problem_create(C.byref(d), 0, C.byref(h))
    if rc == 0:
        N.lib().hea_problem_destroy(h)
    return rc, N.lib().hea_last_error().decode()


@pytest.fixture(scope="module")
def spec():
    return make_problem("FCC", 2, [0.5, 0.25, 0.25], shells="true", target_seed=1)


@pytest.mark.parametrize("over,code,frag", [
    (dict(n_species=1, counts=[32]), N.HEA_EINVAL, "n_species"),
    (dict(n_species=9, counts=[4] * 8), N.HEA_EINVAL, "n_species"),
    (dict(counts=[16, 8, 7]), N.HEA_EINVAL, "sum"),
    (dict(counts=[24, 8, 0]), N.HEA_EINVAL, "at least one"),
    (dict(weight=[4.0, -1.0, 0.5]), N.HEA_EINVAL, "weights"),
    (dict(weight=[4.0, float("nan"), 0.5]), N.HEA_EINVAL, "weights"),
    (dict(n_shells=4), N.HEA_EINVAL, "n_shells"),
    (dict(shell_size=[12, 6, 40]), N.HEA_EINVAL, "shell_size"),
])
def test_invalid_problem_rejected(spec, over, code, frag):
    rc, msg = _create(_desc(spec, **over))
    assert rc == code and frag in msg


def test_out_of_range_neighbour(spec):
    bad = spec.nbr.copy()
    bad[3, 2] = spec.n_sites
    rc, msg = _create(_desc(spec, nbr=bad))
    assert rc == N.HEA_EINVAL and "range" in msg


def test_non_symmetric_neighbours_rejected(spec):
    bad = spec.nbr.copy()
    bad[0, 0] = (bad[0, 0] + 1) % spec.n_sites  # break j in N(i) <=> i in N(j)
    rc, msg = _create(_desc(spec, nbr=bad))
    assert rc == N.HEA_EUNSUPPORTED and "symmetric" in msg


def test_valid_problem_without_gpu_reports_nodev(spec):
    if N.device_count() > 0:
        pytest.skip("GPU present")
    rc, msg = _create(_desc(spec))
    assert rc == N.HEA_ENODEV


def test_python_wrapper_raises_value_error(spec):
    from pyhea_amd.hcs import Problem

    with pytest.raises(ValueError):
        Problem(spec.nbr, spec.shell_sizes, [16, 8, 7], spec.weight, spec.target)


def test_relaxation_row_width_comes_from_the_library():
    """p and q hold 4 elements per site up to 4 species, then 4 + 1, 4 + 2
    and 4 + 4 (columns 0-3, then the rest); the library's hea_relax_cols is
    the one source of that rule (no GPU needed)."""
    from pyhea_amd.relax import relax_cols

    assert [relax_cols(k) for k in range(1, 9)] == [4, 4, 4, 4, 5, 6, 8, 8]
    assert N.lib().hea_relax_cols(0) == -1 and N.lib().hea_relax_cols(9) == -1
    with pytest.raises(ValueError):
        relax_cols(9)
