"""hea_local_mc_device_noalloc: the local search on device buffers with its
workspace in the library's fixed per-device array, called through the C ABI
and compared bit for bit with hea_local_mc_device (and the oracle) on the
same inputs: two calls in a row, the second smaller and on another problem,
so the second runs on what the first left in the shared workspace."""
import numpy as np
import pytest

from oracle import Oracle
from pyhea_amd import _native as N
from pyhea_amd.hcs import Problem
from pyhea_amd.problem import make_problem

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 4 << 20  # HEA_NOALLOC_SCRATCH_BYTES


def _on_device(entry, prob, lat, fit, T, D, seed, it, phase):
    import torch

    dl = torch.from_numpy(np.ascontiguousarray(lat)).cuda()
    df = torch.from_numpy(np.ascontiguousarray(fit)).cuda()
    st = torch.cuda.current_stream()
    N.check(entry(prob.handle, lat.shape[0], dl.data_ptr(), df.data_ptr(), T, D, seed, it, phase,
                  st.cuda_stream))
    torch.cuda.synchronize()
    return dl.cpu().numpy(), df.cpu().numpy()


def _bits(a):
    return np.asarray(a, np.float64).view(np.uint64)


def test_noalloc_matches_device_entry_point(gpu):
    lib = N.lib()
    cases = [
        # (lattice, cell, fractions, solutions, task, depth, seed, iter, phase)
        ("FCC", 4, [0.4, 0.3, 0.2, 0.1], 24, 96, 3, 11, 0, N.PHASE_LS1),
        ("BCC", 4, [0.3, 0.25, 0.2, 0.15, 0.1], 7, 64, 2, 12, 3, N.PHASE_LS2),
    ]
    for lt, cell, fr, L, T, D, seed, it, phase in cases:
        spec = make_problem(lt, cell, fr, shells="true", target_seed=seed)
        prob, o = Problem.from_spec(spec), Oracle.from_spec(spec)
        lat0, fit0 = prob.generate(L, seed=seed + 100)
        print(f"[no_alloc] {lt} {cell}^3 K={spec.K} L={L} T={T} D={D}", flush=True)
        # the fixed-workspace call first, on a problem whose own scratch
        # buffers were never allocated by a local search
        ln, fn = _on_device(lib.hea_local_mc_device_noalloc, prob, lat0, fit0, T, D, seed, it,
                            phase)
        ld, fd = _on_device(lib.hea_local_mc_device, prob, lat0, fit0, T, D, seed, it, phase)
        assert np.array_equal(ln, ld)
        assert np.array_equal(_bits(fn), _bits(fd))
        lo, fo, _ = o.local_mc(lat0, fit0, T, D, seed, it, phase, sort=False)
        assert np.array_equal(ln, lo)
        assert np.array_equal(_bits(fn), _bits(fo))
        assert not np.array_equal(ln, lat0)  # the search moved something
        prob.close()


def test_noalloc_rejects_calls_beyond_the_workspace(gpu):
    import torch

    spec = make_problem("FCC", 4, [0.25, 0.25, 0.25, 0.25], shells="true", target_seed=5)
    prob = Problem.from_spec(spec)
    per = 32 + 4 * spec.S * spec.K * spec.K
    L = SCRATCH_BYTES // per + 1
    dl = torch.zeros((L, spec.n_sites), dtype=torch.uint8, device="cuda")
    df = torch.zeros(L, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError, match="workspace"):
        N.check(N.lib().hea_local_mc_device_noalloc(prob.handle, L, dl.data_ptr(), df.data_ptr(),
                                                    64, 2, 1, 0, N.PHASE_LS1, None))
    prob.close()
