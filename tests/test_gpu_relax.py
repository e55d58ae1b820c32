"""Differentiable relaxation: HIP forward and hand-written HIP backward
against a plain PyTorch FP64 reference, against FP64 central finite
differences (h = 1e-6), and the one-hot limit against integer pair counts."""
import numpy as np
import pytest
import torch

from oracle import Oracle
from pyhea_amd.hcs import Problem
from pyhea_amd.problem import make_problem
from pyhea_amd.relax import RelaxObjective, SoftPairCounts, site_gain

pytestmark = pytest.mark.gpu


def _spec(typ="FCC", cell=3, K=4, S=3, seed=1):
    fr = np.random.default_rng(seed).dirichlet([3.0] * K) + 0.05
    return make_problem(typ, cell, fr, weight=[4.0, 1.0, 0.5][:S], shells="true",
                        target_seed=seed)


def torch_reference_counts(logits64, nbr, shell_sizes):
    """n[s] = P^T Q_s with Q_s[i] = sum_{j in N_s(i)} p_j, all in FP64."""
    p = torch.softmax(logits64, dim=1)
    out, off = [], 0
    for z in shell_sizes:
        idx = torch.as_tensor(nbr[:, off:off + z], dtype=torch.long, device=p.device)
        q = p[idx].sum(1)
        out.append(p.T @ q)
        off += z
    return torch.stack(out)


@pytest.mark.parametrize("typ,cell,K,S", [("FCC", 3, 4, 3), ("BCC", 4, 5, 2), ("FCC", 2, 6, 1)])
def test_forward_matches_torch(gpu, typ, cell, K, S):
    spec = _spec(typ, cell, K, S)
    prob = Problem.from_spec(spec)
    th = torch.randn(spec.n_sites, K, dtype=torch.float64, device="cuda")
    n64 = SoftPairCounts.apply(th, prob)
    ref = torch_reference_counts(th, spec.nbr, spec.shell_sizes)
    assert torch.allclose(n64, ref, rtol=1e-12, atol=1e-12)
    n32 = SoftPairCounts.apply(th.float(), prob)
    assert torch.allclose(n32, ref, rtol=1e-5, atol=1e-5)


def test_one_hot_limit_is_integer_pair_counts(gpu):
    spec = _spec("BCC", 4, 3, 3)
    prob, o = Problem.from_spec(spec), Oracle.from_spec(spec)
    lat = o.permute(3, 0, 0, 0)
    th = torch.full((spec.n_sites, 3), -60.0, dtype=torch.float64, device="cuda")
    th[torch.arange(spec.n_sites), torch.as_tensor(lat, dtype=torch.long)] = 60.0
    n = SoftPairCounts.apply(th, prob).cpu().numpy()
    assert np.allclose(n, o.pair_counts(lat), rtol=0, atol=1e-9)
    # and the relaxed objective equals the reference fitness squared
    f2 = RelaxObjective(prob)(th).item()
    assert abs(f2 - o.fitness(lat)[0] ** 2) < 1e-9 * max(1.0, f2)


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-10), (torch.float32, 2e-4)])
def test_backward_matches_torch_autograd(gpu, dtype, tol):
    spec = _spec("FCC", 3, 5, 3, seed=2)
    prob = Problem.from_spec(spec)
    obj = RelaxObjective(prob)
    th0 = torch.randn(spec.n_sites, 5, dtype=torch.float64, device="cuda")
    th = th0.detach().clone().to(dtype).requires_grad_(True)
    obj(th).backward()
    ref_th = th0.detach().clone().requires_grad_(True)
    obj.from_counts(torch_reference_counts(ref_th, spec.nbr, spec.shell_sizes)).backward()
    g, r = th.grad.double(), ref_th.grad
    assert (g - r).abs().max().item() <= tol * r.abs().max().item()


def test_backward_matches_central_differences(gpu):
    spec = _spec("BCC", 3, 4, 3, seed=3)
    prob = Problem.from_spec(spec)
    obj = RelaxObjective(prob)
    th = torch.randn(spec.n_sites, 4, dtype=torch.float64, device="cuda").requires_grad_(True)
    obj(th).backward()
    grad = th.grad.clone()
    h = 1e-6
    rng = np.random.default_rng(0)
    for _ in range(24):
        i, c = int(rng.integers(spec.n_sites)), int(rng.integers(4))
        tp, tm = th.detach().clone(), th.detach().clone()
        tp[i, c] += h
        tm[i, c] -= h
        fd = (obj(tp).item() - obj(tm).item()) / (2 * h)
        assert abs(fd - grad[i, c].item()) <= 1e-6 * max(1.0, abs(fd)) + 1e-8, (i, c)


def test_site_gain_is_one_hot_gradient(gpu):
    """u[k][b] = dF2/dp[k][b] - dF2/dp[k][a_k] at the one-hot lattice."""
    spec = _spec("FCC", 3, 4, 3, seed=4)
    prob, o = Problem.from_spec(spec), Oracle.from_spec(spec)
    obj = RelaxObjective(prob)
    lats = np.stack([o.permute(8, 0, 0, l) for l in range(3)])
    dn = torch.stack([obj.count_gradient(torch.as_tensor(o.pair_counts(l), dtype=torch.float64,
                                                          device="cuda")) for l in lats])
    u = site_gain(prob, torch.as_tensor(lats, device="cuda"), dn).double()
    for l in range(3):
        p = torch.nn.functional.one_hot(torch.as_tensor(lats[l], dtype=torch.long), 4)
        p = p.double().cuda().requires_grad_(True)
        out, off = [], 0
        for z in spec.shell_sizes:
            idx = torch.as_tensor(spec.nbr[:, off:off + z], dtype=torch.long, device="cuda")
            out.append(p.T @ p[idx].sum(1))
            off += z
        obj.from_counts(torch.stack(out)).backward()
        g = p.grad
        want = g - g.gather(1, torch.as_tensor(lats[l], dtype=torch.long,
                                               device="cuda")[:, None])
        assert torch.allclose(u[l], want, rtol=1e-5, atol=1e-6 * want.abs().max().item())


@pytest.mark.parametrize("typ,cell,K,S", [("BCC", 8, 4, 3), ("FCC", 5, 3, 2), ("FCC", 6, 7, 3)])
def test_forward_backward_many_blocks(gpu, typ, cell, K, S):
    """Lattices spanning several workgroups of sites (4 or 8 lanes per site,
    last workgroup partly empty): counts and gradient in FP64 against the
    PyTorch FP64 reference."""
    spec = _spec(typ, cell, K, S, seed=cell)
    prob = Problem.from_spec(spec)
    obj = RelaxObjective(prob)
    th0 = torch.randn(spec.n_sites, K, dtype=torch.float64, device="cuda")
    ref_th = th0.clone().requires_grad_(True)
    ref = torch_reference_counts(ref_th, spec.nbr, spec.shell_sizes)
    th = th0.clone().requires_grad_(True)
    n = SoftPairCounts.apply(th, prob)
    assert torch.allclose(n, ref, rtol=1e-12, atol=1e-12)
    obj.from_counts(n).backward()
    obj.from_counts(ref).backward()
    g, r = th.grad, ref_th.grad
    assert (g - r).abs().max().item() <= 1e-10 * r.abs().max().item()


@pytest.mark.parametrize("dtype,ntol,gtol", [(torch.float64, 1e-12, 1e-10),
                                             (torch.float32, 2e-5, 2e-4)])
def test_fcc30_full_size(gpu, dtype, ntol, gtol):
    """FCC 30^3 (108000 sites, 5 species, shells 12/6/24), the time-to-fitness
    problem: soft pair counts against the PyTorch FP64 reference, the
    gradient against FP64 autograd, and the same bits on a second call."""
    from pyhea_amd.problem import workload

    spec, _ = workload(3)
    prob = Problem.from_spec(spec)
    obj = RelaxObjective(prob)
    g0 = torch.Generator(device="cpu").manual_seed(30)
    th0 = torch.randn((spec.n_sites, spec.K), generator=g0, dtype=torch.float64).cuda()
    ref_th = th0.clone().requires_grad_(True)
    ref = torch_reference_counts(ref_th, spec.nbr, spec.shell_sizes)
    obj.from_counts(ref).backward()
    th = th0.to(dtype).requires_grad_(True)
    n = SoftPairCounts.apply(th, prob)
    assert ((n - ref).abs() <= ntol * ref.abs().max()).all()
    obj.from_counts(n).backward()
    g, r = th.grad.double(), ref_th.grad
    assert (g - r).abs().max().item() <= gtol * r.abs().max().item()
    n2 = SoftPairCounts.apply(th.detach(), prob)
    assert torch.equal(n, n2)


@pytest.mark.parametrize("typ,cell,K", [("BCC", 8, 5), ("FCC", 5, 2)])
def test_site_gain_many_blocks(gpu, typ, cell, K):
    """Site gains on lattices spanning several workgroups of sites (two
    lattices per call) against the FP64 autograd gradient at the one-hot
    point."""
    spec = _spec(typ, cell, K, 3, seed=11)
    prob, o = Problem.from_spec(spec), Oracle.from_spec(spec)
    obj = RelaxObjective(prob)
    lats = np.stack([o.permute(12, 0, 0, l) for l in range(2)])
    dn = torch.stack([obj.count_gradient(torch.as_tensor(o.pair_counts(l), dtype=torch.float64,
                                                          device="cuda")) for l in lats])
    u = site_gain(prob, torch.as_tensor(lats, device="cuda"), dn).double()
    for l in range(2):
        p = torch.nn.functional.one_hot(torch.as_tensor(lats[l], dtype=torch.long), K)
        p = p.double().cuda().requires_grad_(True)
        obj.from_counts(torch_reference_counts_p(p, spec.nbr, spec.shell_sizes)).backward()
        g = p.grad
        want = g - g.gather(1, torch.as_tensor(lats[l], dtype=torch.long,
                                               device="cuda")[:, None])
        assert torch.allclose(u[l], want, rtol=1e-5, atol=1e-6 * want.abs().max().item())


def torch_reference_counts_p(p, nbr, shell_sizes):
    out, off = [], 0
    for z in shell_sizes:
        idx = torch.as_tensor(nbr[:, off:off + z], dtype=torch.long, device=p.device)
        out.append(p.T @ p[idx].sum(1))
        off += z
    return torch.stack(out)


@pytest.mark.parametrize("typ,cell,K,S", [("FCC", 6, 8, 3), ("BCC", 7, 2, 3)])
def test_forward_backward_extreme_species(gpu, typ, cell, K, S):
    """The most species the kernels take (8) and the fewest (2), several
    workgroups of sites, FP64 against the PyTorch reference."""
    spec = _spec(typ, cell, K, S, seed=K)
    prob = Problem.from_spec(spec)
    obj = RelaxObjective(prob)
    th0 = torch.randn(spec.n_sites, K, dtype=torch.float64, device="cuda")
    ref_th = th0.clone().requires_grad_(True)
    ref = torch_reference_counts(ref_th, spec.nbr, spec.shell_sizes)
    th = th0.clone().requires_grad_(True)
    n = SoftPairCounts.apply(th, prob)
    assert torch.allclose(n, ref, rtol=1e-12, atol=1e-12)
    obj.from_counts(n).backward()
    obj.from_counts(ref).backward()
    g, r = th.grad, ref_th.grad
    assert (g - r).abs().max().item() <= 1e-10 * r.abs().max().item()


def test_side_stream_gives_same_bits(gpu):
    """Forward and backward launched on a side stream (the first relaxation
    call of the problem among them) give the default stream's bits."""
    spec = _spec("FCC", 5, 5, 3, seed=21)
    th0 = torch.randn(spec.n_sites, 5, dtype=torch.float32, device="cuda")
    out = []
    for side in (True, False):
        prob = Problem.from_spec(spec)
        obj = RelaxObjective(prob)
        th = th0.clone().requires_grad_(True)
        s = torch.cuda.Stream() if side else torch.cuda.current_stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            n = SoftPairCounts.apply(th, prob)
            obj.from_counts(n).backward()
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        out.append((n.detach().clone(), th.grad.clone()))
    assert torch.equal(out[0][0], out[1][0]) and torch.equal(out[0][1], out[1][1])


def test_misaligned_buffers_are_rejected(gpu):
    """p and q rows are moved as 16-byte vectors: a q that is not 16-byte
    aligned is an argument error, raised before any kernel runs."""
    from pyhea_amd import _native as Nat
    from pyhea_amd.relax import relax_cols

    spec = _spec("FCC", 3, 5, 3)
    prob = Problem.from_spec(spec)
    N, kp = spec.n_sites, relax_cols(5)
    th = torch.randn(N, 5, device="cuda")
    p = torch.empty((N, kp), device="cuda")
    q = torch.empty(3 * N * kp + 4, device="cuda")
    n = torch.empty((3, 5, 5), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError, match="aligned"):
        Nat.check(Nat.lib().hea_relax_forward(prob.handle, 0, th.data_ptr(), p.data_ptr(),
                                              q.data_ptr() + 4, n.data_ptr(), None))
