"""Differentiable relaxation of the SRO lattice search (torch front end).

``SoftPairCounts`` maps site logits [N, K] to soft pair counts n [S, K, K]
(float64) with the HIP forward (softmax, neighbour gather, deterministic
FP64 reduction) and the hand-written HIP backward (neighbour-gather adjoint)
of ``hea_relax_forward`` / ``hea_relax_backward``; the occupancies and
neighbour sums it keeps for the backward hold ``relax_cols(K)`` elements
per site (the library's row layout; 5 for five species).  ``soft_fitness_sq`` builds the relaxed
objective on top with ordinary torch ops:

    gamma = n / (Z_s N),  alpha = 1 - gamma / (c_a c_b),
    F^2   = sum_s w_s sum_ab (alpha_sab - t_sab)^2

which equals the reference fitness squared (pyhea/cpp/accelerate.cpp:
165-178) when the logits are one-hot.  PyTorch only moves tensors and runs
the K x K algebra; the O(N) work is in the HIP kernels.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _native as Nat


def _dtype_code(t: torch.Tensor) -> int:
    if t.dtype == torch.float32:
        return 0
    if t.dtype == torch.float64:
        return 1
    raise TypeError("logits must be float32 or float64")


def relax_cols(K: int) -> int:
    """Elements per site of the occupancies p [N * KP] and neighbour sums
    q [S * N * KP] that the forward keeps for the backward, as the library
    lays them out (hea_relax_cols)."""
    kp = int(Nat.lib().hea_relax_cols(int(K)))
    if kp < 0:
        raise ValueError(f"no relaxation layout for {K} species")
    return kp


def _stream() -> C.c_void_p:
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class SoftPairCounts(torch.autograd.Function):
    @staticmethod
    def forward(ctx, logits: torch.Tensor, problem) -> torch.Tensor:
        if not logits.is_cuda:
            raise ValueError("logits must be a GPU tensor")
        lg = logits.contiguous()
        N, K, S = problem.n_sites, problem.K, problem.S
        if tuple(lg.shape) != (N, K):
            raise ValueError(f"logits must be [{N}, {K}]")
        kp = relax_cols(K)
        p = torch.empty((N, kp), dtype=lg.dtype, device=lg.device)
        q = torch.empty((S, N, kp), dtype=lg.dtype, device=lg.device)
        n = torch.empty((S, K, K), dtype=torch.float64, device=lg.device)
        Nat.check(Nat.lib().hea_relax_forward(problem.handle, _dtype_code(lg), lg.data_ptr(),
                                              p.data_ptr(), q.data_ptr(), n.data_ptr(),
                                              _stream()))
        ctx.save_for_backward(p, q)
        ctx.problem = problem
        return n

    @staticmethod
    def backward(ctx, dn: torch.Tensor):
        p, q = ctx.saved_tensors
        dn = dn.to(torch.float64).contiguous()
        grad = torch.empty((p.shape[0], ctx.problem.K), dtype=p.dtype, device=p.device)
        Nat.check(Nat.lib().hea_relax_backward(ctx.problem.handle, _dtype_code(p), p.data_ptr(),
                                               q.data_ptr(), dn.data_ptr(), grad.data_ptr(),
                                               _stream()))
        return grad, None


class RelaxObjective:
    """Constants of the relaxed objective for one problem (device tensors)."""

    def __init__(self, problem, device="cuda"):
        self.problem = problem
        S, K, N = problem.S, problem.K, problem.n_sites
        conc = torch.tensor(np.asarray(problem.counts, np.float64) / N, device=device)
        self.cc = conc[:, None] * conc[None, :]
        self.zn = torch.tensor([float(z) * N for z in problem.shell_sizes[:S]], device=device,
                               dtype=torch.float64)
        self.target = torch.tensor(np.asarray(problem.target, np.float64).reshape(S, K, K),
                                   device=device)
        self.weight = torch.tensor(np.asarray(problem.weight, np.float64), device=device)

    def from_counts(self, n: torch.Tensor) -> torch.Tensor:
        """F^2 from (soft or integer) pair counts n [..., S, K, K]."""
        alpha = 1.0 - (n / self.zn[:, None, None]) / self.cc
        return ((alpha - self.target) ** 2).sum((-1, -2)).mul(self.weight).sum(-1)

    def __call__(self, logits: torch.Tensor) -> torch.Tensor:
        return self.from_counts(SoftPairCounts.apply(logits, self.problem))

    def count_gradient(self, n: torch.Tensor) -> torch.Tensor:
        """d F^2 / d n at pair counts n [..., S, K, K] (closed form)."""
        alpha = 1.0 - (n / self.zn[:, None, None]) / self.cc
        return -2.0 * self.weight[:, None, None] * (alpha - self.target) / (
            self.zn[:, None, None] * self.cc)


def soft_fitness_sq(logits: torch.Tensor, problem) -> torch.Tensor:
    return RelaxObjective(problem, logits.device)(logits)


def site_gain(problem, lattices: torch.Tensor, dn: torch.Tensor) -> torch.Tensor:
    """u [L, N, K] float32: first-order F^2 change of turning each site into
    each species, at the lattices (uint8 [L, N]) with dn = dF^2/dn [L,S,K,K]."""
    L = lattices.shape[0]
    u = torch.empty((L, problem.n_sites, problem.K), dtype=torch.float32, device=lattices.device)
    dn = dn.to(torch.float64).contiguous()
    Nat.check(Nat.lib().hea_site_gain(problem.handle, L, lattices.contiguous().data_ptr(),
                                      dn.data_ptr(), u.data_ptr(), _stream()))
    return u
