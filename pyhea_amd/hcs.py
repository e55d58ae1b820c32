"""Python host API over the C ABI (numpy in, numpy out).

``Problem`` wraps ``hea_problem`` (device-resident neighbour table and
constants); its methods map one-to-one onto the entry points of
``include/pyhea_hcs.h``.  Reference counterparts are in
``pyhea/cpp/accelerate.cpp``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from .problem import ProblemSpec

_u8p = C.POINTER(C.c_uint8)
_dp = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)


def _ptr(a: np.ndarray, ct):
    return a.ctypes.data_as(ct)


class Problem:
    """A search problem made resident on one GPU (hea_problem_create)."""

    def __init__(self, nbr, shell_sizes, counts, weight, target, device: int = 0):
        self.nbr = np.ascontiguousarray(np.asarray(nbr, dtype=np.int32))
        self.shell_sizes = [int(z) for z in shell_sizes]
        self.counts = np.ascontiguousarray(np.asarray(counts, dtype=np.int32))
        self.weight = np.ascontiguousarray(np.asarray(weight, dtype=np.float64))
        S, K = len(self.weight), len(self.counts)
        self.target = np.ascontiguousarray(np.asarray(target, dtype=np.float64)[:S].reshape(S, K * K))
        if self.nbr.ndim != 2 or self.nbr.shape[1] != sum(self.shell_sizes[:S]):
            raise ValueError("nbr must be [N][sum of the first len(weight) shell sizes]")
        self.n_sites, self.K, self.S = int(self.nbr.shape[0]), K, S
        d = N.HeaProblemDesc()
        d.n_sites, d.n_species, d.n_shells = self.n_sites, K, S
        for s in range(3):
            d.shell_size[s] = self.shell_sizes[s] if s < S else 0
        d.nbr = _ptr(self.nbr, _i32p)
        d.counts = _ptr(self.counts, _i32p)
        d.weight = _ptr(self.weight, _dp)
        d.target = _ptr(self.target, _dp)
        h = C.c_void_p()
        N.check(N.lib().hea_problem_create(C.byref(d), int(device), C.byref(h)))
        self._h = h

    @classmethod
    def from_spec(cls, spec: ProblemSpec, device: int = 0) -> "Problem":
        return cls(spec.nbr, spec.shell_sizes, spec.counts, spec.weight, spec.target, device)

    def close(self) -> None:
        if getattr(self, "_h", None):
            N.lib().hea_problem_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def _lat(self, lattices) -> np.ndarray:
        a = np.ascontiguousarray(np.asarray(lattices, dtype=np.uint8))
        if a.ndim == 1:
            a = a[None]
        if a.ndim != 2 or a.shape[1] != self.n_sites:
            raise ValueError("lattices must be [L][n_sites] uint8")
        return a

    def fitness(self, lattices) -> np.ndarray:
        """calculate_fitness_of_solutions (accelerate.cpp:118-180)."""
        lat = self._lat(lattices)
        out = np.empty(lat.shape[0], dtype=np.float64)
        N.check(N.lib().hea_fitness(self._h, lat.shape[0], _ptr(lat, _u8p), _ptr(out, _dp)))
        return out

    def pair_counts(self, lattices) -> np.ndarray:
        lat = self._lat(lattices)
        out = np.empty((lat.shape[0], self.S, self.K, self.K), dtype=np.int32)
        N.check(N.lib().hea_pair_counts(self._h, lat.shape[0], _ptr(lat, _u8p), _ptr(out, _i32p)))
        return out

    def generate(self, solutions: int, seed: int, iter: int = 0, phase: int = N.PHASE_INIT):
        """Keyed random lattices + recount fitness (accelerate.cpp:218-252)."""
        lat = np.empty((solutions, self.n_sites), dtype=np.uint8)
        fit = np.empty(solutions, dtype=np.float64)
        N.check(N.lib().hea_generate(self._h, solutions, seed, iter, phase, _ptr(lat, _u8p),
                                     _ptr(fit, _dp)))
        return lat, fit

    def local_mc(self, lattices, fitness, parallel_task: int, converge_depth: int, seed: int,
                 iter: int = 0, phase: int = N.PHASE_LS1, sort: bool = True):
        """local_parallel_monte_carlo (accelerate.cpp:601-694) on copies."""
        lat = self._lat(lattices).copy()
        fit = np.ascontiguousarray(np.asarray(fitness, dtype=np.float64)).copy()
        st = N.HeaStats()
        N.check(N.lib().hea_local_mc(self._h, lat.shape[0], _ptr(lat, _u8p), _ptr(fit, _dp),
                                     parallel_task, converge_depth, seed, iter, phase,
                                     1 if sort else 0, C.byref(st)))
        return lat, fit, st.as_dict()

    def run_hcs(self, solutions: int, total_iter: int, parallel_task: int, converge_depth: int,
                thr: float = 1e-3, seed: int = 0x5EED):
        """run_local_parallel_hcs (accelerate.cpp:713-763)."""
        lat = np.empty((solutions, self.n_sites), dtype=np.uint8)
        fit = np.empty(solutions, dtype=np.float64)
        st = N.HeaStats()
        N.check(N.lib().hea_hcs_run(self._h, solutions, total_iter, parallel_task, converge_depth,
                                    thr, seed, _ptr(lat, _u8p), _ptr(fit, _dp), C.byref(st)))
        return lat, fit, st.as_dict()

    def guided_scores(self, lattices, parallel_task: int, candidates: int, seed: int,
                      iter: int = 0, phase: int = N.PHASE_LS1):
        """The guided search's first candidate phase on each lattice:
        (pairs [L, T*cand, 2], predicted change of F^2 [L, T*cand])."""
        lat = self._lat(lattices)
        L, items = lat.shape[0], parallel_task * candidates
        pairs = np.empty((L, items, 2), dtype=np.int32)
        pred = np.empty((L, items), dtype=np.float32)
        N.check(N.lib().hea_guided_scores(self._h, L, _ptr(lat, _u8p), parallel_task, candidates,
                                          seed, iter, phase, _ptr(pairs, _i32p),
                                          pred.ctypes.data_as(C.POINTER(C.c_float))))
        return pairs, pred

    def session(self, solutions: int, parallel_task: int, converge_depth: int,
                thr: float = 1e-3, seed: int = 0x5EED) -> "Session":
        return Session(self, solutions, parallel_task, converge_depth, thr, seed)


class Session:
    """Device-resident population; one HIP graph per global iteration."""

    def __init__(self, problem: Problem, solutions, parallel_task, converge_depth, thr, seed):
        self.problem = problem
        self.solutions = int(solutions)
        h = C.c_void_p()
        N.check(N.lib().hea_session_create(problem.handle, solutions, parallel_task,
                                           converge_depth, thr, seed, C.byref(h)))
        self._h = h

    def tile(self, tp: int) -> None:
        """Proposals evaluated per pass (launch shape only)."""
        N.check(N.lib().hea_session_tile(self._h, int(tp)))

    def guide(self, candidates: int = 12, tau: float = 0.1, accepts: int = 1) -> None:
        """Switch both local searches to gradient-guided proposals; up to
        `accepts` non-interacting swaps are accepted per round."""
        N.check(N.lib().hea_session_guide(self._h, candidates, tau, accepts))

    def init(self) -> None:
        N.check(N.lib().hea_session_init(self._h))

    def step(self, n_iter: int = 1) -> None:
        N.check(N.lib().hea_session_step(self._h, n_iter))

    def step_timed(self, n_iter: int = 1):
        """Eager iterations with per-launch kernel times: (ms [n,2], evals [n,2])."""
        ms = np.zeros((n_iter, 2), dtype=np.float32)
        ev = np.zeros((n_iter, 2), dtype=np.int64)
        N.check(N.lib().hea_session_step_timed(self._h, n_iter,
                                               ms.ctypes.data_as(C.POINTER(C.c_float)),
                                               ev.ctypes.data_as(C.POINTER(C.c_int64))))
        return ms, ev

    def sync(self) -> None:
        N.check(N.lib().hea_session_sync(self._h))

    def best(self) -> float:
        v = C.c_double()
        N.check(N.lib().hea_session_best(self._h, C.byref(v)))
        return v.value

    def read(self):
        lat = np.empty((self.solutions, self.problem.n_sites), dtype=np.uint8)
        fit = np.empty(self.solutions, dtype=np.float64)
        st = N.HeaStats()
        N.check(N.lib().hea_session_read(self._h, _ptr(lat, _u8p), _ptr(fit, _dp), C.byref(st)))
        return lat, fit, st.as_dict()

    def close(self) -> None:
        if getattr(self, "_h", None):
            N.lib().hea_session_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many problems searched together (hea_batch_*): one persistent
    local-search launch per phase over every (problem, lattice) item.
    Problem p draws from seed ``seed + p``."""

    def __init__(self, problems, solutions: int, parallel_task: int, converge_depth: int,
                 thr: float = 1e-3, seed: int = 0x5EED):
        self.problems = list(problems)
        self.solutions = int(solutions)
        arr = (C.c_void_p * len(self.problems))(*[p.handle for p in self.problems])
        h = C.c_void_p()
        N.check(N.lib().hea_batch_create(arr, len(self.problems), solutions, parallel_task,
                                         converge_depth, thr, seed, C.byref(h)))
        self._h = h

    def guide(self, candidates: int = 12, tau: float = 0.1, accepts: int = 1) -> None:
        N.check(N.lib().hea_batch_guide(self._h, candidates, tau, accepts))

    def tile(self, tp: int) -> None:
        N.check(N.lib().hea_batch_tile(self._h, int(tp)))

    def set_grid(self, grid: int) -> None:
        N.check(N.lib().hea_batch_grid(self._h, int(grid)))

    def init(self) -> None:
        N.check(N.lib().hea_batch_init(self._h))

    def step(self, n_iter: int = 1) -> None:
        N.check(N.lib().hea_batch_step(self._h, n_iter))

    def step_timed(self, n_iter: int = 1):
        """Eager iterations with the duration of every persistent launch:
        (ms [n, 2], counters [n, 2] as dicts summed over the problems)."""
        ms = np.zeros((n_iter, 2), dtype=np.float32)
        st = (N.HeaStats * (2 * n_iter))()
        N.check(N.lib().hea_batch_step_timed(self._h, n_iter,
                                             ms.ctypes.data_as(C.POINTER(C.c_float)), st))
        return ms, [[st[2 * i + h].as_dict() for h in range(2)] for i in range(n_iter)]

    def sync(self) -> None:
        N.check(N.lib().hea_batch_sync(self._h))

    def read(self, p: int):
        prob = self.problems[p]
        lat = np.empty((self.solutions, prob.n_sites), dtype=np.uint8)
        fit = np.empty(self.solutions, dtype=np.float64)
        st = N.HeaStats()
        N.check(N.lib().hea_batch_read(self._h, p, _ptr(lat, _u8p), _ptr(fit, _dp),
                                       C.byref(st)))
        return lat, fit, st.as_dict()

    def best(self, p: int = 0) -> float:
        """Lowest stored fitness of problem p (fitness only, no lattices)."""
        fit = np.empty(self.solutions, dtype=np.float64)
        N.check(N.lib().hea_batch_read(self._h, p, None, _ptr(fit, _dp), None))
        return float(fit.min())

    def close(self) -> None:
        if getattr(self, "_h", None):
            N.lib().hea_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
