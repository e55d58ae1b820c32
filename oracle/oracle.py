"""TEST INFRASTRUCTURE ONLY -- the CPU oracle of the SRO lattice search.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module, and only as the checker or the CPU baseline.  Product code
(pyhea_amd/) never imports it.

``Oracle`` binds oracle/liboracle_hcs.so (oracle/hcs_oracle.cpp), a
function-by-function restatement of pyhea/cpp/accelerate.cpp.
``RefHarness`` runs oracle/_ref/ref_hcs, the reference's own accelerate.cpp
compiled in place by oracle/Makefile (target ``ref``) with the keyed streams
injected (oracle/ref_harness.cpp).

How each oracle function is pinned (tests/test_oracle.py unless noted):

  philox / stream words    Random123 known-answer vectors for Philox4x32-10.
  permute                  bijection + multiset invariants; identical to the
                           harness's permutations (same header) and to the
                           product kernel (GPU tests).
  gamma recount            brute-force Python recount on tiny lattices
                           (k-fold sequential sums), symmetry, sum rules;
                           bit-identical to the reference harness (golden).
  fitness                  textbook Warren-Cowley formula within 1e-12;
                           bit-identical to calculate_fitness_of_solutions
                           through the golden vectors (pins pow(x,2) = x*x).
  incremental chain        drifted gamma and every fitness bit-identical to
                           calculate_sro_coefficient +
                           calculate_fitness_incremental (golden vectors,
                           including swaps of adjacent sites and self
                           neighbours of 1x1x1 cells).
  local_mc                 bit-identical populations, fitness and stream
                           consumption vs local_parallel_monte_carlo (golden
                           vectors, task in {1, 4, 8}, depth 1..4, K 2..6).
  run_hcs                  every stage snapshot bit-identical to the
                           reference loop (golden vectors).
Nothing here is "parity unpinned": every function meets one of the checks.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_hcs.so")
REF_BIN = os.path.join(HERE, "_ref", "ref_hcs")

_i32p = C.POINTER(C.c_int32)
_u32p = C.POINTER(C.c_uint32)
_i64p = C.POINTER(C.c_int64)
_u8p = C.POINTER(C.c_uint8)
_dp = C.POINTER(C.c_double)


def _p(a, ct):
    return a.ctypes.data_as(ct)


class Oracle:
    """ctypes front of liboracle_hcs.so bound to one problem."""

    _lib = None

    @classmethod
    def lib(cls):
        if cls._lib is None:
            if not os.path.exists(LIB_PATH):
                subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)
            cls._lib = C.CDLL(LIB_PATH)
        return cls._lib

    def __init__(self, nbr, shell_sizes, counts, weight, target):
        self.nbr = np.ascontiguousarray(nbr, dtype=np.int32)
        self.S = len(weight)
        self.Z = np.zeros(3, dtype=np.int32)
        self.Z[: self.S] = shell_sizes[: self.S]
        self.counts = np.ascontiguousarray(counts, dtype=np.int32)
        self.K = len(self.counts)
        self.weight = np.ascontiguousarray(weight, dtype=np.float64)
        self.target = np.ascontiguousarray(np.asarray(target, dtype=np.float64)[: self.S]
                                           .reshape(self.S, self.K * self.K))
        self.N = int(self.nbr.shape[0])
        assert self.nbr.shape[1] == int(self.Z.sum())

    @classmethod
    def from_spec(cls, spec):
        return cls(spec.nbr, spec.shell_sizes, spec.counts, spec.weight, spec.target)

    def _prob(self):
        return (self.N, self.K, self.S, _p(self.Z, _i32p), _p(self.nbr, _i32p))

    def _full(self):
        return self._prob() + (_p(self.counts, _i32p), _p(self.weight, _dp), _p(self.target, _dp))

    # ---- stream ------------------------------------------------------------
    @classmethod
    def philox(cls, ctr, key):
        c = np.asarray(ctr, dtype=np.uint32)
        k = np.asarray(key, dtype=np.uint32)
        o = np.zeros(4, dtype=np.uint32)
        cls.lib().oracle_philox(_p(c, _u32p), _p(k, _u32p), _p(o, _u32p))
        return o

    @classmethod
    def stream_words(cls, seed, iter, phase, slot, start, n):
        o = np.zeros(n, dtype=np.uint32)
        f = cls.lib().oracle_stream_words
        f.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64, C.c_int64, _u32p]
        f(seed, iter, phase, slot, start, n, _p(o, _u32p))
        return o

    def permute(self, seed, iter, phase, slot):
        o = np.zeros(self.N, dtype=np.uint8)
        f = self.lib().oracle_permute
        f.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, _i32p, C.c_int, C.c_int, _u8p]
        f(seed, iter, phase, slot, _p(self.counts, _i32p), self.K, self.N, _p(o, _u8p))
        return o

    # ---- fitness -----------------------------------------------------------
    def gamma(self, lat):
        lat = np.ascontiguousarray(lat, dtype=np.uint8)
        g = np.zeros((self.K, self.K, self.S), dtype=np.float64)
        self.lib().oracle_gamma(*self._prob(), _p(lat, _u8p), _p(g, _dp))
        return g

    def pair_counts(self, lat):
        lat = np.ascontiguousarray(lat, dtype=np.uint8)
        n = np.zeros((self.S, self.K, self.K), dtype=np.int64)
        self.lib().oracle_pair_counts(*self._prob(), _p(lat, _u8p), _p(n, _i64p))
        return n

    def fitness(self, lats):
        lats = np.ascontiguousarray(np.atleast_2d(lats), dtype=np.uint8)
        f = np.zeros(lats.shape[0], dtype=np.float64)
        self.lib().oracle_fitness(*self._full(), lats.shape[0], _p(lats, _u8p), _p(f, _dp))
        return f

    def incremental_chain(self, lat, swaps):
        lat = np.ascontiguousarray(lat, dtype=np.uint8).copy()
        sw = np.ascontiguousarray(swaps, dtype=np.int32).reshape(-1, 2)
        a1 = np.ascontiguousarray(sw[:, 0])
        a2 = np.ascontiguousarray(sw[:, 1])
        fits = np.zeros(len(sw), dtype=np.float64)
        g = np.zeros((self.K, self.K, self.S), dtype=np.float64)
        self.lib().oracle_incremental_chain(*self._full(), _p(lat, _u8p), len(sw), _p(a1, _i32p),
                                            _p(a2, _i32p), _p(fits, _dp), _p(g, _dp))
        return fits, g, lat

    # ---- search ------------------------------------------------------------
    def local_mc(self, lats, fit, task, depth, seed, iter, phase, sort=True, per_slot=False):
        lats = np.ascontiguousarray(np.atleast_2d(lats), dtype=np.uint8).copy()
        fit = np.ascontiguousarray(fit, dtype=np.float64).copy()
        L = lats.shape[0]
        st = np.zeros(5, dtype=np.int64)
        ps = np.zeros((L, 4), dtype=np.int64)
        f = self.lib().oracle_local_mc
        f.argtypes = [C.c_int, C.c_int, C.c_int, _i32p, _i32p, _i32p, _dp, _dp, C.c_int, _u8p,
                      _dp, C.c_int, C.c_int, C.c_uint64, C.c_uint32, C.c_uint32, C.c_int, _i64p,
                      _i64p]
        f(*self._full(), L, _p(lats, _u8p), _p(fit, _dp), task, depth, seed, iter, phase,
          1 if sort else 0, _p(st, _i64p), _p(ps, _i64p))
        stats = dict(evals=int(st[0]), rounds=int(st[1]), accepts=int(st[2]), draws=int(st[3]),
                     ties=int(st[4]))
        if per_slot:
            stats["per_slot"] = ps
        return lats, fit, stats

    def local_mc_trace(self, lat, fit, task, depth, seed, iter, phase, slot, cap=1 << 16):
        lat = np.ascontiguousarray(lat, dtype=np.uint8).copy()
        log = np.zeros((cap, 2), dtype=np.int32)
        st = np.zeros(4, dtype=np.int64)
        f = self.lib().oracle_local_mc_trace
        f.restype = C.c_double
        f.argtypes = [C.c_int, C.c_int, C.c_int, _i32p, _i32p, _i32p, _dp, _dp, _u8p, C.c_double,
                      C.c_int, C.c_int, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, _i32p,
                      C.c_int64, _i64p]
        fr = f(*self._full(), _p(lat, _u8p), fit, task, depth, seed, iter, phase, slot,
               _p(log, _i32p), cap, _p(st, _i64p))
        n = int(min(st[2], cap))
        return lat, fr, log[:n], dict(evals=int(st[0]), rounds=int(st[1]), accepts=int(st[2]),
                                      draws=int(st[3]))

    def run_hcs(self, solutions, total_iter, task, depth, seed):
        lat = np.zeros((solutions, self.N), dtype=np.uint8)
        fit = np.zeros(solutions, dtype=np.float64)
        trace = np.zeros(max(total_iter, 1), dtype=np.float64)
        st = np.zeros(5, dtype=np.int64)
        f = self.lib().oracle_run_hcs
        f.argtypes = [C.c_int, C.c_int, C.c_int, _i32p, _i32p, _i32p, _dp, _dp, C.c_int, C.c_int,
                      C.c_int, C.c_int, C.c_uint64, _u8p, _dp, _dp, _i64p]
        f(*self._full(), solutions, total_iter, task, depth, seed, _p(lat, _u8p), _p(fit, _dp),
          _p(trace, _dp), _p(st, _i64p))
        return lat, fit, trace[:total_iter], dict(evals=int(st[0]), rounds=int(st[1]),
                                                  accepts=int(st[2]), draws=int(st[3]),
                                                  ties=int(st[4]))


class RefHarness:
    """Runs the reference engine (oracle/_ref/ref_hcs) on one problem."""

    MAGIC = 0x52414548

    @staticmethod
    def available() -> bool:
        return os.path.exists(REF_BIN)

    def __init__(self, oracle: Oracle):
        self.o = oracle

    def _header(self, mode, L=0, I=0, T=0, D=0, n_swaps=0, seed=0, iter=0, phase=0):
        o = self.o
        h = np.zeros(14, dtype=np.int32)
        h[:] = [self.MAGIC, mode, o.N, o.K, o.S, o.Z[0], o.Z[1], o.Z[2], L, I, T, D, n_swaps, 0]
        tail = np.array([seed], dtype=np.uint64).tobytes() + np.array([iter, phase],
                                                                      dtype=np.uint32).tobytes()
        tgt3 = np.zeros((3, o.K * o.K), dtype=np.float64)
        tgt3[: o.S] = o.target
        return (h.tobytes() + tail + o.nbr.tobytes() + o.counts.tobytes() + o.weight.tobytes()
                + tgt3.tobytes())

    def _run(self, payload: bytes) -> bytes:
        with tempfile.TemporaryDirectory() as d:
            fi, fo = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
            with open(fi, "wb") as f:
                f.write(payload)
            subprocess.run([REF_BIN, fi, fo], check=True)
            with open(fo, "rb") as f:
                return f.read()

    def fitness(self, lats):
        lats = np.ascontiguousarray(np.atleast_2d(lats), dtype=np.uint8)
        out = self._run(self._header(2, L=lats.shape[0]) + lats.tobytes())
        return np.frombuffer(out, dtype=np.float64).copy()

    def incremental_chain(self, lat, swaps):
        o = self.o
        sw = np.ascontiguousarray(swaps, dtype=np.int32).reshape(-1, 2)
        out = self._run(self._header(3, n_swaps=len(sw)) + np.asarray(lat, np.uint8).tobytes()
                        + sw.tobytes())
        n = len(sw)
        fits = np.frombuffer(out[: 8 * n], dtype=np.float64).copy()
        g = np.frombuffer(out[8 * n: 8 * n + 8 * o.K * o.K * o.S], dtype=np.float64)
        lat_out = np.frombuffer(out[8 * n + 8 * o.K * o.K * o.S:], dtype=np.uint8).copy()
        return fits, g.reshape(o.K, o.K, o.S).copy(), lat_out

    def local_mc(self, lats, fit, task, depth, seed, iter, phase):
        lats = np.ascontiguousarray(np.atleast_2d(lats), dtype=np.uint8)
        L = lats.shape[0]
        payload = self._header(1, L=L, T=task, D=depth, seed=seed, iter=iter, phase=phase)
        out = self._run(payload + lats.tobytes() + np.asarray(fit, np.float64).tobytes())
        N = self.o.N
        lat = np.frombuffer(out[: L * N], dtype=np.uint8).reshape(L, N).copy()
        f = np.frombuffer(out[L * N: L * N + 8 * L], dtype=np.float64).copy()
        words = np.frombuffer(out[L * N + 8 * L: -8], dtype=np.int64).copy()
        secs = float(np.frombuffer(out[-8:], dtype=np.float64)[0])
        return lat, f, words, secs

    def run_hcs(self, solutions, total_iter, task, depth, seed):
        """Snapshots: [init, (ls1, merged, ls2) per iteration] as (lat, fit)."""
        L, N = solutions, self.o.N
        out = self._run(self._header(0, L=L, I=total_iter, T=task, D=depth, seed=seed))
        snaps, off = [], 0
        for _ in range(1 + 3 * total_iter):
            lat = np.frombuffer(out[off: off + L * N], dtype=np.uint8).reshape(L, N).copy()
            off += L * N
            f = np.frombuffer(out[off: off + 8 * L], dtype=np.float64).copy()
            off += 8 * L
            snaps.append((lat, f))
        words = np.frombuffer(out[off: -8], dtype=np.int64).copy()
        secs = float(np.frombuffer(out[-8:], dtype=np.float64)[0])
        return snaps, words, secs
