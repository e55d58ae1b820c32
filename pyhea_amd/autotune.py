"""Launch-shape autotuner for the local-search kernels.

Two knobs change only how the work is laid out on the GPU, never the result
(the first-strict-minimum rule and the keyed streams make every launch shape
produce the same bits):

* ``tile`` -- proposals evaluated per pass (lanes per shell; 64..256);
* ``grid`` -- persistent workgroups of a batched launch.

``tune_batch`` times one global iteration for each candidate on the actual
problems (after one untimed iteration, which also captures the HIP graph) and
keeps the fastest; ``tune_session`` does the same for a single problem,
timing two iterations.  Choices are cached in a JSON file keyed by the workload signature
(lattice sizes, species counts, shell sizes, population, parallel_task,
converge_depth), so later runs reuse them without measuring.
"""
from __future__ import annotations

import json
import os
import time

from .hcs import Batch, Problem

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "pyhea_amd", "tune.json")


def _cu_count() -> int:
    try:
        import torch

        return int(torch.cuda.get_device_properties(0).multi_processor_count)
    except Exception:  # pragma: no cover - no torch device
        return 256


def _signature(problems, L, T, D, guided) -> str:
    parts = sorted(f"{p.n_sites}:{p.K}:{'-'.join(map(str, p.shell_sizes[:p.S]))}"
                   for p in problems)
    return f"L{L}/T{T}/D{D}/g{int(bool(guided))}/" + ",".join(parts)


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(path, cache):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def _tiles(T: int, S: int):
    return [t for t in (64, 128, 192, 256) if t <= max(64, T)]


def tune_batch(problems, solutions: int, parallel_task: int, converge_depth: int,
               seed: int = 0x5EED, guide=None, cache_path: str | None = DEFAULT_CACHE,
               grids=None, tiles=None, verbose: bool = False) -> dict:
    """Best {"grid", "tile", "seconds"} for a batch, measured or cached."""
    key = _signature(problems, solutions, parallel_task, converge_depth, guide)
    cache = _load(cache_path) if cache_path else {}
    if key in cache:
        return cache[key]
    cus = _cu_count()
    S = problems[0].S
    grids = grids or [cus, 2 * cus, 4 * cus]
    tiles = tiles or _tiles(parallel_task, S)
    best = None
    for tile in tiles:
        for grid in grids:
            b = Batch(problems, solutions, parallel_task, converge_depth, seed=seed)
            b.tile(tile)
            b.set_grid(grid)
            if guide:
                b.guide(**guide)
            b.init()
            b.step(1)
            b.sync()
            t0 = time.perf_counter()
            b.step(1)
            b.sync()
            dt = time.perf_counter() - t0
            b.close()
            if verbose:
                print(f"[autotune] tile={tile} grid={grid}: {dt * 1e3:.1f} ms", flush=True)
            if best is None or dt < best["seconds"]:
                best = {"grid": grid, "tile": tile, "seconds": dt}
    if cache_path:
        cache[key] = best
        _store(cache_path, cache)
    return best


def tune_session(problem: Problem, solutions: int, parallel_task: int, converge_depth: int,
                 seed: int = 0x5EED, guide=None, cache_path: str | None = DEFAULT_CACHE,
                 tiles=None, verbose: bool = False) -> dict:
    """Best {"tile", "seconds"} for one problem's session."""
    key = "session/" + _signature([problem], solutions, parallel_task, converge_depth, guide)
    cache = _load(cache_path) if cache_path else {}
    if key in cache:
        return cache[key]
    best = None
    for tile in tiles or _tiles(parallel_task, problem.S):
        s = problem.session(solutions, parallel_task, converge_depth, seed=seed)
        s.tile(tile)
        if guide:
            s.guide(**guide)
        s.init()
        s.step(1)
        s.sync()
        t0 = time.perf_counter()
        s.step(2)
        s.sync()
        dt = time.perf_counter() - t0
        s.close()
        if verbose:
            print(f"[autotune] tile={tile}: {dt * 1e3:.1f} ms", flush=True)
        if best is None or dt < best["seconds"]:
            best = {"tile": tile, "seconds": dt}
    if cache_path:
        cache[key] = best
        _store(cache_path, cache)
    return best
