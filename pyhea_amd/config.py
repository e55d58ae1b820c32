"""YAML run configuration (the knob surface of ``pyhea/io/input.py``).

Keeps the reference's keys and meanings (``ALLOWED_KEYS``, input.py:45-48):
``type, element, cell_dim, solutions, device, total_iter, weight,
parallel_task, converge_depth, max_shell_num, structure, output, target_sro``,
plus build-only keys: ``seed`` (stream key), ``shells`` ("pyhea" reference
shells or "true" first three shells) and ``mode`` ("plain", "batched",
"guided").

Differences from the reference, on purpose:
  * the GPU probe asks the HIP runtime (``hea_device_count``) instead of
    nvidia-smi (pyhea/utils/device.py:80-85), which on an AMD machine finds
    nothing and silently demotes ``device: gpu`` to CPU (input.py:147-153);
  * ``structure`` may name the lattice type directly (``FCC``/``BCC``) or a
    POSCAR file whose name contains it; the POSCAR itself is not parsed;
  * the target file may hold any K (the reference stops at K = 4,
    input.py:457-458).
"""
from __future__ import annotations

import os

import numpy as np
import yaml

from .problem import target_from_lower_triangle

ALLOWED_KEYS = {
    "type", "element", "cell_dim", "solutions", "device", "total_iter", "weight",
    "parallel_task", "converge_depth", "max_shell_num", "structure", "output", "target_sro",
    "seed", "shells", "mode",
}


def read_target_sro(path: str, num_types: int) -> np.ndarray:
    """TARGET_SRO file -> [3][K*K] (input.py:350-465 layout, any K)."""
    expected = num_types * (num_types + 1) // 2
    shells, cur = [], []
    with open(path) as f:
        lines = [ln.strip() for ln in f if ln.strip()]
    for ln in lines:
        if ln.startswith("#"):
            continue
        low = ln.lower()
        if low.startswith("shell") or low.endswith("shell") or " shell" in low:
            if cur:
                shells.append(cur)
                cur = []
            continue
        try:
            cur.extend(float(v) for v in ln.split())
        except ValueError:
            continue
    if cur:
        shells.append(cur)
    for i, s in enumerate(shells):
        if len(s) != expected:
            raise ValueError(f"Shell {i + 1} has incorrect number of values. "
                             f"Expected {expected}, got {len(s)}")
    return target_from_lower_triangle(shells[:3], num_types)


class input_config:  # noqa: N801  (reference class name)
    def __init__(self, file_path: str):
        self.file_path = file_path
        with open(file_path) as f:
            self.config = yaml.safe_load(f) or {}
        unknown = set(self.config) - ALLOWED_KEYS
        if unknown:
            raise ValueError(f"Unknown configuration keys: {sorted(unknown)}")
        self._validate()

    def _validate(self) -> None:
        c = self.config
        device = str(c.get("device", "cpu")).lower()
        if device not in ("cpu", "gpu"):
            raise ValueError(f"Invalid device type '{device}'. Must be either 'cpu' or 'gpu'")
        if any(w <= 0 for w in c.get("weight", [])):
            raise ValueError("All weights must be positive values")
        cell = c.get("cell_dim", [])
        if cell and not all(isinstance(d, int) and d > 0 for d in cell):
            raise ValueError("Cell dimensions must be positive integers")
        if cell and len(set(cell)) != 1:
            raise ValueError("Only cubic cell dimensions are supported.")
        if "element" in c and "type" in c and len(c["element"]) != c["type"]:
            raise ValueError("Number of elements should be equal to type")
        if cell and "element" in c and "structure" in c:
            n = (4 if self.latt_type == "FCC" else 2) * cell[0] ** 3
            if sum(c["element"]) != n:
                raise ValueError(f"The sum of elements {sum(c['element'])} does not match "
                                 f"the total number of atoms {n}")
        if self.shells not in ("pyhea", "true"):
            raise ValueError("shells must be 'pyhea' or 'true'")

    type = property(lambda self: self.config.get("type", []))
    element = property(lambda self: self.config.get("element", []))
    cell_dim = property(lambda self: self.config.get("cell_dim", []))
    solutions = property(lambda self: self.config.get("solutions", 0))
    parallel_task = property(lambda self: self.config.get("parallel_task", 0))
    converge_depth = property(lambda self: self.config.get("converge_depth", 0))
    total_iter = property(lambda self: self.config.get("total_iter", 0))
    weight = property(lambda self: self.config.get("weight", []))
    device = property(lambda self: self.config.get("device", "cpu"))
    # None (no `seed` key): the engine draws a fresh key per run, as the
    # reference seeds from std::random_device (pyhea/cpp/accelerate.cpp:241-242)
    seed = property(lambda self: None if self.config.get("seed") is None
                    else int(self.config["seed"]))
    shells = property(lambda self: self.config.get("shells", "pyhea"))
    mode = property(lambda self: self.config.get("mode", "plain"))

    @property
    def max_shell_num(self):
        return self.config.get("max_shell_num", len(self.weight))

    @property
    def latt_type(self) -> str:
        s = str(self.config.get("structure", "FCC")).upper()
        return "BCC" if "BCC" in os.path.basename(s) else "FCC"

    @property
    def output_format(self) -> str:
        return self.config.get("output", {}).get("format", "vasp/poscar")

    @property
    def output_name(self) -> str:
        return self.config.get("output", {}).get("name", "output")

    @property
    def target_sro(self) -> np.ndarray:
        K = self.type
        path = self.config.get("target_sro")
        if path is None:
            return np.zeros((3, K * K))
        # relative to the working directory, as pyhea/io/input.py:373-374 does
        path = os.path.join(os.getcwd(), path)
        if not os.path.exists(path):
            raise FileNotFoundError(f"Target SRO file not found: {path}")
        return read_target_sro(path, K)
