"""Orchestration: ``opt_model`` of ``pyhea/model/opt.py``.

Maps the YAML knobs to the engine arguments exactly as opt.py:37-52 does
(nnet = solutions, step = total_iter, task = parallel_task, depth =
converge_depth, thr = 0.001, weight count <= shells) and dispatches on
``device`` (opt.py:64-96).  Both devices run the MI355X engine, whose doubles
equal the reference CPU engine's on the same streams; errors raise instead of
calling exit().  There is no MPI here: one process drives one GPU.
"""
from __future__ import annotations

import logging

from .cpp import accelerate as acc

logger = logging.getLogger("pyhea_amd")


class opt_model:  # noqa: N801  (reference class name)
    def __init__(self, latt, conf, comm=None):
        self.latt = latt
        self.conf = conf
        self.comm = comm
        self.thr = 0.001
        self.ntyp = conf.type
        self.nnet = conf.solutions
        self.step = conf.total_iter
        self.task = conf.parallel_task
        self.depth = conf.converge_depth
        self.device = conf.device
        self.element = conf.element
        self.target_sro = conf.target_sro
        if len(conf.weight) > latt.shells:
            raise ValueError("The number of weight is larger than the maximum number(shell) "
                             "of neighbors.")
        self.weight = conf.weight
        self.seed = getattr(conf, "seed", None)  # None: a fresh key per run

    def run_optimization(self):
        logger.info("Running Local Parallel HCS optimization...")
        if self.device not in ("cpu", "gpu"):
            raise ValueError("Invalid device type. Please specify either 'cpu' or 'gpu'.")
        if self.device == "gpu" and not acc.cuda_available():
            raise RuntimeError("No GPU is visible to the HIP runtime.")
        run = acc.run_local_parallel_hcs if self.device == "cpu" else \
            acc.run_local_parallel_hcs_cuda
        latts, fitss = run(self.nnet, self.step, self.task, self.depth, self.thr,
                           self.latt.nbor, self.element, self.weight, self.target_sro,
                           seed=self.seed, shell_size=self.latt.shell_sizes)
        logger.info("Best fitness value: %s", fitss[0])
        return latts, fitss
