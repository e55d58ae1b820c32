#!/usr/bin/env python3
"""Benchmark of the SRO lattice search on one MI355X.

Default workload ``ttf-fcc30``, the headline comparison: FCC 30^3 = 108000
sites, 5 equimolar species, shells 12/6/24, weights [4, 1, 0.5], seeded
targets, solutions = 256, parallel_task = 512, converge_depth = 10.  Time to
fitness <= 1e-3 of the gradient-guided search in the batched persistent
launch (replayed as a HIP graph) against the plain GPU search and against the
reference CPU engine.  One step = one global iteration of
run_local_parallel_hcs (pyhea/cpp/accelerate.cpp:741-761): generate, local
search I, merge, local search II.  The timed region starts a fresh population
(so the device clock that stamps the first fitness <= 1e-3 starts inside it)
and runs --steps iterations of the guided + batched search; `value` is its
rate of swap evals (proposals walked and screened: rounds x parallel_task).
The plain search then runs from its own fresh population until it reaches
the threshold (or --budget-s), and the reference engine is timed on a bounded
sample of the same workload on the host cores.

Prints one JSON line: value, time to fitness of both modes and their ratios,
the roofline of the dominant kernel (the guided batched local-search kernel,
each launch timed with HIP events on its stream) and the CPU baseline.

Other workloads (--workload): ttf-fcc50 (the same comparison at FCC 50^3,
500000 sites, 6 species, 64 solutions); plain-bcc16 (plain drop-in search at
BCC 16^3, 512 solutions, 50 iterations); plain-fcc10; batched-64 (64 uneven
problems in one persistent launch per phase); relax-fcc30 / relax-fcc50
(forward and backward of the differentiable relaxation).
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s


def bytes_per_eval(sum_z: int) -> int:
    """Algorithmic bytes of one swap eval: both sites' neighbour indices
    (int32) and neighbour species (uint8), plus the two sites' species."""
    return 2 * sum_z * (4 + 1) + 2


def profiled_traffic(workload: str, kernel: str = "k_local_mc", key: str = "traffic_bytes_per_launch",
                     fname: str | None = None):
    """HBM bytes per launch (or per step) from the committed rocprofv3 byte
    counters (FETCH_SIZE and WRITE_SIZE, scaled by a copy kernel of known
    size), or None when no profile covers the workload."""
    name = workload.replace("-", "_").replace("relax_", "")
    path = os.path.join(ROOT, "profiles", fname or f"{kernel}_{name}.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(key + ":"):
                    return int(line.split(":")[1])
    except OSError:
        pass
    return None


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawTextHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed global iterations (default: 5 for ttf-*, 49 for plain-*)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="ttf-fcc30",
                    choices=["plain-bcc16", "plain-fcc10", "ttf-fcc30", "ttf-fcc50", "batched-64",
                             "relax-fcc30", "relax-fcc50"])
    ap.add_argument("--budget-s", type=float, default=60.0,
                    help="time-to-fitness workloads: wall-clock budget per mode")
    ap.add_argument("--thr", type=float, default=1e-3)
    ap.add_argument("--guide-candidates", type=int, default=12,
                    help="time-to-fitness workloads: candidate pairs per guided proposal")
    ap.add_argument("--guide-accepts", type=int, default=8,
                    help="time-to-fitness workloads: guided accepts per round at most")
    ap.add_argument("--dump-problem", default=None,
                    help="write the workload's problem for pyhea_amd/hea_profile and exit")
    ap.add_argument("--profile-steps", type=int, default=3,
                    help="eager iterations timed per launch for the roofline")
    ap.add_argument("--cpu-seconds", type=float, default=20.0,
                    help="search seconds of each CPU-baseline process (0 disables it)")
    ap.add_argument("--cpu-procs", type=int, default=min(16, os.cpu_count() or 1),
                    help="CPU baseline processes (default: every host CPU, at most 16)")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the launch-shape autotuner (default tile)")
    ap.add_argument("--tile", type=int, default=0,
                    help="plain and batched workloads: proposals per pass (64..256) instead of "
                         "the autotuner's")
    ap.add_argument("--grid", type=int, default=0,
                    help="batched workload: persistent workgroups instead of the autotuner's")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="after the run, copy this many MiB device-to-device (a copy "
                         "kernel of known size for calibrating HBM byte counters)")
    return ap.parse_args()


def _host_cpu() -> str:
    """CPU model and logical CPU count of this host (for the baseline's record)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return f"{model}, {os.cpu_count()} logical CPUs"


def cpu_baseline(wid: int, spec, run, target_secs: float, procs: int, lat0, fit0,
                 start: str) -> dict:
    """The reference engine (oracle/_ref/ref_hcs; the oracle's restatement when
    the binary is absent) running local search I on the lattices lat0, one
    process per core, each process on its own chunk of lattices.  A process
    repeats its chunk with fresh proposal streams (seed + 1000 + r) until
    about target_secs of its own search time have passed; the search time is
    the engine's own clock around the reference's local search (process start
    and file I/O excluded).  Swap evals come from the GPU search of exactly the
    same chunks and streams (bit-identical decisions, so the counts are the
    reference's).  value = sum over processes of evals / search seconds."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import RefHarness

    from pyhea_amd.hcs import Problem

    kind = "reference" if RefHarness.available() else "port"
    T, D = run["parallel_task"], run["converge_depth"]
    chunks = [c for c in np.array_split(np.arange(len(lat0)), procs) if len(c)]
    with tempfile.TemporaryDirectory() as tmp:
        jobs, outs = [], []
        t0 = time.perf_counter()
        for ci, idx in enumerate(chunks):
            f, fo = os.path.join(tmp, f"c{ci}.npz"), os.path.join(tmp, f"o{ci}.npy")
            np.savez(f, lat=lat0[idx], fit=fit0[idx])
            outs.append(fo)
            jobs.append(subprocess.Popen([sys.executable, os.path.abspath(__file__),
                                          "--cpu-worker", f, fo, kind, str(T), str(D),
                                          str(run["seed"]), str(wid), str(target_secs)]))
        for j in jobs:
            if j.wait() != 0:
                raise RuntimeError("cpu baseline process failed")
        wall = time.perf_counter() - t0
        secs = [np.load(fo) for fo in outs]  # per process: search seconds of each repetition
    p = Problem.from_spec(spec)
    rate, evals, reps = 0.0, 0, 0
    for idx, sc in zip(chunks, secs):
        ev = 0
        for r in range(len(sc)):
            _, _, st = p.local_mc(lat0[idx], fit0[idx], T, D, run["seed"] + 1000 + r, 0, 2,
                                  sort=False)
            ev += st["evals"]
        rate += ev / float(sc.sum())
        evals += ev
        reps += len(sc)
    p.close()
    search_s = float(sum(sc.sum() for sc in secs))
    return {
        "value": rate,
        "unit": "swap evals/s",
        "cores": len(chunks),
        "host": _host_cpu(),
        "kind": kind,
        "sample": f"local search I (T={T}, depth={D}) of {len(lat0)} {start} "
                  f"{spec.latt_typ} {spec.cell_dim}^3 lattices in {len(chunks)} processes "
                  f"(one chunk each, {reps} chunk searches with fresh streams): {evals} evals "
                  f"in {search_s:.1f} core-seconds of the engine's own search time, "
                  f"{wall:.1f} s wall",
    }


def cpu_worker(path: str, out: str, kind: str, T: int, D: int, seed: int, wid: int,
               target_secs: float) -> None:
    """One CPU-baseline process: local search I on its chunk of lattices,
    repeated with fresh streams until target_secs of search time."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle, RefHarness

    from pyhea_amd.problem import workload

    spec, _ = workload(wid)
    o = Oracle.from_spec(spec)
    d = np.load(path)
    secs = []
    while not secs or (sum(secs) < target_secs and len(secs) < 1000):
        r = len(secs)
        if kind == "reference":
            secs.append(RefHarness(o).local_mc(d["lat"], d["fit"], T, D, seed + 1000 + r, 0, 2)[3])
        else:
            t0 = time.perf_counter()
            o.local_mc(d["lat"], d["fit"], T, D, seed + 1000 + r, 0, 2, sort=False)
            secs.append(time.perf_counter() - t0)
    np.save(out, np.array(secs, dtype=np.float64))


def calib_copy(mb: int) -> None:
    """Copies mb MiB device to device (a kernel of known size for calibrating
    the HBM byte counters); nothing when mb <= 0."""
    if mb <= 0:
        return
    import torch

    src = torch.ones(mb * (1 << 20) // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()


def guided_launch_bytes(st: dict, T: int, sum_z: int, cand: int) -> int:
    """Algorithmic bytes of a guided local-search launch: every round scores
    T x cand candidate pairs (two 8-byte neighbour-count records and two
    species bytes each), walks its T chosen proposals (bytes_per_eval each;
    re-screens after an accept reuse the walks and read nothing), and every
    accepted swap rewrites the records of both sites' neighbours (8 bytes
    read + 8 written each)."""
    return (st["rounds"] * T * (bytes_per_eval(sum_z) + cand * (2 * 8 + 2))
            + st["accepts"] * 2 * sum_z * 16)


def time_to_fitness(a):
    """Headline: guided + batched vs plain time to fitness <= thr (see the
    module docstring)."""
    import torch

    from pyhea_amd import _native
    from pyhea_amd.hcs import Batch, Problem
    from pyhea_amd.problem import workload

    if _native.device_count() < 1:
        print(json.dumps({"error": "no GPU visible"}))
        sys.exit(1)
    wid = {"ttf-fcc30": 3, "ttf-fcc50": 4}[a.workload]
    spec, run = workload(wid)
    L, T, D, seed = run["solutions"], run["parallel_task"], run["converge_depth"], run["seed"]
    cand, acc = a.guide_candidates, a.guide_accepts
    prob = Problem.from_spec(spec)

    # guided + batched: persistent launch per phase, iteration replayed as a HIP graph
    b = Batch([prob], L, T, D, thr=a.thr, seed=seed)
    b.guide(candidates=cand, tau=0.1, accepts=acc)
    b.init()
    b.step(a.warmup)  # captures the graph; untimed
    b.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.init()  # fresh population; starts the device clock
    b.step(a.steps)
    b.sync()
    t1 = time.perf_counter()
    dt = t1 - t0
    _, gfit, gst = b.read(0)
    walked = gst["rounds"] * T
    # every persistent launch of the same iterations, eagerly, with HIP events
    b.init()
    ms, cnt = b.step_timed(a.steps)
    b.close()
    launches = [c for it in cnt for c in it]
    byts = sum(guided_launch_bytes(c, T, spec.sum_z, cand) for c in launches)
    ach = byts / (float(ms.sum()) * 1e-3) / 1e9

    # plain GPU search from its own fresh population until thr (or the budget)
    s = prob.session(L, T, D, thr=a.thr, seed=seed)
    t2 = time.perf_counter()
    s.init()
    it = 0
    while True:
        s.step(1)
        it += 1
        best = s.best()
        if best <= a.thr or time.perf_counter() - t2 > a.budget_s:
            break
    pwall = time.perf_counter() - t2
    plat, pfit, pst = s.read()
    s.close()
    ttf_g = gst["first_hit_s"] if gst["first_hit_s"] >= 0 else None
    ttf_p = pst["first_hit_s"] if pst["first_hit_s"] >= 0 else None
    plain = {"time_to_thr_s": ttf_p, "iterations": it, "wall_s": round(pwall, 3),
             "best": float(pfit.min()), "swap_evals": pst["evals"],
             "swap_evals_per_s": pst["evals"] / pwall}
    guided = {"time_to_thr_s": ttf_g, "iterations": a.steps, "best": float(gfit.min()),
              "swap_evals": walked, "swap_evals_per_s": walked / dt,
              "screens_incl_rescreens": gst["evals"], "accepts": gst["accepts"],
              "rounds": gst["rounds"], "candidates": cand, "tau": 0.1, "accepts_per_round": acc}
    cpu = None
    if a.cpu_seconds > 0:
        # the reference makes plain mode's decisions, so it needs the evals
        # plain mode needed; its rate is measured on the plain run's final
        # (converged) lattices, repeated with fresh streams
        n0 = min(L, 16 * a.cpu_procs)
        cpu = cpu_baseline(wid, spec, run, a.cpu_seconds, a.cpu_procs, plat[:n0], pfit[:n0],
                           "converged")
        if ttf_p:
            evals_to_thr = pst["evals"] / pwall * ttf_p  # plain's rate is steady
            cpu["time_to_thr_projected_s"] = evals_to_thr / cpu["value"]
            cpu["projection"] = ("plain mode's swap evals up to its first fitness <= thr "
                                 "divided by the measured reference rate; not run to the end")
    ttf = {"thr": a.thr, "guided_batched_s": ttf_g, "plain_s": ttf_p,
           "speedup_vs_plain": (ttf_p / ttf_g) if (ttf_p and ttf_g) else None,
           "cpu_projected_s": cpu.get("time_to_thr_projected_s") if cpu else None}
    ttf["speedup_vs_cpu"] = (ttf["cpu_projected_s"] / ttf_g) if (ttf["cpu_projected_s"]
                                                                  and ttf_g) else None
    prob.close()
    calib_copy(a.calib_copy_mb)
    out = {
        "metric": "swap evals/s; time to fitness<=1e-3: guided+batched vs plain MC, 1 MI355X",
        "value": walked / dt, "unit": "swap evals/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": dt * 1e3 / a.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": a.workload, "lattice": f"{spec.latt_typ} {spec.cell_dim}^3",
                   "n_sites": spec.n_sites, "species": spec.counts, "shells": spec.shell_sizes,
                   "solutions": L, "parallel_task": T, "converge_depth": D, "seed": seed,
                   "mode": "guided + batched (HIP graph; a batch of one problem runs the "
                           "per-lattice kernel)"},
        "time_to_fitness": ttf,
        "guided_batched": guided,
        "plain": plain,
        "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                     "traffic": (profiled_traffic("guided-fcc30", "k_local_mc")
                                 if a.workload == "ttf-fcc30" else
                                 profiled_traffic("", key="guided_traffic_bytes_per_launch",
                                                  fname="k_local_mc_fcc50.md")),
                     "kernel": "k_local_mc (guided, batch of one problem)",
                     "launches": len(launches),
                     "avg_launch_ms": round(float(ms.mean()), 3),
                     "bytes_per_launch": int(byts // len(launches)),
                     "byte_model": "rounds*T*(B_eval + cand*18) + accepts*2*sumZ*16",
                     "bytes_per_eval": bytes_per_eval(spec.sum_z)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def batched(a):
    """Batched workload: 64 problems of uneven size, species count and lattice
    type (cell edges 6-20, FCC/BCC, K 3-6) searched together, 128 solutions
    each, parallel_task 256, converge_depth 10; one persistent local-search
    launch per phase, the iteration replayed as a HIP graph.  Launch shape
    from the autotuner.  Algorithmic bytes per eval follow each problem's
    shell sizes; the rate is over whole iterations (every kernel of the graph
    counts), so it bounds the local-search kernel's rate from below."""
    from pyhea_amd.hcs import Batch, Problem
    from pyhea_amd.problem import batched_workload

    specs = batched_workload()
    probs = [Problem.from_spec(sp) for sp in specs]
    L, T, D, seed = 128, 256, 10, 2002
    shape = {"grid": a.grid or None, "tile": a.tile or None}
    if not a.no_tune and not (a.grid and a.tile):
        from pyhea_amd.autotune import tune_batch

        t = tune_batch(probs, L, T, D, seed=seed, cache_path=None,
                       grids=[a.grid] if a.grid else None, tiles=[a.tile] if a.tile else None)
        shape = {"grid": t["grid"], "tile": t["tile"]}
    b = Batch(probs, L, T, D, seed=seed)
    if shape["tile"]:
        b.tile(shape["tile"])
    if shape["grid"]:
        b.set_grid(shape["grid"])
    b.init()
    b.step(a.warmup)
    b.sync()
    ev0 = [b.read(q)[2]["evals"] for q in range(len(probs))]
    t0 = time.perf_counter()
    b.step(a.steps)
    b.sync()
    dt = time.perf_counter() - t0
    ev1 = [b.read(q)[2]["evals"] for q in range(len(probs))]
    evals = [e1 - e0 for e0, e1 in zip(ev0, ev1)]
    total = sum(evals)
    byts = sum(e * bytes_per_eval(sum(sp.shell_sizes[:sp.S])) for e, sp in zip(evals, specs))
    b.close()
    out = {
        "metric": "swap evals/s", "value": total / dt, "unit": "swap evals/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": dt * 1e3 / a.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "batched-64", "problems": len(specs),
                   "sites": [sp.n_sites for sp in specs], "solutions": L, "parallel_task": T,
                   "converge_depth": D, "seed": seed, "grid": shape["grid"], "tile": shape["tile"]},
        "roofline": {"bound": "hbm", "achieved": round(byts / dt / 1e9, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(byts / dt / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": profiled_traffic(a.workload, "k_local_mc_batch"),
                     "kernel": "k_local_mc_batch (whole iterations)"},
    }
    print(json.dumps(out))


def relax_bytes(n: int, sum_z: int, k: int, elem: int):
    """Algorithmic bytes of one relaxation pass: the forward reads every
    site's neighbour indices (int32) and occupancies; the backward reads the
    same and writes the gradient."""
    fwd = n * sum_z * 4 + n * k * elem
    return fwd, fwd + n * k * elem


def relaxation(a):
    """Relaxation pass on one problem (FCC 30^3: 108000 sites, 5 species,
    shells 12/6/24; or FCC 50^3: 500000 sites, 6 species): forward (softmax,
    neighbour gather, deterministic FP64 pair-count reduction) and backward
    (neighbour-gather adjoint) of the soft pair counts on float32 logits, the
    kernels behind SoftPairCounts.  One step = one forward + one backward at
    fixed synthetic logits and a fixed dn (the closed-form gradient of the
    relaxed fitness at the first forward's counts).  Each call is timed with
    HIP events on the stream it is launched on."""
    import torch

    from pyhea_amd import _native as Nat
    from pyhea_amd.hcs import Problem
    from pyhea_amd.problem import workload
    from pyhea_amd.relax import RelaxObjective

    wid = {"relax-fcc30": 3, "relax-fcc50": 4}[a.workload]
    spec, _ = workload(wid)
    prob = Problem.from_spec(spec)
    N, K, S = spec.n_sites, spec.K, spec.S
    dev = torch.device("cuda", 0)
    g = torch.Generator(device="cpu").manual_seed(wid)
    th = torch.randn((N, K), generator=g, dtype=torch.float32).to(dev)
    from pyhea_amd.relax import relax_cols

    kp = relax_cols(K)
    p = torch.empty((N, kp), dtype=torch.float32, device=dev)
    q = torch.empty((S, N, kp), dtype=torch.float32, device=dev)
    n = torch.empty((S, K, K), dtype=torch.float64, device=dev)
    grad = torch.empty_like(th)
    st = torch.cuda.current_stream()
    h = prob.handle
    lib = Nat.lib()

    def fwd():
        Nat.check(lib.hea_relax_forward(h, 0, th.data_ptr(), p.data_ptr(), q.data_ptr(),
                                        n.data_ptr(), st.cuda_stream))

    def bwd(dn):
        Nat.check(lib.hea_relax_backward(h, 0, p.data_ptr(), q.data_ptr(), dn.data_ptr(),
                                         grad.data_ptr(), st.cuda_stream))

    fwd()
    dn = RelaxObjective(prob, dev).count_gradient(n).contiguous()
    for _ in range(max(a.warmup, 1)):
        fwd()
        bwd(dn)
    steps = max(a.steps, 1)
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for e in ev:
        e[0].record(st)
        fwd()
        e[1].record(st)
        bwd(dn)
        e[2].record(st)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    f_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in ev]))
    b_ms = float(np.mean([e[1].elapsed_time(e[2]) for e in ev]))
    bf, bb = relax_bytes(N, spec.sum_z, K, 4)
    ach = (bf + bb) / ((f_ms + b_ms) * 1e-3) / 1e9
    cpu = None
    if a.cpu_seconds > 0:
        cpu = relax_cpu_baseline(spec, th.cpu(), dn.cpu(), min(a.cpu_seconds, 10.0), a.cpu_procs)
    out = {
        "metric": "relaxation passes/s", "value": steps / dt, "unit": "passes/s",
        "n_gpus": 1, "steps": steps, "warmup": a.warmup, "ms_per_step": dt * 1e3 / steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": a.workload, "lattice": f"{spec.latt_typ} {spec.cell_dim}^3",
                   "n_sites": N, "species": K, "shells": spec.shell_sizes[:S]},
        "forward_ms": round(f_ms, 4), "backward_ms": round(b_ms, 4),
        "forward_gbs": round(bf / (f_ms * 1e-3) / 1e9, 1),
        "backward_gbs": round(bb / (b_ms * 1e-3) / 1e9, 1),
        "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                     "traffic": profiled_traffic(a.workload, "k_relax", "traffic_bytes_per_step"),
                     "kernel": "relaxation forward + backward",
                     "bytes_forward": bf, "bytes_backward": bb},
        "cpu_baseline": cpu,
    }
    prob.close()
    calib_copy(a.calib_copy_mb)
    print(json.dumps(out))


def relax_cpu_baseline(spec, th, dn, target_secs: float, threads: int):
    """The same float32 forward + backward (softmax, neighbour sums,
    n_s = P^T Q_s, gradient of <n, dn>) in PyTorch on the host CPU, repeated
    for about target_secs (the relaxation has no counterpart in the
    reference engine)."""
    import torch

    torch.set_num_threads(max(1, threads))
    idx = []
    off = 0
    for z in spec.shell_sizes[:spec.S]:
        idx.append(torch.as_tensor(np.ascontiguousarray(spec.nbr[:, off:off + z]), dtype=torch.long))
        off += z
    dn64 = dn.to(torch.float64)
    passes, t0 = 0, time.perf_counter()
    while True:
        x = th.clone().requires_grad_(True)
        p = torch.softmax(x, dim=1)
        n = torch.stack([(p.T @ p[ix].sum(1)).double() for ix in idx])
        (n * dn64).sum().backward()
        passes += 1
        dt = time.perf_counter() - t0
        if dt >= target_secs or passes >= 200:
            break
    return {"value": passes / dt, "unit": "passes/s", "cores": torch.get_num_threads(),
            "kind": "port", "host": _host_cpu(),
            "sample": f"{passes} forward + backward passes of the same problem in PyTorch "
                      f"(float32, CPU), {dt:.1f} s"}


def main():
    if len(sys.argv) == 10 and sys.argv[1] == "--cpu-worker":
        cpu_worker(sys.argv[2], sys.argv[3], sys.argv[4], *map(int, sys.argv[5:9]),
                   float(sys.argv[9]))
        return
    a = parse()
    if a.steps is None:
        a.steps = 5 if a.workload.startswith("ttf-") else 49
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                          "(the population would shard one sub-population per GPU)"}))
        sys.exit(2)
    if a.dump_problem:
        from pyhea_amd.problem import workload

        spec, _ = workload({"plain-bcc16": 1, "plain-fcc10": 0, "ttf-fcc30": 3,
                            "ttf-fcc50": 4}[a.workload])
        with open(a.dump_problem, "wb") as f:
            z = list(spec.shell_sizes) + [0] * (3 - spec.S)
            f.write(np.array([spec.n_sites, spec.K, spec.S] + z, np.int32).tobytes())
            f.write(np.ascontiguousarray(spec.nbr, np.int32).tobytes())
            f.write(np.array(spec.counts, np.int32).tobytes())
            f.write(np.array(spec.weight, np.float64).tobytes())
            f.write(np.ascontiguousarray(spec.target, np.float64).tobytes())
        return
    if a.workload.startswith("ttf-"):
        time_to_fitness(a)
        return
    if a.workload == "batched-64":
        batched(a)
        return
    if a.workload.startswith("relax-"):
        relaxation(a)
        return
    from pyhea_amd import _native
    from pyhea_amd.hcs import Problem
    from pyhea_amd.problem import workload

    if _native.device_count() < 1:
        print(json.dumps({"error": "no GPU visible"}))
        sys.exit(1)
    wid = {"plain-bcc16": 1, "plain-fcc10": 0}[a.workload]
    spec, run = workload(wid)
    L, T, D, seed = run["solutions"], run["parallel_task"], run["converge_depth"], run["seed"]
    prob = Problem.from_spec(spec)
    tile = a.tile or None
    if not a.no_tune and not tile:  # launch shape only; results are identical for every tile
        from pyhea_amd.autotune import tune_session

        tile = tune_session(prob, L, T, D, seed=seed, cache_path=None)["tile"]
    sess = prob.session(L, T, D, thr=1e-3, seed=seed)
    if tile:
        sess.tile(tile)
    t_start = time.perf_counter()
    sess.init()
    sess.step(a.warmup)
    sess.sync()
    ev0 = sess.read()[2]["evals"]
    t0 = time.perf_counter()
    sess.step(a.steps)
    sess.sync()
    t1 = time.perf_counter()
    lat, fit, st = sess.read()
    evals = st["evals"] - ev0
    secs = t1 - t0
    value = evals / secs

    # roofline of the dominant kernel: per-launch times with HIP events
    ms, ev = sess.step_timed(a.profile_steps)
    bpe = bytes_per_eval(spec.sum_z)
    achieved = float((ev.sum() * bpe) / (ms.sum() * 1e-3) / 1e9)
    n_launch = ms.size
    roof = {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": profiled_traffic(a.workload),
        "kernel": "k_local_mc",
        "bytes_per_eval": bpe,
        "avg_launch_ms": round(float(ms.mean()), 4),
        "evals_per_launch": int(ev.sum() // n_launch),
    }
    calib_copy(a.calib_copy_mb)
    cpu = None
    if a.cpu_seconds > 0:
        # random lattices of the workload (the product's keyed permutations,
        # stream phase "new population"), four per process
        lat_r, fit_r = prob.generate(4 * a.cpu_procs, seed=seed, iter=0, phase=1)
        cpu = cpu_baseline(wid, spec, run, a.cpu_seconds, a.cpu_procs, lat_r, fit_r, "random")
    out = {
        "metric": "swap evals/s",
        "value": value,
        "unit": "swap evals/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": secs / max(a.steps, 1) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": a.workload,
            "lattice": f"{spec.latt_typ} {spec.cell_dim}^3",
            "n_sites": spec.n_sites,
            "species": spec.counts,
            "shells": spec.shell_sizes,
            "solutions": L,
            "parallel_task": T,
            "converge_depth": D,
            "total_iter": a.warmup + a.steps,
            "seed": seed,
            "tile": tile,
        },
        "best_fitness": float(fit[0]),
        "time_to_fitness_1e-3_s": st["first_hit_s"] if st["first_hit_s"] >= 0 else None,
        "wall_s": time.perf_counter() - t_start,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
