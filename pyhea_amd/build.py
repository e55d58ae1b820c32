"""In-tree build of the native parts.

* ``pyhea_amd/libpyhea_hcs.so``: the HIP kernels and the C ABI of
  ``include/pyhea_hcs.h``, compiled for gfx950 with hipcc.
* ``pyhea_amd/cpp/accelerate*.so``: the pybind11 module that keeps the
  reference's ``pyhea.cpp.accelerate`` names (``pyhea/cpp/bind.cpp``).

Floating-point flags matter: the kernels reproduce the reference CPU engine's
doubles bit for bit, so FP contraction is off (no a*b+c fusion) and fast-math
is never enabled.  Packed FP32 instructions are disabled as required on this
hardware.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sysconfig
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libpyhea_hcs.so")
ARCH = os.environ.get("PYHEA_AMD_ARCH", "gfx950")

HIP_SOURCES = ["hea_mc.hip", "hea_api.hip", "hea_relax.hip"]
HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP extension cannot be built")


def _run(cmd: list[str], cwd: str | None = None) -> None:
    res = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"build step failed ({res.returncode}): {' '.join(cmd)}\n{res.stdout}")


def _stale(obj: str, src: str) -> bool:
    """True when obj is older than its source or any header of csrc/ or include/."""
    if not os.path.exists(obj):
        return True
    deps = [src] + [os.path.join(d, f) for d in (CSRC, os.path.join(ROOT, "include"))
                    for f in os.listdir(d) if f.endswith(".h")]
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(d) > t for d in deps)


def build_hip(verbose: bool = False, incremental: bool = False) -> str:
    """Compiles every HIP source (incremental=True: only those whose object is
    out of date) and links the library."""
    os.makedirs(BUILD, exist_ok=True)
    hipcc = _hipcc()
    sources = [s for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]

    def compile_one(src: str) -> str:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        if incremental and not _stale(obj, os.path.join(CSRC, src)):
            return obj
        _run([hipcc, *HIPCC_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        if verbose:
            print(f"[build] {src}", flush=True)
        return obj

    with ThreadPoolExecutor(max_workers=len(sources)) as ex:
        objs = list(ex.map(compile_one, sources))
    tmp = LIB + ".tmp"
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


def build_bind(verbose: bool = False) -> str:
    import pybind11

    ext = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    out = os.path.join(PKG, "cpp", "accelerate" + ext)
    py_inc = sysconfig.get_paths()["include"]
    cmd = [
        os.environ.get("CXX", "g++"),
        "-O2",
        "-std=c++17",
        "-shared",
        "-fPIC",
        "-fvisibility=hidden",
        f"-I{pybind11.get_include()}",
        f"-I{py_inc}",
        os.path.join(CSRC, "accelerate_bind.cpp"),
        f"-L{PKG}",
        "-lpyhea_hcs",
        "-Wl,-rpath,$ORIGIN/..",
        "-o",
        out + ".tmp",
    ]
    _run(cmd)
    os.replace(out + ".tmp", out)
    if verbose:
        print(f"[build] {os.path.basename(out)}", flush=True)
    return out


def build_profile_driver(verbose: bool = False) -> str:
    """pyhea_amd/hea_profile: a Python-free driver for hardware-counter runs."""
    out = os.path.join(PKG, "hea_profile")
    _run([_hipcc(), "-O2", "-std=c++17", os.path.join(CSRC, "hea_profile_main.cpp"),
          f"-L{PKG}", "-lpyhea_hcs", "-Wl,-rpath,$ORIGIN", "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    if verbose:
        print("[build] hea_profile", flush=True)
    return out


def build_oracle(with_reference: bool | None = None) -> None:
    """Test infrastructure: the CPU oracle and, where the reference tree
    exists, the reference harness (outputs under oracle/)."""
    ref_root = os.environ.get("PYHEA_REF_ROOT", "/root/reference")
    targets = ["oracle"]
    if with_reference is None:
        with_reference = os.path.exists(os.path.join(ref_root, "pyhea", "cpp", "accelerate.cpp"))
    if with_reference:
        targets.append("ref")
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), f"REF_ROOT={ref_root}", *targets])


def build_all(verbose: bool = False, incremental: bool = False) -> None:
    build_hip(verbose, incremental)
    build_bind(verbose)
    build_profile_driver(verbose)


if __name__ == "__main__":
    import sys

    build_all(verbose=True, incremental="--incremental" in sys.argv)
    build_oracle()
