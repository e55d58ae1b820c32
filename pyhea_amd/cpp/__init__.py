"""Drop-in for ``pyhea.cpp`` (pyhea/cpp/__init__.py): the pybind11 module
``accelerate`` with run_local_parallel_hcs, run_local_parallel_hcs_cuda and
cuda_available over the MI355X engine.  Import fails loudly if the native
module has not been built."""
try:
    from . import accelerate  # noqa: F401
    from .accelerate import *  # noqa: F401,F403
except ImportError as e:  # pragma: no cover - message for an unbuilt tree
    raise ImportError(
        f"pyhea_amd.cpp.accelerate is not built ({e}); run python -m pyhea_amd.build"
    ) from e
