"""peridyn_amd — MI355X-native explicit-dynamics hot path of PeriDyn.

The HIP library (libperidyn_hip.so) is loaded on first use (Context creation);
importing the package needs no GPU.
"""
from . import workloads  # noqa: F401

__all__ = ["Context", "Env", "run_batch", "workloads"]


def __getattr__(name):
    if name in ("Context", "Env", "run_batch"):
        from . import core
        return getattr(core, name)
    raise AttributeError(name)
