"""MI355X-native DG residual and SSP-RK3 engine for the 2D nozzle Euler solver.

The package directory name starts with a digit, so import it with
``importlib.import_module("2d_nozzle_amd")`` (or load it under an alias).

``get_backend("hip", ops, flow, opts)`` is the registry branch the reference's
``src/backends/__init__.py:41-57`` gains; ``HipBackend`` is the drop-in.
"""

from __future__ import annotations

__version__ = "0.1.0"

BACKENDS = ("hip",)


def available_backends() -> tuple[str, ...]:
    """``("hip",)`` when a ROCm GPU is visible, else ``()``."""
    try:
        import torch
    except ImportError:  # pragma: no cover
        return ()
    return ("hip",) if torch.cuda.is_available() else ()


def get_backend(name: str, ops, flow, opts):
    """Instantiate a backend by name (``src/backends/__init__.py:41-57``)."""
    if name == "hip":
        from .backend import HipBackend

        return HipBackend(ops, flow, opts)
    raise ValueError(f"unknown backend {name!r}; choose from {BACKENDS}")
