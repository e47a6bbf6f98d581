"""Backend selection (the reference's backends/__init__.py:12-33)."""

from __future__ import annotations

from typing import Any

from .base import Array, ArrayBackend, BackendUnavailableError


def create_backend(name: str, mesh: Any, *, device: str = "auto", dtype: str = "auto",
                   torch_threads: int | None = None) -> ArrayBackend:
    """Create the HIP backend.

    ``"hip"`` is the only compute backend of this package; ``"torch"`` /
    ``"pytorch"`` on a ROCm device are accepted as aliases because PyTorch's
    ``cuda`` device is the HIP device here.  ``"numpy"`` and CPU devices raise
    ``BackendUnavailableError``: there is no CPU fallback.
    """
    normalized = name.lower()
    if torch_threads is not None:
        raise BackendUnavailableError("torch_threads is only valid for a CPU backend")
    if normalized in {"hip", "torch", "pytorch"}:
        from .hip_backend import HipBackend

        return HipBackend(mesh, device=device, dtype=dtype)
    if normalized == "numpy":
        raise BackendUnavailableError(
            "this package has no NumPy backend; the time step runs only on the HIP device")
    raise ValueError("backend must be 'hip' (or its aliases 'torch'/'pytorch')")


__all__ = ["Array", "ArrayBackend", "BackendUnavailableError", "create_backend"]
