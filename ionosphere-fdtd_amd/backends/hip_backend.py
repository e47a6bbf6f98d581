"""HIP backend: torch ROCm tensors for state, libgeofdtd for the step.

Mirrors the allocation / host-exchange half of the reference's TorchBackend
(torch_backend.py:13-222) on a gfx950 device.  The surface operators of the
reference (edge_difference, face_circulation, ...) are not exposed: the fused
HIP step kernels implement them, and there is no second (eager) GPU path.
"""

from __future__ import annotations

from typing import Any

import numpy as np

from .base import ArrayBackend, BackendUnavailableError

_DTYPES = ("float32", "float64", "float16", "bfloat16")


class HipBackend(ArrayBackend):
    name = "hip"

    def __init__(self, mesh: Any, *, device: str = "auto", dtype: str = "auto"):
        try:
            import torch
        except ImportError as error:  # pragma: no cover - torch is part of the image
            raise BackendUnavailableError("the HIP backend needs PyTorch for device memory") \
                from error
        self.torch = torch
        if dtype == "auto":
            dtype = "float32"
        if dtype not in _DTYPES:
            raise ValueError("dtype must be 'auto', 'float32', 'float64', 'float16' or 'bfloat16'")
        self.dtype_name = dtype
        self.storage_dtype = getattr(torch, dtype)
        self.compute_dtype = torch.float64 if dtype == "float64" else torch.float32
        # asarray/zeros follow the arithmetic type so host exchange of
        # coefficients and traces is never narrowed.
        self.dtype = self.compute_dtype
        self.torch_device = self._resolve_device(device)
        self.device = str(self.torch_device)
        self.n_vertices = mesh.n_vertices

    def _resolve_device(self, requested: str):
        torch = self.torch
        requested = requested.lower()
        if requested in {"auto", "gpu", "hip"}:
            requested = "cuda"
        if not requested.startswith("cuda"):
            raise BackendUnavailableError("the HIP backend runs on a ROCm device (cuda[:N])")
        if not torch.cuda.is_available():
            raise BackendUnavailableError("no ROCm device is visible to PyTorch")
        device = torch.device(requested)
        if device.index is None:
            device = torch.device("cuda", torch.cuda.current_device())
        if device.index >= torch.cuda.device_count():
            raise BackendUnavailableError(f"device index {device.index} is unavailable")
        arch = torch.cuda.get_device_properties(device).gcnArchName
        if not str(arch).startswith("gfx950"):
            raise BackendUnavailableError(f"libgeofdtd is built for gfx950, device is {arch}")
        return device

    def current_stream_handle(self) -> int:
        return int(self.torch.cuda.current_stream(self.torch_device).cuda_stream)

    def synchronize(self) -> None:
        self.torch.cuda.synchronize(self.torch_device)

    def asarray(self, values: Any):
        return self.torch.tensor(np.asarray(values), dtype=self.dtype, device=self.torch_device)

    def index_array(self, values: Any):
        return self.torch.tensor(np.asarray(values), dtype=self.torch.long,
                                 device=self.torch_device)

    def zeros(self, shape: tuple[int, ...]):
        return self.torch.zeros(shape, dtype=self.dtype, device=self.torch_device)

    def empty_like(self, values: Any):
        return self.torch.empty_like(values)

    def to_numpy(self, values: Any) -> np.ndarray:
        if not self.torch.is_tensor(values):
            return np.asarray(values)
        values = values.detach()
        if values.dtype in (self.torch.float16, self.torch.bfloat16):
            values = values.float()
        return values.cpu().numpy()

    def scalar(self, value: Any) -> float:
        return float(value.detach().float().item() if value.dtype == self.torch.bfloat16
                     else value.detach().item())

    def max_abs(self, values: Any) -> float:
        return float(self.torch.max(self.torch.abs(values)).detach().float().item())

    def nbytes(self, values: Any) -> int:
        return int(values.numel() * values.element_size())
