"""HIP backend: torch ROCm tensors for state, libgeofdtd for the step.

Mirrors the reference's TorchBackend (torch_backend.py:13-222) on a gfx950
device: allocation and host exchange through torch, and the surface
operators of the backend interface (diff, edge_difference,
dual_edge_difference, face_circulation, dual_cell_circulation;
backends/base.py:38-56) as libgeofdtd kernels with the reference's order of
additions.  The time step itself runs only in the fused step kernels
(``compile_step``); there is no eager per-operator step.
"""

from __future__ import annotations

import ctypes as C
import weakref
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
        self.mesh = mesh
        self._ops = None

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

    # ------------------------------------------------ surface operators
    def _mesh_ops(self):
        if self._ops is None:
            from .. import _native
            from ..mesh import vertex_incidence

            lib = _native.load_library()
            mesh = self.mesh
            ve, vs = vertex_incidence(mesh)
            keep = [np.ascontiguousarray(a, dtype=np.int64) for a in
                    (mesh.edges, mesh.face_edges, mesh.face_edge_signs, mesh.edge_left_faces,
                     mesh.edge_right_faces, ve, vs)]
            desc = _native.GeoMeshDesc(
                n_vertices=mesh.n_vertices, n_edges=mesh.n_edges, n_faces=mesh.n_faces,
                edges=_native.i64_ptr(keep[0]), face_edges=_native.i64_ptr(keep[1]),
                face_edge_signs=_native.i64_ptr(keep[2]), edge_left_faces=_native.i64_ptr(keep[3]),
                edge_right_faces=_native.i64_ptr(keep[4]), vertex_edges=_native.i64_ptr(keep[5]),
                vertex_edge_signs=_native.i64_ptr(keep[6]), max_degree=ve.shape[1])
            handle = C.c_void_p()
            _native.check(lib.geo_mesh_ops_create(C.byref(desc), C.byref(handle)), lib)
            self._ops = (lib, handle)
            weakref.finalize(self, lib.geo_mesh_ops_destroy, handle)
        return self._ops

    def _operand(self, values: Any):
        torch = self.torch
        x = values if torch.is_tensor(values) else self.asarray(values)
        if x.device != self.torch_device:
            raise ValueError(f"operator input must be on {self.torch_device}")
        if x.dtype not in (torch.float32, torch.float64):
            raise ValueError("surface operators take float32 or float64 values")
        return x.contiguous()

    def _mesh_op(self, op: int, values: Any, in_rows: int, out_rows: int):
        from .. import _native

        x = self._operand(values)
        if x.dim() < 1 or x.shape[0] != in_rows:
            raise ValueError(f"operator input must have {in_rows} rows, got {tuple(x.shape)}")
        lib, handle = self._mesh_ops()
        cols = int(np.prod(x.shape[1:])) if x.dim() > 1 else 1
        out = self.torch.empty((out_rows,) + tuple(x.shape[1:]), dtype=x.dtype,
                               device=self.torch_device)
        if out.numel():
            _native.check(lib.geo_mesh_op(handle, op, x.element_size(), x.data_ptr(), cols,
                                          out.data_ptr(), self.current_stream_handle()), lib)
        return out

    def edge_difference(self, vertex_values: Any):
        """Head-minus-tail values on primal edges (numpy_backend.py:93-94)."""
        return self._mesh_op(0, vertex_values, self.mesh.n_vertices, self.mesh.n_edges)

    def dual_edge_difference(self, face_values: Any):
        """Left-minus-right values across primal edges (numpy_backend.py:96-97)."""
        return self._mesh_op(1, face_values, self.mesh.n_faces, self.mesh.n_edges)

    def face_circulation(self, edge_values: Any):
        """Oriented edge values summed around primal faces (numpy_backend.py:99-104)."""
        return self._mesh_op(2, edge_values, self.mesh.n_edges, self.mesh.n_faces)

    def dual_cell_circulation(self, edge_values: Any):
        """Oriented dual-edge values accumulated around dual cells, slot by slot
        (numpy_backend.py:106-114)."""
        return self._mesh_op(3, edge_values, self.mesh.n_edges, self.mesh.n_vertices)

    def diff(self, values: Any, axis: int):
        """Adjacent differences along an axis of a 1-D or 2-D array (np.diff)."""
        from .. import _native

        x = self._operand(values)
        if x.dim() not in (1, 2):
            raise ValueError("diff takes a 1-D or 2-D array")
        axis = axis + x.dim() if axis < 0 else axis
        if not 0 <= axis < x.dim():
            raise ValueError(f"axis {axis} is out of range for a {x.dim()}-D array")
        rows, cols = (1, x.shape[0]) if x.dim() == 1 else (x.shape[0], x.shape[1])
        ax = 1 if x.dim() == 1 else axis
        shape = list(x.shape)
        shape[axis] -= 1
        out = self.torch.empty(shape, dtype=x.dtype, device=self.torch_device)
        if out.numel():
            lib = _native.load_library()
            _native.check(lib.geo_array_diff(x.element_size(), x.data_ptr(), rows, cols, ax,
                                             out.data_ptr(), self.current_stream_handle()), lib)
        return out
