"""Array-backend contract (the reference's backends/base.py:15-89).

The time step itself never goes through these operator methods: the HIP
backend runs the fused step kernels behind ``compile_step``.  The methods
exist so code written against the reference's ``ArrayBackend`` surface
(allocation, host exchange, synchronisation) keeps working.
"""

from __future__ import annotations

from abc import ABC, abstractmethod
from collections.abc import Callable
from typing import Any

import numpy as np
from numpy.typing import NDArray

Array = Any


class BackendUnavailableError(RuntimeError):
    """Raised when a requested compute backend or device cannot be used."""


class ArrayBackend(ABC):
    """Minimal tensor operations required around the FDTD update loop."""

    name: str
    device: str
    dtype_name: str
    threads: int | None = None

    def compile_step(self, step: Callable[[Array], None]) -> Callable[[Array], None]:
        raise BackendUnavailableError(
            f"the {self.name} backend does not support compiled field steps")

    def synchronize(self) -> None:
        """Wait for queued device work."""

    @abstractmethod
    def asarray(self, values: Any) -> Array: ...

    @abstractmethod
    def index_array(self, values: Any) -> Array: ...

    @abstractmethod
    def zeros(self, shape: tuple[int, ...]) -> Array: ...

    @abstractmethod
    def empty_like(self, values: Array) -> Array: ...

    @abstractmethod
    def to_numpy(self, values: Array) -> NDArray[np.generic]: ...

    @abstractmethod
    def scalar(self, value: Array) -> float: ...

    @abstractmethod
    def max_abs(self, values: Array) -> float: ...

    @abstractmethod
    def nbytes(self, values: Array) -> int: ...
