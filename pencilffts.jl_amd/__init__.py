"""MI355X-native PencilFFTs transform pipeline (one GPU).

Drop-in for PencilFFTs.jl's PencilFFTPlan with Transforms.FFT/BFFT/RFFT/
BRFFT/R2R/NoTransform: plans are built by the C++ plan layer and executed by
hand-written HIP kernels for gfx950 through the C ABI of include/pfft.h.
"""
from . import _lib
from . import transforms as Transforms
from . import autodiff, batched, spectral, workloads
from ._lib import ArgumentError, DeviceError, PfftError, UnsupportedError
from .plans import (ManyPencilArray, Pencil, PencilArray, PencilFFTPlan, Val, allocate_input,
                    allocate_output, bmul_, execute_batch, execute_many, export_wisdom, forget_wisdom,
                    import_wisdom, ldiv_, mul_, scale_factor)

__all__ = ["Transforms", "PencilFFTPlan", "PencilArray", "ManyPencilArray", "Pencil",
           "allocate_input",
           "allocate_output", "mul_", "ldiv_", "bmul_", "scale_factor", "Val",
           "execute_batch", "execute_many", "export_wisdom", "import_wisdom", "forget_wisdom",
           "ArgumentError", "UnsupportedError", "DeviceError", "PfftError",
           "autodiff", "batched", "spectral", "workloads"]

__version__ = _lib.lib.pfft_version().decode()
