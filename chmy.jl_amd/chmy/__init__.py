"""chmy: Chmy.jl's host API over the MI355X (gfx950) stencil library.

Mirrors the reference's exported names (src/Chmy.jl:91-99) for the hot path:
Arch / UniformGrid / Field / VectorField / TensorField / Launcher / bc! /
batch / Dirichlet / Neumann / set! / FunctionField / interior, with Julia's `!` suffix
written as a trailing underscore (bc_, set_).  Kernels are the HIP
counterparts of the reference's example kernels (chmy.kernels); every call
goes through libchmy_mi355x.so's C ABI.
"""
from . import _lib, kernels
from ._lib import ChmyError
from .architectures import (Arch, MI355XBackend, SingleDeviceArchitecture, activate, current_stream,
                            get_backend, get_device, heuristic_groupsize, set_device, synchronize)
from .boundary_conditions import (BatchSet, BoundaryFunction, Dirichlet, EmptyBatch, FieldBatch, FirstOrderBC,
                                  Neumann, batch, bc_)
from .fields import (Field, TensorField, VectorField, halo, interior, location, set_, set_storage,
                     tensor_locations, vector_location)
from .field_mask import FieldMask
from .function_field import FunctionField, TraceError, ifelse
from .grids import (Center, FunctionAxis, StructuredGrid, UniformAxis, UniformGrid, Vertex, coord, flip,
                    inv_spacing, spacing)
from .kernel_launch import Launcher, Offset

__all__ = [
    "Arch", "MI355XBackend", "SingleDeviceArchitecture", "activate", "current_stream", "get_backend",
    "get_device", "heuristic_groupsize", "set_device", "synchronize", "BatchSet", "BoundaryFunction", "Dirichlet", "EmptyBatch",
    "FieldBatch", "FirstOrderBC", "Neumann", "batch", "bc_", "Field", "TensorField", "VectorField", "halo",
    "interior", "location", "set_", "set_storage", "tensor_locations", "vector_location", "Center",
    "UniformGrid", "Vertex", "coord", "flip", "inv_spacing", "spacing", "Launcher", "Offset", "kernels",
    "ChmyError", "FunctionField", "TraceError", "ifelse", "FieldMask", "FunctionAxis", "UniformAxis",
    "StructuredGrid",
]
