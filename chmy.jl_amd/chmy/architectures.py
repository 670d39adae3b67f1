"""Architecture for the MI355X backend (src/Architectures.jl,
ext/ChmyAMDGPUExt/ChmyAMDGPUExt.jl hook set).

One process drives one GPU; device memory and streams come from
PyTorch-ROCm (plumbing), kernels from libchmy_mi355x.so.
"""
from __future__ import annotations

import torch

from . import _lib


class MI355XBackend:
    """KernelAbstractions-backend stand-in: names the device type of Fields."""

    def __repr__(self):
        return "MI355XBackend()"

    def __eq__(self, other):
        return isinstance(other, MI355XBackend)

    def __hash__(self):
        return hash("MI355XBackend")


class SingleDeviceArchitecture:
    """SingleDeviceArchitecture{B,D} (Architectures.jl:21-28)."""

    def __init__(self, backend, device):
        self.backend = backend
        self.device = device
        set_device(device)

    def __repr__(self):
        return f"SingleDeviceArchitecture({self.backend}, {self.device})"


def Arch(backend=None, *, device_id=1):
    """Arch(backend; device_id=1) (Architectures.jl:46-49); device_id is 1-based."""
    backend = backend or MI355XBackend()
    return SingleDeviceArchitecture(backend, get_device(backend, device_id))


def get_device(backend, device_id=1):
    return torch.device("cuda", device_id - 1)


def set_device(device):
    """set_device!: select the HIP device for torch and for the library."""
    if not torch.cuda.is_available():
        raise _lib.ChmyError("no HIP device visible: the MI355X backend needs a GPU")
    torch.cuda.set_device(device)
    _lib.call("chmy_init", device.index or 0)


def get_backend(arch):
    return arch.backend


def activate(arch, priority="normal"):
    """activate!(arch; priority) (Architectures.jl:71-74).  Streams stay at the
    default priority on this backend; the argument is accepted for API parity."""
    set_device(arch.device)


def heuristic_groupsize(backend, nd):
    """Work-group shape of the HIP kernels: one wave64 along x, 4 rows in y."""
    return {1: (256,), 2: (64, 4), 3: (64, 4, 1)}[nd]


def current_stream():
    """hipStream_t of torch's current stream, as a void* for the C ABI."""
    return torch.cuda.current_stream().cuda_stream


def synchronize(backend=None):
    """KernelAbstractions.synchronize(backend) (KernelLaunch.jl:117)."""
    _lib.call("chmy_stream_sync", current_stream())
