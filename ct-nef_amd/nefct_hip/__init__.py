"""nefct_hip -- MI355X-native (gfx950) cone-beam CT operators.

A drop-in for the projector path of bill52547/CT-NEF (``nefct``): the Siddon
forward projector and its adjoint, the FDK weighted backprojection and the
trilinear volume warp, as hand-written HIP kernels behind a C ABI
(include/nefct_hip.h), exposed as PyTorch-ROCm custom ops.
"""
from ._lib import NefctError, REF_ADJOINT, REF_FORWARD, load  # noqa: F401
from .geometry import ConeGeometry, VolumeGeometry  # noqa: F401
from . import ops  # noqa: F401
from .siddon import (SiddonBackprojection, SiddonProjection, back_project, backproject,  # noqa: F401
                     count_intersections, forward_project, project)

__all__ = [
    "NefctError", "ConeGeometry", "VolumeGeometry", "project", "backproject", "count_intersections",
    "forward_project", "back_project", "SiddonProjection", "SiddonBackprojection", "REF_FORWARD",
    "REF_ADJOINT", "load",
]
