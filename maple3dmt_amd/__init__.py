"""MI355X-native octree FV curl-curl forward / adjoint solve for 3-D MT.

The package exposes the host producer and the device solve of
``lib/libm3dmt.so`` through ctypes (``native``) and the workload definitions
(``configs``).
"""
from .native import (Context, DeviceOperator, ForwardParams, MeshParams, Model, NativeError,
                     SolveParams, MODE_FORWARD, MODE_INNER, MODE_INTERLEAVED, MODE_SYSTEM,
                     PREC_FP64, PREC_MIXED, lib)

__all__ = ["Context", "DeviceOperator", "ForwardParams", "MeshParams", "Model", "NativeError",
           "SolveParams", "MODE_FORWARD", "MODE_INNER", "MODE_INTERLEAVED", "MODE_SYSTEM",
           "PREC_FP64", "PREC_MIXED", "lib"]
