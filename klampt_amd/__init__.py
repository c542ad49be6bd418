"""MI355X-native batched collision, distance and feasibility engine behind
Klampt's planning hot path (RobotModel config updates, Geometry3D collide /
distance, WorldModel collision, RobotCSpace feasibility and edge checks).

The compute path is ``libklampt_gpu.so`` (HIP kernels for gfx950 behind a C
ABI, include/klampt_gpu.h); this package binds it with ctypes and mirrors the
Klampt Python API on top.
"""
from ._lib import KlamptGPUError, LIB_PATH  # noqa: F401

__version__ = "0.1.0"
