"""MI355X-native lidar -> TSDF integration engine.

The compute path is the HIP library built from ros_tsdf_amd/csrc (loaded by
ros_tsdf_amd.abi); this package holds the Python host interface over its C
ABI, workload generation and the batched / serving / autotuning drivers.
The reference-shaped C++ handler (tsdf::TSDFHandler) is built from
csrc/tsdf_handler.cpp into libtsdf_package.so.
"""
from . import abi  # noqa: F401
from .engine import BatchGraph, DeviceBuffer, Engine, Map, derived_params, serve  # noqa: F401
