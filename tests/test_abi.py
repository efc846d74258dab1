"""C ABI and drop-in library checks that need no GPU: the libraries load,
export every declared entry point, and host-only calls behave."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
from ros_tsdf_amd import _build, abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "tsdf_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(tsdf_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_entry_point():
    L = abi.load()
    names = declared_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), n
    # and the Python binding knows every one of them
    bound = set(abi.SIGNATURES) | set(abi.SPECIAL_RESTYPE)
    assert set(names) == bound, set(names) ^ bound


@pytest.mark.parametrize("params", [(0.5, 0.1, 1000, 2500, 250000), (0.1, 0.4, 1000, 2500, 22500),
                                    (1.0, 4.0, 10000, 2500, 1e6), (0.25, 0.3, 50, 100, 100)])
def test_derived_params_match_oracle(params):
    from ros_tsdf_amd import derived_params
    assert derived_params(params) == oracle.derived(list(params))


def test_engine_create_fails_loudly_without_gpu():
    import torch  # noqa: F401  (device count only; no GPU initialisation is needed)
    L = abi.load()
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is present")
    cfg = abi.EngineConfig(0, 0, 1024, 1024, 1024)
    h = C.c_void_p()
    r = L.tsdf_engine_create(C.byref(cfg), C.byref(h))
    assert r != abi.TSDF_OK and h.value is None
    assert L.tsdf_last_error()


def test_invalid_arguments_rejected_before_any_device_call():
    L = abi.load()
    assert L.tsdf_engine_create(None, None) == abi.TSDF_ERR_INVALID_ARG
    assert L.tsdf_integrate(None, None, 0, None) == abi.TSDF_ERR_INVALID_ARG
    assert L.tsdf_publish(None, None, None, None, 0, None) == abi.TSDF_ERR_INVALID_ARG
    assert L.tsdf_gc(None, None, None) == abi.TSDF_ERR_INVALID_ARG
    assert b"invalid map" in L.tsdf_last_error()
    bad = abi.EngineConfig(0, 7, 1024, 1024, 1024)
    h = C.c_void_p()
    assert L.tsdf_engine_create(C.byref(bad), C.byref(h)) == abi.TSDF_ERR_INVALID_ARG
    with pytest.raises(abi.TsdfError):
        abi.check(L.tsdf_map_num_blocks(None, None))


def test_dropin_library_exports_reference_symbols():
    so = _build.PACKAGE_SO
    assert os.path.exists(so)
    out = subprocess.run(["nm", "-DC", so], capture_output=True, text=True, check=True).stdout
    for sig in [
        "tsdf::initGlobalVars(float, float, float, float, float)",
        "tsdf::TSDFHandler::TSDFHandler()",
        "tsdf::TSDFHandler::~TSDFHandler()",
        "tsdf::TSDFHandler::processPointCloudAndUpdateVoxels(std::shared_ptr<pcl::PointCloud<pcl::PointXYZ> >, "
        "Vector3f*, Vector3f*, int*, tsdf::Voxel*)",
        "tsdf::TSDFHandler::allocateVoxelBlocksAndUpdateVoxels(pcl::PointXYZ*, Vector3f*, int*, int, "
        "tsdf::HashTable*, tsdf::BlockHeap*)",
        "tsdf::TSDFHandler::retrievePublishableVoxels(Vector3f*, Vector3f*, int*, tsdf::Voxel*, tsdf::HashTable*, "
        "tsdf::BlockHeap*)",
        "tsdf::TSDFHandler::garbageCollectDistantBlocks(Vector3f*, tsdf::HashTable*, tsdf::BlockHeap*)",
        "tsdf::TSDFContainer::getCudaHashTable()",
        "tsdf::TSDFContainer::getCudaBlockHeap()",
    ]:
        assert sig in out, sig


def test_layouts():
    # scan descriptors and requests are plain C structs
    assert C.sizeof(abi.ScanDesc) == 32
    assert C.sizeof(abi.Request) == 32
    assert C.sizeof(abi.ScanStats) == 16 * 8
    assert np.dtype(np.float32).itemsize * 4 == 16  # pcl::PointXYZ rows


def test_calibration_library_exports_its_entry_points():
    from ros_tsdf_amd import calib
    src = open(os.path.join(ROOT, "include", "tsdf_calib.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = sorted(set(re.findall(r"\b(tsdf_[a-z0-9_]+)\s*\(", src)))
    assert names == ["tsdf_calib_atomics", "tsdf_calib_last_error"]
    L = calib.load()
    for n in names:
        assert hasattr(L, n), n
    # arguments are checked before any device call
    ops, secs = C.c_int64(), C.c_double()
    assert L.tsdf_calib_atomics(0, 9, 1, 1, C.byref(ops), C.byref(secs)) == -1
    assert b"unknown pattern" in L.tsdf_calib_last_error()
    assert L.tsdf_calib_atomics(0, 0, 0, 1, C.byref(ops), C.byref(secs)) == -1
