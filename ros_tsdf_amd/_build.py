"""Build recipes of the native libraries (gfx950 HIP engine and the C++ drop-in).

Outputs go to ros_tsdf_amd/_lib/ in the source tree so that they travel with
the repository to the GPU host; nothing is installed into site-packages.
"""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_DIR = os.path.join(HERE, "_lib")
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")

ENGINE_SO = os.path.join(LIB_DIR, "libtsdf_hip.so")
PACKAGE_SO = os.path.join(LIB_DIR, "libtsdf_package.so")
CALIB_SO = os.path.join(LIB_DIR, "libtsdf_calib.so")
HANDLER_TEST_BIN = os.path.join(LIB_DIR, "handler_dropin_test")

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: no FMA contraction (bit-exact FP32 geometry);
# -packed-fp32-ops off: no v_pk_*_f32 instructions.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-value", "-Wno-unused-result",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return r


def build_engine() -> str:
    os.makedirs(LIB_DIR, exist_ok=True)
    tmp = ENGINE_SO + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-I", INCLUDE, "-o", tmp, os.path.join(CSRC, "tsdf_engine.hip")])
    os.replace(tmp, ENGINE_SO)
    return ENGINE_SO


def build_calib() -> str:
    """The bench's device-rate calibration kernels (include/tsdf_calib.h)."""
    os.makedirs(LIB_DIR, exist_ok=True)
    tmp = CALIB_SO + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-I", INCLUDE, "-o", tmp, os.path.join(CSRC, "calib.hip")])
    os.replace(tmp, CALIB_SO)
    return CALIB_SO


def build_package() -> str:
    """The C++ drop-in (tsdf::TSDFHandler & co.) over the C ABI, plus its test program."""
    os.makedirs(LIB_DIR, exist_ok=True)
    tmp = PACKAGE_SO + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I", INCLUDE,
          "-o", tmp, os.path.join(CSRC, "tsdf_handler.cpp"),
          "-L", LIB_DIR, "-ltsdf_hip", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, PACKAGE_SO)
    tmpb = HANDLER_TEST_BIN + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I", INCLUDE, "-I", os.path.join(ROCM, "include"),
          "-o", tmpb, os.path.join(ROOT, "tests", "cpp", "handler_dropin_test.cpp"),
          "-L", LIB_DIR, "-ltsdf_package", "-ltsdf_hip", "-L", os.path.join(ROCM, "lib"), "-lamdhip64",
          "-Wl,-rpath,$ORIGIN"])
    os.replace(tmpb, HANDLER_TEST_BIN)
    return PACKAGE_SO


def build_all() -> None:
    build_engine()
    build_package()
    build_calib()


if __name__ == "__main__":
    build_all()
    print(ENGINE_SO)
