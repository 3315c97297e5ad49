"""Build libpiper_hip.so (gfx950) in-tree with hipcc.

Every HIP translation unit gets -ffp-contract=off (the camera, the watertight triangle test and
the Sobol conversion must match the CPU restatement bit for bit; fused multiply-adds are written
out explicitly where they are wanted) and packed-FP32 instructions are disabled.
"""
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpiper_hip.so")
SOURCES = ["pg_api.cpp", "pg_kernels.hip", "pg_guide.hip", "pg_bvh_gpu.hip", "pg_bvh.cpp", "pg_sobol.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, src + ".o")
    lang = [] if src.endswith(".hip") else ["-x", "hip"]
    cmd = [HIPCC, *FLAGS, *lang, "-c", os.path.join(CSRC, src), "-o", obj]
    subprocess.run(cmd, check=True)
    return obj


PLUGIN_SRC = os.path.join(HERE, "plugin", "PiperHIP.cpp")
PLUGIN_LIB = os.path.join(HERE, "libpiper_hip_plugin.so")
FRONTEND_SRC = os.path.join(HERE, "frontend", "piper_render.cpp")
FRONTEND_BIN = os.path.join(HERE, "piper_render")
CXX = os.environ.get("CXX", "g++")


def build() -> str:
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp], check=True)
    os.replace(tmp, LIB)
    # C++ component pair over the C ABI (host code only: plain C++17, no HIP headers)
    tmp = PLUGIN_LIB + ".tmp"
    subprocess.run([CXX, "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wextra", PLUGIN_SRC, "-L" + HERE,
                    "-lpiper_hip", "-Wl,-rpath,$ORIGIN", "-o", tmp], check=True)
    os.replace(tmp, PLUGIN_LIB)
    # scene front-end (the Render operator over the component pair): JSON scene -> EXR frames
    tmp = FRONTEND_BIN + ".tmp"
    subprocess.run([CXX, "-O2", "-std=c++17", "-Wall", "-Wextra", FRONTEND_SRC, "-I" + os.path.join(HERE, "plugin"),
                    "-L" + HERE, "-lpiper_hip_plugin", "-lpiper_hip", "-Wl,-rpath,$ORIGIN", "-o", tmp], check=True)
    os.replace(tmp, FRONTEND_BIN)
    return LIB


if __name__ == "__main__":
    print(build())
