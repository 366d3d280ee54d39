"""Build libdpmring.so (gfx950) in-tree with hipcc.

Every HIP translation unit is compiled with:
  --offload-arch=gfx950            the MI355X ISA only
  -ffp-contract=off                no FMA contraction, so the kernels evaluate the
                                   same expression order as the host restatement
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                   no packed FP32 instructions
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libdpmring.so")

SOURCES = ["dpm_kernels.hip", "dpm_adjoint.hip", "dpm_observe.hip", "dpm_api.hip"]

# C++ class layer (DPM2D / FIRE / Simulator / ioDPMFile) and the drivers built on it: plain host
# C++ over the C ABI, linked against libdpmring.so (found next to bin/ through the rpath)
HOST = os.path.join(HERE, "host")
DRIVERS = os.path.join(HERE, "drivers")
BIN = os.path.join(HERE, "bin")
HOST_SOURCES = ["DPM2D.cpp", "FIRE.cpp", "Simulator.cpp"]
DRIVER_SOURCES = ["compressFIRE.cpp", "makeDPM.cpp", "compressNVE.cpp", "runDPM.cpp", "runRigid.cpp", "testInteraction.cpp",
                  "runNVE.cpp", "runNVT.cpp", "runActive.cpp", "runActiveShape.cpp"]
CXX = os.environ.get("CXX", "g++")
CXXFLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-variable"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-parameter", "-Wno-unused-value",
]


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def _cxx(src, obj):
    r = subprocess.run([CXX, *CXXFLAGS, "-c", src, "-o", obj], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{CXX} failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build_host(verbose=False):
    """Compile the class layer and link one executable per driver into bin/."""
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(BIN, exist_ok=True)
    jobs = [(os.path.join(HOST, s), os.path.join(BUILD, "host_" + os.path.splitext(s)[0] + ".o"))
            for s in HOST_SOURCES]
    jobs += [(os.path.join(DRIVERS, s), os.path.join(BUILD, "drv_" + os.path.splitext(s)[0] + ".o"))
             for s in DRIVER_SOURCES]
    with ThreadPoolExecutor(max_workers=len(jobs)) as ex:
        objs = list(ex.map(lambda j: _cxx(*j), jobs))
    host_objs = objs[:len(HOST_SOURCES)]
    for src, obj in zip(DRIVER_SOURCES, objs[len(HOST_SOURCES):]):
        exe = os.path.join(BIN, os.path.splitext(src)[0])
        cmd = [CXX, "-o", exe, obj, *host_objs, f"-L{HERE}", "-ldpmring", "-Wl,-rpath,$ORIGIN/.."]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"link failed for {src}:\n{r.stdout}\n{r.stderr}")
        if verbose:
            print(f"built {exe}")
    return BIN


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    build_host(verbose=True)
    sys.exit(0)
