"""Build the TurboNeRF MI355X native libraries in-tree.

* ``turbonerf_amd/lib/libturbonerf_kernels.so`` -- every hand-written gfx950
  HIP kernel behind the ``tn_*`` C ABI (include/turbonerf/tn_c_api.h).
* ``turbonerf_amd/lib/libturbonerf.so`` -- the C++ host layer (Trainer,
  NerfNetwork, OccupancyGrid, Renderer, controllers) over that C ABI.
* ``turbonerf_amd/PyTurboNeRF*.so`` -- the pybind11 module mirroring the
  reference's Python surface.

Run ``python -m turbonerf_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
import sysconfig
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "lib"
OBJ = PKG / "_build"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-mcode-object-version=5",
    # packed fp32 VALU forms are kept out of generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-command-line-argument",
]

HOST_FLAGS = [
    "-O2",
    "-fPIC",
    "-std=c++17",
    "-D__HIP_PLATFORM_AMD__",
    f"-I{ROCM / 'include'}",
    f"-I{ROOT / 'include'}",
    "-Wall",
]

# per-source code generation options: the fused MLP kernels keep their MFMA
# results in ordinary VGPRs, so the VALU work between layers reads them
# without a v_accvgpr_read per value
KERNEL_FLAGS = {"mlp": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}

KERNEL_SOURCES = sorted((CSRC / "kernels").glob("*.hip"))
HOST_SOURCES = sorted((CSRC / "host").glob("*.cpp"))
API_SOURCES = sorted((CSRC / "api").glob("*.cpp"))


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + res.stdout + res.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} ... {cmd[-1]}")


def _hipcc() -> str:
    exe = shutil.which("hipcc") or str(ROCM / "bin" / "hipcc")
    if not Path(exe).exists():
        raise RuntimeError("hipcc not found; the TurboNeRF kernels need ROCm's hipcc")
    return exe


def build_kernels(jobs: int = 8) -> Path:
    OBJ.mkdir(parents=True, exist_ok=True)
    LIB.mkdir(parents=True, exist_ok=True)
    hipcc = _hipcc()
    objs = []
    cmds = []
    for src in KERNEL_SOURCES:
        obj = OBJ / (src.stem + ".o")
        objs.append(obj)
        cmds.append([hipcc, *HIP_FLAGS, *KERNEL_FLAGS.get(src.stem, []), "-c", str(src), "-o", str(obj)])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    out = LIB / "libturbonerf_kernels.so"
    tmp = LIB / "libturbonerf_kernels.so.tmp"
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def build_host(jobs: int = 8) -> Path | None:
    if not HOST_SOURCES:
        return None
    OBJ.mkdir(parents=True, exist_ok=True)
    cxx = shutil.which("g++") or "g++"
    objs, cmds = [], []
    for src in HOST_SOURCES:
        obj = OBJ / ("host_" + src.stem + ".o")
        objs.append(obj)
        cmds.append([cxx, *HOST_FLAGS, "-c", str(src), "-o", str(obj)])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    out = LIB / "libturbonerf.so"
    tmp = LIB / "libturbonerf.so.tmp"
    _run([cxx, "-shared", "-fPIC", *map(str, objs), "-o", str(tmp),
          f"-L{LIB}", "-lturbonerf_kernels", f"-L{ROCM / 'lib'}", "-lamdhip64", "-lz",
          "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{ROCM / 'lib'}"])
    os.replace(tmp, out)
    return out


def build_python_module() -> Path | None:
    if not API_SOURCES:
        return None
    import pybind11

    cxx = shutil.which("g++") or "g++"
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    out = PKG / f"PyTurboNeRF{suffix}"
    tmp = PKG / f"PyTurboNeRF{suffix}.tmp"
    py_inc = sysconfig.get_paths()["include"]
    _run([cxx, *HOST_FLAGS, "-shared", f"-I{pybind11.get_include()}", f"-I{py_inc}",
          f"-I{CSRC / 'host'}", *map(str, API_SOURCES), "-o", str(tmp),
          f"-L{LIB}", "-lturbonerf", "-lturbonerf_kernels", f"-L{ROCM / 'lib'}", "-lamdhip64",
          "-Wl,-rpath,$ORIGIN/lib", f"-Wl,-rpath,{ROCM / 'lib'}"])
    os.replace(tmp, out)
    return out


def build_oracle() -> Path:
    """The oracle is test infrastructure; it is built beside the product."""
    _run(["make", "-s", "-C", str(ROOT / "oracle")])
    return ROOT / "oracle" / "build" / "libtn_oracle.so"


def build_all(jobs: int = 8) -> None:
    build_kernels(jobs)
    build_host(jobs)
    build_python_module()


if __name__ == "__main__":
    build_all()
    build_oracle()
    print("built", LIB)
