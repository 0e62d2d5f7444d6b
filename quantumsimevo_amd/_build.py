"""In-tree build of the native parts of quantumsimevo_amd.

Products (all written next to the sources, never installed):
  quantumsimevo_amd/libqse_hip.so        HIP kernels for gfx950 + the C ABI (include/qse_hip.h)
  quantumsimevo_amd/_quantumsimevo*.so   pybind11 module: the C++ Circuit / StateVector /
                                         Noise / CircuitUnitaryOperationHIP host classes
  oracle/liboracle.so (+ oracle/_ref/)   the CPU oracle and, where the reference tree is present, its
                                         CPU backends and drop-in module (test infrastructure, via
                                         oracle/Makefile)

Every HIP source is compiled with hipcc --offload-arch=gfx950 and with packed
FP32 instructions disabled (-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
from __future__ import annotations

import os
import subprocess
import sysconfig
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
HOST = CSRC / "host"
INCLUDE = ROOT / "include"
BUILD = CSRC / "_build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value",
    f"-I{INCLUDE}", f"-I{CSRC}",
]
LIB_SOURCES = ["qse_capi.cpp", "qse_run.cpp", "qse_gates.hip", "qse_readout.hip", "qse_fused.hip",
               "qse_plan.cpp", "qse_adjoint.hip", "qse_jit.cpp", "qse_jit_gen.cpp", "qse_check.hip",
               "qse_adjoint_plan.cpp", "qse_batch_plan.cpp", "qse_batch.hip", "qse_dense.hip"]
HOST_SOURCES = ["StateVector.cpp", "Noise.cpp", "CircuitUnitaryOperationHIP.cpp",
                "CircuitUnitaryOperationFactory.cpp", "Circuit.cpp", "HardwareCheck.cpp", "GateListFile.cpp"]


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def lib_path() -> Path:
    return PKG / "libqse_hip.so"


def build_lib(verbose: bool = False, jobs: int = 8) -> Path:
    BUILD.mkdir(parents=True, exist_ok=True)
    srcs = [s for s in LIB_SOURCES if (CSRC / s).exists()]
    objs = [BUILD / (s.rsplit(".", 1)[0] + ".o") for s in srcs]

    def compile_one(i: int) -> None:
        src = CSRC / srcs[i]
        cmd = [HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(objs[i])]
        if src.suffix == ".cpp":
            cmd.insert(1, "-x")
            cmd.insert(2, "hip")
        _run(cmd, verbose)

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(compile_one, range(len(srcs))))
    out = lib_path()
    tmp = out.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs),
          "-L/opt/rocm/lib", "-lhiprtc", "-Wl,-rpath,/opt/rocm/lib"], verbose)
    os.replace(tmp, out)
    return out


def build_pymodule(verbose: bool = False, jobs: int = 8) -> Path:
    import pybind11

    BUILD.mkdir(parents=True, exist_ok=True)
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    out = PKG / f"_quantumsimevo{suffix}"
    py_inc = sysconfig.get_paths()["include"]
    cxx = os.environ.get("CXX", "g++")
    flags = ["-std=c++17", "-O2", "-fPIC", "-Wall", "-Wno-unused-result", f"-I{INCLUDE}",
             f"-I{pybind11.get_include()}", f"-I{py_inc}"]
    srcs = [HOST / s for s in HOST_SOURCES] + [CSRC / "pymodule.cpp"]
    objs = [BUILD / ("host_" + s.stem + ".o") for s in srcs]

    def compile_one(i: int) -> None:
        _run([cxx, *flags, "-c", str(srcs[i]), "-o", str(objs[i])], verbose)

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(compile_one, range(len(srcs))))
    tmp = out.with_name(out.name + ".tmp")
    _run([cxx, "-shared", "-o", str(tmp), *map(str, objs), f"-L{PKG}", "-lqse_hip",
          "-Wl,-rpath,$ORIGIN"], verbose)
    os.replace(tmp, out)
    return out


def build_oracle(verbose: bool = False) -> None:
    """The CPU checker: oracle/liboracle.so, and where the reference tree is
    present oracle/_ref (its CPU backends, and the drop-in module built from
    its Circuit/Noise/pybind sources against include/quantumsimevo and
    libqse_hip.so, which must be built first)."""
    _run(["make", "-s", "-C", str(ROOT / "oracle"), "all", "dropin"], verbose)


def build_test_emulator(verbose: bool = False) -> Path:
    """tests/_build/libfused_emu.so: CPU emulation of the fused kernel over the
    product planner (qse_plan.cpp); test infrastructure only."""
    out_dir = ROOT / "tests" / "_build"
    out_dir.mkdir(parents=True, exist_ok=True)
    out = out_dir / "libfused_emu.so"
    _run([os.environ.get("CXX", "g++"), "-std=c++17", "-O2", "-fPIC", "-shared", "-D__HIP_PLATFORM_AMD__",
          "-I/opt/rocm/include", f"-I{INCLUDE}", f"-I{CSRC}", "-o", str(out),
          str(ROOT / "tests" / "fused_emu.cpp"), str(CSRC / "qse_plan.cpp"), str(CSRC / "qse_jit_gen.cpp"),
          str(CSRC / "qse_adjoint_plan.cpp"), str(CSRC / "qse_batch_plan.cpp")], verbose)
    return out


def build_all(verbose: bool = False) -> None:
    build_lib(verbose)
    if (CSRC / "pymodule.cpp").exists():
        build_pymodule(verbose)
    build_oracle(verbose)
    build_test_emulator(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
