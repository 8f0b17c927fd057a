"""Build the native parts of the package in-tree.

* libcupf_hip.so  -- HIP kernels for gfx950 + C++ host driver + C ABI
                     (hipcc, one object per source, compiled in parallel)
* _cupf*.so       -- pybind11 module over the C ABI (g++, no HIP/torch headers)
* cupf_bench      -- dump-case command-line benchmark over the C ABI (g++)

Both land in gpu-powerflow_amd/cupf/.  Run:  python gpu-powerflow_amd/build.py
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
PKG = HERE / "cupf"
ROOT = HERE.parent
BUILD = HERE / "_build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
]

# source -> extra flags
SOURCES = {
    "nr_kernels.hip": ["-ffp-contract=off"],
    "lu_kernels.hip": [],
    "newton_solver.cpp": [],
    "batched_lu.cpp": [],
    "analyze.cpp": [],
    "capi.cpp": [],
    "mlp_kernels.hip": [],
    "warm_start.cpp": [],
    "static_workspace.hip": [],
    "case_io.cpp": [],
}


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout)
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd[:4])} ...")
    # keep warnings visible but short
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l
                    and "not a recognized feature" not in l)
    if out.strip():
        sys.stderr.write(out + "\n")


def _compile(src: str, flags: list[str]) -> Path:
    obj = BUILD / (src.replace(".", "_") + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    _run([HIPCC, *COMMON, *flags, *lang, "-c", str(CSRC / src), "-o", str(obj)])
    return obj


def build_hip_library() -> Path:
    BUILD.mkdir(exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(lambda kv: _compile(*kv), SOURCES.items()))
    out = PKG / "libcupf_hip.so"
    tmp = PKG / "libcupf_hip.so.tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-Wl,--no-undefined", "-o", str(tmp),
          *map(str, objs)])
    os.replace(tmp, out)
    return out


def build_pybind_module() -> Path:
    import pybind11
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    out = PKG / f"_cupf{suffix}"
    tmp = PKG / f"_cupf{suffix}.tmp"
    cxx = shutil.which("g++") or "g++"
    _run([cxx, "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall",
          f"-I{pybind11.get_include()}", f"-I{sysconfig.get_paths()['include']}",
          f"-I{ROOT / 'include'}", f"-I{CSRC}",
          str(CSRC / "pybind_cupf.cpp"), "-o", str(tmp),
          f"-L{PKG}", "-lcupf_hip", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, out)
    return out


def build_cli() -> Path:
    """cupf_bench: the dump-case command-line benchmark, over the C ABI."""
    out = PKG / "cupf_bench"
    tmp = PKG / "cupf_bench.tmp"
    cxx = shutil.which("g++") or "g++"
    rocm_lib = Path(HIPCC).resolve().parent.parent / "lib"
    _run([cxx, "-O2", "-std=c++17", "-Wall", f"-I{ROOT / 'include'}",
          str(CSRC / "cupf_bench.cpp"), "-o", str(tmp),
          f"-L{PKG}", "-lcupf_hip", "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath-link,{rocm_lib}"])
    os.replace(tmp, out)
    return out


def build() -> None:
    build_hip_library()
    build_pybind_module()
    build_cli()


if __name__ == "__main__":
    build()
    print("built", PKG / "libcupf_hip.so")
