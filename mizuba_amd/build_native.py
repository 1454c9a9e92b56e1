"""Build the native parts of mizuba_amd in-tree.

- libmizuba_amd.so: the gfx950 kernels and the C ABI (include/mizuba_amd.h),
  compiled with hipcc for --offload-arch=gfx950;
- _core.<ext>.so: the pybind11 front end (host C++ only), linked against
  libmizuba_amd.so with an $ORIGIN rpath.

hipcc cross-compiles gfx950 code without a GPU, so this runs anywhere the ROCm
toolchain is installed. Objects go to mizuba_amd/_build/.
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import pathlib
import shutil
import subprocess
import sys
import sysconfig

PKG = pathlib.Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libmizuba_amd.so"

HIP_SOURCES = ["capi.hip", "batch.hip", "screen_kernels.hip", "narrow_kernels.hip", "eval_kernels.hip", "sgp4_kernels.hip"]
HOST_SOURCES = ["host/module.cpp", "host/synth.cpp", "host/pj_store.cpp"]

ARCH = os.environ.get("MIZUBA_AMD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # every FP multiply and add separately rounded (bit parity with the CPU oracle)
    "-ffp-contract=off",
    # no packed FP32 instructions
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and pathlib.Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the ROCm toolchain is required to build mizuba_amd")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def _narrow_units():
    """Split the order-templated narrow-phase kernels over several objects so
    that they compile in parallel."""
    groups = [(2, 5), (6, 7), (8, 8), (9, 10), (11, 13), (14, 16), (17, 20)]
    return groups


def _stale(obj: pathlib.Path, deps) -> bool:
    if not obj.exists():
        return True
    t = obj.stat().st_mtime
    return any(d.stat().st_mtime > t for d in deps)


def build(verbose: bool = False, jobs: int | None = None, force: bool = True) -> pathlib.Path:
    """Compile everything (force=True) or only what is older than its sources."""
    hipcc = _hipcc()
    headers = list(CSRC.glob("*.hpp")) + [ROOT / "include" / "mizuba_amd.h"]
    BUILD.mkdir(exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 1)
    cmds = []
    objs = []
    for src in HIP_SOURCES:
        if src == "narrow_kernels.hip":
            for lo, hi in _narrow_units():
                obj = BUILD / f"narrow_{lo}_{hi}.o"
                extra = [f"-DMZH_NARROW_LO={lo}", f"-DMZH_NARROW_HI={hi}"]
                if lo == 2:
                    extra.append("-DMZH_NARROW_HOST")
                if force or _stale(obj, [CSRC / src, *headers]):
                    cmds.append([hipcc, *HIPCC_FLAGS, *extra, "-c", str(CSRC / src), "-o", str(obj)])
                objs.append(obj)
        else:
            obj = BUILD / (pathlib.Path(src).stem + ".o")
            if force or _stale(obj, [CSRC / src, *headers]):
                cmds.append([hipcc, *HIPCC_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
            objs.append(obj)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for r in ex.map(_run, cmds):
            if verbose and r.stderr.strip():
                print(r.stderr, file=sys.stderr)
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", str(LIB), *map(str, objs)])

    # pybind11 front end (host compiler only)
    import pybind11

    ext = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    core = PKG / f"_core{ext}"
    py_inc = sysconfig.get_paths()["include"]
    host_cmd = [
        os.environ.get("CXX", "g++"),
        "-O2",
        "-std=c++17",
        "-fPIC",
        "-shared",
        "-fopenmp",
        "-fvisibility=hidden",
        f"-I{pybind11.get_include()}",
        f"-I{py_inc}",
        *[str(CSRC / s) for s in HOST_SOURCES],
        f"-L{PKG}",
        "-lmizuba_amd",
        "-Wl,-rpath,$ORIGIN",
        "-o",
        str(core),
    ]
    _run(host_cmd)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv, force="--incremental" not in sys.argv)
    print("built", LIB)
