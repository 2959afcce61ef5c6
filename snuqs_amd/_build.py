"""Build recipe for the engine (run by ``__graft_entry__.build()``).

Produces, in-tree:
  snuqs_amd/libsnuqs_hip.so   HIP kernels + C ABI (include/snuqs_hip.h), gfx950
  snuqs_amd/_C<ext-suffix>    pybind11 module over the C ABI (host C++ only)

Every HIP translation unit is compiled with the packed-FP32 target feature
switched off (`-Xclang -target-feature -Xclang -packed-fp32-ops`); the host
pass warns that it does not know the feature, which is expected.
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "snuqs_amd")
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(ROOT, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_SOURCES = [
    "gate_plan.cpp",
    "kernels_gate.hip",
    "kernels_reduce.hip",
    "kernels_sample.hip",
    "kernels_qaoa.hip",
    "kernels_layer.hip",
    "kernels_cbatch.hip",
    "capi.hip",
    "host_tier.cpp",
]
HOST_SOURCES = ["host/pymodule.cc", "host/gates.cc"]

HIP_FLAGS = [
    "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]


def _run(cmd, quiet_warn=("packed-fp32-ops",)):
    p = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in p.stderr.splitlines()
                    if not any(q in l for q in quiet_warn))
    if p.returncode != 0:
        raise RuntimeError(f"build step failed: {' '.join(cmd)}\n{err}\n{p.stdout}")
    return err


def _ext_suffix():
    return sysconfig.get_config_var("EXT_SUFFIX") or ".so"


def build_hip(jobs: int = 8) -> str:
    os.makedirs(BUILD, exist_ok=True)

    def one(src):
        obj = os.path.join(BUILD, os.path.splitext(os.path.basename(src))[0] + ".o")
        _run([HIPCC, *HIP_FLAGS, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj])
        return obj

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(one, HIP_SOURCES))
    out = os.path.join(PKG, "libsnuqs_hip.so")
    tmp = out + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_pymodule() -> str:
    import pybind11

    out = os.path.join(PKG, "_C" + _ext_suffix())
    tmp = out + ".tmp"
    py_inc = sysconfig.get_paths()["include"]
    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-fvisibility=hidden",
           "-I" + pybind11.get_include(), "-I" + py_inc,
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(CSRC, "host"),
           *[os.path.join(CSRC, s) for s in HOST_SOURCES],
           "-L" + PKG, "-lsnuqs_hip", "-Wl,-rpath,$ORIGIN", "-o", tmp]
    _run(cmd)
    os.replace(tmp, out)
    return out


def build_oracle() -> None:
    """Test infrastructure: the CPU oracle and, where the reference tree is
    present, the reference's own CPU gate path (oracle/_ref)."""
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


def build_all() -> None:
    build_hip()
    build_pymodule()
    build_oracle()


if __name__ == "__main__":
    build_all()
    print("built", file=sys.stderr)
