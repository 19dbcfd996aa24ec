"""Builds ocotillo_amd/libocotillo.so (HIP, gfx950) and, when amdflang is
present, the Fortran drop-in host ocotillo_amd/fortran/flux_feautrier_gpu.x.

Usage: python -m ocotillo_amd.build  (or ocotillo_amd.build.build()).
Always compiles (no staleness check), so a fresh call reflects the sources.
"""
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libocotillo.so")
ARCH = os.environ.get("OCO_OFFLOAD_ARCH", "gfx950")

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
CXX = os.path.join(ROCM, "llvm", "bin", "clang++")
# -ffp-contract=off: every FP64 operation rounds exactly as the reference's;
# packed FP32 instructions are disabled for all device code.
COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wno-unused-function"]
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def _run(cmd, cwd=None):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd, cwd=cwd)


def build_lib(out=LIB, defines=(), build_dir=BUILD):
    """Compile and link the library.  `defines` are extra -D tuning macros
    (e.g. OCO_SWEEP_WAVES_PER_EU=4) for the kernels."""
    os.makedirs(build_dir, exist_ok=True)
    objs = []
    for src, device in (("oco_kernels.hip", True), ("oco_api.cpp", False), ("oco_serve.cpp", False),
                        ("oco_output.cpp", False)):
        obj = os.path.join(build_dir, os.path.splitext(src)[0] + ".o")
        if device:
            cmd = [HIPCC] + COMMON + DEVICE + ["-D" + d for d in defines] + ["-c", os.path.join(CSRC, src), "-o", obj]
        else:   # host-only C++ (the C ABI): no device pass
            cmd = [CXX] + COMMON + ["-pthread", "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROCM, "include"),
                                    "-c", os.path.join(CSRC, src), "-o", obj]
        _run(cmd)
        objs.append(obj)
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-pthread", "-o", tmp] + objs)
    os.replace(tmp, out)
    return out


def build_fortran():
    fc = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang" if os.path.exists("/opt/rocm/bin/amdflang") else None)
    fdir = os.path.join(PKG, "fortran")
    if fc is None or not os.path.exists(os.path.join(fdir, "Makefile")):
        return None
    _run(["make", "-s", "-C", fdir, "FC=" + fc])
    return os.path.join(fdir, "flux_feautrier_gpu.x")


def build():
    build_lib()
    build_fortran()


if __name__ == "__main__":
    sys.exit(build())
