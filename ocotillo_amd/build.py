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
# packed FP32 instructions are disabled for all device code (the device pass
# takes the feature; the host pass warns that it does not know it), and
# check_no_packed_fp32 verifies that the built code object holds none.
COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wno-unused-function"]
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]
PACKED_FP32 = ("v_pk_fma_f32", "v_pk_mul_f32", "v_pk_add_f32")


def check_no_packed_fp32(obj):
    """Disassemble the gfx950 code object this build put into `obj` (its
    .hip_fatbin section's offload bundle) and fail if it holds any packed FP32
    arithmetic instruction."""
    llvm = os.path.join(ROCM, "llvm", "bin")
    fat, co = obj + ".fatbin", obj + ".co"
    _run(["objcopy", "-O", "binary", "--only-section=.hip_fatbin", obj, fat])
    _run([os.path.join(llvm, "clang-offload-bundler"), "--unbundle", "--type=o",
          "--targets=hipv4-amdgcn-amd-amdhsa--" + ARCH, "--input=" + fat, "--output=" + co])
    dis = subprocess.run([os.path.join(llvm, "llvm-objdump"), "-d", co], check=True, capture_output=True,
                         text=True).stdout
    os.remove(fat)
    os.remove(co)
    found = {p: dis.count(p) for p in PACKED_FP32 if p in dis}
    if found:
        raise RuntimeError("packed FP32 instructions in %s: %s" % (obj, found))
    if "v_fma_f64" not in dis:
        raise RuntimeError("no device code found in %s" % obj)


def _run(cmd, cwd=None):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd, cwd=cwd)


# Host sanitizers for the C-ABI layer (the library's host code: the C++
# files and the host side of the HIP file; device code is never
# instrumented).  On a hipcc line each flag goes after -Xarch_host.
SANITIZE = ("address", "undefined")
SAN_LIB = os.path.join(PKG, "libocotillo_san.so")
SAN_BUILD = os.path.join(PKG, "_build_san")


def _san_host_flags():
    return ["-fsanitize=" + s for s in SANITIZE] + ["-fno-omit-frame-pointer", "-fno-sanitize-recover=all"]


def _san_hip_flags():
    out = []
    for f in _san_host_flags():
        out += ["-Xarch_host", f]
    return out


def build_lib(out=LIB, defines=(), build_dir=BUILD, sanitize=False):
    """Compile and link the library.  `defines` are extra -D tuning macros
    (e.g. OCO_SWEEP_WAVES_PER_EU=4) for the kernels; `sanitize` builds the
    host code with AddressSanitizer and UndefinedBehaviorSanitizer (a program
    that loads such a library must itself be built with them)."""
    os.makedirs(build_dir, exist_ok=True)
    objs = []
    for src, device in (("oco_kernels.hip", True), ("oco_api.cpp", False), ("oco_serve.cpp", False),
                        ("oco_output.cpp", False)):
        obj = os.path.join(build_dir, os.path.splitext(src)[0] + ".o")
        if device:
            cmd = ([HIPCC] + COMMON + DEVICE + (_san_hip_flags() if sanitize else []) + ["-D" + d for d in defines] +
                   ["-c", os.path.join(CSRC, src), "-o", obj])
        else:   # host-only C++ (the C ABI): no device pass
            cmd = ([CXX] + COMMON + (_san_host_flags() if sanitize else []) +
                   ["-pthread", "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROCM, "include"),
                    "-c", os.path.join(CSRC, src), "-o", obj])
        _run(cmd)
        if device:
            check_no_packed_fp32(obj)
        objs.append(obj)
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-pthread"] + (_san_hip_flags() if sanitize else []) +
         ["-o", tmp] + objs)
    os.replace(tmp, out)
    return out


SAN_DRIVER_SRC = os.path.join(os.path.dirname(PKG), "tests", "sanitize", "abi_check.c")
SAN_DRIVER = os.path.join(os.path.dirname(PKG), "tests", "sanitize", "abi_check")


def build_sanitized():
    """The sanitized library (libocotillo_san.so) and the C program that
    drives its entry points under the sanitizers (tests/sanitize/abi_check,
    test infrastructure)."""
    build_lib(out=SAN_LIB, build_dir=SAN_BUILD, sanitize=True)
    if os.path.exists(SAN_DRIVER_SRC):
        _run([os.path.join(ROCM, "llvm", "bin", "clang"), "-O1", "-g", "-std=gnu11", "-Wall"] + _san_host_flags() +
             ["-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROCM, "include"), SAN_DRIVER_SRC, "-o", SAN_DRIVER,
              "-L" + PKG, "-locotillo_san", "-Wl,-rpath," + PKG, "-Wl,-rpath,$ORIGIN/../../ocotillo_amd",
              "-L" + os.path.join(ROCM, "lib"), "-lamdhip64", "-Wl,-rpath," + os.path.join(ROCM, "lib"), "-lm"])
    return SAN_LIB


def build_fortran():
    fc = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang" if os.path.exists("/opt/rocm/bin/amdflang") else None)
    fdir = os.path.join(PKG, "fortran")
    if fc is None or not os.path.exists(os.path.join(fdir, "Makefile")):
        return None
    _run(["make", "-s", "-C", fdir, "FC=" + fc])
    return os.path.join(fdir, "flux_feautrier_gpu.x")


def build():
    build_lib()
    build_sanitized()
    build_fortran()


if __name__ == "__main__":
    sys.exit(build())
