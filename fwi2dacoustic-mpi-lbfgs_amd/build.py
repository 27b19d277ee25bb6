"""Builds libfwi_hip.so in-tree (gfx950) from csrc/.

    python fwi2dacoustic-mpi-lbfgs_amd/build.py

Device code: hipcc --offload-arch=gfx950, -ffp-contract=off (every
expression rounds as written, like the C reference) and packed FP32 ops
disabled.  The C host driver (grad4_1shot.c), the input/resampling layer
(fwi_io.c), the trace conditioning (fwi_sigproc.c), the misfit pool
(fwi_msf.c), the gradient post-processing (fwi_gradprec.c) and the L-BFGS
loop (fwi_lbfgs.c) are compiled with gcc.

also builds libfwi_hip_san.so, the host code under AddressSanitizer and
UndefinedBehaviorSanitizer, for the CPU tests (--no-sanitize skips it).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfwi_hip.so")
LIB_TORCH = os.path.join(HERE, "libfwi_hip_torch.so")
INC = os.path.join(os.path.dirname(HERE), "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FWI_HIP_ARCH", "gfx950")
HIP_FLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-variable", "-Wno-unused-function", f"-I{INC}"]
C_FLAGS = ["-O2", "-fPIC", "-std=gnu11", "-ffp-contract=off", "-pthread", "-Wall", f"-I{INC}"]
# device objects: (source, object, extra flags).  The tile kernels are built
# twice, with 8-row strips and with 4-row strips (launchers suffixed _s4).
DEVICE_OBJS = (
    ("fwi_kernels.hip", "fwi_kernels.hip.o", []),
    ("fwi_kernels.hip", "fwi_kernels_s4.hip.o",
     ["-DFWI_TZ2=16", "-DFWI_REC_ROWS=4", "-DFWI_VARIANT=_s4", "-DFWI_SECONDARY"]),
    ("fwi_noalloc.hip", "fwi_noalloc.hip.o", []),
)


def run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src, name, extra in DEVICE_OBJS + (("fwi_hip.cpp", "fwi_hip.cpp.o", []),):
        obj = os.path.join(BUILD, name)
        run([HIPCC, *HIP_FLAGS, *extra, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in ("grad4_1shot.c", "fwi_io.c", "fwi_sigproc.c", "fwi_msf.c", "fwi_msf_rank.c", "fwi_gradprec.c", "fwi_lbfgs.c"):
        path = os.path.join(CSRC, src)
        if os.path.exists(path):
            obj = os.path.join(BUILD, src + ".o")
            run(["gcc", *C_FLAGS, "-c", path, "-o", obj])
            objs.append(obj)
    tmp = LIB + ".tmp"
    run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"])
    os.replace(tmp, LIB)
    # Second link of the same objects against the HIP runtime that PyTorch
    # bundles, so that one process holds a single HIP runtime when the
    # library works on torch tensors (fwihip.autograd); C callers use LIB.
    tlib = torch_hip_dir()
    if tlib:
        tmp = LIB_TORCH + ".tmp"
        run(["g++", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread", f"-L{tlib}", "-l:libamdhip64.so",
             f"-Wl,-rpath,{tlib}", "-Wl,--no-undefined"])
        os.replace(tmp, LIB_TORCH)
    return LIB


# Sanitized build of the host code (CPU tests only): the host C files and the
# host side of fwi_hip.cpp with AddressSanitizer and UndefinedBehaviorSanitizer
# (every finding fatal), linked with the ordinary device objects into
# libfwi_hip_san.so.  One toolchain (ROCm's clang and its shared sanitizer
# runtime) for every instrumented object, so that one runtime serves the
# process; the device code is not instrumented.
CLANG = os.environ.get("FWI_CLANG", "/opt/rocm/llvm/bin/clang")
LIB_SAN = os.path.join(HERE, "libfwi_hip_san.so")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g"]
HOST_C = ("grad4_1shot.c", "fwi_io.c", "fwi_sigproc.c", "fwi_msf.c", "fwi_msf_rank.c", "fwi_gradprec.c",
          "fwi_lbfgs.c")


def sanitizer_runtime():
    """the shared sanitizer runtime the instrumented objects link against"""
    out = subprocess.run([CLANG, "-print-file-name=libclang_rt.asan-x86_64.so"], check=True,
                         capture_output=True, text=True).stdout.strip()
    return out if os.path.exists(out) else None


def build_sanitized():
    sb = os.path.join(BUILD, "san")
    os.makedirs(sb, exist_ok=True)
    objs = []
    for src, name, extra in DEVICE_OBJS:
        obj = os.path.join(BUILD, name)
        if not os.path.exists(obj):
            run([HIPCC, *HIP_FLAGS, *extra, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    host_san = [f for flag in SAN for f in ("-Xarch_host", flag)]
    obj = os.path.join(sb, "fwi_hip.cpp.o")
    run([HIPCC, *HIP_FLAGS, *host_san, "-c", os.path.join(CSRC, "fwi_hip.cpp"), "-o", obj])
    objs.append(obj)
    for src in HOST_C:
        obj = os.path.join(sb, src + ".o")
        run([CLANG, "-O1", "-fPIC", "-std=gnu11", "-ffp-contract=off", "-pthread", *SAN, f"-I{INC}", "-c",
             os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB_SAN + ".tmp"
    run([CLANG + "++", "-shared", "-fPIC", *SAN, "-shared-libsan", "-o", tmp, *objs, "-L/opt/rocm/lib",
         "-lamdhip64", "-lpthread", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, LIB_SAN)
    return LIB_SAN


def torch_hip_dir():
    """Directory of torch's bundled libamdhip64.so (None without torch)."""
    try:
        import importlib.util
        spec = importlib.util.find_spec("torch")
    except Exception:
        return None
    if not spec or not spec.origin:
        return None
    d = os.path.join(os.path.dirname(spec.origin), "lib")
    return d if os.path.exists(os.path.join(d, "libamdhip64.so")) else None


if __name__ == "__main__":
    build()
    if "--no-sanitize" not in sys.argv[1:]:
        build_sanitized()
    sys.exit(0)
