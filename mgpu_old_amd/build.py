"""Build recipe of the native parts: gfx950 code objects and libmgpusort.so.

Outputs go to mgpu_old_amd/lib/ (git-ignored, shipped with the tree):
  mgpusort_radix.hsaco   digit_count, digit_scan, rank_scatter kernels
  libmgpusort.so         the C ABI of include/mgpusort.h (host C++ over HIP)

python -m mgpu_old_amd.build --ubsan DIR builds DIR/libmgpusort.so with the
host code under UndefinedBehaviorSanitizer (load it with MGPUSORT_LIBRARY).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.environ.get("HIPCC", os.path.join(ROCM, "bin", "hipcc"))
CXX = os.environ.get("CXX", "g++")
ARCH = "gfx950"
# Packed FP32 instructions stay off: none of this code wants them, and some of
# their forms are unreliable on the target boxes.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                # Work counters are taken by one lane whose reply is used much
                # later; the atomic optimizer's wave-wide rewrite would wait for
                # the reply (and every earlier load and store) at once.
                "-mllvm", "-amdgpu-atomic-optimizer-strategy=None"]

CODE_OBJECTS = {
    "mgpusort_radix.hsaco": ["radixsort.hip"],
    "mgpusort_seg.hsaco": ["segsort.hip"],
}
HOST_SOURCES = ["engine.cpp", "loader.cpp", "segmented.cpp", "server.cpp"]
# Host code with device code of its own, linked into libmgpusort.so: the
# fixed scratch array of sortArrayNoAlloc (and its clear kernel).
HIP_HOST_SOURCES = ["noalloc.hip"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=True):
    os.makedirs(LIB, exist_ok=True)
    for obj, srcs in CODE_OBJECTS.items():
        # A plain gfx950 ELF code object (no offload bundle) for hipModuleLoad.
        _run([HIPCC, f"--offload-arch={ARCH}", "-O3", "-std=c++17", "--offload-device-only",
              "--no-gpu-bundle-output", *DEVICE_FLAGS,
              "-I", CSRC, *[os.path.join(CSRC, s) for s in srcs],
              "-o", os.path.join(LIB, obj)])
    objects = []
    for src in HIP_HOST_SOURCES:
        obj = os.path.join(LIB, os.path.splitext(src)[0] + ".o")
        _run([HIPCC, f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-c",
              "-fvisibility=hidden", *DEVICE_FLAGS,
              "-I", CSRC, "-I", INCLUDE, os.path.join(CSRC, src), "-o", obj])
        objects.append(obj)
    # Host code only: plain g++ against the HIP runtime, no device pass.
    _link_host(os.path.join(LIB, "libmgpusort.so"), objects, ["-O2"])


def _link_host(out, objects, flags):
    _run([CXX, *flags, "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wextra",
          "-fvisibility=hidden", "-D__HIP_PLATFORM_AMD__",
          "-I", CSRC, "-I", INCLUDE, "-I", os.path.join(ROCM, "include"),
          *[os.path.join(CSRC, s) for s in HOST_SOURCES], *objects,
          "-o", out,
          "-L", os.path.join(ROCM, "lib"), f"-Wl,-rpath,{os.path.join(ROCM, 'lib')}",
          "-lamdhip64", "-ldl", "-pthread"])


# Host code under sanitizers (the kernels are the same code objects); both
# need the normal build's noalloc.o.
#   ubsan: UndefinedBehaviorSanitizer, aborting at the first report: a shared
#          library that needs no preloaded runtime, so any process that loads
#          it through MGPUSORT_LIBRARY checks the host side of what it runs;
#   asan:  AddressSanitizer + UBSan, for executables linked with
#          -fsanitize=address (which load the ASan runtime first), such as the
#          C++ client's tests/cpp/client_asan.
SANITIZER_FLAGS = {
    "ubsan": ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=undefined",
              "-fno-sanitize-recover=undefined"],
    "asan": ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
             "-fno-sanitize-recover=undefined"],
}


def build_sanitized(out_dir, kind):
    os.makedirs(out_dir, exist_ok=True)
    objects = [os.path.join(LIB, os.path.splitext(src)[0] + ".o") for src in HIP_HOST_SOURCES]
    for o in objects:
        if not os.path.exists(o):
            raise FileNotFoundError(f"{o}: run the normal build first")
    out = os.path.join(out_dir, "libmgpusort.so")
    _link_host(out, objects, SANITIZER_FLAGS[kind])
    return out


def build_ubsan(out_dir):
    return build_sanitized(out_dir, "ubsan")


if __name__ == "__main__":
    if len(sys.argv) == 3 and sys.argv[1] == "--ubsan":
        build_ubsan(sys.argv[2])
    else:
        build()
    sys.exit(0)
