"""Build libpetl.so (HIP kernels + C ABI) for gfx950 in-tree.

Usage: python -m petl_amd.build [--jobs N] [--sanitize-host]

Every .hip source in csrc/ is compiled to an object with hipcc and linked into petl_amd/libpetl.so, the
library the ctypes bindings load.  Packed-FP32 instructions are disabled for the device code.
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpetl.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PETL_OFFLOAD_ARCH", "gfx950")

CFLAGS = [
    "-x", "hip", f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-fopenmp",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-I", os.path.join(HERE, "..", "include"),
]
LDFLAGS = ["-shared", "-fopenmp", f"--offload-arch={ARCH}", "-L/opt/rocm/lib", "-lhipfft",
           "-Wl,-rpath,/opt/rocm/lib"]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


# host-only AddressSanitizer + UndefinedBehaviorSanitizer variant (the device code is built as usual): every
# -fsanitize flag applies to the host compilation only
SAN_FLAGS = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
             "-Xarch_host", "-fno-omit-frame-pointer", "-Xarch_host", "-g",
             # its kernels are never run (the sanitizer tests drive host code only): no device optimisation
             "-Xarch_device", "-O0"]
SAN_LIB = os.path.join(HERE, "_san", "libpetl_san.so")


def _compile(src, build_dir=BUILD, defines=(), extra=()):
    obj = os.path.join(build_dir, os.path.basename(src)[:-4] + ".o")
    cmd = [HIPCC] + CFLAGS + list(extra) + [f"-D{d}" for d in defines] + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs=None, verbose=True, lib=LIB, build_dir=BUILD, defines=(), sanitize_host=False):
    """Compiles every source and links `lib`.  `defines` (NAME=VALUE strings) select compile-time kernel
    variants for tuning builds; the default build uses none.  sanitize_host: the host code built with
    AddressSanitizer and UndefinedBehaviorSanitizer (load it with the sanitizer runtime preloaded)."""
    os.makedirs(build_dir, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(8, os.cpu_count() or 4, len(srcs))
    extra = SAN_FLAGS if sanitize_host else ()
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda src: _compile(src, build_dir, defines, extra), srcs))
    tmp = lib + ".tmp"
    ld = LDFLAGS + (["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined"] if sanitize_host else [])
    r = subprocess.run([HIPCC] + ld + objs + ["-o", tmp], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, lib)
    if verbose:
        print(f"built {lib}")
    return lib


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--sanitize-host", action="store_true",
                    help="build petl_amd/_san/libpetl_san.so with host-only ASan + UBSan instead")
    a = ap.parse_args()
    try:
        if a.sanitize_host:
            build(a.jobs, lib=SAN_LIB, build_dir=os.path.join(HERE, "_san", "obj"), sanitize_host=True)
        else:
            build(a.jobs)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
