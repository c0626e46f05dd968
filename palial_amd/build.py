"""Build libpalial_hip.so (gfx950) and the palial driver, in-tree.

    python -m palial_amd.build          # or palial_amd.build.build()

hipcc cross-compiles for gfx950 without a GPU present.  Every HIP source is
compiled with packed FP32 instructions disabled (-packed-fp32-ops off).
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpalial_hip.so")
DRIVER_SRC = os.path.join(HERE, "driver", "palial.c")
DRIVER = os.path.join(HERE, "bin", "palial")

ARCH = os.environ.get("PALIAL_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-I" + os.path.join(ROOT, "include"),
]


def _sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


# per-file flags: the generators must round like plain C (no FMA contraction)
EXTRA = {"k_misc.hip": ["-ffp-contract=off"]}


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC, *HIP_FLAGS, *EXTRA.get(os.path.basename(src), []), "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{out}")
    return obj


def build_lib(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = _sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_driver(verbose=False):
    if not os.path.exists(DRIVER_SRC):
        return None
    os.makedirs(os.path.dirname(DRIVER), exist_ok=True)
    cc = shutil.which("gcc") or "cc"
    cmd = [cc, "-O2", "-std=gnu11", "-Wall", "-I" + os.path.join(ROOT, "include"), DRIVER_SRC,
           "-o", DRIVER, "-L" + HERE, "-lpalial_hip", "-Wl,-rpath,$ORIGIN/.."]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("driver build failed:\n" + r.stdout + r.stderr)
    if verbose:
        print("built", DRIVER)
    return DRIVER


def build(verbose=False):
    build_lib(verbose)
    build_driver(verbose)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
