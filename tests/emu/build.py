"""Build the CPU-executed copy of libpalial_hip.so (tests only).

Every palial_amd/csrc/*.hip file is compiled unchanged as host C++ by the
ROCm clang against the emulator header (tests/emu/hipemu.hpp, reached through
the stand-in <hip/hip_runtime.h> in tests/emu/include).  The one source edit
is mechanical and made on the fly: a dynamic LDS declaration
`extern __shared__ T name[];` becomes `T *name = (T *)emu::g_dyn_lds;`, the
launch's LDS block.  api.hip's translation unit also gets the test entry
points of tests/emu/api_hooks.inc, which call the library's internal
routines.  Output: tests/emu/_build/{plain,asan,race}/libpalial_emu.so,
rebuilt only when a source or a flag changes.  The race variant compiles the
library's sources with the thread-sanitizer instrumentation and the emulator
runtime without it: the runtime's own callbacks check every device-memory
access against the order of the stream operations (hipemu.cpp).

    python tests/emu/build.py [--asan | --race]
"""
import concurrent.futures as cf
import hashlib
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "palial_amd", "csrc")
OUT = os.path.join(HERE, "_build")
CXX = os.environ.get("HIPEMU_CXX", "/opt/rocm/llvm/bin/clang++")

DYN_LDS = re.compile(r"extern\s+__shared__\s+(?:__attribute__\(\([^;]*?\)\)\s*)?([A-Za-z_][\w:]*)\s+(\w+)\s*\[\s*\]\s*;")

FLAGS = ["-std=c++17", "-fPIC", "-g0", "-I" + os.path.join(HERE, "include"), "-I" + CSRC,
         "-I" + os.path.join(ROOT, "include"), "-Wno-unused-result", "-Wno-deprecated-declarations",
         "-ffp-contract=off"]
# test entry points compiled into the translation unit of a source file
HOOKS = {"api.hip": "api_hooks.inc"}
VARIANTS = {
    "plain": ["-O2"],
    "asan": ["-O1", "-fsanitize=address", "-fno-omit-frame-pointer"],
    # stream race checking: the library's sources carry the thread-sanitizer
    # instrumentation, whose callbacks the emulator runtime defines itself
    # (no sanitizer runtime is linked)
    "race": ["-O1", "-fsanitize=thread", "-DHIPEMU_RACE"],
}
# flags of the emulator runtime itself, where they differ from the sources'
RUNTIME = {"race": ["-O1", "-DHIPEMU_RACE"]}


def transform(text):
    return DYN_LDS.sub(lambda m: f"{m.group(1)} *{m.group(2)} = ({m.group(1)} *)emu::g_dyn_lds;", text)


def _sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))


def _digest(variant):
    h = hashlib.sha256()
    h.update(" ".join([CXX, *FLAGS, *VARIANTS[variant], "|", *RUNTIME.get(variant, [])]).encode())
    names = [os.path.join(CSRC, f) for f in _sources()]
    names += [os.path.join(CSRC, f) for f in sorted(os.listdir(CSRC)) if f.endswith((".hpp", ".h"))]
    names += [os.path.join(HERE, h) for h in sorted(HOOKS.values())]
    names += [os.path.join(HERE, "hipemu.hpp"), os.path.join(HERE, "hipemu.cpp"),
              os.path.join(HERE, "include", "hip", "hip_runtime.h"), os.path.join(ROOT, "include", "palial_hip.h"),
              os.path.abspath(__file__)]
    for n in names:
        h.update(n.encode())
        with open(n, "rb") as f:
            h.update(f.read())
    return h.hexdigest()


def _compile(args):
    src, obj, flags = args
    r = subprocess.run([CXX, *flags, "-c", src, "-o", obj], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipemu build failed on {src}:\n{r.stdout}{r.stderr}")
    return obj


def build(variant="plain", verbose=False):
    """Return the path of libpalial_emu.so for `variant` ("plain", "asan" or "race")."""
    d = os.path.join(OUT, variant)
    lib = os.path.join(d, "libpalial_emu.so")
    stamp = os.path.join(d, "stamp")
    dig = _digest(variant)
    if os.path.exists(lib) and os.path.exists(stamp) and open(stamp).read() == dig:
        return lib
    os.makedirs(os.path.join(d, "src"), exist_ok=True)
    flags = [*FLAGS, *VARIANTS[variant]]
    jobs = []
    for f in _sources():
        with open(os.path.join(CSRC, f)) as fh:
            text = fh.read()
        gen = os.path.join(d, "src", f[:-4] + ".cpp")
        with open(gen, "w") as fh:
            fh.write(f'#line 1 "{os.path.join(CSRC, f)}"\n' + transform(text))
            if f in HOOKS:
                fh.write(f'\n#include "{os.path.join(HERE, HOOKS[f])}"\n')
        jobs.append((gen, os.path.join(d, f[:-4] + ".o"), flags))
    jobs.append((os.path.join(HERE, "hipemu.cpp"), os.path.join(d, "hipemu.o"),
                 [*FLAGS, *RUNTIME[variant]] if variant in RUNTIME else flags))
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(_compile, jobs))
    tmp = lib + ".tmp"
    link = [CXX, "-shared", "-fPIC", "-o", tmp, *objs, "-Wl,-Bsymbolic", "-ldl"]
    if variant == "asan":
        link[1:1] = ["-fsanitize=address", "-shared-libasan"]
    r = subprocess.run(link, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipemu link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, lib)
    with open(stamp, "w") as fh:
        fh.write(dig)
    if verbose:
        print("built", lib)
    return lib


def asan_runtime():
    """Path of the ASan runtime that an asan-variant process must preload."""
    r = subprocess.run([CXX, "-print-file-name=libclang_rt.asan-x86_64.so"], capture_output=True, text=True)
    return r.stdout.strip()


if __name__ == "__main__":
    build("asan" if "--asan" in sys.argv else "race" if "--race" in sys.argv else "plain", verbose=True)
