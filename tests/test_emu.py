"""The library's HIP sources executed on the CPU (tests/emu), under
AddressSanitizer.

tests/emu/build.py compiles every palial_amd/csrc/*.hip file unchanged as
host C++ against an emulator of the HIP device model: the lanes of a
workgroup are fibers that meet at every barrier and every cross-lane
operation (shuffles, readlane, DPP, MFMA with the CDNA operand layouts), and
every hipMalloc block and every launch's dynamic LDS block has exactly its
requested size.  The library's own host code (schedules, solve, refinement)
drives the emulated kernels through the same C ABI, so the GPU test files
run unchanged against it (PALIAL_HIPEMU=asan in their environment).

What passes here: results that meet the GPU tests' own bars, and no load or
store of any kernel outside the buffers, for the paths run -- among them the
two that have not yet run on a GPU, paired trailing updates and blocked
correction sweeps.  What it does not show: timing, effects of asynchronous
loads or of concurrency between streams (streams run in issue order here),
and code that relies on a lane reading what a lower lane of its wave wrote
earlier with no cross-lane operation in between (the emulator runs the lanes
of a wave in descending order between synchronization points).

Default: a compact subset plus the correction-sweep bounds check at N=1024
(every admissible tile size, step width, right-hand-side count, factor
precision and algorithm).  PALIAL_EMU_FULL=1 adds the whole GPU test files of
the paths not yet run on a GPU and the bounds check at N=2048 (with the block
inverses) and N=32768.
"""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
FULL = os.environ.get("PALIAL_EMU_FULL") == "1"
ASAN_OPTIONS = "detect_leaks=0:detect_stack_use_after_return=0"


@pytest.fixture(scope="module")
def emu_lib():
    from emu import build as emu_build
    return emu_build.build("asan"), emu_build.asan_runtime()


def _env(rt):
    env = dict(os.environ, PALIAL_HIPEMU="asan", ASAN_OPTIONS=ASAN_OPTIONS)
    env["LD_PRELOAD"] = rt + (" " + os.environ["LD_PRELOAD"] if os.environ.get("LD_PRELOAD") else "")
    return env


def _start(emu_lib, targets, kexpr=None):
    _, rt = emu_lib
    cmd = [sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "gpu", *targets]
    if kexpr:
        cmd += ["-k", kexpr]
    return subprocess.Popen(cmd, cwd=ROOT, env=_env(rt), stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)


# group: (GPU test node ids of the default subset, (targets, -k) of the full run)
GROUPS = {
    "kernels": (["tests/test_gpu_parity.py::test_tile_tasks[32]", "tests/test_gpu_parity.py::test_dpotrf_tile[128]",
                 "tests/test_gpu_parity.py::test_dgetrf_tile[128]", "tests/test_gpu_parity.py::test_qr_tiles[64]",
                 "tests/test_gpu_parity.py::test_cholesky[512-128]", "tests/test_gpu_parity.py::test_lu[512-128]",
                 "tests/test_gpu_parity.py::test_qr[256-64]"],
                (["tests/test_gpu_parity.py"], "not 2048")),
    "solves": (["tests/test_gpu_solve.py::test_solve_diag_dominant[512-128-1-2-cholesky]",
                "tests/test_gpu_solve.py::test_solve_diag_dominant[512-128-1-4-lu]"],
               (["tests/test_gpu_solve.py::test_solve_diag_dominant"], "512")),
    "paired_updates": (["tests/wip_gpu_pair.py::test_paired_updates_match[1280-128-0-cholesky-2]"],
                       (["tests/wip_gpu_pair.py"], None)),
    "blocked_sweeps": (["tests/wip_gpu_sweep_block.py::test_sweep_blocks_agree[1-2-cholesky]",
                        "tests/wip_gpu_sweep_block.py::test_sweep_block_refused"],
                       (["tests/wip_gpu_sweep_block.py"], "not tuner")),
}


def test_gpu_tests_on_emulator(emu_lib):
    """Every group's GPU tests pass on the AddressSanitizer build (the groups
    run side by side, one process each)."""
    procs = {name: _start(emu_lib, targets, kexpr) if FULL else _start(emu_lib, subset)
             for name, (subset, (targets, kexpr)) in GROUPS.items()}
    bad = {}
    for name, p in procs.items():
        out, err = p.communicate(timeout=4 * 3600 if FULL else 1200)
        if p.returncode != 0 or " passed" not in out or "failed" in out:
            bad[name] = (out[-3000:], err[-3000:])
    assert not bad, bad


def sweep_configs(N, blkinv):
    """Every (algo, N, b, q, nrhs, f32) the solve accepts for correction sweeps
    at this N with b in {128, 256, 512}: q | N/b with N/b >= 2q, at most 4
    right-hand sides and q b (x4 above one right-hand side) <= 4096 for q > 1;
    nrhs 16 takes the MFMA sweep kernels (q = 1)."""
    out = []
    for b in (128, 256, 512):
        MT = N // b
        for q in (1, 2, 4):
            if MT % q or (q > 1 and MT < 2 * q):
                continue
            for nrhs in (1, 3, 16):
                if q > 1 and (nrhs > 4 or q * b * (1 if nrhs == 1 else 4) > 4096):
                    continue
                for f32 in (0, 1):
                    for algo in (0, 1):
                        out.append((algo, N, b, q, nrhs, f32, int(blkinv and q > 1)))
    return out


_RUNNER = r"""
import ctypes as C, sys
lib = C.CDLL(sys.argv[1])
f = lib.hipemu_sweep_check
f.argtypes = [C.c_int] * 7
for a in sys.argv[2:]:
    args = tuple(int(v) for v in a.split(","))
    rc = f(*args)
    print(a, rc, flush=True)
    if rc != 0:
        sys.exit(3)
"""


def _sweep_check(emu_lib, configs, workers=1, timeout=900):
    lib, rt = emu_lib
    chunks = [configs[i::workers] for i in range(workers)]
    procs = [subprocess.Popen([sys.executable, "-c", _RUNNER, lib, *(",".join(map(str, c)) for c in ch)],
                              cwd=ROOT, env=_env(rt), stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
             for ch in chunks if ch]
    done = 0
    for p in procs:
        out, err = p.communicate(timeout=timeout)
        assert p.returncode == 0, (out[-2000:], err[-6000:])
        assert "AddressSanitizer" not in err, err[-6000:]
        done += sum(1 for line in out.splitlines() if line.endswith(" 0"))
    assert done == len(configs)


def test_sweep_bounds_n1024(emu_lib):
    """Every load and store of the correction-sweep kernels (one-tile and
    blocked steps, FP64 and FP32 factors, 1, 3 and 16 right-hand sides) and of
    the block-inverse construction stays inside its buffer at N=1024."""
    cfg = sweep_configs(1024, blkinv=True)
    assert len(cfg) == 60
    _sweep_check(emu_lib, cfg, workers=2)


@pytest.mark.skipif(not FULL, reason="PALIAL_EMU_FULL=1: about 15 minutes on 6 cores")
def test_sweep_bounds_large(emu_lib):
    """The same at N=2048 with the block inverses, and at N=32768 (sweeps
    only: the block-inverse GEMMs are too slow to emulate at that size; their
    tile maps are checked on the host at every size by the plan tests)."""
    _sweep_check(emu_lib, sweep_configs(2048, blkinv=True), workers=6)
    _sweep_check(emu_lib, sweep_configs(32768, blkinv=False), workers=6, timeout=4 * 3600)


def test_emulator_reports_out_of_bounds(emu_lib):
    """The check has teeth: a sweep whose factor buffer is b*b + 1 elements
    short (the harness's planted mistake: the last element of U(MT-2, MT-1),
    which the sweeps read, falls outside it) is reported by AddressSanitizer."""
    lib, rt = emu_lib
    code = ("import ctypes as C, sys\nlib = C.CDLL(sys.argv[1])\n"
            "lib.hipemu_sweep_check_short.argtypes = [C.c_int] * 4\n"
            "print(lib.hipemu_sweep_check_short(0, 1024, 128, 1))\n")
    r = subprocess.run([sys.executable, "-c", code, lib], cwd=ROOT, env=_env(rt), capture_output=True, text=True,
                       timeout=300)
    assert r.returncode != 0 and "AddressSanitizer" in r.stderr and "heap-buffer-overflow" in r.stderr, r.stderr[-3000:]
