"""Stream and workgroup race check of the library's schedules on the CPU.

The race variant of the CPU build (tests/emu, PALIAL_HIPEMU=race) compiles
every palial_amd/csrc/*.hip file with the compiler's thread-sanitizer
instrumentation and gives the instrumentation callbacks to the emulator
runtime, so that each load and store the real kernels make is checked against
the happens-before order of the stream operations: stream order, event
record -> wait, host synchronisation and the null stream.  A read or write of
device memory that conflicts with an unordered access of another operation,
or with another workgroup's access in the same launch, fails the test
(tests/conftest.py); so does an LDS access that conflicts with another wave's
access in the same barrier interval, a read of LDS the workgroup has not
written, a value passed between lanes of one wave with no wave operation
between whose outcome depends on the emulator's lane order (the device runs
the wave in lockstep), a read of LDS that a global_load_lds still in flight
is filling (s_waitcnt vmcnt not yet reached), and a vector access whose
address is not a multiple of its size.  This covers what the issue-order emulation cannot: a
missing wait, a workspace shared by two streams, a tile read by one stream
while another rewrites it -- for every schedule the GPU test files exercise,
including paired trailing updates and blocked correction sweeps.

Default: a subset of GPU tests (one per kind of schedule), run in parallel
processes, plus planted mistakes that the check must report.
PALIAL_EMU_FULL=1 runs whole GPU test files instead.
"""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
FULL = os.environ.get("PALIAL_EMU_FULL") == "1"

# group: (GPU test node ids of the default subset, (targets, -k) of the full run)
GROUPS = {
    "factorizations": (["tests/test_gpu_parity.py::test_graph_and_streams_bitwise[cholesky]"],
                       (["tests/test_gpu_parity.py"], "not 2048 and not test_qr")),
    "qr": (["tests/test_gpu_parity.py::test_qr[512-128]", "tests/test_gpu_parity.py::test_qr_tiles[64]"],
           (["tests/test_gpu_parity.py", "tests/test_gpu_qr_apply.py"],
            "(test_qr or qr_apply or form_q or qr_solve) and not 2048")),
    "solves": (["tests/test_gpu_solve.py::test_solve_diag_dominant[512-128-1-2-cholesky]",
                "tests/test_gpu_solve.py::test_solve_diag_dominant[512-128-1-4-lu]",
                "tests/test_gpu_solve.py::test_solve_diag_dominant[512-128-1-3-lu]"],
               (["tests/test_gpu_solve.py"], "not 1024")),
    "paired_updates": (["tests/wip_gpu_pair.py::test_paired_updates_match[1280-128-0-cholesky-2]"],
                       (["tests/wip_gpu_pair.py"], "1280 or rejects")),
    "blocked_sweeps": (["tests/wip_gpu_sweep_block.py::test_sweep_blocks_agree[1-2-cholesky]"],
                       (["tests/wip_gpu_sweep_block.py"], "not tuner")),
    "batched_and_tiles": (["tests/test_gpu_batched.py::test_batched_single_kind"],
                          (["tests/test_gpu_batched.py", "tests/test_gpu_update.py", "tests/test_zz_no_alloc.py"],
                           None)),
    # the same schedules on other streams: the diagonal chain on its own CUs
    # with paired updates, and per-task-class CU policies
    "chain_cus": (["tests/test_gpu_solve.py::test_solve_diag_dominant[512-128-1-2-cholesky]",
                   "tests/wip_gpu_pair.py::test_paired_updates_match[1280-128-0-lu-2]"],
                  (["tests/test_gpu_solve.py", "tests/wip_gpu_pair.py"], "512-128 or 1280")),
    "cu_policy": (["tests/test_gpu_solve.py::test_solve_diag_dominant[512-128-1-2-lu]"],
                  (["tests/test_gpu_solve.py"], "512-128")),
}
# environment of a group's process (the options are read when the library starts)
GROUP_ENV = {"chain_cus": {"PALIAL_CHAIN_CUS": "1"},
             "cu_policy": {"PALIAL_CU_POLICY": "bulk=8-255,aux=8-71,chain=0-7"}}


@pytest.fixture(scope="module")
def race_lib():
    sys.path.insert(0, HERE)
    from emu import build as emu_build
    return emu_build.build("race")


def _env(extra=None):
    return dict(os.environ, PALIAL_HIPEMU="race", **(extra or {}))


def _start(targets, kexpr=None, env=None):
    cmd = [sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", "-m", "gpu", *targets]
    if kexpr:
        cmd += ["-k", kexpr]
    return subprocess.Popen(cmd, cwd=ROOT, env=_env(env), stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                            text=True)


def test_schedules_race_free(race_lib):
    """Every group's GPU tests pass on the race-checking build: their results
    meet the GPU bars and no access races.  Beside them, tests/emu/fuzz.py
    runs seeded random factorizations and solves (every algorithm, schedule,
    precision tier, sweep block and pairing setting, tile sizes 32 to 256)
    with the same checks.  Everything runs side by side, one process each."""
    procs = {}
    for name, (subset, (targets, kexpr)) in GROUPS.items():
        env = GROUP_ENV.get(name)
        procs[name] = _start(targets, kexpr, env) if FULL else _start(subset, env=env)
    fuzz = subprocess.Popen([sys.executable, os.path.join(HERE, "emu", "fuzz.py"), "7", "150" if FULL else "12"],
                            cwd=ROOT, env=_env(), stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    bad = {}
    for name, p in procs.items():
        out, err = p.communicate(timeout=4 * 3600 if FULL else 1200)
        if p.returncode != 0 or " passed" not in out or "failed" in out or "error" in out:
            bad[name] = out[-3000:] + err[-1000:]
    out, err = fuzz.communicate(timeout=4 * 3600 if FULL else 1200)
    if fuzz.returncode != 0 or "fails 0" not in out:
        bad["fuzz"] = "\n".join(l for l in out.splitlines() if not l.startswith("ok")) + err[-2000:]
    assert not bad, bad


_PLANT = r"""
import ctypes as C, sys
sys.path.insert(0, sys.argv[1])
import numpy as np
import palial_amd as P
import oracle
P.LIB_PATH = sys.argv[2]
P.init(0)
L = P.lib()
L.hipemu_races.restype = C.c_long
N, b = 512, 128
rhs = np.random.default_rng(1).standard_normal((N, 1))
def run(algo, drop, prec=None):
    graph = algo.endswith("-graph")            # the FP64 factorization as captured and replayed
    algo = algo.replace("-graph", "")
    a = oracle.gen("sym" if algo == "cholesky" else "dense", N, b, seed=3, shift=0.0 if algo == "qr" else N)
    L.hipemu_race_reset()
    L.hipemu_drop_wait(drop)
    if prec is not None:
        P.solve(a, rhs, algo=algo, precision=prec, b=b, factor_tile=b, check=False)
    elif algo == "qr":
        P.qr(a, np.zeros(N * N), b=b, no_graph=True)
    else:
        getattr(P, algo)(a, b=b, no_graph=not graph)
    L.hipemu_drop_wait(0)
    return L.hipemu_races()
for algo, prec, drops in (("cholesky", None, (3, 5)), ("lu", None, (3,)), ("qr", None, (17, 18)),
                          ("cholesky", 2, (6,)), ("cholesky-graph", None, (3,))):
    print(algo, prec, run(algo, 0, prec), *[run(algo, d, prec) for d in drops], flush=True)
L.hipemu_race_reset()
assert L.hipemu_plant_wg_race() == 0
assert L.hipemu_plant_lds_race() == 0
buf = C.create_string_buffer(1 << 14)
L.hipemu_race_report(buf, len(buf))
print("WG", L.hipemu_races(), buf.value.decode().replace("\n", " | "))
"""


def test_checker_reports_planted_races(race_lib):
    """The check has teeth.  With one event wait of a schedule left out
    (hipemu_drop_wait: the harness's planted mistake) the FP64 Cholesky, LU
    and QR factorizations, the BF16 Cholesky solve and the FP64 Cholesky
    captured into a graph (the wait left out of the capture, so the replay
    lacks that edge) report racing accesses, and the same runs with every
    wait in place report none; a launch in which
    two workgroups store to one element, or one reads another's store, is
    reported as a race between workgroups; a wave that reads what another
    wave stored to LDS with no barrier between, a workgroup that reads LDS it
    never stored to, lanes of one wave that exchange values with no wave
    operation between in an order the emulator does not reproduce, a read of
    a global_load_lds destination before its s_waitcnt, and a 16-byte LDS load
    at an address that is a multiple of 8 only, are reported too."""
    r = subprocess.run([sys.executable, "-c", _PLANT, ROOT, race_lib], cwd=ROOT, env=_env(), capture_output=True,
                       text=True, timeout=900)
    assert r.returncode == 0, r.stderr[-3000:]
    lines = [l.split() for l in r.stdout.splitlines()]
    runs = [l for l in lines if l[0] != "WG"]
    assert len(runs) == 5, r.stdout
    for l in runs:
        clean, dropped = int(l[2]), [int(x) for x in l[3:]]
        assert clean == 0 and all(d > 0 for d in dropped), l
    wg = r.stdout[r.stdout.index("WG "):]
    assert "write over another workgroup's write in one launch" in wg, wg
    assert "read of another workgroup's write in one launch" in wg, wg
    assert "LDS write over another wave's read with no barrier between" in wg, wg
    assert "LDS read before the workgroup wrote it" in wg, wg
    assert "read of what another lane of the wave wrote, with no wave operation between" in wg, wg
    assert "write over what another lane of the wave read, with no wave operation between" in wg, wg
    assert "LDS read of a global_load_lds destination before its s_waitcnt vmcnt" in wg, wg
    # a misaligned vector access ends the run with its report
    code = "import ctypes as C, sys\nprint(C.CDLL(sys.argv[1]).hipemu_plant_misaligned())\n"
    r = subprocess.run([sys.executable, "-c", code, race_lib], cwd=ROOT, env=_env(), capture_output=True, text=True,
                       timeout=300)
    assert r.returncode != 0 and "vector access at an address not aligned to its size: 16 bytes at LDS" in r.stderr, \
        r.stderr[-2000:]


_TRAFFIC = r"""
import ctypes as C, json, sys
sys.path.insert(0, sys.argv[1])
import palial_amd as P
P.LIB_PATH = sys.argv[2]
P.init(0)
L = P.lib()
L.hipemu_ops.restype = C.c_long
L.palial_hip_update.argtypes = [C.c_int, P.MATRIX_desc, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)]
N, b = 1024, 256
A = P.DeviceMatrix(N, b)
A.generate(P.GEN_SYM, seed=1, shift=float(N))
for algo in (0, 1):
    for prec in (1, 2, 3, 4):
        n0 = L.hipemu_ops()
        assert L.palial_hip_update(algo, A.desc, 0, prec | 8, 1, C.byref(C.c_double())) == 0
        out, name = (C.c_double * 4)(), C.create_string_buffer(256)
        for i in range(n0 + 1, L.hipemu_ops() + 1):
            L.hipemu_op_stats(i, out, name, 256)
            if name.value.startswith(b"(gemm_"):
                print(json.dumps({"algo": algo, "prec": prec, "kernel": name.value.decode(),
                                  "rd": out[0], "wr": out[1], "fp_rd": out[2], "fp_wr": out[3]}))
"""


def test_packed_update_moves_the_algorithmic_bytes(race_lib):
    """The byte count that the roofline of the packed trailing update uses
    (each FP32 C tile read and written once, each packed panel tile read
    once) is what the kernel's own loads and stores touch.  One update of
    step 0 at N=1024, b=256 (3 trailing tile rows) per tier and algorithm,
    traffic counted by the race build:
      * every C element read is written exactly once, and nothing is
        written that was not read first;
      * the rest of the read footprint is the packed panels: T tiles of
        b^2 elements at 4 (FP32), 2 (BF16), 1 (FP8, plus one 4-byte scale
        exponent per tile) or 4 (split BF16: hi and lo) bytes, twice for LU
        (row and column panel);
      * Cholesky writes the whole GEMM tiles and, in its diagonal (SYRK)
        tiles, at most their upper 128-blocks -- never more than the whole
        tiles the roofline counts."""
    import json
    r = subprocess.run([sys.executable, "-c", _TRAFFIC, ROOT, race_lib], cwd=ROOT, env=_env(), capture_output=True,
                       text=True, timeout=900)
    assert r.returncode == 0, r.stderr[-3000:]
    rows = [json.loads(l) for l in r.stdout.splitlines() if l.startswith("{")]
    assert len(rows) == 8, r.stdout
    b, T = 256, 3
    tile = 4 * b * b                               # an FP32 C tile
    esz = {1: 4, 2: 2, 3: 1, 4: 4}
    for x in rows:
        panels = (1 if x["algo"] == 0 else 2) * T * (esz[x["prec"]] * b * b + (4 if x["prec"] == 3 else 0))
        assert x["fp_wr"] == 0, x
        assert x["fp_rd"] - panels == x["wr"], x
        if x["algo"] == 1:
            assert x["wr"] == T * T * tile, x
        else:
            gemm, syrk = T * (T - 1) // 2, T
            nb = b // 128
            upper = gemm * tile + syrk * tile * (nb * (nb + 1) // 2) // (nb * nb)
            assert gemm * tile + syrk * tile // 2 <= x["wr"] <= upper, x
        assert x["rd"] > x["fp_rd"]                   # panel slices are read again by other blocks


_SWEEPS = r"""
import ctypes as C, json, sys
L = C.CDLL(sys.argv[1])
L.hipemu_ops.restype = C.c_long
for algo in (0, 1):
    for f32 in (0, 1):
        for q, nrhs in ((1, 1), (2, 1), (1, 3)):
            n0 = L.hipemu_ops()
            assert L.hipemu_sweep_check(algo, 1024, 128, q, nrhs, f32, 0) == 0
            tot = [0.0, 0.0]
            out, name = (C.c_double * 4)(), C.create_string_buffer(256)
            for i in range(n0 + 1, L.hipemu_ops() + 1):
                L.hipemu_op_stats(i, out, name, 256)
                tot[0] += out[0]
                tot[1] += out[2]
            print(json.dumps({"algo": algo, "f32": f32, "q": q, "nrhs": nrhs, "rd": tot[0], "fp_rd": tot[1]}))
"""


def test_correction_sweeps_read_the_factor_once(race_lib):
    """The bandwidth floor of a refinement step's two triangular sweeps is
    one read of each factor triangle (tile inverses in place of the
    diagonal tiles).  On the CPU build, the sweep launches of one correction
    (N=1024, b=128; Cholesky and LU; FP64 and FP32 factors; one-tile and
    two-tile steps; 1 and 3 right-hand sides) touch those bytes plus the
    vectors, and ask for little more: the factor is not read twice."""
    import json
    r = subprocess.run([sys.executable, "-c", _SWEEPS, race_lib], cwd=ROOT, env=_env(), capture_output=True,
                       text=True, timeout=900)
    assert r.returncode == 0, r.stderr[-3000:]
    rows = [json.loads(l) for l in r.stdout.splitlines() if l.startswith("{")]
    assert len(rows) == 12, r.stdout
    N, b = 1024, 128
    MT = N // b
    for x in rows:
        tri = MT * (MT + 1) // 2 * b * b * (4 if x["f32"] else 8)    # one triangle of tiles
        vec = 3 * N * x["nrhs"] * 8                                     # r, y and x
        assert 2 * tri <= x["fp_rd"] <= 2 * tri + 4 * vec, x
        assert x["rd"] <= 1.3 * x["fp_rd"], x


_RESIDUAL = r"""
import ctypes as C, json, sys
sys.path.insert(0, sys.argv[1])
import numpy as np
import palial_amd as P
import oracle
P.LIB_PATH = sys.argv[2]
P.init(0)
L = P.lib()
L.hipemu_ops.restype = C.c_long
N, b = 1024, 128
for algo in ("cholesky", "lu"):
    a = oracle.gen("sym" if algo == "cholesky" else "dense", N, b, seed=1, shift=N)
    n0 = L.hipemu_ops()
    X, rep, rc = P.solve(a, np.ones((N, 1)), algo=algo, b=b, factor_tile=b)
    out, name = (C.c_double * 4)(), C.create_string_buffer(256)
    for i in range(n0 + 1, L.hipemu_ops() + 1):
        L.hipemu_op_stats(i, out, name, 256)
        if "residual_sym_kernel" in name.value.decode() or "residual_partial_kernel" in name.value.decode():
            print(json.dumps({"algo": algo, "rd": out[0], "fp_rd": out[2]}))
"""


def test_residual_reads_the_matrix_once(race_lib):
    """The FP64 residual of a refinement step reads A once: the upper tiles
    for Cholesky (the symmetric matrix the factorization sees, 4 N^2 bytes),
    every tile for LU (8 N^2 bytes), plus the solution vector -- the HBM
    figure the refinement's time model uses.  Counted on the CPU build for
    the first residual of an FP64 solve at N=1024, b=128."""
    import json
    r = subprocess.run([sys.executable, "-c", _RESIDUAL, ROOT, race_lib], cwd=ROOT, env=_env(), capture_output=True,
                       text=True, timeout=900)
    assert r.returncode == 0, r.stderr[-3000:]
    rows = [json.loads(l) for l in r.stdout.splitlines() if l.startswith("{")]
    N, b = 1024, 128
    MT = N // b
    seen = set()
    for x in rows:
        mat = (MT * (MT + 1) // 2 * b * b if x["algo"] == "cholesky" else N * N) * 8
        assert mat <= x["fp_rd"] <= mat + 4 * N * 8, x
        assert x["rd"] <= 1.1 * x["fp_rd"], x
        seen.add(x["algo"])
    assert seen == {"cholesky", "lu"}, r.stdout
