"""CPU tests of the drop-in CLI's argument handling (palial_src/main.c:62-105
of the reference): messages on stdout and exit status 0 for missing
arguments, an unknown option and N % B != 0; an unknown algorithm runs
nothing and prints nothing.  None of these paths reaches the GPU."""
import os
import subprocess

import pytest


def run(*args):
    import palial_amd
    exe = palial_amd.DRIVER_PATH
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built (python -m palial_amd.build)")
    return subprocess.run([exe, *args], capture_output=True, text=True, timeout=60)


def test_missing_arguments():
    r = run()
    assert r.returncode == 0 and r.stdout == "Missing arguments.\nAborting...\n"


def test_tile_must_divide_matrix():
    r = run("-a", "cholesky", "-m", "100", "-b", "30")
    assert r.returncode == 0 and r.stdout == "Tile size does not divide the matrix size. Aborting...\n"


def test_unknown_option():
    r = run("-a", "lu", "-m", "64", "-b", "32", "-x")
    assert r.returncode == 0 and r.stdout == "Invalid arguments.\nAborting...\n"


def test_unknown_algorithm_is_silent():
    r = run("-a", "svd", "-m", "64", "-b", "32")
    assert r.returncode == 0 and r.stdout == "" and r.stderr == ""
