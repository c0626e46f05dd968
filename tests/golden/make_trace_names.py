"""Writes cholesky_trace_names_mt16.txt: the task names, in creation order,
of the reference tracer's output for a Cholesky of N=8192, b=512 (16 x 16
tiles), taken from the reference's recorded run
exp/data/tasks_times_8192_512_31_ondemand (lines `name, ms`; only the names
are kept: the times are CPU times of that machine)."""
import os
import sys

SRC = "/root/reference/exp/data/tasks_times_8192_512_31_ondemand"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cholesky_trace_names_mt16.txt")

if __name__ == "__main__":
    names = [line.split(",")[0].strip() for line in open(SRC) if line.strip()]
    with open(OUT, "w") as f:
        f.write("\n".join(names) + "\n")
    print(f"{len(names)} names -> {OUT}")
    sys.exit(0)
