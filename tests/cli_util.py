"""Helpers for the tests of the drop-in programs: paths, config files, text inputs."""
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "wrt-project_amd", "bin")


def run(prog, args, cwd, timeout=120):
    return subprocess.run([os.path.join(BIN, prog)] + list(args), cwd=cwd, capture_output=True,
                          text=True, timeout=timeout)


def write_config(path, values, labels=("a", "b", "c", "d")):
    with open(path, "w") as f:
        for v, l in zip(values, labels):
            f.write(f"{v} {l}\n")


def write_column(path, values):
    with open(path, "w") as f:
        for v in np.asarray(values).ravel():
            f.write(repr(float(v)) + "\n")


def read_table(path, cols):
    a = np.loadtxt(path, delimiter="," if cols > 1 else None, ndmin=2)
    assert a.shape[1] == cols, a.shape
    return a


def lf(values):
    """the %lf text of values, one per line (what the programs print)"""
    return ["%f" % v for v in np.asarray(values).ravel()]
