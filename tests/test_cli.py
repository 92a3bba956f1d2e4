"""CPU tests of the drop-in programs' host side: options, the 4-line config format
(radon_grid/src/init.c:17-82: "%d %[^\\n]\\n" per line) and the text readers.  Every case here
stops before the first device call, so no GPU is needed."""
import os

import pytest

from cli_util import BIN, run, write_column, write_config

PROGS = ["radon_grid", "ray_grid", "radon_reduction", "radon3d_adjoint", "wrt_serve"]


def test_programs_are_built():
    for p in PROGS:
        assert os.access(os.path.join(BIN, p), os.X_OK), p


@pytest.mark.parametrize("prog", PROGS[:4])
def test_help_and_missing_arguments(prog, tmp_path):
    r = run(prog, ["-h"], tmp_path)
    assert r.returncode == 0 and "Usage" in r.stdout
    r = run(prog, ["--help"], tmp_path)  # long options work (the reference's never reach its switch)
    assert r.returncode == 0 and "Usage" in r.stdout
    r = run(prog, [], tmp_path)
    assert r.returncode != 0 and "arguments required" in r.stderr
    r = run(prog, ["-p", "cfg.txt"], tmp_path)
    assert r.returncode != 0
    r = run(prog, ["-x"], tmp_path)
    assert r.returncode != 0


@pytest.mark.parametrize("prog", PROGS[:4])
def test_config_errors(prog, tmp_path):
    # a missing file, too few lines, a line without its label, a non-integer
    r = run(prog, ["-p", "nope.txt", "-i", "in.txt", "-o", "out.txt"], tmp_path)
    assert r.returncode != 0 and "nope.txt" in r.stderr
    with open(tmp_path / "short.txt", "w") as f:
        f.write("8 a\n8 b\n")
    r = run(prog, ["-p", "short.txt", "-i", "in.txt", "-o", "out.txt"], tmp_path)
    assert r.returncode != 0 and "end of file" in r.stderr
    with open(tmp_path / "nolabel.txt", "w") as f:
        f.write("8\n8\n9\n8\n")
    r = run(prog, ["-p", "nolabel.txt", "-i", "in.txt", "-o", "out.txt"], tmp_path)
    assert r.returncode != 0
    with open(tmp_path / "bad.txt", "w") as f:
        f.write("x a\n8 b\n9 c\n8 d\n")
    r = run(prog, ["-p", "bad.txt", "-i", "in.txt", "-o", "out.txt"], tmp_path)
    assert r.returncode != 0 and "matching failure" in r.stderr


def test_readers_report_short_and_malformed_inputs(tmp_path):
    write_config(tmp_path / "cfg.txt", [4, 4, 5, 4])  # radon_grid: nphi ntheta nshift ngrid
    write_column(tmp_path / "short.txt", range(10))  # needs 4^3 values
    r = run("radon_grid", ["-p", "cfg.txt", "-i", "short.txt", "-o", "out.txt"], tmp_path)
    assert r.returncode != 0 and "end of file" in r.stderr
    with open(tmp_path / "bad.txt", "w") as f:
        f.write("1.0\nabc\n")
    r = run("radon_grid", ["-p", "cfg.txt", "-i", "bad.txt", "-o", "out.txt"], tmp_path)
    assert r.returncode != 0 and "matching failure" in r.stderr
    # the 4-column reader of radon_reduction / radon3d_adjoint
    write_config(tmp_path / "rcfg.txt", [5, 4, 2, 4])  # nshift nphi ntheta ngrid
    with open(tmp_path / "rays.txt", "w") as f:
        f.write("0.1, 0.2, 0.3\n")
    r = run("radon_reduction", ["-p", "rcfg.txt", "-i", "rays.txt", "-o", "out.txt"], tmp_path)
    assert r.returncode != 0 and "matching failure" in r.stderr


AN_PROGS = ["radon_analytic", "ray_analytic"]


@pytest.mark.parametrize("prog", AN_PROGS)
def test_analytic_programs_options(prog, tmp_path):
    assert os.access(os.path.join(BIN, prog), os.X_OK)
    r = run(prog, ["-h"], tmp_path)
    assert r.returncode == 0 and "exponential_bump" in r.stdout
    r = run(prog, [], tmp_path)
    assert r.returncode != 0 and "arguments required" in r.stderr
    # the test function: exactly one of -F / -f
    r = run(prog, ["-p", "cfg.txt", "-o", "out.txt"], tmp_path)
    assert r.returncode != 0 and "-F" in r.stderr
    r = run(prog, ["-p", "cfg.txt", "-o", "out.txt", "-F", "exponential_bump", "-f", "t.txt"], tmp_path)
    assert r.returncode != 0
    r = run(prog, ["-p", "cfg.txt", "-o", "out.txt", "-F", "no_such_bump"], tmp_path)
    assert r.returncode != 0 and "no_such_bump" in r.stderr


@pytest.mark.parametrize("prog", AN_PROGS)
def test_analytic_term_files_and_configs(prog, tmp_path):
    with open(tmp_path / "bad.txt", "w") as f:
        f.write("# comment\nexp_bump 0.2 0 0 0.5\n")  # one number short
    r = run(prog, ["-p", "cfg.txt", "-o", "out.txt", "-f", "bad.txt"], tmp_path)
    assert r.returncode != 0 and "takes 5 numbers" in r.stderr
    with open(tmp_path / "unk.txt", "w") as f:
        f.write("cube 0 0 0 1\n")
    r = run(prog, ["-p", "cfg.txt", "-o", "out.txt", "-f", "unk.txt"], tmp_path)
    assert r.returncode != 0 and "unknown term" in r.stderr
    with open(tmp_path / "ok.txt", "w") as f:
        f.write("exp_bump 0.2 0 0 0.5 0.01  # the reference's example\n\n"
                "sine_bump 0 0.2 0 1.0 0.4 0.6\n")
    # a config with too few lines fails after the test function was read
    with open(tmp_path / "short.txt", "w") as f:
        f.write("8 a\n")
    r = run(prog, ["-p", "short.txt", "-o", "out.txt", "-f", "ok.txt"], tmp_path)
    assert r.returncode != 0 and "2 terms" in r.stdout and "end of file" in r.stderr
