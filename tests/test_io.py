"""CPU tests of the file formats (host/wrt_io.c): the parallel %lf writer produces the C
library's bytes, the parallel reader the values fscanf reads (with its error messages), and the
binary format round-trips."""
import ctypes as C
import os

import numpy as np
import pytest

import wrt

_libc = C.CDLL(None)


def c_lf(v):
    """the C library's "%lf" of v"""
    buf = C.create_string_buffer(400)
    _libc.snprintf(buf, 400, b"%lf", C.c_double(v))
    return buf.value.decode()


def _values():
    rng = np.random.default_rng(0)
    v = [0.0, -0.0, 1e-300, -1e-300, 5e-324, -5e-7, 5e-7, 4.9999999e-7, 1.0, -1.0, 0.5, 1e14,
         -9.99e15, 2.0 ** 43, 2.0 ** 43 - 0.5, 1.7976931348623157e308, -1e200, float("inf"),
         float("-inf"), 123456.7890125, 0.0000005, 0.0000015, 0.0000025]
    # exact ties of the 6th decimal: odd multiples of 2^-7 (x * 10^6 ends in .5)
    v += [k / 128.0 for k in range(-300, 300)]
    v += [k / 2.0 ** 20 for k in range(-200, 200, 3)]
    for e in range(-12, 15):
        v += list(rng.uniform(-1, 1, 300) * 10.0 ** e)
    v += list(rng.standard_normal(2000))
    v += [float(np.nextafter(k / 128.0, 1.0)) for k in range(1, 60)]
    v += [float(np.nextafter(k / 128.0, -1.0)) for k in range(1, 60)]
    return np.array(v)


def test_lf_writer_matches_the_c_library(tmp_path):
    v = _values()
    wrt.write_text_values(tmp_path / "v.txt", v)
    want = "".join(c_lf(x) + "\n" for x in v)
    got = open(tmp_path / "v.txt").read()
    assert got == want


def test_lf_writer_nan_and_blocks(tmp_path):
    # NaN and more rows than one formatting block
    v = np.concatenate([[float("nan"), -float("nan")], np.random.default_rng(1).standard_normal(70000)])
    wrt.write_text_values(tmp_path / "v.txt", v)
    got = open(tmp_path / "v.txt").read().split("\n")[:-1]
    assert len(got) == v.size
    assert got[:2] == [c_lf(v[0]), c_lf(v[1])]
    assert all(g == c_lf(x) for g, x in zip(got[2::997], v[2::997]))


def test_grid_writer_rows(tmp_path):
    a, b, c = np.array([-1.0, 0.5]), np.array([0.0, 1.25, 3.5]), np.array([0.1, 0.2])
    vals = np.arange(12) / 7.0
    wrt.write_text_grid(tmp_path / "g.txt", a, b, c, vals)
    lines = open(tmp_path / "g.txt").read().split("\n")[:-1]
    k = 0
    for x in a:
        for y in b:
            for z in c:
                assert lines[k] == f"{c_lf(x)}, {c_lf(y)}, {c_lf(z)}, {c_lf(vals[k])}"
                k += 1


def test_reader_values_and_formats(tmp_path):
    v = np.random.default_rng(2).standard_normal(5000) * 1e3
    with open(tmp_path / "one.txt", "w") as f:
        for x in v:
            f.write(repr(float(x)) + "\n")
    assert np.array_equal(wrt.read_text(tmp_path / "one.txt", v.size, 1), v)
    with open(tmp_path / "four.txt", "w") as f:
        for i, x in enumerate(v):
            f.write(f"{i * 0.5}, {-i}, 1e-3, {float(x)!r}\n")
    assert np.array_equal(wrt.read_text(tmp_path / "four.txt", v.size, 4), v)
    # fscanf's white-space rules: blank lines, CRLF, leading blanks, a record across lines,
    # trailing data beyond the count
    with open(tmp_path / "ws.txt", "w") as f:
        f.write("  1.5\r\n\n2.5\n\t-3e2\n4\n5\n")
    assert list(wrt.read_text(tmp_path / "ws.txt", 4, 1)) == [1.5, 2.5, -300.0, 4.0]
    with open(tmp_path / "ws4.txt", "w") as f:
        f.write("0, 0, 0, 1.0\n0,\n  0, 0,   2.0\n0, 0, 0, 3.0")
    assert list(wrt.read_text(tmp_path / "ws4.txt", 3, 4)) == [1.0, 2.0, 3.0]


def test_reader_errors(tmp_path, capfd):
    with open(tmp_path / "short.txt", "w") as f:
        f.write("1\n2\n")
    with pytest.raises(wrt.WrtError) as e:
        wrt.read_text(tmp_path / "short.txt", 3, 1)
    assert e.value.code == wrt.WRT_E_IO
    assert "end of file" in capfd.readouterr().err
    with open(tmp_path / "bad.txt", "w") as f:
        f.write("1\nabc\n3\n")
    with pytest.raises(wrt.WrtError):
        wrt.read_text(tmp_path / "bad.txt", 3, 1)
    assert "matching failure" in capfd.readouterr().err
    with open(tmp_path / "sp.txt", "w") as f:
        f.write("0 , 0, 0, 1\n")  # a blank before the comma: fscanf stops there
    with pytest.raises(wrt.WrtError):
        wrt.read_text(tmp_path / "sp.txt", 1, 4)
    assert "matching failure" in capfd.readouterr().err
    with pytest.raises(wrt.WrtError):
        wrt.read_text(tmp_path / "missing.txt", 1, 1)


def test_binary_round_trip_and_checks(tmp_path):
    v = np.random.default_rng(3).standard_normal(1001)
    wrt.write_f64(tmp_path / "v.f64", v)
    assert os.path.getsize(tmp_path / "v.f64") == 16 + 8 * v.size
    assert np.array_equal(wrt.read_f64(tmp_path / "v.f64", v.size), v)
    for count in (1000, 1002):
        with pytest.raises(wrt.WrtError):
            wrt.read_f64(tmp_path / "v.f64", count)
    with open(tmp_path / "t.txt", "w") as f:
        f.write("1.0\n" * 100)
    with pytest.raises(wrt.WrtError):
        wrt.read_f64(tmp_path / "t.txt", 1)
