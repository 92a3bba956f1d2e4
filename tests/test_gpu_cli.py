"""GPU tests of the drop-in programs end to end: each reads the reference's config and text
formats, runs on the GPU and prints %lf text in the reference's element order.  The text is
compared with the oracle's values printed the same way: every value within the %lf quantum
(5e-7) plus 1e-9 of the oracle, and at least 99.9% of the lines identical (a line can flip
in the last digit where the value lies on a rounding boundary).  Chains go through files, as
the reference's programs are chained: radon_grid -> radon3d_adjoint, ray_grid ->
radon_reduction (4-column and one-column inputs)."""
import numpy as np
import pytest

import orc
import wrt
from cli_util import lf, read_table, run, write_column, write_config

pytestmark = pytest.mark.gpu


def _same_text(got, want, frac=0.999):
    assert len(got) == len(want)
    g, w = np.array([float(v) for v in got]), np.array([float(v) for v in want])
    assert np.max(np.abs(g - w)) <= 5e-7 + 1e-9, np.max(np.abs(g - w))
    same = np.mean([a == b for a, b in zip(got, want)])
    assert same >= frac, same


def _col(path, cols):
    lines = open(path).read().split("\n")[:-1]
    return [l.split(", ")[cols - 1] if cols > 1 else l for l in lines]


def test_radon_grid_and_adjoint_chain(tmp_path):
    n, nphi, nth, ns = 16, 8, 4, 17
    vol = wrt.phantom(n, True, 10, 4, 0.01)
    write_config(tmp_path / "cfg.txt", [nphi, nth, ns, n])
    write_column(tmp_path / "vol.txt", vol)
    r = run("radon_grid", ["-p", "cfg.txt", "-i", "vol.txt", "-o", "radon.txt", "-n", "4"], tmp_path)
    assert r.returncode == 0, r.stderr
    t = read_table(tmp_path / "radon.txt", 4)
    phi = orc.phi_grid(nphi)
    th, w = orc.theta_grid(nth)
    s = orc.shift_grid(ns)
    S, P, T = np.meshgrid(s, phi, th, indexing="ij")
    assert np.allclose(t[:, 0], S.ravel(), atol=5e-7) and np.allclose(t[:, 1], P.ravel(), atol=5e-7)
    assert np.allclose(t[:, 2], T.ravel(), atol=5e-7)
    want = orc.radon_grid(vol, phi, th, s)
    _same_text(_col(tmp_path / "radon.txt", 4), lf(want))
    readme = open(tmp_path / "readme_radon.txt").read()
    assert "[s], [phi], [theta], [Rf]" in readme and f": {n}," in readme

    # the adjoint reads radon_grid's output file (values quantised by %lf, as in the reference)
    write_config(tmp_path / "acfg.txt", [nphi, nth, ns, n])
    r = run("radon3d_adjoint", ["-p", "acfg.txt", "-i", "radon.txt", "-o", "adj.txt"], tmp_path)
    assert r.returncode == 0, r.stderr
    rq = t[:, 3].reshape(ns, nphi, nth)
    want = orc.radon_adjoint(rq, phi, th, w, s, n)
    _same_text(_col(tmp_path / "adj.txt", 1), lf(want))


def test_ray_grid_and_reduction_chain(tmp_path):
    n, nz, ns, nphi, nth = 17, 17, 9, 8, 4
    vol = wrt.phantom(n, False, 10, 6, 0.0)
    write_config(tmp_path / "cfg.txt", [n, nz, ns, nphi])
    write_column(tmp_path / "vol.txt", vol)
    r = run("ray_grid", ["-p", "cfg.txt", "-i", "vol.txt", "-o", "rays.txt"], tmp_path)
    assert r.returncode == 0, r.stderr
    t = read_table(tmp_path / "rays.txt", 4)
    s, phi = orc.shift_grid(ns), orc.phi_grid(nphi)
    z = orc.zslice_grid(nz)
    want = orc.ray_grid(vol, z, s, phi)
    _same_text(_col(tmp_path / "rays.txt", 4), lf(want))

    write_config(tmp_path / "rcfg.txt", [ns, nphi, nth, nz])
    th, _ = orc.theta_grid(nth)
    rays_q = t[:, 3].reshape(nz, ns, nphi)
    want = orc.radon_reduction(rays_q, s, phi, th)
    r = run("radon_reduction", ["-p", "rcfg.txt", "-i", "rays.txt", "-o", "red.txt"], tmp_path)
    assert r.returncode == 0, r.stderr
    _same_text(_col(tmp_path / "red.txt", 1), lf(want))
    # one value per line with -s
    write_column(tmp_path / "rays1.txt", rays_q)
    r = run("radon_reduction", ["-p", "rcfg.txt", "-i", "rays1.txt", "-o", "red1.txt", "-s"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert open(tmp_path / "red1.txt").read() == open(tmp_path / "red.txt").read()


def test_precision_option(tmp_path):
    n, nphi, nth, ns = 32, 8, 4, 33
    vol = wrt.phantom(n, True, 10, 8, 0.01)
    write_config(tmp_path / "cfg.txt", [nphi, nth, ns, n])
    write_column(tmp_path / "vol.txt", vol)
    for prec in ("f64", "f32", "bf16"):
        r = run("radon_grid", ["-p", "cfg.txt", "-i", "vol.txt", "-o", f"{prec}.txt", "-P", prec],
                tmp_path)
        assert r.returncode == 0, r.stderr
    a = read_table(tmp_path / "f64.txt", 4)[:, 3]
    b = read_table(tmp_path / "f32.txt", 4)[:, 3]
    c = read_table(tmp_path / "bf16.txt", 4)[:, 3]
    assert np.max(np.abs(a - b)) <= 1e-5 * np.max(np.abs(a)) + 1e-6
    assert orc.rel_l2(c, a) <= 2e-3


def test_serve_program(tmp_path):
    r = run("wrt_serve", ["-q", "200", "-r", "12", "-e", "32,48", "-f", "3"], tmp_path)
    assert r.returncode == 0, r.stderr
    import json
    rep = json.loads(r.stdout.strip().split("\n")[-1])
    assert rep["requests"] == 12 and rep["p50_ms"] <= rep["p99_ms"] and rep["samples"] > 0


def test_radon_analytic_program(tmp_path):
    from golden_inputs import example_terms

    nphi, nth, ns, g = 8, 4, 9, 17
    write_config(tmp_path / "cfg.txt", [nphi, nth, ns, g])
    r = run("radon_analytic", ["-p", "cfg.txt", "-o", "r.txt", "-F", "sinusoidal_bump"], tmp_path)
    assert r.returncode == 0, r.stderr
    t = read_table(tmp_path / "r.txt", 4)
    phi = orc.phi_grid(nphi)
    th, _ = orc.theta_grid(nth)
    s = orc.shift_grid(ns)
    S, P, T = np.meshgrid(s, phi, th, indexing="ij")
    assert np.allclose(t[:, 0], S.ravel(), atol=5e-7) and np.allclose(t[:, 2], T.ravel(), atol=5e-7)
    want = orc.analytic_radon_grid(example_terms("sinusoidal_bump"), g, phi, th, s)
    _same_text(_col(tmp_path / "r.txt", 4), lf(want))
    # the same function from a term file
    with open(tmp_path / "terms.txt", "w") as f:
        for kind, c, p in example_terms("sinusoidal_bump"):
            f.write("sine_bump " + " ".join(repr(v) for v in list(c) + list(p)) + "\n")
    r = run("radon_analytic", ["-p", "cfg.txt", "-o", "r2.txt", "-f", "terms.txt"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert open(tmp_path / "r2.txt").read() == open(tmp_path / "r.txt").read()


def test_ray_analytic_program(tmp_path):
    from golden_inputs import analytic_zslices, example_terms

    ns, nphi, g = 9, 8, 17
    write_config(tmp_path / "cfg.txt", [ns, nphi, g])
    r = run("ray_analytic", ["-p", "cfg.txt", "-o", "p.txt", "-F", "exponential_bump"], tmp_path)
    assert r.returncode == 0, r.stderr
    t = read_table(tmp_path / "p.txt", 4)
    z = analytic_zslices(g)
    s, phi = orc.shift_grid(ns), orc.phi_grid(nphi)
    Z, S, P = np.meshgrid(z, s, phi, indexing="ij")
    assert np.allclose(t[:, 0], Z.ravel(), atol=5e-7) and np.allclose(t[:, 2], P.ravel(), atol=5e-7)
    want = orc.analytic_ray_grid(example_terms("exponential_bump"), g, z, s, phi)
    _same_text(_col(tmp_path / "p.txt", 4), lf(want))
    assert "[z], [s], [phi], [Pf]" in open(tmp_path / "readme_p.txt").read()


def test_binary_chain_matches_library_calls(tmp_path):
    # ray_grid -B -> radon_reduction -b -B -> radon3d_adjoint -b -B: full binary64 values
    # between the programs (no %lf quantisation), equal to the library's own results
    n, ns, nphi, nth = 17, 9, 8, 4
    vol = wrt.phantom(n, False, 10, 6, 0.0)
    wrt.write_f64(tmp_path / "vol.f64", vol)
    write_config(tmp_path / "rcfg.txt", [n, n, ns, nphi])  # ngrid nslices nshift nphi
    r = run("ray_grid", ["-p", "rcfg.txt", "-i", "vol.f64", "-b", "-o", "rays.f64", "-B"], tmp_path)
    assert r.returncode == 0, r.stderr
    s, phi = wrt.ray_grid_init(ns, nphi)
    z = wrt.zslice_grid(n)
    with wrt.Context(0) as ctx:
        rays, _ = ctx.ray_grid(vol, z, s, phi)
        assert np.array_equal(wrt.read_f64(tmp_path / "rays.f64", rays.size), rays.ravel())
        write_config(tmp_path / "dcfg.txt", [ns, nphi, nth, n])  # nshift nphi ntheta nslices
        r = run("radon_reduction", ["-p", "dcfg.txt", "-i", "rays.f64", "-b", "-o", "radon.f64", "-B"],
                tmp_path)
        assert r.returncode == 0, r.stderr
        _, th, w, _ = ctx.radon_grid_init(nphi, nth, ns)
        radon, _ = ctx.radon_reduction(rays, s, phi, th)
        assert np.array_equal(wrt.read_f64(tmp_path / "radon.f64", radon.size), radon.ravel())
        write_config(tmp_path / "acfg.txt", [nphi, nth, ns, 12])  # nphi ntheta nshift ngrid
        r = run("radon3d_adjoint", ["-p", "acfg.txt", "-i", "radon.f64", "-b", "-o", "adj.f64", "-B"],
                tmp_path)
        assert r.returncode == 0, r.stderr
        adj, _ = ctx.radon_adjoint(radon, phi, th, w, s, 12)
        assert np.array_equal(wrt.read_f64(tmp_path / "adj.f64", adj.size), adj.ravel())
    # a binary file of the wrong size is refused
    r = run("radon3d_adjoint", ["-p", "acfg.txt", "-i", "rays.f64", "-b", "-o", "x.f64"], tmp_path)
    assert r.returncode != 0 and "values" in r.stderr
