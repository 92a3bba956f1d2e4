"""GPU tests at the full workload shapes and at the reference's own example configurations.

  C1  ray_grid 128^3, 128 z x 180 phi in [0, pi) x 257 shifts: the dot-product test
      <A f, g> = <f, A^T g> of the FP32 and the FP64 pairs on seeded Gaussian f and g;
  C2  radon_grid 256^3 (phantom + N(0, 0.01^2) texture), 256 phi x 128 theta x 257 shifts:
      the FP32 and BF16 forms against the plain FP64 form on every 16th phi, the FP32 form
      and the plain form against the oracle on 64 seeded planes (a quarter of them at phi a
      multiple of pi/2; every shift of this grid lies on a cell face), and the FP32
      transpose's dot-product test;
  C3  48 volumes with edges drawn from {32, ..., 192}, forward + backprojection each, in one
      persistent batch: every result equals its alone run;
  C4  a served Poisson stream with edges {64, 96, 128, 192, 256} and the 50/30/20 mix: every
      result equals its alone run;
  the reference's binary/config.txt of radon_grid (256 phi x 128 theta x 129 shifts, n 129),
  ray_grid (129 slices x 129 shifts x 256 angles, n 129), radon_reduction (129 shifts x 128
  phi x 128 theta from 129 slices) and radon3d_ajoint (129 phi x 128 theta x 256 shifts ->
  128^3): the plain forms against the oracle on seeded subsets, the fast forms against the
  plain ones on the whole grid.

Tolerances: plain FP64 against the oracle 1e-10 (relative L2 and max relative); FP32 1e-5;
BF16 bricks relative L2 1e-4 (8 significand bits per voxel); batched and served results
against their alone runs as in test_gpu_batch.py / test_gpu_serve.py; dot-product tests
1e-6 (FP32 pairs) and 1e-12 (FP64 pairs) relative to |A f| |g|."""
import os

import numpy as np
import pytest

import orc
import wrt

pytestmark = pytest.mark.gpu


def _dot_rel(Af, g, f, ATg):
    lhs = float(np.dot(np.ravel(Af), np.ravel(g)))
    rhs = float(np.dot(np.ravel(f), np.ravel(ATg)))
    return abs(lhs - rhs) / (np.linalg.norm(Af) * np.linalg.norm(g))


# ------------------------------------------------------------------------------ C1 ---
@pytest.mark.parametrize("prec, tol", [("f32", 1e-6), ("f64", 1e-12)])
def test_c1_ray_transpose_dot(ctx, prec, tol):
    n = 128
    z = wrt.zslice_grid(128)
    s, phi2 = wrt.ray_grid_init(257, 360)
    phi = np.ascontiguousarray(phi2[:180])
    f = wrt.gaussian((n, n, n), seed=21)
    g = wrt.gaussian((len(z), len(s), len(phi)), seed=22)
    Af, _ = ctx.ray_grid(f, z, s, phi, prec=prec)
    ATg, _ = ctx.ray_grid_T(g, n, z, s, phi, prec=prec)
    err = _dot_rel(Af, g, f, ATg)
    assert err <= tol, (prec, err)


# ------------------------------------------------------------------------------ C2 ---
@pytest.fixture(scope="module")
def c2(ctx):
    n = 256
    vol = wrt.phantom(n, True, 10, 0, 0.01)
    phi, th, _, s = ctx.radon_grid_init(256, 128, 257)
    return vol, phi, th, s


def _c2_planes(phi, th, s, m=64, seed=7):
    rng = np.random.default_rng(seed)
    iP = rng.integers(0, len(phi), m)
    iP[: m // 4] = rng.choice([0, 64, 128, 192], m // 4)  # phi = k pi/2
    iT = rng.integers(0, len(th), m)
    iS = rng.integers(0, len(s), m)
    return iS, iP, iT


def test_c2_fast_forms_against_plain_and_oracle(ctx, c2):
    vol, phi, th, s = c2
    got = {}
    for prec in ("f32", "bf16"):
        out, st = ctx.radon_grid(vol, phi, th, s, prec=prec)
        assert st["logical_samples"] == 16.0 * 256 ** 2 * out.size
        got[prec] = out
    sub = np.arange(0, 256, 16)
    ref, _ = ctx.radon_grid(vol, phi[sub], th, s, prec="f64")
    f32, b16 = got["f32"][:, sub, :], got["bf16"][:, sub, :]
    assert orc.rel_l2(f32, ref) <= 1e-5, orc.rel_l2(f32, ref)
    assert orc.max_rel(f32, ref) <= 1e-5, orc.max_rel(f32, ref)
    assert orc.rel_l2(b16, ref) <= 1e-4, orc.rel_l2(b16, ref)
    iS, iP, iT = _c2_planes(phi, th, s)
    want = orc.radon_planes(vol, phi[iP], th[iT], s[iS])
    assert np.count_nonzero(want) > 40
    fast = got["f32"][iS, iP, iT]
    assert orc.rel_l2(fast, want) <= 1e-5, orc.rel_l2(fast, want)
    assert np.max(np.abs(fast - want)) <= 1e-5 * np.max(np.abs(want))


def test_c2_plain_against_oracle(ctx, c2):
    vol, phi, th, s = c2
    iS, iP, iT = _c2_planes(phi, th, s, seed=8)
    # the plain form on the planes' own grids (one plane per (shift, phi, theta) triple)
    for k in range(0, len(iS), 16):
        sl = slice(k, k + 16)
        out, _ = ctx.radon_grid(vol, phi[iP[sl]], th[iT[sl]], s[iS[sl]], prec="f64")
        m = np.arange(len(iS[sl]))
        got = out[m, m, m]
        want = orc.radon_planes(vol, phi[iP[sl]], th[iT[sl]], s[iS[sl]])
        assert orc.rel_l2(got, want) <= 1e-10
        assert orc.max_rel(got, want) <= 1e-10


def test_c2_transpose_dot(ctx, c2):
    vol, phi, th, s = c2
    n = vol.shape[0]
    fast, _ = ctx.radon_grid(vol, phi, th, s, prec="f32")
    g = wrt.gaussian((len(s), len(phi), len(th)), seed=11)
    ATg, _ = ctx.radon_grid_T(g, n, phi, th, s, prec="f32")
    err = _dot_rel(fast, g, vol, ATg)
    assert err <= 1e-6, err


# ------------------------------------------------------------------------------ C3 ---
def test_c3_batch_equals_alone(ctx, tmp_path, monkeypatch):
    # bit for bit: the per-job launches walk the persistent kernel's 8 x 8 bands
    monkeypatch.setenv("WRT_BAND_ROWS", "8")
    old_cache = os.environ.get("WRT_TUNE_CACHE", "")
    os.environ["WRT_TUNE_CACHE"] = str(tmp_path / "tune.txt")
    try:
        rng = np.random.default_rng(3)
        edges = [int(e) for e in rng.choice([32, 48, 64, 96, 128, 160, 192], size=48)]
        assert max(edges) == 192 and min(edges) == 32
        jobs = []
        for i, e in enumerate(edges):
            phi, th, w, s = ctx.radon_grid_init(e // 2, e // 4, e + 1)
            jr = ctx.job("radon", "f32", n=e, phi=phi, theta=th, shift=s)
            jr.load(wrt.phantom(e, True, 10, 1000 + i, 0.0))
            ja = ctx.job("adjoint", "f32", n=e, phi=phi, theta=th, theta_w=w, shift=s)
            ja.load(wrt.gaussian(len(s) * len(phi) * len(th), 2000 + i))
            jobs += [("radon", jr), ("adjoint", ja)]
        want = []
        for _, j in jobs:
            j.run()
            want.append(j.fetch())
        st = ctx.batch_run([j for _, j in jobs])
        assert st["kernel_seconds"] > 0
        for (op, j), w in zip(jobs, want):
            got = j.fetch()
            assert np.array_equal(got, w), op
        for _, j in jobs:
            j.close()
    finally:
        os.environ["WRT_TUNE_CACHE"] = old_cache


# ------------------------------------------------------------------------------ C4 ---
def test_c4_served_equals_alone(ctx, monkeypatch):
    # bit for bit: the alone runs walk the serve rounds' 8 x 8 bands
    monkeypatch.setenv("WRT_BAND_ROWS", "8")
    cfg = wrt.serve_cfg(qps=24.0, nreq=20, seed=5, mix=(0.5, 0.3, 0.2),
                        edges=(64, 96, 128, 192, 256), ndatasets=2, max_inflight=4)
    reqs = [wrt.serve_describe(cfg, i) for i in range(cfg.nreq)]
    assert {r.edge for r in reqs} >= {256, 192}
    rep, lat, outs = ctx.serve(cfg, keep=range(cfg.nreq))
    assert rep["completed"] == cfg.nreq and rep["peak_inflight"] >= 2
    for i, got in outs.items():
        j = ctx.serve_job(reqs[i])
        j.run()
        want = j.fetch()
        j.close()
        assert np.array_equal(got, want), (i, reqs[i].op)


# ------------------------------------------------------ the reference's example configs ---
def test_reference_config_radon_grid(ctx):
    # radon_grid/binary/config.txt: 256 phi, 128 theta, 129 shifts, n 129
    n = 129
    vol = wrt.phantom(n, True, 10, 4, 0.0)
    phi, th, _, s = ctx.radon_grid_init(256, 128, 129)
    ref, _ = ctx.radon_grid(vol, phi, th, s, prec="f64")
    got, _ = ctx.radon_grid(vol, phi, th, s, prec="f32")
    assert orc.rel_l2(got, ref) <= 1e-5 and orc.max_rel(got, ref) <= 1e-5
    iS, iP, iT = _c2_planes(phi, th, s, m=48, seed=9)
    want = orc.radon_planes(vol, phi[iP], th[iT], s[iS])
    assert orc.rel_l2(ref[iS, iP, iT], want) <= 1e-10
    assert orc.max_rel(ref[iS, iP, iT], want) <= 1e-10


def test_reference_config_ray_grid(ctx):
    # ray_grid/binary/config.txt (in its parse order): 129 slices, 129 shifts, 256 angles,
    # n 129
    n = 129
    vol = wrt.phantom(n, False, 10, 5, 0.0)
    z = wrt.zslice_grid(129)
    s, phi = wrt.ray_grid_init(129, 256)
    ref, _ = ctx.ray_grid(vol, z, s, phi, prec="f64")
    got, _ = ctx.ray_grid(vol, z, s, phi, prec="f32")
    assert orc.rel_l2(got, ref) <= 1e-5 and orc.max_rel(got, ref) <= 1e-5
    rng = np.random.default_rng(10)
    iz, i_s, ip = rng.integers(0, 129, 200), rng.integers(0, 129, 200), rng.integers(0, 256, 200)
    want = orc.rays(vol, z[iz], s[i_s], phi[ip])
    assert orc.rel_l2(ref[iz, i_s, ip], want) <= 1e-10
    assert orc.max_rel(ref[iz, i_s, ip], want) <= 1e-10


def test_reference_config_reduction(ctx):
    # radon_reduction/binary/config.txt: 129 shifts, 128 phi, 128 theta, 129 slices
    nz = 129
    rs, rphi = wrt.ray_grid_init(129, 128)
    z = wrt.zslice_grid(nz)
    vol = wrt.phantom(128, False, 10, 6, 0.0)
    rays, _ = ctx.ray_grid(vol, z, rs, rphi, prec="f64")
    _, th, _, _ = ctx.radon_grid_init(128, 128, 129)
    R, _ = ctx.radon_reduction(rays, rs, rphi, th)
    sub = np.arange(0, 128, 16)  # 8 azimuths, every theta and shift
    want = orc.radon_reduction(np.ascontiguousarray(rays[:, :, sub]), rs, rphi[sub], th)
    assert orc.rel_l2(R[:, sub, :], want) <= 1e-10
    assert orc.max_rel(R[:, sub, :], want) <= 1e-10


def test_reference_config_adjoint(ctx):
    # radon_inversion/backprojection/C/binary/config.txt: 129 phi, 128 theta, 256 shifts,
    # 128^3 voxels
    n = 128
    phi, th, w, s = ctx.radon_grid_init(129, 128, 256)
    g = wrt.gaussian((len(s), len(phi), len(th)), seed=31)
    ref, _ = ctx.radon_adjoint(g, phi, th, w, s, n, prec="f64")
    got, _ = ctx.radon_adjoint(g, phi, th, w, s, n, prec="f32")
    assert orc.rel_l2(got, ref) <= 1e-5, orc.rel_l2(got, ref)
    xs = -1.0 + 1.0 / n + np.arange(n) * (2.0 / n)
    rng = np.random.default_rng(12)
    for iz, ix, iy in rng.integers(0, n, (12, 3)):
        want = orc.radon_adjoint_point(g, phi, th, w, s, xs[ix], xs[iy], xs[iz])
        assert abs(ref[iz, ix, iy] - want) <= 1e-10 * max(1.0, abs(want))
