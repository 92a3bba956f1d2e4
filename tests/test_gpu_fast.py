"""GPU tests of the fast plane-integral form (FP32 values / BF16 bricks, FP32 lane and wave
sums added into exact 64-bit fixed-point plane sums, the reference's index decisions: fixed
point with an exact binary64 path near cell boundaries) against the plain FP64 form and the
oracle; integer
volumes pin every index decision (test_fast_f32_index_decisions_exact), for every band shape
the autotuner can pick.

Tolerances: FP32 relative L2 <= 1e-5 and max |error| <= 1e-5 * max|R| (the target of the
form); BF16 bricks relative L2 <= 2e-3 (BF16 keeps 8 significand bits per voxel)."""
import numpy as np
import pytest

import orc
import wrt

pytestmark = pytest.mark.gpu


def _phantom_case(ctx, n, nphi, nth, ns, seed=0, sigma=0.0):
    vol = wrt.phantom(n, True, 10, seed, sigma)
    phi, th, w, s = ctx.radon_grid_init(nphi, nth, ns)
    return vol, phi, th, s


@pytest.mark.parametrize("n", [64, 48, 33, 96])
def test_fast_f32_matches_plain(ctx, n):
    vol, phi, th, s = _phantom_case(ctx, n, 32, 16, n + 1, seed=n, sigma=0.01)
    ref, _ = ctx.radon_grid(vol, phi, th, s, prec="f64")
    got, st = ctx.radon_grid(vol, phi, th, s, prec="f32")
    assert orc.rel_l2(got, ref) <= 1e-5, orc.rel_l2(got, ref)
    assert orc.max_rel(got, ref) <= 1e-5, orc.max_rel(got, ref)
    assert st["logical_samples"] == 16.0 * n * n * ref.size


@pytest.mark.parametrize("n", [300, 520])
def test_fast_f32_large_odd_edges(ctx, n):
    # edges past the benchmark's 256 that are not multiples of the 32^3 brick (partial face
    # bricks on three faces), on a coarse direction grid
    vol, phi, th, s = _phantom_case(ctx, n, 8, 4, 33, seed=n, sigma=0.01)
    ref, _ = ctx.radon_grid(vol, phi, th, s, prec="f64")
    got, _ = ctx.radon_grid(vol, phi, th, s, prec="f32")
    assert orc.rel_l2(got, ref) <= 1e-5, orc.rel_l2(got, ref)
    assert orc.max_rel(got, ref) <= 1e-5, orc.max_rel(got, ref)


def test_fast_f32_c0_full_grid(ctx):
    vol, phi, th, s = _phantom_case(ctx, 64, 64, 32, 129)
    ref, _ = ctx.radon_grid(vol, phi, th, s, prec="f64")
    got, _ = ctx.radon_grid(vol, phi, th, s, prec="f32")
    assert orc.rel_l2(got, ref) <= 1e-5
    # structured directions (phi multiple of pi/2) where index decisions hinge on rounding
    for ip in (0, 16, 32, 48):
        assert orc.rel_l2(got[:, ip, :], ref[:, ip, :]) <= 1e-5
    rng = np.random.default_rng(3)
    idx = rng.choice(ref.size, 64, replace=False)
    iS, iP, iT = np.unravel_index(idx, ref.shape)
    want = orc.radon_planes(vol, phi[iP], th[iT], s[iS])
    assert orc.rel_l2(got.ravel()[idx], want) <= 1e-5


def test_fast_f32_random_volume_and_odd_grids(ctx):
    # a volume with no zero bricks and values of both signs everywhere in the cube
    rng = np.random.default_rng(5)
    n = 40
    vol = rng.standard_normal((n, n, n))
    phi = rng.uniform(0, 2 * np.pi, 7)
    th = np.arccos(rng.uniform(-1, 1, 5))
    s = np.sort(rng.uniform(-1, 1, 11))
    ref, _ = ctx.radon_grid(vol, phi, th, s, prec="f64")
    got, _ = ctx.radon_grid(vol, phi, th, s, prec="f32")
    assert orc.rel_l2(got, ref) <= 1e-5


def test_fast_f32_planes_past_the_support(ctx):
    # Non-zero values only within radius 0.4 of the centre (support radius rho about 0.42):
    # each direction's bisected shift range is clipped to |s| <= rho, so the planes farther
    # out are never walked and must come back exactly zero, as the plain form's sums of
    # zero cells are; the planes inside must match the plain form.
    n = 48
    c = -1.0 + (2.0 * np.arange(n) + 1.0) / n
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    rng = np.random.default_rng(17)
    vol = np.where(X * X + Y * Y + Z * Z <= 0.16, 1.0 + rng.standard_normal((n, n, n)), 0.0)
    phi, th, w, s = ctx.radon_grid_init(16, 8, n + 1)
    ref, _ = ctx.radon_grid(vol, phi, th, s, prec="f64")
    got, _ = ctx.radon_grid(vol, phi, th, s, prec="f32")
    far = np.abs(s) > 0.5
    assert far.any() and (~far).any()
    assert np.all(ref[far] == 0.0) and np.all(got[far] == 0.0)
    assert orc.rel_l2(got, ref) <= 1e-5, orc.rel_l2(got, ref)
    assert orc.max_rel(got, ref) <= 1e-5, orc.max_rel(got, ref)


def test_fast_bf16_bricks(ctx):
    vol, phi, th, s = _phantom_case(ctx, 64, 32, 16, 65, seed=2, sigma=0.01)
    ref, _ = ctx.radon_grid(vol, phi, th, s, prec="f64")
    got, _ = ctx.radon_grid(vol, phi, th, s, prec="bf16")
    err = orc.rel_l2(got, ref)
    assert err <= 2e-3, err


def test_fast_job_repeatable(ctx):
    vol, phi, th, s = _phantom_case(ctx, 64, 16, 8, 65, seed=4)
    j = ctx.job("radon", "f32", n=64, phi=phi, theta=th, shift=s)
    j.load(vol)
    j.run()
    a = j.fetch()
    j.run()
    b = j.fetch()
    j.close()
    # the same bits: plane sums are exact 64-bit fixed-point sums, whatever the order in
    # which the concurrent interior and face launches add their sections
    assert np.array_equal(a, b)


def _adjoint_case(ctx, n, nphi, nth, ns, seed):
    vol = wrt.phantom(n, True, 10, seed, 0.01)
    phi, th, w, s = ctx.radon_grid_init(nphi, nth, ns)
    radon, _ = ctx.radon_grid(vol, phi, th, s, prec="f64")
    return radon, phi, th, w, s


@pytest.mark.parametrize("n", [32, 48, 64])
def test_fast_adjoint_f32_matches_plain(ctx, n):
    radon, phi, th, w, s = _adjoint_case(ctx, n, 32, 16, n + 1, seed=n)
    ref, _ = ctx.radon_adjoint(radon, phi, th, w, s, n, prec="f64")
    got, st = ctx.radon_adjoint(radon, phi, th, w, s, n, prec="f32")
    assert orc.rel_l2(got, ref) <= 1e-5, orc.rel_l2(got, ref)
    assert np.max(np.abs(got - ref)) <= 1e-5 * np.max(np.abs(ref))
    assert st["logical_samples"] == float(len(phi) * len(th) * n ** 3)


def test_fast_adjoint_f32_oracle_points(ctx):
    # noisy rows (every interval's cubic matters) against the oracle at a voxel sample
    n = 24
    phi, th, w, s = ctx.radon_grid_init(12, 8, 49)
    radon = wrt.gaussian((len(s), len(phi), len(th)), seed=11)
    got, _ = ctx.radon_adjoint(radon, phi, th, w, s, n, prec="f32")
    want = orc.radon_adjoint(radon, phi, th, w, s, n)
    assert orc.rel_l2(got, want) <= 1e-5, orc.rel_l2(got, want)


def test_fast_adjoint_job_reuses_rows(ctx):
    radon, phi, th, w, s = _adjoint_case(ctx, 32, 16, 8, 33, seed=9)
    j = ctx.job("adjoint", "f32", n=32, phi=phi, theta=th, theta_w=w, shift=s)
    j.load(radon)
    j.run()
    a = j.fetch()
    j.load(2.0 * radon)
    j.run()
    b = j.fetch()
    j.close()
    assert orc.rel_l2(b, 2.0 * a) <= 1e-6


@pytest.mark.parametrize("n,nphi,nth", [(128, 32, 16), (64, 128, 64)])
def test_fast_adjoint_f32_cut_decisions(ctx, n, nphi, nth):
    # rows that do not vanish at s = +-1: the |s| <= 1 cut must follow the FP64 decision
    phi, th, w, s = ctx.radon_grid_init(nphi, nth, n + 1)
    radon = wrt.gaussian((len(s), nphi, nth), seed=1)
    ref, _ = ctx.radon_adjoint(radon, phi, th, w, s, n, prec="f64")
    got, _ = ctx.radon_adjoint(radon, phi, th, w, s, n, prec="f32")
    assert orc.rel_l2(got, ref) <= 1e-5, orc.rel_l2(got, ref)
    assert np.max(np.abs(got - ref)) <= 1e-4 * np.max(np.abs(ref)), np.max(np.abs(got - ref))


@pytest.mark.parametrize("ns", [65, 2049, 2050])
def test_fast_adjoint_f32_row_lengths(ctx, ns):
    # The per-job launch stages chunks of the directions' cubic rows in LDS, a multiple of 4
    # directions whose rows fit 8192 float4: 65 shifts give chunks of 128 directions (35
    # directions: one partial chunk, 3 single directions past the last four), 2049 shifts
    # chunks of exactly 4 (9 chunks), and 2050 rows too long for LDS (the gathers read the
    # table in memory, as the persistent batch kernel's always do).  Noisy rows against the
    # batch kernel, bit for bit; smooth rows against the plain FP64 kernel (noisy rows on a
    # fine grid have slopes of order 1/ds, where the FP32 s alone errs by 1e-4 relative).
    n = 20
    phi, th, w, s = ctx.radon_grid_init(7, 5, ns)
    j = ctx.job("adjoint", "f32", n=n, phi=phi, theta=th, theta_w=w, shift=s)
    j.load(wrt.gaussian((len(s), len(phi), len(th)), seed=ns))
    j.run()
    alone = j.fetch()
    ctx.batch_run([j])
    assert np.array_equal(j.fetch(), alone)
    j.close()
    radon, _ = ctx.radon_grid(wrt.phantom(n, True, 10, ns, 0.0), phi, th, s, prec="f64")
    ref, _ = ctx.radon_adjoint(radon, phi, th, w, s, n, prec="f64")
    got, _ = ctx.radon_adjoint(radon, phi, th, w, s, n, prec="f32")
    assert orc.rel_l2(got, ref) <= 1e-5, orc.rel_l2(got, ref)
    assert np.max(np.abs(got - ref)) <= 1e-5 * np.max(np.abs(ref)), np.max(np.abs(got - ref))


@pytest.mark.parametrize("n", [32, 40, 64])
def test_fast_f32_structured_directions_on_faces(ctx, n):
    # Non-zero values up to the cube's faces, axis-aligned and diagonal directions, and
    # shifts on cell faces and on the cube's faces: the index and cube decisions that hinge
    # on the last bit must follow the plain form's (a flipped decision is an O(1) error).
    rng = np.random.default_rng(n)
    vol = rng.standard_normal((n, n, n))
    phi = np.array([0.0, np.pi / 4, np.pi / 2, np.pi, 1.5 * np.pi, 0.3])
    _, th_gl, _, _ = ctx.radon_grid_init(4, 6, 3)
    th = np.concatenate([th_gl, [np.pi / 2, np.pi / 4, 0.0]])
    s = np.unique(np.concatenate([-1.0 + np.arange(0, n + 1, 3) * (2.0 / n), [-1.0, 1.0, 0.0]]))
    ref, _ = ctx.radon_grid(vol, phi, th, s, prec="f64")
    got, _ = ctx.radon_grid(vol, phi, th, s, prec="f32")
    scale = np.max(np.abs(ref))
    assert orc.rel_l2(got, ref) <= 1e-5, orc.rel_l2(got, ref)
    assert np.max(np.abs(got - ref)) <= 1e-5 * scale, np.max(np.abs(got - ref)) / scale


@pytest.mark.parametrize("n", [33, 64, 128, 200])
def test_fast_ray_f32_matches_plain(ctx, n):
    # node-centred phantom (n = 33: power-of-two spacing; 200: the plane does not fit LDS),
    # phi grid with 0 and pi/2 (the trilinear index quirk's rays), slices outside [-1, 1]
    vol = wrt.phantom(n, False, 10, n, 0.01)
    s, phi = wrt.ray_grid_init(n + 1, 16)
    z = np.concatenate([wrt.zslice_grid(24), [1.5, -1.0000001]])
    ref, _ = ctx.ray_grid(vol, z, s, phi, prec="f64")
    got, st = ctx.ray_grid(vol, z, s, phi, prec="f32")
    assert orc.rel_l2(got, ref) <= 1e-5, orc.rel_l2(got, ref)
    assert np.max(np.abs(got - ref)) <= 1e-5 * np.max(np.abs(ref))
    assert np.all(got[-2:] == 0.0)
    assert st["logical_samples"] == (4.0 * n + 1) * got.size


def test_fast_ray_f32_above_fixed_point_edge(ctx):
    # n = 520 is past the edge (500) below which ray points step in 32.32 fixed point; here
    # every point takes the binary64 form of the reference's index arithmetic
    n = 520
    vol = wrt.phantom(n, False, 10, 7, 0.01)
    s, phi = wrt.ray_grid_init(33, 8)
    phi = np.concatenate([phi, [0.3, 2.0]])
    z = np.array([-0.51, 0.0, 0.37])
    ref, _ = ctx.ray_grid(vol, z, s, phi, prec="f64")
    got, _ = ctx.ray_grid(vol, z, s, phi, prec="f32")
    assert orc.rel_l2(got, ref) <= 1e-5, orc.rel_l2(got, ref)
    assert np.max(np.abs(got - ref)) <= 1e-5 * np.max(np.abs(ref))


def test_fast_ray_f32_random_volume_and_angles(ctx):
    rng = np.random.default_rng(9)
    n = 48
    vol = rng.standard_normal((n, n, n))
    phi = np.concatenate([rng.uniform(0, np.pi, 9), [0.0, np.pi / 2, np.pi / 4]])
    s = np.sort(rng.uniform(-1.2, 1.2, 13))
    z = rng.uniform(-1, 1, 7)
    ref, _ = ctx.ray_grid(vol, z, s, phi, prec="f64")
    got, _ = ctx.ray_grid(vol, z, s, phi, prec="f32")
    assert orc.rel_l2(got, ref) <= 1e-5, orc.rel_l2(got, ref)
    assert np.max(np.abs(got - ref)) <= 1e-5 * np.max(np.abs(ref))


@pytest.mark.parametrize("n", [17, 40])
def test_fast_ray_f32_rays_on_cube_faces(ctx, n):
    # Regression: rays lying on a face of the cube (s = +-1 along phi = k pi/2) have a
    # direction component of ~1e-17, and x = s + sx*i rounds to exactly +-1 for dozens of
    # points beyond the continuous bound; all of them pass the reference's cube test.  A
    # volume that is non-zero on the faces makes every such point count.
    rng = np.random.default_rng(n)
    vol = rng.standard_normal((n, n, n))
    s, phi = wrt.ray_grid_init(n + 1, 16)  # s from -1 to 1, phi multiples of pi/8
    z = wrt.zslice_grid(7)
    ref, _ = ctx.ray_grid(vol, z, s, phi, prec="f64")
    got, _ = ctx.ray_grid(vol, z, s, phi, prec="f32")
    assert orc.rel_l2(got, ref) <= 1e-5, orc.rel_l2(got, ref)
    assert np.max(np.abs(got - ref)) <= 1e-5 * np.max(np.abs(ref))
    # the face rays themselves
    for js in (0, len(s) - 1):
        assert orc.rel_l2(got[:, js, :], ref[:, js, :]) <= 1e-5


@pytest.mark.parametrize("n", [32, 33, 40, 48, 64])
@pytest.mark.parametrize("support", ["cube", "ball"])
def test_fast_f32_index_decisions_exact(ctx, n, support):
    _exact_case(ctx, n, support)


@pytest.mark.parametrize("rows", [4, 8, 16])
@pytest.mark.parametrize("n", [33, 64])
def test_fast_f32_band_shapes_exact(ctx, n, rows, monkeypatch):
    # every band shape the autotuner may pick (rows x columns 4 x 16, 8 x 8, 16 x 4) walks
    # exactly the reference's samples
    monkeypatch.setenv("WRT_BAND_ROWS", str(rows))
    _exact_case(ctx, n, "ball")
    _exact_case(ctx, n, "cube")


def test_fast_f32_autotune_caches_per_shape(ctx, tmp_path, monkeypatch):
    # the first launch of a shape times the band shapes and direction groups and records the
    # winner; a second load of the same shape finds it (no new line) and gives the same sums
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("WRT_TUNE_CACHE", str(cache))
    monkeypatch.delenv("WRT_AUTOTUNE", raising=False)
    vol, phi, th, s = _phantom_case(ctx, 96, 64, 32, 97, seed=5, sigma=0.01)
    ref, _ = ctx.radon_grid(vol, phi, th, s, prec="f64")
    got1, _ = ctx.radon_grid(vol, phi, th, s, prec="f32")
    lines = cache.read_text().splitlines()
    assert len(lines) == 1
    _, rows, group, sec = lines[0].split()
    assert int(rows) in (4, 8, 16) and int(group) >= 16 and float(sec) > 0
    got2, _ = ctx.radon_grid(vol, phi, th, s, prec="f32")
    assert cache.read_text().splitlines() == lines
    for got in (got1, got2):
        assert orc.rel_l2(got, ref) <= 1e-5, orc.rel_l2(got, ref)


def _exact_case(ctx, n, support):
    # Small integer values make every sum exact (FP32 lane and wave sums, FP64 totals), so the
    # fast form must give the plain form's integer sum on every plane: a single sample put
    # into another cell changes it (unless the two cells hold equal values).  Random and
    # structured directions, shifts on cell faces and on the cube's faces.
    rng = np.random.default_rng(100 + n)
    vol = rng.integers(-8, 9, size=(n, n, n)).astype(np.float64)
    if support == "ball":
        # zero outside radius 0.7: rows are clipped to the support sphere, whose edge then
        # cuts through non-zero cells on every plane that meets it
        c = -1.0 + (np.arange(n) + 0.5) * (2.0 / n)
        X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
        vol[X ** 2 + Y ** 2 + Z ** 2 > 0.49] = 0.0
    phi = np.concatenate([[0.0, np.pi / 4, np.pi / 2, np.pi, 1.5 * np.pi], rng.uniform(0, 2 * np.pi, 5)])
    _, th_gl, _, _ = ctx.radon_grid_init(4, 6, 3)
    th = np.concatenate([th_gl, [np.pi / 2, np.pi / 4, 0.0], np.arccos(rng.uniform(-1, 1, 3))])
    s = np.unique(np.concatenate([-1.0 + np.arange(0, n + 1, 2) * (2.0 / n), [-1.0, 1.0, 0.0],
                                  rng.uniform(-1, 1, 9)]))
    ref, _ = ctx.radon_grid(vol, phi, th, s, prec="f64")
    got, _ = ctx.radon_grid(vol, phi, th, s, prec="f32")
    d2 = (1.0 / n) ** 2
    assert np.array_equal(np.rint(got / d2), np.rint(ref / d2))
    assert np.max(np.abs(ref / d2 - np.rint(ref / d2))) < 1e-6


@pytest.mark.parametrize("bad", [np.nan, np.inf, -np.inf, 1e39])
def test_fast_forms_refuse_values_not_finite_in_fp32(ctx, bad):
    # the fixed-point plane sums need a finite largest |value|: the FP32 and BF16 forms and
    # the entry point that allocates nothing return WRT_E_ARG (1e39 is finite in binary64
    # but not in FP32); the plain FP64 form keeps the reference's arithmetic
    vol, phi, th, s = _phantom_case(ctx, 40, 8, 4, 17, seed=6)
    vol[20, 21, 22] = bad
    for prec in ("f32", "bf16"):
        with pytest.raises(wrt.WrtError) as e:
            ctx.radon_grid(vol, phi, th, s, prec=prec)
        assert e.value.code == wrt.WRT_E_ARG, prec
    with pytest.raises(wrt.WrtError) as e:
        ctx.radon_grid_noalloc(vol, phi, th, s)
    assert e.value.code == wrt.WRT_E_ARG
    # the context still works afterwards
    vol[20, 21, 22] = 0.5
    got, _ = ctx.radon_grid(vol, phi, th, s, prec="f32")
    ref, _ = ctx.radon_grid(vol, phi, th, s, prec="f64")
    assert orc.rel_l2(got, ref) <= 1e-5
