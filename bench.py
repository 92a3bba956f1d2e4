"""Benchmark of the plane-integral hot path (radon_grid) on one MI355X.

Workload (the metric's configuration): a 256^3 seeded ellipsoid phantom with N(0, 0.01^2)
texture, FP32 bricks, Gauss-uniform plane grid 256 phi x 128 Gauss-Legendre theta x 257
shifts -- 8,421,376 planes, 16 n^2 = 1,048,576 lattice samples each (8.83e12 plane samples
per step).  A step is one pass of the FP32 form over that grid with the volume resident in
HBM.  Printed: one JSON line with plane samples/s, the dominant kernel's roofline (VALU
instruction issue: SQ_INSTS_VALU per pass from the committed counter summary over the pass
time measured here; the measured HBM traffic and a byte model beside it), the FP32 form's relative L2 and
max error against the plain FP64 form, the plain FP64 form's rate on a subset of the same
grid, and a CPU baseline (the reference's own plane_integral, compiled from its sources, on
the host cores over a bounded sample of planes; the oracle's restatement when that build is
absent).

The "serve" object is the served form (C request loop): p50/p99 latency of a seeded
Poisson stream (8 requests/s, 64 requests, mix 50% plane integrals FP32 / 30% backprojection
FP32 / 20% reduction FP64, edges {64,96,128,192,256}), and the throughput of the first 32
requests of that stream arriving at once, served (overlapping, resident data) and in the
plain per-request FP64 form (one at a time).

The "batch" object is the batched form: 48 seeded volumes with edges drawn from
{32,48,64,96,128,160,192}, each with a forward pass (FP32 bricks, (e/2) phi x (e/4) GL theta
x (e+1) shifts) and a backprojection (FP32) of a seeded Gaussian sinogram on the same grid,
all in one persistent graph-captured launch (direction groups autotuned on the first call);
the replay's time is compared with the same 96 jobs launched one by one, with all 96 launched
at once on their own streams, and with their plain FP64 forms.  Its "small_volumes" object
is the same for 96 volumes with edges from {16,24,32,48,64}, where per-job launches are
bound by launch and tail latency.

The "bf16" object is the same pass from BF16 bricks (time, rate, its byte model, relL2
against the plain FP64 subset): an accuracy-reduced storage form, not a faster one, since the
pass is bound by instruction issue and not by the bricks' bytes.  The "transpose" object is the FP32 exact transpose of the plane integrals on the same grid
(rate, error against the plain FP64 transpose on a phi subset, and the dot-product test of
the FP32 pair at full size); "rays" is the line-integral workload at 128^3 (128 z x 180 phi
x 257 shifts: FP32 against FP64, the FP32 transpose with its dot-product test, and the FP64
ray -> Radon reduction chain); "analytic" is radon_analytic on its own example configuration
(256 phi x 128 theta x 129 shifts, ngrid 129, the exponential-bump example function) beside
the reference's compiled plane_integral on the host cores; "laplace" is the spectral
Laplacian of a 256^3 volume; "adjoint" is the backprojection at the reference's default
configuration (129 phi x 128 theta x 256 s -> 128^3) in FP32 and FP64, beside the oracle's C
restatement on the host cores; "fourier" is the Chang-type Fourier-slice inversion (64 phi x
32 theta x 129 shifts -> 64^3, FP64) with its rate against the FP64 peak; "programs" runs the radon_grid drop-in program end to end at
256^3 with text files and with the binary format.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--band-rows R --dgroup G]
    python bench.py --traffic-probe --band-rows R --dgroup G
                                        (run under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE)

The brick kernel's tile shape (rows per band, directions per task) is autotuned by the
library at the job's first launch, which the bench runs before its warmup; the bench line
reports it in config.band_rows / config.dgroup, and the profiling runs fix it to those values
so that their counters hold no tuning pass.

The roofline's "traffic" (HBM bytes per pass) comes from those two counter runs: the probe
runs a plain conversion kernel of known size (reads 8 B, writes 4 B per element) and then
one pass of the workload; each counter is scaled by what it reads for the conversion
kernel.  The committed summary profiles/radon_fast_256_traffic.md holds the result, which
the bench reports.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "wrt-project_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK = 8.0e12  # MI355X HBM3E, B/s
FP64_PEAK = 78.6e12  # MI355X dense FP64, flop/s
# VALU issue: 256 CUs x 4 SIMDs x 2.4 GHz, one wave64 instruction per SIMD per 4 cycles
VALU_PEAK = 256 * 4 * 2.4e9 / 4


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


REF_RADON_LIB = os.path.join(ROOT, "oracle", "_ref", "libref_radon_grid.so")
REF_RAY_LIB = os.path.join(ROOT, "oracle", "_ref", "libref_ray_grid.so")
# the same sources at -O2 (oracle/Makefile): the stronger CPU figure
REF_RADON_LIB_O2 = os.path.join(ROOT, "oracle", "_ref", "libref_radon_grid_O2.so")
REF_RAY_LIB_O2 = os.path.join(ROOT, "oracle", "_ref", "libref_ray_grid_O2.so")


def _seeded_planes(rng, phi, th, s, m):
    return phi[rng.integers(len(phi), size=m)], th[rng.integers(len(th), size=m)], \
        s[rng.integers(len(s), size=m)]


def _ref_batch(path, fn, vol, threads, seconds, draw):
    """Time the reference's own per-element function `fn` (plane_integral or ray_transform,
    compiled from its sources by oracle/Makefile) through the OpenMP batch driver linked
    beside it (oracle/ref_batch.c), one thread per core as the reference's OpenMP drivers run
    it (radon_grid/src/radon_grid.c:229-264, ray_grid/src/ray_grid.c:224-262), over seeded
    elements draw(m) -> (a, b, c), growing m until the sample takes about `seconds`.
    Returns (m, seconds taken)."""
    import ctypes as C

    lib = C.CDLL(path)
    f = getattr(lib, fn)
    f.restype = None
    f.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_long, C.c_int,
                  C.c_void_p]
    n = vol.shape[0]
    v = np.ascontiguousarray(vol, dtype=np.float64)
    rows, top = _jagged(v)
    m = 4 * threads
    while True:
        a, b, c = (np.ascontiguousarray(x, dtype=np.float64) for x in draw(m))
        out = np.empty(m)
        t0 = time.perf_counter()
        f(top, n, a.ctypes.data, b.ctypes.data, c.ctypes.data, m, threads, out.ctypes.data)
        dt = time.perf_counter() - t0
        if dt >= 0.8 * seconds or m >= 1 << 24:
            return m, dt
        m = max(m + threads, int(m * min(16.0, 1.1 * seconds / max(dt, 1e-3))))


def cpu_baseline(vol, phi, th, s, seconds=10.0):
    """CPU rate of the plain plane integrals on the host cores, on a bounded seeded sample of
    the grid's planes: the reference's own plane_integral built from its sources at -O2
    (value) and, beside it, at -O0 as its Makefile builds it, when oracle/_ref is built;
    else the oracle's restatement (plain C, -O2, OpenMP)."""
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    rng = np.random.default_rng(1234)
    n = vol.shape[0]
    if os.path.exists(REF_RADON_LIB_O2) and os.path.exists(REF_RADON_LIB):
        def draw(m):
            return _seeded_planes(rng, phi, th, s, m)

        m, dt = _ref_batch(REF_RADON_LIB_O2, "ref_plane_batch", vol, threads, seconds, draw)
        m0, dt0 = _ref_batch(REF_RADON_LIB, "ref_plane_batch", vol, threads, 0.6 * seconds, draw)
        return {"value": m * 16.0 * n * n / dt, "unit": "plane samples/s", "cores": threads,
                "kind": "reference",
                "sample": f"{m} seeded uniform planes of the {n}^3 grid, {dt:.1f} s; the "
                          "reference's plane_integral built from its sources at -O2, "
                          "called per plane from an OpenMP loop, one thread per core",
                "reference_O0": {"value": m0 * 16.0 * n * n / dt0, "unit": "plane samples/s",
                                 "sample": f"{m0} seeded planes, {dt0:.1f} s; the same sources "
                                           "at -O0 (its Makefile's flags)"}}
    import orc

    m = max(threads, 8)
    while True:
        p, t, sh = _seeded_planes(rng, phi, th, s, m)
        t0 = time.perf_counter()
        orc.radon_planes(vol, p, t, sh)
        dt = time.perf_counter() - t0
        if dt >= seconds or m >= 1 << 22:
            break
        m = int(m * min(16.0, max(1.2, 1.1 * seconds / max(dt, 1e-3))))
    return {"value": m * 16.0 * n * n / dt, "unit": "plane samples/s", "cores": threads,
            "kind": "port",
            "sample": f"{m} seeded uniform planes of the {n}^3 grid, {dt:.1f} s; the oracle's "
                      "C restatement at -O2 with OpenMP"}


def _jagged(v):
    """the reference's double*** view of a contiguous n^3 buffer (kept alive by the caller)"""
    import ctypes as C

    n = v.shape[0]
    rows = (C.c_void_p * (n * n))(*[v.ctypes.data + 8 * n * r for r in range(n * n)])
    top = (C.c_void_p * n)(*[C.addressof(rows) + C.sizeof(C.c_void_p) * n * x for x in range(n)])
    return rows, top


def rays_cpu(vol, z, s, phi, rays_in, rs, rphi, th, seconds):
    """CPU rates beside the rays leg on bounded seeded samples: the reference's own
    ray_transform (ray_grid/src/integral.c:11-50, built from its sources at -O2 and at -O0 by
    oracle/Makefile) over seeded rays of the grid, one thread per core as its OpenMP driver
    runs it (ray_grid/src/ray_grid.c:242-255); and the oracle's reduction (its spline rows
    solved once, OpenMP over planes) on a subset of theta of the reduction chain."""
    import orc

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    rng = np.random.default_rng(4321)
    n = vol.shape[0]
    out = {}
    for path, flags, key in ((REF_RAY_LIB_O2, "-O2", "ray_grid"), (REF_RAY_LIB, "-O0", "ray_grid_O0")):
        if not os.path.exists(path):
            continue

        def draw(m):
            return (a[rng.integers(len(a), size=m)] for a in (z, s, phi))

        m, dt = _ref_batch(path, "ref_ray_batch", vol, threads, seconds, draw)
        out[key] = {"value": m * (4.0 * n + 1.0) / dt, "unit": "ray samples/s",
                    "cores": threads, "kind": "reference",
                    "sample": f"{m} seeded rays of the {n}^3 grid, {dt:.1f} s; the "
                              f"reference's ray_transform built from its sources at {flags}, "
                              "called per ray from an OpenMP loop"}
    k = 1
    while True:
        sub = np.ascontiguousarray(th[:: max(1, len(th) // k)][:k])
        t0 = time.perf_counter()
        orc.radon_reduction(rays_in, rs, rphi, sub)
        dt = time.perf_counter() - t0
        if dt >= 0.5 * seconds or k >= len(th):
            break
        k = min(len(th), max(k + 1, int(k * 0.9 * seconds / max(dt, 1e-3))))
    nz = rays_in.shape[0]
    out["reduction"] = {"value": (4.0 * nz + 1.0) * len(rs) * len(rphi) * k / dt,
                        "unit": "tau-steps/s", "cores": threads, "kind": "port",
                        "sample": f"{k} of {len(th)} theta of the reduction chain, {dt:.1f} s; "
                                  "the oracle's C restatement at -O2 with OpenMP (spline rows "
                                  "solved once, where the reference solves 2-3 per tau-step)"}
    return out


def batch_leg(ctx, wrt, nvol, seed=3, pool=(32, 48, 64, 96, 128, 160, 192)):
    rng = np.random.default_rng(seed)
    edges = [int(e) for e in rng.choice(list(pool), size=nvol)]
    specs = []
    for i, e in enumerate(edges):
        phi, th, w, s = ctx.radon_grid_init(e // 2, e // 4, e + 1)
        specs.append(("radon", e, dict(n=e, phi=phi, theta=th, shift=s),
                      lambda e=e, i=i: wrt.phantom(e, True, 10, 1000 + i, 0.0)))
        specs.append(("adjoint", e, dict(n=e, phi=phi, theta=th, theta_w=w, shift=s),
                      lambda e=e, i=i, n=len(s) * len(phi) * len(th): wrt.gaussian(n, 2000 + i)))
    jobs = []
    for op, e, kw, gen in specs:
        j = ctx.job(op, "f32", **kw)
        j.load(gen())
        jobs.append(j)
    st0 = ctx.batch_run(jobs)  # plan, autotune, capture
    st = ctx.batch_run(jobs)   # graph replay
    dgroup = ctx.batch_dgroup()
    for j in jobs:             # the per-job launches' own tuning runs outside the timing
        j.run()
    # one job at a time (each on the GPU alone), and every job at once on its own stream
    per_job = sum(j.run()["seconds"] for j in jobs)
    t0 = time.perf_counter()
    for j in jobs:
        j.launch()
    for j in jobs:
        j.wait()
    concurrent = time.perf_counter() - t0
    for j in jobs:
        j.close()
    plain = 0.0
    for op, e, kw, gen in specs:
        j = ctx.job(op, "f64", **kw)
        j.load(gen())
        plain += j.run()["seconds"]
        j.close()
    return {"volumes": nvol, "jobs": len(jobs), "edges": edges, "dgroup": dgroup,
            "first_call_s": st0["kernel_seconds"], "graph_s": st["kernel_seconds"],
            "samples": st["logical_samples"],
            "samples_per_s": st["logical_samples"] / st["kernel_seconds"],
            "per_job_fast_s": per_job, "per_job_concurrent_s": concurrent,
            "plain_f64_s": plain,
            "speedup_vs_per_job": per_job / st["kernel_seconds"],
            "speedup_vs_concurrent": concurrent / st["kernel_seconds"],
            "speedup_vs_plain": plain / st["kernel_seconds"]}


def transpose_leg(ctx, wrt, vol, fast, phi, th, s, stride):
    """The FP32 exact transpose of the plane integrals at the same grid: its rate, its error
    against the plain FP64 transpose on every stride-th phi, and the dot-product test
    <A f, g> = <f, A^T g> of the FP32 pair at full size (f = the phantom, A f = the FP32
    pass above, g = seeded Gaussian plane values)."""
    n = vol.shape[0]
    g = wrt.gaussian((len(s), len(phi), len(th)), seed=11)
    j = ctx.job("radon_T", "f32", n=n, phi=phi, theta=th, shift=s)
    j.load(g)
    j.run()
    st = j.run()
    ATg = j.fetch().reshape(n, n, n)
    j.close()
    lhs = float(np.dot(fast.ravel().astype(np.float64), g.ravel()))
    rhs = float(np.dot(vol.ravel(), ATg.ravel()))
    dot_rel = abs(lhs - rhs) / (np.linalg.norm(fast) * np.linalg.norm(g))
    sub = phi[::stride]
    gs = np.ascontiguousarray(g[:, ::stride, :])
    out = {}
    for prec in ("f32", "f64"):
        js = ctx.job("radon_T", prec, n=n, phi=sub, theta=th, shift=s)
        js.load(gs)
        sst = js.run()
        out[prec] = (js.fetch(), sst["kernel_seconds"])
        js.close()
    return {"seconds": st["kernel_seconds"],
            "samples_per_s": st["logical_samples"] / st["kernel_seconds"],
            "dot_lhs": lhs, "dot_rhs": rhs, "dot_rel": dot_rel,
            "subset": f"every {stride}th phi",
            "rel_l2_vs_f64": rel_l2(out["f32"][0], out["f64"][0]),
            "plain_f64_subset_s": out["f64"][1], "f32_subset_s": out["f32"][1]}


def rays_leg(ctx, wrt, cpu_seconds=4.0):
    """The line-integral workload at 128^3: 128 z-slices x 180 phi in [0, pi) x 257 shifts,
    FP32 against the plain FP64 form; the FP32 exact transpose and the dot-product test of
    the FP32 pair on seeded Gaussian f and g; the FP64 reduction chain (rays on 128
    slices x 64 phi in [0, 2 pi) x 129 shifts -> Radon 64 phi x 32 theta x 129 shifts); and
    the same Radon grid reduced from the FP64 throughput rays themselves."""
    n = 128
    z = wrt.zslice_grid(128)
    s, phi2 = wrt.ray_grid_init(257, 360)
    phi = np.ascontiguousarray(phi2[:180])
    vol = wrt.phantom(n, False, 10, 0, 0.0)
    res = {}
    for prec in ("f32", "f64"):
        j = ctx.job("ray", prec, n=n, z=z, shift=s, phi=phi)
        j.load(vol)
        j.run()
        st = j.run()
        res[prec] = (j.fetch(), st)
        j.close()
    f = wrt.gaussian((n, n, n), seed=21)
    g = wrt.gaussian((len(z), len(s), len(phi)), seed=22)
    j = ctx.job("ray", "f32", n=n, z=z, shift=s, phi=phi)
    j.load(f)
    j.run()
    Af = j.fetch()
    j.close()
    j = ctx.job("ray_T", "f32", n=n, z=z, shift=s, phi=phi)
    j.load(g)
    j.run()
    tst = j.run()
    ATg = j.fetch()
    j.close()
    lhs, rhs = float(np.dot(Af, g.ravel())), float(np.dot(f.ravel(), ATg))
    # reduction chain (FP64 by design): rays on the Radon output's phi and shift grids
    rs, rphi = wrt.ray_grid_init(129, 64)
    rz = wrt.zslice_grid(128)
    rj = ctx.job("ray", "f64", n=n, z=rz, shift=rs, phi=rphi)
    rj.load(vol)
    rst = rj.run()
    rays = rj.fetch()
    rj.close()
    _, th, _, _ = ctx.radon_grid_init(64, 32, 129)
    dj = ctx.job("reduction", "f64", phi=rphi, theta=th, shift=rs, z=rz, nz=len(rz))
    dj.load(rays)
    dst = dj.run()
    red = dj.fetch()
    dj.close()
    # the same Radon grid straight from the throughput rays (180 angles in [0, pi) x 257
    # shifts): rows at each azimuth by P(z, s, phi + pi) = P(z, -s, phi) and cubic Lagrange
    # interpolation in phi, splines over the 257 ray shifts
    trays = res["f64"][0].reshape(len(z), len(s), len(phi))
    for _ in range(2):  # the second call is the one reported
        gred, gst = ctx.radon_reduction_rays(trays, s, True, rs, rphi, th)
    cpu = rays_cpu(vol, z, s, phi, rays.reshape(len(rz), len(rs), len(rphi)), rs, rphi, th,
                   cpu_seconds) if cpu_seconds > 0 else None
    f32, st32 = res["f32"]
    f64, st64 = res["f64"]
    return {"workload": "ray_grid 128^3, 128 z x 180 phi in [0,pi) x 257 shifts",
            "f32_s": st32["kernel_seconds"], "f64_s": st64["kernel_seconds"],
            "f32_samples_per_s": st32["logical_samples"] / st32["kernel_seconds"],
            "f64_samples_per_s": st64["logical_samples"] / st64["kernel_seconds"],
            "rel_l2_vs_f64": rel_l2(f32, f64),
            "max_err_vs_f64": float(np.max(np.abs(f32 - f64)) / np.max(np.abs(f64))),
            "transpose_f32_s": tst["kernel_seconds"],
            "dot_rel": abs(lhs - rhs) / (np.linalg.norm(Af) * np.linalg.norm(g)),
            "reduction": {"grid": "rays 128 z x 64 phi x 129 s -> 64 phi x 32 theta x 129 s",
                          "rays_f64_s": rst["kernel_seconds"], "reduction_f64_s": dst["kernel_seconds"],
                          "tau_steps_per_s": dst["logical_samples"] / dst["kernel_seconds"]},
            "reduction_from_throughput_rays": {
                "grid": "rays 128 z x 180 phi in [0,pi) x 257 s -> 64 phi x 32 theta x 129 s",
                "seconds": gst["seconds"], "reduction_f64_s": gst["kernel_seconds"],
                "rel_l2_vs_parity_chain": rel_l2(gred.ravel(), red)},
            "cpu": cpu}


def analytic_leg(ctx, wrt, seconds=6.0):
    """radon_analytic on the reference's example configuration (radon_analytic/binary/
    config.txt: 256 phi x 128 theta x 129 shifts, ngrid 129) with its exponential-bump example
    function, FP64 on the GPU; beside it the reference's own plane_integral compiled from its
    sources with that example function (oracle/_ref), on seeded planes over the host cores,
    whose values also give the GPU's error against the reference."""
    import ctypes as C
    from concurrent.futures import ThreadPoolExecutor

    g = 129
    phi, th, w, s = ctx.radon_grid_init(256, 128, 129)
    terms = wrt.fn_example("exponential_bump")
    ctx.radon_analytic(terms, g, phi, th, s)
    out, st = ctx.radon_analytic(terms, g, phi, th, s)
    res = {"workload": "radon_analytic, exponential bump, 256 phi x 128 theta x 129 s, ngrid 129",
           "seconds": st["kernel_seconds"],
           "samples_per_s": st["logical_samples"] / st["kernel_seconds"]}
    lib_path = os.path.join(ROOT, "oracle", "_ref", "libref_radon_analytic_exponential_bump.so")
    if not os.path.exists(lib_path):
        return res
    lib = C.CDLL(lib_path)
    lib.plane_integral.restype = C.c_double
    lib.plane_integral.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double]
    fptr = C.cast(lib.test_function, C.c_void_p)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    rng = np.random.default_rng(77)

    def run(idx):
        return [lib.plane_integral(fptr, g, float(phi[i]), float(th[j]), float(s[k])) for k, i, j in idx]

    m = 2 * threads
    while True:
        idx = list(zip(rng.integers(len(s), size=m), rng.integers(len(phi), size=m),
                       rng.integers(len(th), size=m)))
        parts = [idx[t::threads] for t in range(threads)]
        with ThreadPoolExecutor(threads) as ex:
            t0 = time.perf_counter()
            vals = list(ex.map(run, parts))
            dt = time.perf_counter() - t0
        if dt >= 0.8 * seconds or m >= 1 << 20:
            break
        m = max(m + threads, int(m * min(16.0, 1.1 * seconds / max(dt, 1e-3))) // threads * threads)
    ref = np.array([v for t in range(threads) for v in vals[t]])
    got = np.array([out[k, i, j] for t in range(threads) for (k, i, j) in parts[t]])
    res.update({"rel_l2_vs_reference": rel_l2(got, ref),
                "max_err_vs_reference": float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))})
    if seconds <= 0:  # the error check only: no CPU rate from a token sample
        return res
    cpu_rate = m * 16.0 * g * g / dt
    res.update({"cpu_baseline": {"value": cpu_rate, "unit": "plane samples/s", "cores": threads,
                                 "kind": "reference",
                                 "sample": f"{m} seeded planes, {dt:.1f} s; the reference's "
                                           "plane_integral with its example exponential_bump, "
                                           "built from its sources at -O0"},
                "speedup_vs_cpu": res["samples_per_s"] / cpu_rate})
    return res


def laplace_leg(ctx, wrt, n=256):
    """the spectral Laplacian of an n^3 backprojection-sized volume (FP64): time, and the
    HBM bytes its three passes must move (real volume in, spectrum out and in, volume out)"""
    u = wrt.gaussian((n, n, n), seed=31)
    ctx.laplace(u, 2.0)
    ts = [ctx.laplace(u, 2.0)[1]["kernel_seconds"] for _ in range(3)]
    t = float(np.median(ts))
    spec = n * n * (n // 2 + 1) * 16
    alg = 2 * n ** 3 * 8 + 2 * spec  # the two real-complex transforms, each read and written once
    return {"workload": f"spectral Laplacian {n}^3 FP64 (hipFFT D2Z, factor, Z2D)", "seconds": t,
            "alg_bytes": alg, "achieved_GBps": alg / t / 1e9, "frac_of_hbm": alg / t / HBM_PEAK}


def adjoint_leg(ctx, wrt, seconds=5.0):
    """The backprojection radon3d_adjoint at the reference's default configuration (129 phi x
    128 GL theta x 256 shifts -> 128^3 voxel centres; radon_inversion/backprojection/binary/
    config.txt) on the Radon data of a phantom: the FP32 form against the plain FP64 form
    (time, relative L2), in (voxel, direction) evaluations per second, and the oracle's C
    restatement (splines solved once per row, -O2, OpenMP) timed on the host cores on a
    smaller voxel grid of the same directions and shifts."""
    import orc

    n, nphi, ntheta, nshift = 128, 129, 128, 256
    phi, th, w, s = ctx.radon_grid_init(nphi, ntheta, nshift)
    radon, _ = ctx.radon_grid(wrt.phantom(n, True, 10, 5, 0.0), phi, th, s, prec="f32")
    evals = float(n) ** 3 * nphi * ntheta
    ctx.radon_adjoint(radon, phi, th, w, s, n, prec="f32")
    runs = [ctx.radon_adjoint(radon, phi, th, w, s, n, prec="f32") for _ in range(3)]
    t32 = float(np.median([r[1]["kernel_seconds"] for r in runs]))
    got = runs[-1][0]
    ref, st64 = ctx.radon_adjoint(radon, phi, th, w, s, n, prec="f64")
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    m = 8
    while True:
        t0 = time.perf_counter()
        orc.radon_adjoint(radon, phi, th, w, s, m)
        dt = time.perf_counter() - t0
        if dt >= 0.5 * seconds or m >= 128:
            break
        m = min(128, max(m + 4, int(m * min(2.0, (seconds / max(dt, 1e-3)) ** (1 / 3)))))
    cpu = float(m) ** 3 * nphi * ntheta / dt
    return {"workload": f"radon3d_adjoint, {nphi} phi x {ntheta} theta x {nshift} s -> {n}^3",
            "f32_s": t32, "evals_per_s": evals / t32, "f64_s": st64["kernel_seconds"],
            "f64_evals_per_s": evals / st64["kernel_seconds"],
            "speedup_vs_plain": st64["kernel_seconds"] / t32, "rel_l2_vs_f64": rel_l2(got, ref),
            "cpu_port": {"value": cpu, "unit": "(voxel, direction) evaluations/s", "cores": threads,
                         "kind": "port", "sample": f"{m}^3 voxel centres, same directions and "
                         f"shifts, {dt:.1f} s; the oracle's C restatement (-O2, OpenMP)"},
            "speedup_vs_cpu_port": evals / t32 / cpu}


def fourier_leg(ctx, wrt, nphi=64, ntheta=32, nshift=129, ngrid=64):
    """the Fourier-slice inversion (FP64) of the plane integrals of an off-centre Gaussian:
    time of the two kernels, the reference's adjoint-sum terms (voxel x direction x frequency)
    per second, and the FP64 rate of the exact sum as evaluated: 4 fused multiply-adds (8 flops)
    per voxel, direction and folded frequency (Horner's scheme)"""
    phi = (2 * np.pi / nphi) * np.arange(nphi)
    x, w = np.polynomial.legendre.leggauss(ntheta)
    theta = np.arccos(x)
    s = -1.0 + np.arange(nshift) * (2.0 / (nshift - 1))
    P, T = np.meshgrid(phi, theta, indexing="ij")
    c, s2 = (0.2, -0.12, 0.15), 0.04  # centre and sigma^2: below 1e-5 outside the unit ball
    cu = np.sin(T) * np.cos(P) * c[0] + np.sin(T) * np.sin(P) * c[1] + np.cos(T) * c[2]
    radon = np.pi * s2 * np.exp(-((s[:, None, None] - cu[None]) ** 2) / s2)
    out, st = ctx.fourier_inversion(radon, phi, theta, w, ngrid)
    ts = [ctx.fourier_inversion(radon, phi, theta, w, ngrid)[1]["kernel_seconds"] for _ in range(3)]
    t = float(np.median(ts))
    g = (np.arange(ngrid) - ngrid // 2) * (2.0 / ngrid)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    want = np.exp(-((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2) / s2)
    ntotal = 4 * ((nshift - 1) // 2)  # padding 4 * floor(1 / dshift)
    # the sum folds the nodes at -rho and +rho: ntotal/2 + 1 Horner terms per voxel and direction
    flops = 8.0 * ngrid ** 3 * nphi * ntheta * (ntotal // 2 + 1)
    return {"workload": f"Chang-type inversion, {nphi} phi x {ntheta} theta x {nshift} s -> {ngrid}^3, FP64",
            "seconds": t, "adjoint_terms": st["logical_samples"],
            "terms_per_s": st["logical_samples"] / t, "fp64_TFLOPs": flops / t / 1e12,
            "frac_of_fp64_peak": flops / t / FP64_PEAK,
            "max_err_vs_gaussian": float(np.max(np.abs(out - want)))}


def cli_leg(tmp):
    """end to end radon_grid at 256^3 (FP32 form), text files as the reference's programs read
    and write them, and the binary format; the time of the program from start to exit"""
    import subprocess

    import wrt

    bindir = os.path.join(ROOT, "wrt-project_amd", "bin")
    n = 256
    vol = wrt.phantom(n, True, 10, 0, 0.01)
    cfg = os.path.join(tmp, "cfg.txt")
    with open(cfg, "w") as f:
        f.write("256 phi\n128 theta\n257 shifts\n256 ngrid\n")
    wrt.write_text_values(os.path.join(tmp, "vol.txt"), vol)
    wrt.write_f64(os.path.join(tmp, "vol.f64"), vol)
    res = {"workload": "radon_grid program, 256^3, 256 phi x 128 theta x 257 s, -P f32"}
    for name, args in (("text", ["-i", "vol.txt", "-o", "radon.txt"]),
                       ("binary", ["-i", "vol.f64", "-b", "-o", "radon.f64", "-B"])):
        t0 = time.perf_counter()
        r = subprocess.run([os.path.join(bindir, "radon_grid"), "-p", cfg, "-P", "f32"] + args,
                           cwd=tmp, capture_output=True, text=True, timeout=600)
        dt = time.perf_counter() - t0
        if r.returncode != 0:
            res[name] = {"error": r.stderr[-300:]}
            continue
        res[name] = {"seconds": dt, "output_bytes": os.path.getsize(os.path.join(tmp, args[-2 if name == "binary" else -1]))}
    for f in ("vol.txt", "vol.f64", "radon.txt", "radon.f64"):
        try:
            os.remove(os.path.join(tmp, f))
        except OSError:
            pass
    return res


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "radon_fast_256_traffic.md")
CALIB_COUNT = 1 << 27  # elements of the calibration conversion: 1 GiB read, 512 MiB written


def traffic_probe(args):
    """Calibration kernel of known size, then one pass of the workload (for rocprofv3 --pmc),
    with the tile shape fixed (--band-rows, --dgroup: the values the bench reports), so that no
    tuning pass adds to the counters."""
    import ctypes as C

    if not (args.band_rows and args.dgroup):
        sys.exit("bench.py --traffic-probe needs --band-rows and --dgroup (the tile shape the "
                 "bench line reports in config)")
    import wrt

    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(src), C.c_size_t(CALIB_COUNT * 8)) == 0
    assert hip.hipMalloc(C.byref(dst), C.c_size_t(CALIB_COUNT * 4)) == 0
    L = wrt.lib()
    L.wrtk_fill_f64.argtypes = [C.c_void_p, C.c_double, C.c_size_t, C.c_void_p]
    L.wrtk_f64_to_f32.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]
    assert L.wrtk_fill_f64(src, 1.0, CALIB_COUNT, None) == 0
    assert L.wrtk_f64_to_f32(src, dst, CALIB_COUNT, None) == 0
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    ctx = wrt.Context(0)
    vol = wrt.phantom(args.n, True, 10, 0, 0.01)
    phi, th, w, s = ctx.radon_grid_init(args.nphi, args.ntheta, args.nshift)
    job = ctx.job("radon", args.prec, n=args.n, phi=phi, theta=th, shift=s)
    job.load(vol)
    st = job.run()
    job.close()
    ctx.close()
    print(f"traffic probe: calibration {CALIB_COUNT} elements, pass {st['kernel_seconds']:.3f} s",
          flush=True)


VALU_FILE = os.path.join(ROOT, "profiles", "radon_fast_256_valu_pmc.md")
# the sources the plane-integral kernel is built from, and the build flags: a counter
# summary holds their hash, and its figures are used only while they are unchanged
KERNEL_SOURCES = [os.path.join(ROOT, "wrt-project_amd", "csrc", f)
                  for f in ("radon_brick.hpp", "k_fast.hip", "wrt_fast.h")] + [
    os.path.join(ROOT, "wrt-project_amd", "Makefile")]


def kernel_sources_sha():
    import hashlib

    h = hashlib.sha256()
    for f in KERNEL_SOURCES:
        with open(f, "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()


def _profile_figure(path, key, tile=None):
    """the figure `key:` of a committed counter summary, if it was taken on the current
    kernel sources and, given tile = (band_rows, dgroup), with that tile (None otherwise).
    The summary of the default tile (8 rows, 4096 directions) is `path`; that of another tile
    the tuner may pick carries the tile in its name (`..._8x2048.md`)."""
    paths = [path]
    if tile is not None:
        paths.append(path[:-3] + f"_{tile[0]}x{tile[1]}.md")
    for p in paths:
        vals = {}
        try:
            for line in open(p):
                if ":" in line and not line.startswith((" ", "#", "|")):
                    k, v = line.split(":", 1)
                    vals[k.strip()] = v.strip()
        except OSError:
            continue
        if vals.get("kernel_sources_sha256") != kernel_sources_sha() or key not in vals:
            continue
        if tile is not None and (vals.get("band_rows"), vals.get("dgroup")) != tuple(map(str, tile)):
            continue
        return float(vals[key])
    return None


def traffic_from_profile(tile):
    """HBM bytes per pass from the committed counter summary of this tile (None if absent
    or stale)."""
    return _profile_figure(TRAFFIC_FILE, "traffic_bytes_per_pass", tile)


def valu_from_profile(tile):
    """VALU wavefront instructions per pass from the committed counter summary of this tile
    (None if absent or stale)"""
    return _profile_figure(VALU_FILE, "valu_insts_per_pass", tile)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--nphi", type=int, default=256)
    ap.add_argument("--ntheta", type=int, default=128)
    ap.add_argument("--nshift", type=int, default=257)
    ap.add_argument("--prec", default="f32")
    ap.add_argument("--plain-phi-stride", type=int, default=16,
                    help="plain FP64 reference on every k-th phi (rate and relL2)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-qps", type=float, default=8.0)
    ap.add_argument("--serve-requests", type=int, default=64)
    ap.add_argument("--serve-burst", type=int, default=32)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-transpose", action="store_true")
    ap.add_argument("--no-rays", action="store_true")
    ap.add_argument("--no-analytic", action="store_true")
    ap.add_argument("--no-laplace", action="store_true")
    ap.add_argument("--no-fourier", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-bf16", action="store_true")
    ap.add_argument("--no-programs", action="store_true")
    ap.add_argument("--batch-volumes", type=int, default=48)
    ap.add_argument("--traffic-probe", action="store_true")
    ap.add_argument("--band-rows", type=int, default=0,
                    help="fix the brick kernel's rows per band (4, 8, 16) instead of tuning it "
                         "(profiling runs pass the tuned value the bench reports)")
    ap.add_argument("--dgroup", type=int, default=0,
                    help="fix the brick kernel's directions per task instead of tuning it")
    args = ap.parse_args()
    if args.band_rows:
        os.environ["WRT_BAND_ROWS"] = str(args.band_rows)
    if args.dgroup:
        os.environ["WRT_RADON_DGROUP"] = str(args.dgroup)
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the path is built and measured on a single GPU; multi-GPU sharding is a "
              "design only)", file=sys.stderr)
        sys.exit(2)

    if args.traffic_probe:
        traffic_probe(args)
        return

    import wrt

    ctx = wrt.Context(0)
    n = args.n
    log(f"bench: phantom {n}^3")
    vol = wrt.phantom(n, True, 10, 0, 0.01)
    phi, th, w, s = ctx.radon_grid_init(args.nphi, args.ntheta, args.nshift)
    nplanes = len(phi) * len(th) * len(s)
    logical = 16.0 * n * n * nplanes

    job = ctx.job("radon", args.prec, n=n, phi=phi, theta=th, shift=s)
    job.load(vol)
    # the first launch settles the tile shape (autotuned and cached, or fixed by the options)
    job.run()
    band_rows, dgroup = job.tuning()
    log(f"bench: brick kernel tile: {band_rows} rows x {64 // band_rows} columns per band, "
        f"{dgroup} directions per task")
    for i in range(args.warmup):
        st = job.run()
        log(f"bench: warmup {i} {st['seconds']:.3f} s")
    ks = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        st = job.run()
        ks.append(st["kernel_seconds"])
    wall = time.perf_counter() - t0
    log(f"bench: {args.steps} steps {wall:.3f} s, kernel avg {np.mean(ks):.4f} s")
    fast = job.fetch().reshape(len(s), len(phi), len(th))
    job.close()

    # plain FP64 form on a phi subset: its rate, and the FP32 form's error against it
    sub = phi[::args.plain_phi_stride]
    pj = ctx.job("radon", "f64", n=n, phi=sub, theta=th, shift=s)
    pj.load(vol)
    pst = pj.run()
    plain = pj.fetch().reshape(len(s), len(sub), len(th))
    pj.close()
    plain_rate = 16.0 * n * n * plain.size / pst["kernel_seconds"]
    fs = fast[:, ::args.plain_phi_stride, :]
    err_l2 = rel_l2(fs, plain)
    err_max = float(np.max(np.abs(fs - plain)) / np.max(np.abs(plain)))
    log(f"bench: plain f64 subset {pst['kernel_seconds']:.3f} s, relL2 {err_l2:.3e}")

    incube = ctx.radon_count_incube(n, phi, th, s)
    bf16 = None
    if args.prec == "f32" and not args.no_bf16:
        # the same pass from BF16 bricks (half the brick bytes; FP32 in LDS)
        log("bench: bf16 bricks")
        bj = ctx.job("radon", "bf16", n=n, phi=phi, theta=th, shift=s)
        bj.load(vol)
        bj.run()
        bst = bj.run()
        bout = bj.fetch().reshape(len(s), len(phi), len(th))
        bj.close()
        balg = incube * 2 + nplanes * 4 + n ** 3 * 2
        bf16 = {"seconds": bst["kernel_seconds"], "samples_per_s": logical / bst["kernel_seconds"],
                "alg_bytes_model": balg, "alg_bytes_model_per_s": balg / bst["kernel_seconds"],
                "rel_l2_vs_f64": rel_l2(bout[:, ::args.plain_phi_stride, :], plain)}
        log(f"bench: bf16 {bst['kernel_seconds']:.3f} s, relL2 {bf16['rel_l2_vs_f64']:.2e}")
    eb = 2 if args.prec == "bf16" else 4
    alg_bytes = incube * eb + nplanes * 4 + n ** 3 * eb
    kavg = float(np.mean(ks))
    # The pass is bound by VALU instruction issue (the gathers come from the brick staged in
    # LDS): achieved = VALU wavefront instructions per pass (rocprofv3 SQ_INSTS_VALU, committed
    # with the hash of the kernel sources it was taken on) over the pass time measured here.
    # The HBM figures beside it are the measured traffic and a byte model that prices every
    # in-cube sample's voxel read as an HBM read, which it is not.
    valu = valu_from_profile((band_rows, dgroup))
    traffic = traffic_from_profile((band_rows, dgroup))
    roofline = {
        "bound": "valu",
        "achieved": valu / kavg / 1e9 if valu else None,
        "peak": VALU_PEAK / 1e9,
        "unit": "G VALU wavefront instructions/s",
        "frac": valu / kavg / VALU_PEAK if valu else None,
        "traffic": traffic,
        "hbm_frac": traffic / kavg / HBM_PEAK if traffic else None,
        "valu_insts_per_pass": valu,
        "valu_lane_ops_per_incube_sample": valu * 64 / incube if valu else None,
        "kernel": "k_radon_fast (interior and face-brick launches, concurrent; their span)",
        "kernel_seconds": kavg,
        "alg_bytes_model": {
            "bytes": alg_bytes, "per_second": alg_bytes / kavg,
            "note": "every in-cube sample's voxel priced at the brick's bytes as if read "
                    "from HBM, plus outputs and one volume read; the kernel gathers from "
                    "LDS, so this is a work measure, not a roofline"},
    }
    ms = wall / args.steps * 1e3
    value = logical / (wall / args.steps)

    transpose = None
    if not args.no_transpose:
        log("bench: transpose")
        transpose = transpose_leg(ctx, wrt, vol, fast, phi, th, s, args.plain_phi_stride)
        log(f"bench: transpose {transpose['seconds']:.3f} s, dot rel {transpose['dot_rel']:.2e}, "
            f"relL2 {transpose['rel_l2_vs_f64']:.2e}")
    rays = None
    if not args.no_rays:
        log("bench: rays")
        rays = rays_leg(ctx, wrt, cpu_seconds=0.0 if args.no_cpu else 4.0)
        log(f"bench: rays f32 {rays['f32_s'] * 1e3:.2f} ms, f64 {rays['f64_s'] * 1e3:.2f} ms, "
            f"dot rel {rays['dot_rel']:.2e}")
    analytic = None
    if not args.no_analytic:
        log("bench: analytic")
        analytic = analytic_leg(ctx, wrt, seconds=0.0 if args.no_cpu else 6.0)
        log(f"bench: analytic {analytic['seconds']:.3f} s")
    adjoint = None
    if not args.no_adjoint:
        log("bench: adjoint")
        adjoint = adjoint_leg(ctx, wrt)
        log(f"bench: adjoint {adjoint['f32_s'] * 1e3:.1f} ms")
    fourier = None
    if not args.no_fourier:
        log("bench: fourier")
        fourier = fourier_leg(ctx, wrt)
        log(f"bench: fourier {fourier['seconds'] * 1e3:.2f} ms")
    cpu = None
    if not args.no_cpu:
        log("bench: cpu baseline")
        cpu = cpu_baseline(vol, phi, th, s)
    serve = None
    if not args.no_serve:
        log("bench: serve stream")
        cfg = wrt.serve_cfg(qps=args.serve_qps, nreq=args.serve_requests, seed=1)
        rep, _, _ = ctx.serve(cfg)
        bcfg = wrt.serve_cfg(qps=0.0, nreq=args.serve_burst, seed=1)
        brep, _, _ = ctx.serve(bcfg)
        log("bench: plain per-request FP64 stream")
        pcfg = wrt.serve_cfg(qps=0.0, nreq=args.serve_burst, seed=1, plain=True)
        prep, _, _ = ctx.serve(pcfg)
        serve = {"qps": args.serve_qps, "requests": rep["completed"], "p50_ms": rep["p50_ms"],
                 "p99_ms": rep["p99_ms"], "mean_ms": rep["mean_ms"],
                 "mix": "50% radon f32, 30% adjoint f32, 20% reduction f64; edges 64-256",
                 "burst_requests": brep["completed"],
                 "served_samples_per_s": brep["samples_per_s"],
                 "plain_f64_samples_per_s": prep["samples_per_s"],
                 "speedup_vs_plain": brep["samples_per_s"] / prep["samples_per_s"]}
        log(f"bench: serve p50 {rep['p50_ms']:.1f} ms p99 {rep['p99_ms']:.1f} ms, burst "
            f"{serve['speedup_vs_plain']:.2f}x plain")
    batch = None
    if not args.no_batch:
        log("bench: batch")
        batch = batch_leg(ctx, wrt, args.batch_volumes)
        log(f"bench: batch {batch['graph_s']:.3f} s, {batch['speedup_vs_per_job']:.2f}x per-job, "
            f"{batch['speedup_vs_concurrent']:.2f}x concurrent, "
            f"{batch['speedup_vs_plain']:.2f}x plain")
        # many small volumes, where per-job launches are bound by launch and tail latency
        batch["small_volumes"] = batch_leg(ctx, wrt, 2 * args.batch_volumes, seed=4,
                                           pool=(16, 24, 32, 48, 64))
        sv = batch["small_volumes"]
        log(f"bench: batch of small volumes {sv['graph_s'] * 1e3:.2f} ms, "
            f"{sv['speedup_vs_per_job']:.2f}x per-job, "
            f"{sv['speedup_vs_concurrent']:.2f}x concurrent")
    # The spectral Laplacian last: once the process has made a hipFFT plan, the streams it
    # makes afterwards overlap less (the per-job face-brick side stream costs 5% per pass and
    # the serve rounds lose their overlap), so every other leg runs before it.
    laplace = None
    if not args.no_laplace:
        log("bench: laplace")
        laplace = laplace_leg(ctx, wrt)
        log(f"bench: laplace {laplace['seconds'] * 1e3:.2f} ms")
    ctx.close()
    programs = None
    if not args.no_programs:
        import tempfile

        log("bench: drop-in program end to end")
        # beside the repository (the program's text output is about 340 MB), removed after
        with tempfile.TemporaryDirectory(dir=ROOT, prefix=".bench_programs_") as tmp:
            programs = cli_leg(tmp)
        log(f"bench: programs {programs}")

    out = {
        "metric": "plane samples/s at 256^3 FP32 (relL2<=1e-5 vs FP64) vs plain; p99 served latency",
        "value": value,
        "unit": "plane samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.prec,
        "data": "synthetic (seeded 10-ellipsoid phantom + N(0,0.01^2) texture)",
        "config": {"workload": f"radon_grid {n}^3 {args.prec} bricks, {len(phi)} phi x {len(th)} "
                               f"GL theta x {len(s)} shifts", "n": n, "planes": nplanes,
                   "logical_samples": logical, "incube_samples": incube,
                   "band_rows": band_rows, "dgroup": dgroup,
                   "index_arithmetic": "binary64, reference rounding"},
        "rel_l2_vs_f64": err_l2,
        "max_err_vs_f64": err_max,
        "plain_f64": {"value": plain_rate, "unit": "plane samples/s",
                      "subset": f"every {args.plain_phi_stride}th phi ({plain.size} planes)"},
        "speedup_vs_plain": value / plain_rate,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "serve": serve,
        "batch": batch,
        "transpose": transpose,
        "rays": rays,
        "analytic": analytic,
        "laplace": laplace,
        "bf16": bf16,
        "adjoint": adjoint,
        "fourier": fourier,
        "programs": programs,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
