"""GPU tests of the batch runner: many fast jobs (plane integrals of uneven sizes in FP32
and BF16, FP32 backprojections) in one persistent graph launch, beside plain FP64 jobs on
their own streams.  Each job's batched output equals its own run bit for bit: the per-job
launches walk the 8 x 8 bands the persistent kernel walks (WRT_BAND_ROWS=8; another band
shape sums a section's lanes in another FP32 order), and plane sums are exact 64-bit
fixed-point sums whatever the order of their sections.  Replays, reloads and the
autotuner's cache are covered too."""
import os

import numpy as np
import pytest

import wrt

pytestmark = pytest.mark.gpu


def _jobs(ctx, seed=0):
    jobs = []
    # BF16 at a power-of-two edge and at an edge that is not one (the persistent kernel
    # takes both as run-time task flags), edges that leave partial bricks (40) and partial
    # backprojection tiles (37)
    for k, (op, prec, e) in enumerate([("radon", "f32", 32), ("radon", "f32", 48),
                                        ("radon", "f32", 64), ("radon", "bf16", 64),
                                        ("adjoint", "f32", 32), ("adjoint", "f32", 48),
                                        ("radon", "f64", 32), ("adjoint", "f64", 24),
                                        ("radon", "bf16", 40), ("adjoint", "f32", 37)]):
        phi, th, w, s = ctx.radon_grid_init(e // 2, e // 4, e + 1)
        if op == "radon":
            j = ctx.job("radon", prec, n=e, phi=phi, theta=th, shift=s)
            j.load(wrt.phantom(e, True, 10, 100 + k + seed, 0.01))
        else:
            j = ctx.job("adjoint", prec, n=e, phi=phi, theta=th, theta_w=w, shift=s)
            j.load(wrt.gaussian((len(s), len(phi), len(th)), 200 + k + seed))
        jobs.append((op, prec, j))
    return jobs


def _alone(jobs):
    out = []
    for _, _, j in jobs:
        j.run()
        out.append(j.fetch())
    return out


def _check(jobs, want):
    for (op, prec, j), w in zip(jobs, want):
        assert np.array_equal(j.fetch(), w), (op, prec)


@pytest.fixture(autouse=True)
def _eight_row_bands(monkeypatch):
    monkeypatch.setenv("WRT_BAND_ROWS", "8")


def test_batch_equals_alone_and_replays(ctx, tmp_path):
    old_cache = os.environ.get("WRT_TUNE_CACHE", "")
    os.environ["WRT_TUNE_CACHE"] = str(tmp_path / "tune.txt")
    try:
        jobs = _jobs(ctx)
        want = _alone(jobs)
        st = ctx.batch_run([j for _, _, j in jobs])
        assert st["kernel_seconds"] > 0 and st["logical_samples"] > 0
        _check(jobs, want)
        g = ctx.batch_dgroup()
        assert g in (16, 32, 64, 128, 256, 512)
        # the planner's entry in the tuning cache: "<signature> <G> 0 <seconds>" (the per-job
        # launches of the alone runs add their own "<signature> <rows> <group> <seconds>")
        lines = [l.split() for l in open(tmp_path / "tune.txt").read().splitlines()]
        assert any(int(l[1]) == g and int(l[2]) == 0 for l in lines), lines
        st2 = ctx.batch_run([j for _, _, j in jobs])  # graph replay
        _check(jobs, want)
        assert st2["logical_samples"] == st["logical_samples"]
        # a reload invalidates the plan: the next batch uses the new data
        _, _, j0 = jobs[2]
        j0.load(wrt.phantom(64, True, 10, 999, 0.0))
        j0.run()
        want[2] = j0.fetch()
        ctx.batch_run([j for _, _, j in jobs])
        _check(jobs, want)
        for _, _, j in jobs:
            j.close()
    finally:
        os.environ["WRT_TUNE_CACHE"] = old_cache


def test_batch_after_jobs_closed_and_recreated(ctx, tmp_path):
    # A plan lists its jobs by address and load sequence; jobs closed and recreated with the
    # same count (a new job may reuse a freed job's address) and loaded once must get a new
    # plan, not a replay of kernels that write into the freed buffers.
    old_cache = os.environ.get("WRT_TUNE_CACHE", "")
    os.environ["WRT_TUNE_CACHE"] = str(tmp_path / "tune.txt")
    try:
        for rnd in range(3):
            jobs = _jobs(ctx, seed=1000 * rnd)  # new data, each job loaded once
            want = _alone(jobs)
            ctx.batch_run([j for _, _, j in jobs])
            _check(jobs, want)
            for _, _, j in jobs:
                j.close()
    finally:
        os.environ["WRT_TUNE_CACHE"] = old_cache


def test_batch_without_autotune_uses_default(ctx):
    os.environ["WRT_AUTOTUNE"] = "0"
    try:
        jobs = _jobs(ctx)[:3]
        want = _alone(jobs)
        ctx.batch_run([j for _, _, j in jobs])
        assert ctx.batch_dgroup() == 64
        _check(jobs, want)
        for _, _, j in jobs:
            j.close()
    finally:
        del os.environ["WRT_AUTOTUNE"]
