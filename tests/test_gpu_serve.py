"""GPU tests of the serve loop: every request's result is the same in the mix as alone, bit
for bit (the alone runs walk the rounds' 8 x 8 bands, WRT_BAND_ROWS=8; plane sums are exact
64-bit fixed-point sums whatever the order of their sections), admission under a small HBM
budget evicts and still serves every request, the latency report is consistent, and the
continuous batching runs a small request beside a large one that arrived before it."""
import numpy as np
import pytest

import wrt

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _eight_row_bands(monkeypatch):
    monkeypatch.setenv("WRT_BAND_ROWS", "8")


def _alone(ctx, req):
    j = ctx.serve_job(req)
    j.run()
    out = j.fetch()
    j.close()
    return out


def test_serve_results_alone_equal_mix(ctx):
    cfg = wrt.serve_cfg(qps=400.0, nreq=30, seed=11, edges=(32, 48, 64), ndatasets=2,
                        max_inflight=4)
    rep, lat, outs = ctx.serve(cfg, keep=range(30))
    assert rep["completed"] == 30 and np.all(lat > 0)
    assert rep["peak_inflight"] >= 2  # requests overlapped on the device
    for i, got in outs.items():
        req = wrt.serve_describe(cfg, i)
        want = _alone(ctx, req)
        assert np.array_equal(got, want), (i, req.op)


def test_serve_admission_evicts_under_budget(ctx):
    cfg = wrt.serve_cfg(qps=0.0, nreq=24, seed=5, edges=(48, 64), ndatasets=3, max_inflight=2)
    sizes = []
    for i in range(cfg.nreq):
        r = wrt.serve_describe(cfg, i)
        phi, th, w, s, z = ctx.serve_grids(r.edge)
        kw = dict(phi=phi, theta=th, shift=s)
        name = {v: k for k, v in wrt.OP.items()}[r.op]
        if name == "reduction":
            kw["z"] = z
        else:
            kw["n"] = r.edge
        if name == "adjoint":
            kw["theta_w"] = w
        prec = {v: k for k, v in wrt.PREC.items()}[r.prec]
        sizes.append(wrt.job_bytes(name, prec, **kw))
    budget = 2 * max(sizes) + 1
    cfg = wrt.serve_cfg(qps=0.0, nreq=24, seed=5, edges=(48, 64), ndatasets=3, max_inflight=2,
                        hbm_budget=budget)
    rep, lat, _ = ctx.serve(cfg)
    assert rep["completed"] == 24
    assert rep["evictions"] > 0 and rep["peak_bytes"] <= budget
    assert 0 < rep["p50_ms"] <= rep["p99_ms"] <= rep["max_ms"]
    assert abs(rep["p99_ms"] - np.sort(lat)[int(np.ceil(0.99 * 24)) - 1]) < 1e-9


def test_serve_rounds_interleave_small_with_large(ctx):
    """Continuous batching: a 64^3 plane-integral request queued right behind a 256^3 one (both
    at t = 0, the large one first) runs in the rounds beside it and finishes long before it,
    where a first-come, first-served queue would finish it after; both equal their alone
    runs."""
    seed = next(sd for sd in range(1, 400)
                if [wrt.serve_describe(wrt.serve_cfg(qps=0.0, nreq=2, seed=sd, edges=(64, 256),
                                                     mix=(1.0, 0.0, 0.0)), i).edge
                    for i in range(2)] == [256, 64])
    cfg = wrt.serve_cfg(qps=0.0, nreq=2, seed=seed, edges=(64, 256), mix=(1.0, 0.0, 0.0))
    rep, lat, outs = ctx.serve(cfg, keep=range(2))
    assert rep["completed"] == 2 and rep["batched"] == 2 and rep["rounds"] >= 2
    assert lat[1] < 0.5 * lat[0], lat
    for i in range(2):
        want = _alone(ctx, wrt.serve_describe(cfg, i))
        assert np.array_equal(outs[i], want), i
