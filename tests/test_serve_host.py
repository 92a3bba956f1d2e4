"""CPU tests of the serve loop's request stream: seeded, Poisson arrivals at the stated rate,
the stated mix of operations, edges and datasets, and dataset inputs equal to the seeded
generators'.  (The loop itself runs on the GPU: tests/test_gpu_serve.py.)"""
import numpy as np
import pytest

import wrt


def _stream(cfg):
    return [wrt.serve_describe(cfg, i) for i in range(cfg.nreq)]


def test_stream_is_seeded_and_poisson():
    cfg = wrt.serve_cfg(qps=50.0, nreq=1500, seed=7)
    a = _stream(cfg)
    b = _stream(wrt.serve_cfg(qps=50.0, nreq=1500, seed=7))
    assert [r.as_dict() for r in a] == [r.as_dict() for r in b]
    t = np.array([r.arrival_s for r in a])
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert np.all(gaps > 0)
    # exponential gaps: mean 1/qps, coefficient of variation 1 (4-sigma bounds)
    assert abs(gaps.mean() * 50.0 - 1.0) < 4 / np.sqrt(len(gaps))
    assert abs(gaps.std() / gaps.mean() - 1.0) < 0.15
    c = _stream(wrt.serve_cfg(qps=50.0, nreq=1500, seed=8))
    assert [r.arrival_s for r in c] != list(t)


def test_stream_mix_edges_and_datasets():
    edges = (64, 96, 128, 192, 256)
    cfg = wrt.serve_cfg(qps=10.0, nreq=3000, seed=3, mix=(0.5, 0.3, 0.2), edges=edges, ndatasets=3)
    reqs = _stream(cfg)
    ops = np.array([r.op for r in reqs])
    for op, w in ((wrt.OP["radon"], 0.5), (wrt.OP["adjoint"], 0.3), (wrt.OP["reduction"], 0.2)):
        frac = np.mean(ops == op)
        assert abs(frac - w) < 4 * np.sqrt(w * (1 - w) / len(reqs)), (op, frac)
    for r in reqs:
        assert r.edge in edges and 0 <= r.dataset < 3
        assert r.prec == (wrt.PREC["f64"] if r.op == wrt.OP["reduction"] else wrt.PREC["f32"])
    # one seed per (op, edge, dataset): requests of a dataset share its input
    seen = {}
    for r in reqs:
        key = (r.op, r.edge, r.dataset)
        assert seen.setdefault(key, r.data_seed) == r.data_seed
    assert len(set(seen.values())) == len(seen)


def test_burst_and_invalid_configs():
    cfg = wrt.serve_cfg(qps=0.0, nreq=10)
    assert all(r.arrival_s == 0.0 for r in _stream(cfg))
    for bad in (dict(edges=(63,)), dict(edges=(4,)), dict(mix=(0, 0, 0)), dict(ndatasets=0),
                dict(max_inflight=0)):
        cfg = wrt.serve_cfg(nreq=4, **bad)
        with pytest.raises(wrt.WrtError):
            wrt.serve_describe(cfg, 0)


def test_dataset_inputs_are_the_generators():
    cfg = wrt.serve_cfg(qps=1.0, nreq=40, seed=5, edges=(16, 24))
    done = set()
    for r in _stream(cfg):
        if r.op in done:
            continue
        done.add(r.op)
        x = wrt.serve_input(r)
        e = r.edge
        if r.op == wrt.OP["radon"]:
            assert np.array_equal(x, wrt.phantom(e, True, 10, r.data_seed, 0.01).ravel())
        else:
            assert np.array_equal(x, wrt.gaussian(x.shape, r.data_seed))
    assert len(done) == 3
