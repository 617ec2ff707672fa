"""Prove-request server: results in a concurrent mix are byte-identical to the
same request proven alone and to the oracle; admission keeps the in-flight
device footprint within the HBM budget while requests overlap."""
import numpy as np
import pytest

import oracle as O
import p3gpu as G

pytestmark = pytest.mark.gpu


def _requests(seed, count):
    rng = np.random.default_rng(seed)
    out = []
    for r in range(count):
        kind = G.HASH_POSEIDON2 if r % 3 else G.HASH_KECCAK
        k = int(rng.integers(1, 5))
        logs = sorted(int(x) for x in rng.integers(5, 12, size=k))
        ws = [int(x) for x in rng.integers(1, 40, size=k)]
        lb = int(rng.integers(1, 4))
        arity = int(rng.integers(1, 4))
        params = (lb, arity, int(rng.integers(0, 3)), 0, int(r % 2), 20 + r, int(r % 4 == 1))
        traces = [O.rand_matrix(1000 * r + j, 1 << l, w) for j, (l, w) in enumerate(zip(logs, ws))]
        out.append((kind, params, traces))
    return out


def _same(a, b):
    return (a.root.tolist() == b.root.tolist() and a.eval_points == b.eval_points
            and np.array_equal(a.fields, b.fields) and np.array_equal(a.commitments, b.commitments))


def test_served_results_identical_alone_and_in_mix():
    reqs = _requests(7, 12)
    srv = G.Server(0, workers=4, hbm_budget_bytes=8 << 30)
    tickets = [srv.submit(tr, G.PcsParams(*p), kind) for kind, p, tr in reqs]
    mixed = [srv.result(t) for t in tickets]
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(reqs) and st["peak_inflight_bytes"] <= st["budget_bytes"]
    alone_srv = G.Server(0, workers=1, hbm_budget_bytes=8 << 30)
    for (kind, p, tr), m in zip(reqs, mixed):
        alone = alone_srv.result(alone_srv.submit(tr, G.PcsParams(*p), kind))
        assert _same(alone, m)
        root_o, z_o, f_o, c_o, _ = O.pcs_trace_prove(O.PcsParams(*p), kind, tr)
        assert m.root.tolist() == root_o.tolist() and m.eval_points == tuple(z_o)
        assert np.array_equal(m.fields, f_o) and np.array_equal(m.commitments, c_o)
    alone_srv.close()


def test_admission_budget_and_overlap():
    reqs = [(G.HASH_POSEIDON2, (2, 2, 2, 0, 0, 30, 0), [O.rand_matrix(50 + i, 1 << 14, 24)]) for i in range(8)]
    fp = G.request_footprint(G.HASH_POSEIDON2, G.PcsParams(*reqs[0][1]), reqs[0][2])
    budget = int(2.5 * fp)  # two requests fit at once, three do not
    srv = G.Server(0, workers=4, hbm_budget_bytes=budget)
    tickets = [srv.submit(tr, G.PcsParams(*p), kind) for kind, p, tr in reqs]
    res = [srv.result(t) for t in tickets]
    st = srv.stats()
    srv.close()
    assert st["completed"] == 8
    assert st["peak_inflight_bytes"] <= budget
    assert st["peak_running"] == 2
    # every proof of the same-shaped requests is the plain path's
    dev = G.Device(0)
    cfg = G.LmcsConfig(dev, G.HASH_POSEIDON2)
    for (kind, p, tr), r in zip(reqs[:2], res[:2]):
        root, z, f, c = G.prove(cfg, G.PcsParams(*p), tr)
        assert root.tolist() == r.root.tolist() and np.array_equal(f, r.fields) and np.array_equal(c, r.commitments)
    dev.close()


def test_oversized_request_rejected_and_bad_request_reported():
    srv = G.Server(0, workers=2, hbm_budget_bytes=1 << 20)
    with pytest.raises(G.P3GError) as e:
        srv.submit([O.rand_matrix(1, 1 << 12, 8)], G.PcsParams(1, 1, 0, 0, 0, 4, 0))
    assert e.value.code == 8
    srv.close()
    srv = G.Server(0, workers=2, hbm_budget_bytes=1 << 30)
    t = srv.submit([O.rand_matrix(1, 64, 3), O.rand_matrix(2, 32, 3)], G.PcsParams(1, 1, 0, 0, 0, 4, 0))
    with pytest.raises(G.P3GError) as e:
        srv.result(t)
    assert e.value.code == 3  # unsorted heights, reported through the ticket
    ok = srv.result(srv.submit([O.rand_matrix(3, 64, 3)], G.PcsParams(1, 1, 0, 0, 0, 4, 0)))
    assert ok.fields.size > 0
    srv.close()


@pytest.mark.parametrize("shapes,params", [
    ([(1 << 14, 24)], (2, 2, 2, 0, 0, 30, 0)),
    ([(1 << 10, 5), (1 << 12, 64), (1 << 16, 17)], (1, 1, 3, 0, 0, 50, 0)),
    ([(1 << 15, 100), (1 << 15, 7)], (3, 3, 0, 0, 0, 20, 0)),
    ([(1 << 18, 16)], (1, 2, 4, 0, 0, 100, 0)),
])
def test_footprint_covers_peak_device_memory(device, shapes, params):
    tr = [O.rand_matrix(9 + i, h, w) for i, (h, w) in enumerate(shapes)]
    for kind in (G.HASH_POSEIDON2, G.HASH_KECCAK):
        device.synchronize()
        device.trim()
        base = device.memory(reset_peak=True)
        G.prove(G.LmcsConfig(device, kind), G.PcsParams(*params), tr)
        used = device.memory()["peak_live"] - base["live"]
        fp = G.request_footprint(kind, G.PcsParams(*params), tr)
        assert used <= fp, (kind, used, fp)


def test_server_batches_same_shape_requests():
    # same-shaped requests that queue up while the workers are busy are proven
    # together in lock-step; every result is still the single-proof one
    reqs = [O.rand_matrix(900 + i, 1 << 10, 16) for i in range(32)]
    params = (1, 2, 2, 0, 0, 30, 0)
    srv = G.Server(0, workers=2, hbm_budget_bytes=8 << 30, max_batch=8)
    tickets = [srv.submit([m], G.PcsParams(*params)) for m in reqs]
    res = [srv.result(t) for t in tickets]
    st = srv.stats()
    srv.close()
    assert st["completed"] == 32 and st["batched_requests"] > 0, st
    dev = G.Device(0)
    cfg = G.LmcsConfig(dev, G.HASH_POSEIDON2)
    for m, r in zip(reqs, res):
        root, z, f, c = G.prove(cfg, G.PcsParams(*params), [m])
        assert root.tolist() == r.root.tolist() and z == r.eval_points
        assert np.array_equal(f, r.fields) and np.array_equal(c, r.commitments)
    dev.close()


def test_server_batch_window_gathers_requests_arriving_together():
    # with a batch window longer than the submission takes, the first request waits
    # for the rest of its shape: every request is proven in lock-step, and each
    # result is still the single-proof one
    reqs = [O.rand_matrix(1300 + i, 1 << 9, 8) for i in range(24)]
    params = (1, 1, 2, 0, 0, 20, 0)
    srv = G.Server(0, workers=2, hbm_budget_bytes=8 << 30, max_batch=32, batch_window_ms=500.0)
    tickets = [srv.submit([m], G.PcsParams(*params)) for m in reqs]
    res = [srv.result(t) for t in tickets]
    st = srv.stats()
    srv.close()
    assert st["completed"] == 24 and st["batched_requests"] == 24 and st["batches"] <= 2, st
    dev = G.Device(0)
    cfg = G.LmcsConfig(dev, G.HASH_POSEIDON2)
    for m, r in zip(reqs, res):
        root, z, f, c = G.prove(cfg, G.PcsParams(*params), [m])
        assert root.tolist() == r.root.tolist() and np.array_equal(f, r.fields) and np.array_equal(c, r.commitments)
    dev.close()
