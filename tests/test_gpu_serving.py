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


def _check_oracle(proof, kind, params, traces):
    root_o, z_o, f_o, c_o, _ = O.pcs_trace_prove(O.PcsParams(*params), kind, traces)
    assert proof.root.tolist() == root_o.tolist() and proof.eval_points == tuple(z_o)
    assert np.array_equal(proof.fields, f_o) and np.array_equal(proof.commitments, c_o)


def _wait_running(srv, n=1, timeout_s=30.0):
    import time
    t0 = time.perf_counter()
    while srv.stats()["running"] < n:
        assert time.perf_counter() - t0 < timeout_s, "request never started"
        time.sleep(0.0005)


def test_canceled_requests_leave_the_others_bit_exact(device):
    # one worker: a large proof runs (2^22 x 64 at blowup 8, about 0.2 s of GPU work) while two
    # small ones wait behind it.  Canceling the queued one completes it at once; canceling the
    # running one stops it at its next phase boundary; the remaining request is untouched.
    big = device.alloc_matrix(1 << 22, 64)
    big.fill_random(77)
    small = [O.rand_matrix(500 + i, 1 << 10, 12) for i in range(2)]
    params = (1, 1, 2, 0, 0, 24, 0)
    srv = G.Server(0, workers=1, hbm_budget_bytes=64 << 30, max_batch=1)
    t_big = srv.submit([big], G.PcsParams(3, 1, 4, 0, 0, 40, 0))
    t_keep = srv.submit([small[0]], G.PcsParams(*params))
    t_drop = srv.submit([small[1]], G.PcsParams(*params))
    _wait_running(srv)
    print("cancel: queued request, then the running 2^22 x 64 proof", flush=True)
    srv.cancel(t_drop)
    srv.cancel(t_big)
    for t in (t_drop, t_big):
        with pytest.raises(G.P3GError) as e:
            srv.result(t)
        assert e.value.code == 9, e.value  # P3G_E_CANCELED
    keep = srv.result(t_keep)
    _check_oracle(keep, G.HASH_POSEIDON2, params, [small[0]])
    st = srv.stats()
    assert st["running"] == 0 and st["inflight_bytes"] == 0 and st["queued"] == 0
    # the server keeps serving after the cancellations
    again = srv.result(srv.submit([small[1]], G.PcsParams(*params)))
    _check_oracle(again, G.HASH_POSEIDON2, params, [small[1]])
    srv.close()
    big.free()


def test_canceled_batch_member_leaves_its_batch_mates_bit_exact(device):
    # same-shaped requests queue behind a running proof and would batch together; one is
    # canceled while queued: the others are proven in lock-step and stay bit-exact
    big = device.alloc_matrix(1 << 21, 64)
    big.fill_random(78)
    params = (1, 2, 1, 0, 0, 16, 0)
    jobs = [[O.rand_matrix(600 + i, 1 << 11, 9)] for i in range(5)]
    srv = G.Server(0, workers=1, hbm_budget_bytes=64 << 30, max_batch=8)
    t_big = srv.submit([big], G.PcsParams(2, 1, 4, 0, 0, 40, 0))
    _wait_running(srv)
    ts = [srv.submit(j, G.PcsParams(*params)) for j in jobs]
    srv.cancel(ts[2])
    with pytest.raises(G.P3GError) as e:
        srv.result(ts[2])
    assert e.value.code == 9
    srv.result(t_big)
    for i, t in enumerate(ts):
        if i != 2:
            _check_oracle(srv.result(t), G.HASH_POSEIDON2, params, jobs[i])
    assert srv.stats()["batches"] >= 1
    srv.close()
    big.free()


def test_injected_allocation_failure_is_isolated():
    # A request whose device allocation fails in mid-proof reports P3G_E_NOMEM; its lock-step
    # batch-mates (same shape, proven together) and an unrelated request still finish
    # bit-exact, and the server keeps serving.
    params = (1, 1, 2, 0, 0, 20, 0)
    jobs = [[O.rand_matrix(700 + i, 1 << 10, 10)] for i in range(4)]
    other = [O.rand_matrix(800, 1 << 9, 30), O.rand_matrix(801, 1 << 11, 5)]
    srv = G.Server(0, workers=2, hbm_budget_bytes=8 << 30, max_batch=8, batch_window_ms=200)
    ts = [srv.submit(j, G.PcsParams(*params), inject_alloc_failure=(i == 1)) for i, j in enumerate(jobs)]
    t_other = srv.submit(other, G.PcsParams(2, 2, 0, 0, 0, 12, 0))
    t_alone = srv.submit([O.rand_matrix(900, 1 << 8, 7)], G.PcsParams(1, 1, 0, 0, 0, 8, 0), inject_alloc_failure=True)
    for i, t in enumerate(ts):
        if i == 1:
            with pytest.raises(G.P3GError) as e:
                srv.result(t)
            assert e.value.code == 4, e.value  # P3G_E_NOMEM
        else:
            _check_oracle(srv.result(t), G.HASH_POSEIDON2, params, jobs[i])
    with pytest.raises(G.P3GError) as e:
        srv.result(t_alone)
    assert e.value.code == 4
    _check_oracle(srv.result(t_other), G.HASH_POSEIDON2, (2, 2, 0, 0, 0, 12, 0), other)
    again = srv.result(srv.submit(jobs[1], G.PcsParams(*params)))
    _check_oracle(again, G.HASH_POSEIDON2, params, jobs[1])
    st = srv.stats()
    assert st["running"] == 0 and st["inflight_bytes"] == 0
    srv.close()


def test_submit_rejects_null_trace_data_and_bad_params():
    import ctypes as C
    srv = G.Server(0, workers=1, hbm_budget_bytes=1 << 30)
    m = (G._Matrix * 1)(G._Matrix(None, 64, 4))
    req = G._Request(G.HASH_POSEIDON2, G.PcsParams(1, 1, 0, 0, 0, 4, 0), C.cast(m, C.POINTER(G._Matrix)), 1, 0)
    t = np.zeros(1, np.uint64)
    assert G.lib().p3g_server_submit(srv.h, C.byref(req), G._p(t)) == 1  # P3G_E_INVAL
    tr = O.rand_matrix(1, 64, 4)
    m2 = (G._Matrix * 1)(G._Matrix(tr.ctypes.data_as(G._u64p), 64, 4))
    bad = G.PcsParams.__new__(G.PcsParams)
    C.Structure.__init__(bad, 1, 1, 0, 0, 40, 4, 0)  # 40 DEEP grinding bits
    req2 = G._Request(G.HASH_POSEIDON2, bad, C.cast(m2, C.POINTER(G._Matrix)), 1, 0)
    assert G.lib().p3g_server_submit(srv.h, C.byref(req2), G._p(t)) == 1
    srv.close()


@pytest.mark.parametrize("workers,max_batch", [(1, 1), (3, 1), (2, 8)])
def test_upload_ahead_results_identical_to_the_oracle(device, workers, max_batch):
    # the upload stage copies queued host traces to the device ahead of admission; device-
    # input requests skip it.  Every result must be the oracle's, and the budget must hold.
    reqs = _requests(11, 10)
    # two requests with device inputs among the host ones
    dev_reqs = [(G.HASH_POSEIDON2, (1, 1, 2, 0, 0, 20, 0), [O.rand_matrix(950 + i, 1 << 11, 10)]) for i in range(2)]
    dev_in = [device.upload(tr[0]) for _, _, tr in dev_reqs]
    reqs = reqs[:3] + dev_reqs[:1] + reqs[3:6] + dev_reqs[1:] + reqs[6:]
    srv = G.Server(0, workers=workers, hbm_budget_bytes=8 << 30, max_batch=max_batch, upload_ahead=3)
    tickets = []
    for kind, p, tr in reqs:
        d = [m for (_, _, t), m in zip(dev_reqs, dev_in) if t is tr]
        tickets.append(srv.submit(d if d else tr, G.PcsParams(*p), kind))
    # same-shaped host requests, which batch together when max_batch > 1
    same = [O.rand_matrix(900 + i, 1 << 10, 16) for i in range(6)]
    sp = (1, 1, 2, 0, 0, 20, 0)
    tickets += [srv.submit([t], G.PcsParams(*sp)) for t in same]
    print(f"upload ahead: workers={workers} max_batch={max_batch}", flush=True)
    res = [srv.result(t) for t in tickets]
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(tickets) and st["peak_inflight_bytes"] <= st["budget_bytes"]
    assert st["inflight_bytes"] == 0
    for (kind, p, tr), r in zip(reqs, res):
        _check_oracle(r, kind, p, tr)
    for t, r in zip(same, res[len(reqs):]):
        _check_oracle(r, G.HASH_POSEIDON2, sp, [t])
    for m in dev_in:
        m.free()


def test_upload_ahead_within_a_tight_budget_and_with_cancels():
    # a budget for two requests at a time and a lookahead of four: staged requests reserve
    # their footprint, so the in-flight bytes never pass the budget; canceled queued requests
    # (staged or not) give their reservation back and the rest finish bit-exact
    reqs = [(G.HASH_POSEIDON2, (2, 2, 2, 0, 0, 30, 0), [O.rand_matrix(70 + i, 1 << 14, 24)]) for i in range(8)]
    fp = G.request_footprint(G.HASH_POSEIDON2, G.PcsParams(*reqs[0][1]), reqs[0][2])
    budget = int(2.5 * fp)
    srv = G.Server(0, workers=2, hbm_budget_bytes=budget, max_batch=1, upload_ahead=4)
    tickets = [srv.submit(tr, G.PcsParams(*p), kind) for kind, p, tr in reqs]
    srv.cancel(tickets[5])
    srv.cancel(tickets[7])
    out = {}
    for i, t in enumerate(tickets):
        if i in (5, 7):
            try:
                srv.result(t)
                out[i] = "finished"  # it may have completed before the cancel
            except G.P3GError as e:
                assert e.code == 9, e
                out[i] = "canceled"
        else:
            out[i] = srv.result(t)
    st = srv.stats()
    srv.close()
    assert st["peak_inflight_bytes"] <= budget and st["inflight_bytes"] == 0 and st["queued"] == 0
    assert st["peak_running"] <= 2
    for i, (kind, p, tr) in enumerate(reqs):
        if i not in (5, 7):
            _check_oracle(out[i], kind, p, tr)
