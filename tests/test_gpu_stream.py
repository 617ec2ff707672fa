"""A context on a caller's stream (p3g_ctx_create_on_stream): the library's
work is ordered after the caller's own work on that stream, results match
the oracle, and destroying the context leaves the stream usable.

The caller's stream comes from the HIP runtime the library links
(libamdhip64.so.7, resolved to the instance already loaded in this process):
a stream handle belongs to one runtime instance, and a second runtime bundled
elsewhere (inside a Python wheel, for instance) does not share handles."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
import p3gpu as G

pytestmark = pytest.mark.gpu

H2D, D2D = 1, 3


def _hip():
    G.lib()  # the library's runtime is loaded first, so the SONAME resolves to it
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    hip.hipStreamSynchronize.argtypes = [C.c_void_p]
    hip.hipStreamDestroy.argtypes = [C.c_void_p]
    hip.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
    for f in (hip.hipStreamCreate, hip.hipStreamSynchronize, hip.hipStreamDestroy, hip.hipMemcpyAsync):
        f.restype = C.c_int
    return hip


def test_commit_and_prove_on_a_caller_stream():
    hip = _hip()
    s = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(s)) == 0
    dev = G.Device(0, stream=s.value)
    try:
        cfg = G.LmcsConfig(dev, G.HASH_POSEIDON2)
        h, w = 1 << 12, 16
        tr = O.rand_matrix(5, h, w)
        a, b = dev.alloc_matrix(h, w), dev.alloc_matrix(h, w)
        zero = np.zeros((h, w), np.uint64)
        G.lib().p3g_upload(dev.h, b.ptr, zero.ctypes.data_as(C.c_void_p), zero.nbytes)
        # the caller's copies, queued on its stream; the commit follows them with no host wait
        assert hip.hipMemcpyAsync(a.ptr, tr.ctypes.data, tr.nbytes, H2D, s) == 0
        assert hip.hipMemcpyAsync(b.ptr, a.ptr, tr.nbytes, D2D, s) == 0
        t = G.commit_traces(cfg, [b], 1)
        want = O.commit_traces(O.HASH_P2, [tr], 1)
        assert np.array_equal(t.root(), want.root())
        # a whole proof on this context equals the oracle's
        tr2 = O.rand_matrix(17, 1 << 8, 6)
        params = (1, 1, 2, 0, 0, 11, 0)
        r, _, fields, comms = G.prove(cfg, G.PcsParams(*params), [tr2])
        r2, _, f2, c2, _ = O.pcs_trace_prove(O.PcsParams(*params), O.HASH_P2, [tr2])
        assert np.array_equal(r, r2) and np.array_equal(fields, f2) and np.array_equal(comms, c2)
    finally:
        dev.close()
    # the stream outlives the context
    assert hip.hipStreamSynchronize(s) == 0
    assert hip.hipStreamDestroy(s) == 0


def test_create_on_stream_rejects_the_null_stream():
    h = C.c_void_p()
    assert G.lib().p3g_ctx_create_on_stream(0, None, C.byref(h)) == 1
    assert not h.value
