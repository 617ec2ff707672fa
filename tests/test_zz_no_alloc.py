"""p3g_coset_lde_batch_noalloc: the coset LDE with its tables in the library's
static per-GPU scratch array (no device allocation).  Called through the C ABI
and compared word for word with p3g_coset_lde_batch on the same device buffers
(P3G_DEVICE_PTRS | P3G_BITREV_ROWS), on consecutive calls whose second input is
smaller than the first, so stale table contents from the larger call would
show; the small case is also checked against the oracle.  Shapes beyond the
array give P3G_E_RANGE before anything is launched.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
import p3gpu as G

pytestmark = pytest.mark.gpu

FLAGS = G.DEVICE_PTRS | G.BITREV_ROWS
P3G_E_RANGE = 8


def _both(device, n, w, added_bits, shift, seed):
    lib = G.lib()
    tr = device.alloc_matrix(n, w)
    tr.fill_random(seed)
    ref = device.alloc_matrix(n << added_bits, w)
    got = device.alloc_matrix(n << added_bits, w)
    print(f"noalloc LDE: n={n} w={w} added_bits={added_bits} shift={shift}", flush=True)
    rc = lib.p3g_coset_lde_batch(device.h, C.cast(tr.ptr, G._u64p), n, w, added_bits, shift, FLAGS,
                                 C.cast(ref.ptr, G._u64p))
    assert rc == 0, G.lib().p3g_last_error()
    rc = lib.p3g_coset_lde_batch_noalloc(device.h, tr.ptr, n, w, added_bits, shift, got.ptr)
    assert rc == 0, G.lib().p3g_last_error()
    device.synchronize()
    a, b, t = ref.download(), got.download(), tr.download()
    for m in (tr, ref, got):
        m.free()
    return a, b, t


def test_noalloc_lde_equals_allocating_lde_on_consecutive_calls(device):
    mem0 = device.memory()
    # larger first, then smaller shapes: each call rewrites only the tables it uses
    cases = [(1 << 16, 24, 2, 7), (1 << 10, 8, 1, 49), (1 << 6, 3, 3, 7), (1 << 12, 16, 0, 1)]
    for i, (n, w, ab, shift) in enumerate(cases):
        a, b, t = _both(device, n, w, ab, shift, 100 + i)
        assert a.shape == b.shape == (n << ab, w)
        assert np.array_equal(a, b), (n, w, ab)
        if n <= 1 << 10:
            assert np.array_equal(b, O.coset_lde_batch(t, ab, shift)), (n, w, ab)
    # the no-alloc calls held no device memory of their own beyond the matrices freed above
    assert device.memory()["live"] == mem0["live"]


def test_noalloc_lde_rejects_shapes_beyond_the_scratch_array(device):
    lib = G.lib()
    tr = device.alloc_matrix(1 << 4, 2)
    out = device.alloc_matrix(1 << 8, 2)
    # heights are never read before the size check: 2^26 rows needs more table words than the array holds
    assert lib.p3g_coset_lde_batch_noalloc(device.h, tr.ptr, 1 << 26, 2, 1, 7, out.ptr) == P3G_E_RANGE
    assert b"scratch" in lib.p3g_last_error()
    assert lib.p3g_coset_lde_batch_noalloc(device.h, tr.ptr, 1 << 4, 2, 4, 7, out.ptr) == P3G_E_RANGE
    tr.free()
    out.free()
