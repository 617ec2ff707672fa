"""GPU parity: the HIP library, called through the C ABI, against the CPU
oracle on the same seeded inputs.  Bit-exact (integer field arithmetic).

Sizes stay where the oracle finishes in seconds; the full-size configuration
is covered in test_gpu_full_size.py through size-independent properties.
"""
import numpy as np
import pytest

import oracle as O
import p3gpu as G

pytestmark = pytest.mark.gpu


def _mats(seed, shapes):
    return [O.rand_matrix(seed + i, h, w) for i, (h, w) in enumerate(shapes)]


# ---------------------------------------------------------------- coset LDE
@pytest.mark.parametrize("logn,w,bits", [(0, 1, 1), (0, 5, 3), (1, 1, 1), (2, 3, 2), (3, 16, 1), (5, 17, 2),
                                         (8, 4, 1), (9, 33, 1), (10, 64, 2), (12, 8, 3), (14, 32, 1), (16, 3, 2)])
def test_coset_lde_matches_oracle(device, logn, w, bits):
    tr = O.rand_matrix(1000 + logn, 1 << logn, w)
    shift = O.gl_pow(7, 1 + logn % 3)
    want = O.coset_lde_batch(tr, bits, shift)
    got = G.Dft(device).coset_lde_batch(tr, bits, shift)
    assert np.array_equal(got, want)


def test_coset_lde_natural_order(device):
    tr = O.rand_matrix(5, 256, 7)
    want = O.coset_lde_batch(tr, 2, 7, bit_reverse=False)
    got = G.Dft(device).coset_lde_batch(tr, 2, 7, bit_reverse_rows=False)
    assert np.array_equal(got, want)


# ---------------------------------------------------------------- LMCS
SCEN_P2 = [[(1, 1)], [(1, 7)], [(2, 3), (4, 5), (8, 8)], [(1, 5), (1, 3), (2, 7), (4, 1), (8, 9)],
           [(2, 7), (4, 8), (8, 11)], [(4, 13), (8, 25)], [(1, 16), (2, 15), (4, 16), (8, 22)],
           [(4, 7), (4, 8), (8, 11), (8, 16)], [(8, 7)], [(1024, 64)], [(64, 3), (4096, 40), (4096, 1)]]
SCEN_KECCAK = [[(1, 1)], [(2, 16), (4, 17), (8, 18)], [(4, 34), (8, 35)], [(512, 100)], [(16, 2600)]]


@pytest.mark.parametrize("kind,scen", [(G.HASH_POSEIDON2, s) for s in SCEN_P2] + [(G.HASH_KECCAK, s) for s in SCEN_KECCAK])
def test_lmcs_tree_matches_oracle(device, kind, scen):
    mats = _mats(77, scen)
    want = O.Tree(kind, mats, alignment=O.hash_alignment(kind))
    cfg = G.LmcsConfig(device, kind)
    got = cfg.build_aligned_tree(mats)
    assert got.root().tolist() == want.root().tolist()
    H = got.height()
    for l in range(H.bit_length()):
        assert np.array_equal(got.layer(l), want.layer(l)), l
    for idx in {0, H - 1, H // 3}:
        assert got.rows(idx).tolist() == want.rows(idx).tolist()
    assert got.widths() == [O.aligned_len(w, O.hash_alignment(kind)) for _, w in scen]
    # prove_batch hints: identical transcript streams
    idx = [H - 1, 0, H // 2, H // 2, H // 5]
    gt = G.ProverTranscript(0 if kind == G.HASH_POSEIDON2 else 1)
    ot = O.Transcript(0 if kind == G.HASH_POSEIDON2 else 1)
    got.prove_batch(idx, gt)
    want.prove_batch(idx, ot)
    gf, gc = gt.data()
    of, oc = ot.data()
    assert np.array_equal(gf, of) and np.array_equal(gc, oc)


def test_lmcs_unaligned_and_errors(device):
    cfg = G.LmcsConfig(device, G.HASH_POSEIDON2)
    mats = _mats(3, [(4, 3), (8, 5)])
    t = cfg.build_tree(mats)
    assert t.root().tolist() == O.Tree(O.HASH_P2, mats).root().tolist()
    assert t.widths() == [3, 5]
    with pytest.raises(G.P3GError) as e:
        cfg.build_tree(_mats(3, [(8, 3), (4, 5)]))
    assert e.value.code == 3  # unsorted
    with pytest.raises(G.P3GError) as e:
        cfg.build_tree(_mats(3, [(6, 3)]))
    assert e.value.code == 2  # not a power of two
    with pytest.raises(G.P3GError) as e:
        cfg.build_tree([])
    assert e.value.code == 6  # empty
    with pytest.raises(G.P3GError):
        t.rows(8)


def test_device_resident_inputs(device):
    mats = _mats(11, [(256, 24), (1024, 9)])
    dms = [device.upload(m) for m in mats]
    cfg = G.LmcsConfig(device, G.HASH_POSEIDON2)
    t = cfg.build_aligned_tree(dms)
    assert t.root().tolist() == O.Tree(O.HASH_P2, mats, alignment=8).root().tolist()


# ---------------------------------------------------------------- commit_traces
@pytest.mark.parametrize("kind,shapes,lb", [
    (G.HASH_POSEIDON2, [(1 << 10, 8)], 1),
    (G.HASH_POSEIDON2, [(16, 3), (256, 10), (1024, 50), (1024, 16)], 2),
    (G.HASH_POSEIDON2, [(2, 1), (8, 20)], 3),
    (G.HASH_KECCAK, [(64, 40), (256, 17)], 1),
])
def test_commit_traces_matches_oracle(device, kind, shapes, lb):
    tr = _mats(31, shapes)
    want = O.commit_traces(kind, tr, lb)
    got = G.commit_traces(G.LmcsConfig(device, kind), tr, lb)
    assert got.root().tolist() == want.root().tolist()
    H = got.height()
    assert got.rows(H - 3).tolist() == want.rows(H - 3).tolist()


# ---------------------------------------------------------------- full PCS opening
CASES = [
    # (kind, shapes, PcsParams args)  — C1 first: 2^14 x 32, blowup 1, arity 2, final degree 8, 64 queries
    (G.HASH_POSEIDON2, [(1 << 14, 32)], (1, 1, 3, 0, 0, 64, 0)),
    (G.HASH_POSEIDON2, [(1 << 10, 16)], (2, 2, 4, 0, 0, 100, 0)),
    (G.HASH_POSEIDON2, [(1 << 9, 9)], (1, 3, 2, 0, 0, 30, 0)),
    (G.HASH_POSEIDON2, [(8, 2), (64, 11), (256, 7), (256, 30)], (1, 1, 1, 0, 0, 40, 0)),
    (G.HASH_POSEIDON2, [(1 << 8, 5)], (3, 2, 0, 1, 2, 20, 3)),
    (G.HASH_KECCAK, [(1 << 9, 40)], (1, 1, 2, 0, 0, 50, 0)),
    (G.HASH_KECCAK, [(32, 18), (128, 17)], (2, 2, 1, 0, 1, 25, 2)),
]


@pytest.mark.parametrize("kind,shapes,params", CASES)
def test_prove_matches_oracle_bit_exact(device, kind, shapes, params):
    tr = _mats(2025, shapes)
    op = O.PcsParams(*params)
    root_o, z_o, f_o, c_o, tree_o = O.pcs_trace_prove(op, kind, tr)
    gp = G.PcsParams(*params)
    root_g, z_g, f_g, c_g = G.prove(G.LmcsConfig(device, kind), gp, tr)
    assert root_g.tolist() == root_o.tolist()
    assert z_g == tuple(z_o)
    assert f_g.size == f_o.size and c_g.shape == c_o.shape
    assert np.array_equal(f_g, f_o)
    assert np.array_equal(c_g, c_o)
    assert G.proof_bytes(f_g, c_g) == G.proof_bytes(f_o, c_o)
    logL = tree_o.height().bit_length() - 1
    assert O.pcs_trace_verify(op, kind, root_g, [w for _, w in shapes], logL, f_g, c_g) == 0


def test_open_with_channel_multi_tree_and_points(device):
    # two commitment groups (main + aux style) opened at three points
    cfg = G.LmcsConfig(device, G.HASH_POSEIDON2)
    tr_a = _mats(5, [(64, 6), (128, 10)])
    tr_b = _mats(9, [(128, 3)])
    ga, gb = G.commit_traces(cfg, tr_a, 1), G.commit_traces(cfg, tr_b, 1)
    oa, ob = O.commit_traces(O.HASH_P2, tr_a, 1), O.commit_traces(O.HASH_P2, tr_b, 1)
    z = [(11, 12), (13, 14), (15, 16)]
    params = (1, 2, 1, 0, 0, 17, 0)
    gt, ot = G.ProverTranscript(), O.Transcript()
    G.open_with_channel(G.PcsParams(*params), cfg, 8, z, [ga, gb], gt)
    O.open_with_channel(O.PcsParams(*params), O.HASH_P2, [oa, ob], 8, z, ot)
    gf, gc = gt.data()
    of, oc = ot.data()
    assert np.array_equal(gf, of) and np.array_equal(gc, oc)
    v = O.Transcript()
    v.load(gf, gc)
    rc = O.verify(O.PcsParams(*params), O.HASH_P2, [ga.root(), gb.root()], [[8, 16], [8]], 8, z, v)
    assert rc == 0


def test_repeated_proofs_are_identical(device):
    # the allocation cache recycles buffers between proofs: results must not depend on it
    cfg = G.LmcsConfig(device, G.HASH_POSEIDON2)
    tr = _mats(44, [(1 << 11, 12)])
    p = G.PcsParams(1, 2, 2, 0, 0, 30, 0)
    a = G.prove(cfg, p, tr)
    b = G.prove(cfg, p, tr)
    assert np.array_equal(a[2], b[2]) and np.array_equal(a[3], b[3])
