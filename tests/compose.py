"""The opening composed from the library's step-level calls, in the order of
open_with_channel (p3-miden-lifted-fri/src/prover.rs:35-116): DEEP
evaluations sent per point, grind, alpha and beta, the DEEP quotient, FRI
rounds (commit, send root, grind, beta, fold), the final polynomial, query
grinding and sampling, then the batch openings of every trace tree and every
round tree.  Shared by the step-level parity tests."""
import numpy as np

import oracle as O
import p3gpu as G


def bitrev(i, bits):
    return int(format(i, f"0{bits}b")[::-1], 2) if bits else 0


def aligned_per_point(ev, widths, alignment):
    """evals[column][point][2] -> per point the flat aligned slice the prover sends."""
    out = []
    for j in range(ev.shape[1]):
        flat, off = [], 0
        for w in widths:
            flat.append(ev[off:off + w, j, :].reshape(-1))
            pad = O.aligned_len(w, alignment) - w
            if pad:
                flat.append(np.zeros(2 * pad, np.uint64))
            off += w
        out.append(np.concatenate(flat) if flat else np.zeros(0, np.uint64))
    return out


def compose_open(cfg, params, trees, widths, z, tx):
    """Runs the opening of `trees` (unaligned matrix widths `widths`) at points z on the library
    transcript tx, one step call at a time."""
    lb, la, _, fold_pow, deep_pow, nq, query_pow = params
    logL = int(trees[0].height()).bit_length() - 1
    ev = G.batch_eval_lifted(trees, z, lb)
    for flat in aligned_per_point(ev, widths, trees[0].alignment()):
        tx.send_field_slice(flat)
    tx.grind(deep_pow)
    alpha = tx.sample_algebra_element()
    beta = tx.sample_algebra_element()
    cur = G.deep_quotient(trees, z, lb, ev, alpha, beta)
    fdeg = O.PcsParams(*params).final_degree(logL)
    keep, ftrees = [], []
    while cur.height > fdeg << lb:
        t = G.fri_commit(cfg, cur, la)
        tx.send_commitment(t.root())
        tx.grind(fold_pow)
        b = tx.sample_algebra_element()
        keep.append(cur)
        ftrees.append(t)
        cur = G.fri_fold(cur, la, b)
    tx.send_field_slice(G.fri_final_poly(cur.download()[:fdeg]).reshape(-1))
    tx.grind(query_pow)
    idx = [bitrev(tx.sample_bits(logL), logL) for _ in range(nq)]
    for t in trees:
        t.prove_batch(idx, tx)
    for r, t in enumerate(ftrees):
        t.prove_batch([i >> (la * (r + 1)) for i in idx], tx)
