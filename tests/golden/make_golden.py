"""Writes the small golden fixtures under tests/golden/.

Every expected output here comes from a closed form or a brute-force
computation in numpy, never from the oracle or the HIP library:

* philox_kat.json  -- Philox4x32-10 known-answer vectors published with
  Random123 (Salmon et al., SC'11; kat_vectors, "philox4x32 10" lines).
* cos_gram.npz     -- LowRankTensor::test_gram_matrix_computation's KAT
  (low_rank_tensor.hpp:44-54): U[i][j] = cos(i*R + j) (distmat.hpp:140-151)
  and its Gram U^T U, for I = 37, R = 8.
* tiny_mttkrp.npz  -- a 7 x 6 x 5 sparse tensor, fixed samples, a design
  matrix and the sampled MTTKRP Y = sum over each sampled fiber of x*design,
  by materialising the dense tensor.
* tiny_leverage.npz -- factors of a 5 x 4 x 3 tensor at R = 3 and the exact
  KRP leverage scores of every (i0, i2) row for target mode 1, from a QR of
  the materialised Khatri-Rao product (the definition used by
  experimental/exact_leverage_scores.py:147-156).

Run from the repository root:  python tests/golden/make_golden.py
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def philox():
    vecs = [
        {"ctr": [0, 0, 0, 0], "key": [0, 0], "out": [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]},
        {"ctr": [0xffffffff] * 4, "key": [0xffffffff] * 2, "out": [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]},
        {"ctr": [0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], "key": [0xa4093822, 0x299f31d0],
         "out": [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]},
    ]
    with open(os.path.join(HERE, "philox_kat.json"), "w") as f:
        json.dump(vecs, f, indent=1)


def cos_gram():
    I, R = 37, 8
    i = np.arange(I)[:, None]
    j = np.arange(R)[None, :]
    U = np.cos((i * R + j).astype(np.float64))
    np.savez(os.path.join(HERE, "cos_gram.npz"), U=U, G=U.T @ U)


def tiny_mttkrp():
    rng = np.random.default_rng(2024)
    dims = (7, 6, 5)
    dense = np.zeros(dims)
    mask = rng.random(dims) < 0.35
    dense[mask] = rng.random(mask.sum()) + 0.5
    idx = np.argwhere(dense != 0).astype(np.uint32)
    vals = dense[tuple(idx.T)]
    R = 4
    out = {"dims": np.array(dims), "idx": idx, "vals": vals}
    for mode in range(3):
        others = [k for k in range(3) if k != mode]
        # all distinct tuples of the other two modes, sorted (dedup order)
        tup = np.array([(a, b) for a in range(dims[others[0]]) for b in range(dims[others[1]])])
        pick = rng.choice(len(tup), size=len(tup) // 2, replace=False)
        tup = tup[np.sort(pick)]
        samples = np.zeros((len(tup), 3), np.uint32)
        samples[:, others[0]] = tup[:, 0]
        samples[:, others[1]] = tup[:, 1]
        design = rng.standard_normal((len(tup), R))
        Y = np.zeros((dims[mode], R))
        hits = 0
        for c, (a, b) in enumerate(tup):
            sl = [slice(None)] * 3
            sl[others[0]] = a
            sl[others[1]] = b
            fiber = dense[tuple(sl)]
            hits += int((fiber != 0).sum())
            Y += np.outer(fiber, design[c])
        out[f"samples{mode}"] = samples
        out[f"design{mode}"] = design
        out[f"Y{mode}"] = Y
        out[f"hits{mode}"] = np.array(hits)
    np.savez(os.path.join(HERE, "tiny_mttkrp.npz"), **out)


def tiny_leverage():
    rng = np.random.default_rng(7)
    dims, R, j = (5, 4, 3), 3, 1
    U = [rng.standard_normal((d, R)) for d in dims]
    # KRP rows ordered (i0, i2), i0 major
    K = np.einsum("ar,br->abr", U[0], U[2]).reshape(-1, R)
    Q, _ = np.linalg.qr(K)
    lev = (Q ** 2).sum(axis=1)
    np.savez(os.path.join(HERE, "tiny_leverage.npz"), U0=U[0], U1=U[1], U2=U[2], j=np.array(j), lev=lev)


if __name__ == "__main__":
    philox()
    cos_gram()
    tiny_mttkrp()
    tiny_leverage()
    print("fixtures written to", HERE)
