"""Writes the small golden fixtures of tests/test_golden.py with the FP64 oracle.

Run from the repository root:  python tests/golden/make_golden.py
Inputs are seeded synthetic blob particles; outputs are the oracle's.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "gpu-isac_amd"), ROOT]

from isac_amd import synth  # noqa: E402
from oracle import isac_oracle as O  # noqa: E402


def mref():
    ds = synth.make_dataset(12, 3, 32, 13.0, 0.3, 2, True, seed_refs=41, seed_particles=42, seed_noise=43)
    acc = np.random.default_rng(44).uniform(-1, 1, (12, 2)).astype(np.float32)
    r = O.multi_ref_align(ds.sbj, ds.ref, acc, 13, 2.0, 2.0, 1.0)
    np.savez_compressed(os.path.join(HERE, "golden_mref.npz"), sbj=ds.sbj, ref=ds.ref, acc=acc,
                        ring_num=13, shift_rng=2.0, shift_step=1.0,
                        **{k: np.asarray(v) for k, v in r.items()})


def prealign():
    ds = synth.make_dataset(10, 1, 32, 13.0, 0.5, 1, True, seed_refs=51, seed_particles=52, seed_noise=53)
    acc = np.random.default_rng(54).uniform(-1, 1, (10, 2)).astype(np.float32)
    sh = O.shift_grid(1.0, 1.0, 1.0)
    tab = O.single_ref_table(ds.sbj, ds.ref, np.zeros(10, int), acc.astype(np.float64), 13, sh)
    p = O.decode_single_ref(tab, acc.astype(np.float64), sh, 32, 13)
    np.savez_compressed(os.path.join(HERE, "golden_prealign.npz"), sbj=ds.sbj, ref=ds.ref, acc=acc,
                        ring_num=13, shift_rng=1.0, **{k: np.asarray(v) for k, v in p.items()})


def reffree():
    blobs = synth.make_blobs(2, 13.0, 61)
    rng = np.random.default_rng(62)
    cid = np.repeat(np.arange(2), 6)
    sbj = synth.render_particles(blobs, 32, cid, rng.uniform(0, 360, 12), rng.uniform(-1, 1, (12, 2)),
                                 rng.integers(0, 2, 12), 2.0, 13.0, 63)
    avg = np.stack([sbj[cid == k].mean(0) for k in range(2)]).astype(np.float32)
    sh = O.shift_grid(1.0, 1.0, 1.0)
    acc = np.zeros((12, 2))
    refs = avg.astype(np.float64)
    n_iter, fh, ff = 2, 0.3, 0.2
    for i in range(n_iter):
        refs = O.tangent_filter(refs, 0.1 + (fh - 0.1) * i / float(n_iter - 1), ff)
        prm, refs = O.ref_free_iteration(sbj, refs, cid, acc, 13, sh, 32)
        acc = np.stack([prm["shift_x"], prm["shift_y"]], 1)
    np.savez_compressed(os.path.join(HERE, "golden_reffree.npz"), sbj=sbj, cid=cid, avg=avg,
                        ring_num=13, shift_rng=1.0, n_iter=n_iter, cutoff=fh, falloff=ff,
                        refs_out=refs, **{k: np.asarray(v) for k, v in prm.items()})


if __name__ == "__main__":
    mref()
    prealign()
    reffree()
    print("written:", sorted(f for f in os.listdir(HERE) if f.endswith(".npz")))
