"""Batched uneven subsets: one captured graph with a persistent work-queue kernel
must reproduce, bit for bit, one isac_mref_align_device call per subset."""
import os

import numpy as np
import pytest

from isac_amd import binding, synth
from oracle import coracle
from parity import compare_mref

pytestmark = pytest.mark.gpu

# (P, D, N, shift range); rings out to D/2 - 4
SHAPES = [(37, 64, 9, 2.0), (20, 76, 23, 3.0), (11, 96, 5, 4.0), (29, 128, 13, 2.0),
          (1, 64, 1, 0.0), (45, 76, 17, 3.0)]


def _subsets(torch, seed=0):
    subs = []
    for i, (P, D, N, rng) in enumerate(SHAPES):
        sbj, ref, _ = synth.make_dataset_torch(P, N, D, 0.42 * D, 0.1, max(rng, 1.0), True,
                                               seed_refs=10 + i, seed_particles=20 + i + seed,
                                               seed_noise=30 + i + seed)
        acc = None
        if i % 2 == 0:
            g = np.random.default_rng(i + seed).uniform(-1.5, 1.5, (P, 2)).astype(np.float32)
            acc = torch.from_numpy(g).cuda()
        subs.append({"sbj": sbj, "ref": ref, "ring_num": D // 2 - 4, "shift_rng": rng,
                     "acc_xy": acc})
    return subs


def _per_subset(subs):
    out = []
    for s in subs:
        peak, d = binding.mref_align_device(s["sbj"], s["ref"], s["ring_num"], 1.0, s["shift_rng"],
                                            s["shift_rng"], acc_xy=s["acc_xy"])
        out.append((peak.clone(), d.clone()))
    return out


def _assert_same(a, b, what):
    import torch
    torch.cuda.synchronize()
    for i, ((pa, da), (pb, db)) in enumerate(zip(a, b)):
        assert torch.equal(pa, pb), f"{what}: subset {i} peak_list differs"
        assert torch.equal(da, db), f"{what}: subset {i} d differs"


def test_batched_graph_matches_per_subset(gpu):
    import torch
    subs = _subsets(torch)
    want = _per_subset(subs)
    bm = binding.BatchedMref(subs)
    got = [(p.clone(), d.clone()) for p, d in bm.run(binding.BATCH_GRAPH)]
    rep = bm.report()
    print("batched report", rep)
    assert rep["groups"] == 1 and not rep["replayed"] and rep["waves_per_workgroup"] in (1, 2, 4)
    _assert_same(got, want, "graph")
    got2 = [(p.clone(), d.clone()) for p, d in bm.run(binding.BATCH_PER_SUBSET)]
    _assert_same(got2, want, "per-subset mode")
    # replay after new data is written into the same input tensors
    fresh = _subsets(torch, seed=100)
    for s, f in zip(subs, fresh):
        s["sbj"].copy_(f["sbj"])
    want_fresh = _per_subset(subs)
    got3 = [(p.clone(), d.clone()) for p, d in bm.run(binding.BATCH_GRAPH)]
    assert bm.report()["replayed"]
    _assert_same(got3, want_fresh, "graph replay")


def test_batched_oracle_anchor(gpu):
    """One small subset of the batch against the FP64-referenced oracle."""
    import torch
    subs = _subsets(torch)[:2]
    bm = binding.BatchedMref(subs)
    outs = bm.run(binding.BATCH_GRAPH)
    torch.cuda.synchronize()
    for s, (peak, d) in zip(subs, outs):
        acc = s["acc_xy"].cpu().numpy() if s["acc_xy"] is not None else \
            np.zeros((s["sbj"].shape[0], 2), np.float32)
        ref = coracle.mref(s["sbj"].cpu().numpy(), s["ref"].cpu().numpy(), acc, s["ring_num"],
                           s["shift_rng"], s["shift_rng"], 1.0)
        print(compare_mref(peak.cpu().numpy(), d.cpu().numpy(), ref, "batched"))


def test_batched_many_groups(gpu, monkeypatch):
    """A 4 MB workspace cuts subsets into slices over several groups."""
    import torch
    monkeypatch.setenv("ISAC_BATCH_WORKSPACE_MB", "4")
    subs = _subsets(torch)
    want = _per_subset(subs)
    bm = binding.BatchedMref(subs)
    got = [(p.clone(), d.clone()) for p, d in bm.run(binding.BATCH_GRAPH)]
    rep = bm.report()
    print("batched report (4 MB groups)", rep)
    assert rep["groups"] > 3
    _assert_same(got, want, "graph, many groups")
