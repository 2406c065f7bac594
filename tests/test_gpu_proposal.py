"""Narrowed search around proposals: exactness when a window holds the
exhaustive optimum, the exhaustive fallback, and the oracle of the windows."""
import numpy as np
import pytest

from isac_amd import binding, synth
from oracle import isac_oracle as O

pytestmark = pytest.mark.gpu

R, RNG = 35, 3.0


def _data(P=300, N=30, seed=1):
    import torch
    sbj, ref, _ = synth.make_dataset_torch(P, N, 76, 34.0, 0.05, 3.0, False, seed_refs=seed,
                                           seed_particles=seed + 1, seed_noise=seed + 2)
    return sbj, ref


def _exhaustive(sbj, ref, acc=None):
    peak, d = binding.mref_align_device(sbj, ref, R, 1.0, RNG, RNG, acc_xy=acc)
    return d, peak


def _labels_prop(d, peak):
    from isac_amd import posenet
    lab = posenet.labels_from_search(d, peak, RNG)
    m = lab[:, 1] // 72
    b = lab[:, 1] % 72
    import torch
    return torch.stack([lab[:, 0], m, b, lab[:, 2]], 1).to(torch.int32)


def test_window_on_the_optimum_is_bit_identical(gpu):
    import torch
    sbj, ref = _data()
    acc = torch.from_numpy(np.random.default_rng(0).uniform(-1, 1, (300, 2)).astype(np.float32)).cuda()
    d, peak = _exhaustive(sbj, ref, acc)
    bref, bval, bpeak = binding.best_of_exhaustive(d, peak)
    prop = _labels_prop(d, peak)
    # accumulated shifts move the grid: rebuild the shift index from the peak shift
    sx = bpeak[:, 1] - acc[:, 0]
    sy = bpeak[:, 2] - acc[:, 1]
    prop[:, 3] = (torch.round(sx + RNG) * 7 + torch.round(sy + RNG)).to(torch.int32)
    # pad with junk proposals that must not matter (out-of-range and duplicate)
    junk = torch.tensor([-1, 0, 0, 0], dtype=torch.int32, device="cuda").expand(300, 1, 4)
    props = torch.cat([prop[:, None, :], junk, prop[:, None, :]], 1).contiguous()
    o = binding.mref_proposal_device(sbj, ref, R, 1.0, RNG, RNG, props, threshold=-3.0e38,
                                     acc_xy=acc)
    torch.cuda.synchronize()
    # a maximum on its window's edge falls back to the exhaustive search, which
    # returns the same optimum: every output is still the exhaustive one
    assert o["n_fallback"] < 300
    assert torch.equal(o["best_ref"], bref)
    assert torch.equal(o["best_val"], bval)
    assert torch.equal(o["best_peak"], bpeak)


def test_fallback_is_the_exhaustive_search(gpu):
    import torch
    sbj, ref = _data(P=200, N=20, seed=5)
    d, peak = _exhaustive(sbj, ref)
    bref, bval, bpeak = binding.best_of_exhaustive(d, peak)
    prop = _labels_prop(d, peak)
    prop[:, 0] = (prop[:, 0] + 1) % 20        # wrong references everywhere
    o = binding.mref_proposal_device(sbj, ref, R, 1.0, RNG, RNG, prop[:, None, :].contiguous(),
                                     threshold=3.0e38)
    torch.cuda.synchronize()
    assert o["n_fallback"] == 200 and int(o["fallback"].sum()) == 200
    assert torch.equal(o["best_ref"], bref)
    assert torch.equal(o["best_val"], bval)
    assert torch.equal(o["best_peak"], bpeak)
    # a calibrated threshold sends exactly the particles below it
    thr = float(torch.quantile(bval, 0.5))
    o2 = binding.mref_proposal_device(sbj, ref, R, 1.0, RNG, RNG, prop[:, None, :].contiguous(),
                                      threshold=thr)
    torch.cuda.synchronize()
    fb = o2["fallback"].bool()
    assert o2["n_fallback"] == int(fb.sum())
    assert torch.equal(o2["best_val"][fb], bval[fb])
    assert bool(torch.all(o2["best_val"][~fb] >= thr))


def test_mfma_fallback_equals_the_exhaustive_search(gpu):
    # level 3 on MFMA: every particle falls back; the 8 best references by the
    # FP16 search are re-scored with the FP32 search's own arithmetic, so the
    # best reference, value and peak are the exhaustive search's, except where the
    # runner-up is within 1e-5 (there FP64 decides)
    import torch
    sbj, ref = _data(P=200, N=20, seed=5)
    d, peak = _exhaustive(sbj, ref)
    bref, bval, bpeak = binding.best_of_exhaustive(d, peak)
    prop = _labels_prop(d, peak)
    prop[:, 0] = (prop[:, 0] + 1) % 20        # wrong references everywhere
    for form in ("fp16", "bf16"):
        o = binding.mref_proposal_device(sbj, ref, R, 1.0, RNG, RNG, prop[:, None, :].contiguous(),
                                         threshold=3.0e38, fallback_form=form)
        torch.cuda.synchronize()
        assert o["n_fallback"] == 200 and int(o["fallback"].sum()) == 200
        top2 = torch.topk(d, 2, dim=0).values
        tie = (top2[0] - top2[1]) <= 1e-5 * top2[0].abs()
        ok = ~tie
        print(form, "near-ties", int(tie.sum()), flush=True)
        assert torch.equal(o["best_ref"][ok], bref[ok])
        assert torch.equal(o["best_val"][ok], bval[ok])
        assert torch.equal(o["best_peak"][ok], bpeak[ok])


def test_narrowed_search_matches_window_oracle(gpu):
    """Random proposals (near and far from the optimum) against the FP64 oracle
    of the windows: same reference, mirror and shift, psi within 0.5 degrees,
    value within 1e-4 relative."""
    import torch
    P, N = 24, 6
    ds = synth.make_dataset(P, N, 76, 34.0, 0.1, 3.0, False, seed_refs=3)
    rng = np.random.default_rng(9)
    prop = np.stack([rng.integers(0, N, (P, 4)), rng.integers(0, 2, (P, 4)),
                     rng.integers(0, 72, (P, 4)), rng.integers(0, 49, (P, 4))], 2).astype(np.int32)
    acc = np.zeros((P, 2), np.float32)
    w = O.windowed_best(ds.sbj, ds.ref, acc, R, RNG, RNG, 1.0, prop)
    s, r = torch.from_numpy(ds.sbj).cuda(), torch.from_numpy(ds.ref).cuda()
    pt = torch.from_numpy(prop).cuda()
    # edge maxima flagged like the oracle's (those particles then fall back)
    oe = binding.mref_proposal_device(s, r, R, 1.0, RNG, RNG, pt, threshold=-3.0e38)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(oe["fallback"].cpu().numpy().astype(bool), ~w["interior"])
    # the windowed maxima themselves, with the edge test off
    o = binding.mref_proposal_device(s, r, R, 1.0, RNG, RNG, pt, threshold=-3.0e38,
                                     edge_fallback=False)
    torch.cuda.synchronize()
    assert o["n_fallback"] == 0
    g = {k: v.cpu().numpy() for k, v in o.items() if torch.is_tensor(v)}
    np.testing.assert_array_equal(g["best_ref"], w["ref"])
    np.testing.assert_allclose(g["best_val"], w["d"], rtol=1e-4)
    np.testing.assert_array_equal(g["best_peak"][:, 3], w["mirror"])
    np.testing.assert_allclose(g["best_peak"][:, 1:3], np.stack([w["sx"], w["sy"]], 1), atol=1e-5)
    dpsi = np.abs((g["best_peak"][:, 0] - w["angle"] + 180.0) % 360.0 - 180.0)
    assert dpsi.max() <= 0.5, dpsi.max()


def test_inputs_written_by_queued_work_are_seen(gpu):
    """The library queues its work on the caller's stream (here torch's default
    stream) behind the work that produces its inputs: proposals computed at the
    end of a long queued computation, with no synchronisation before the call,
    must give the same result as proposals that were ready."""
    import torch
    sbj, ref = _data(P=2000, N=20, seed=9)
    d, peak = _exhaustive(sbj, ref)
    torch.cuda.synchronize()
    bref, bval, bpeak = binding.best_of_exhaustive(d, peak)
    prop = _labels_prop(d, peak)[:, None, :].contiguous()
    torch.cuda.synchronize()
    big = torch.randn(8192, 8192, device="cuda")
    for _ in range(4):
        big = big @ big
    late = (prop + (big.abs() < 0).sum().to(torch.int32)).contiguous()   # equal to prop, written late
    o = binding.mref_proposal_device(sbj, ref, R, 1.0, RNG, RNG, late, threshold=-3.0e38)
    torch.cuda.synchronize()
    assert torch.equal(o["best_val"], bval) and torch.equal(o["best_peak"], bpeak)



@pytest.mark.parametrize("K", [3, 4, 8])
def test_level2_is_the_exhaustive_search_over_proposed_refs(gpu, K):
    """Particles whose windowed maximum is on an edge get every shift and psi of
    their proposed references: bit-identical to the exhaustive search's best
    over those references; the others keep the windowed maximum.  K = 4 and 8
    run the candidate form of the exhaustive kernel (one and two wavefronts per
    list), K = 3 the per-item kernel.  K = 8 lists the distinct references of 8
    proposals among 6 references, so some lists have empty slots and some an
    empty second group of four."""
    import torch
    P, N = 24, 6
    ds = synth.make_dataset(P, N, 76, 34.0, 0.1, 3.0, True, seed_refs=4)
    rng = np.random.default_rng(12)
    prop = np.stack([rng.integers(0, N, (P, K)), rng.integers(0, 2, (P, K)),
                     rng.integers(0, 72, (P, K)), rng.integers(0, 49, (P, K))], 2).astype(np.int32)
    acc = np.zeros((P, 2), np.float32)
    s, r = torch.from_numpy(ds.sbj).cuda(), torch.from_numpy(ds.ref).cuda()
    d, peak = _exhaustive(s, r)
    # every other particle also gets a proposal on its exhaustive optimum
    lab = _labels_prop(d, peak).cpu().numpy()
    prop[::2, 0] = lab[::2]
    w = O.windowed_best(ds.sbj, ds.ref, acc, R, RNG, RNG, 1.0, prop)
    o = binding.mref_proposal_device(s, r, R, 1.0, RNG, RNG, torch.from_numpy(prop).cuda(),
                                     threshold=-3.0e38, level2=True)
    torch.cuda.synchronize()
    edge = ~w["interior"]
    assert o["n_level2"] == int(edge.sum()) and o["n_fallback"] == 0
    assert 0 < edge.sum() < P
    d, pk = d.cpu().numpy(), peak.cpu().numpy().reshape(N, P, 4)
    g = {k: v.cpu().numpy() for k, v in o.items() if torch.is_tensor(v)}
    for p in np.nonzero(edge)[0]:
        refs = sorted(set(int(x) for x in prop[p, :, 0]))
        best = max(refs, key=lambda n: (d[n, p], -n))
        assert g["best_ref"][p] == best
        assert g["best_val"][p] == d[best, p]
        np.testing.assert_array_equal(g["best_peak"][p], pk[best, p])
    keep = ~edge
    np.testing.assert_array_equal(g["best_ref"][keep], w["ref"][keep])
    np.testing.assert_allclose(g["best_val"][keep], w["d"][keep], rtol=1e-4)
    # everything flagged and below an infinite threshold: the exhaustive search
    o2 = binding.mref_proposal_device(s, r, R, 1.0, RNG, RNG, torch.from_numpy(prop).cuda(),
                                      threshold=3.0e38, level2=True)
    torch.cuda.synchronize()
    assert o2["n_level2"] == P and o2["n_fallback"] == P
    bref, bval, _ = binding.best_of_exhaustive(torch.from_numpy(d).cuda(),
                                               torch.from_numpy(pk.reshape(N, 4 * P)).cuda())
    assert torch.equal(o2["best_ref"], bref) and torch.equal(o2["best_val"], bval)


def test_levels_across_particle_batches(gpu):
    """100 000 particles x 100 references at 76 px: levels 1 and 2 run in several
    particle batches and level 3 gathers the fallback particles into batches of
    its own.  With every particle sent to level 2 and then to the FP32 fallback
    (threshold above every value), the result is the exhaustive search's, bit for
    bit, for every particle."""
    import torch
    from bench import WORKLOADS
    _, N, D, radius, snr, max_shift, int_shifts, _, _ = WORKLOADS["p20k_n100_d76"]
    P = 100000
    sbj, ref, _ = synth.make_dataset_torch(P, N, D, radius, snr, max_shift, int_shifts,
                                           torch.device("cuda:0"))
    d, peak = _exhaustive(sbj, ref)
    bref, bval, bpeak = binding.best_of_exhaustive(d, peak)
    g = torch.Generator(device="cpu").manual_seed(3)
    prop = torch.stack([torch.randint(0, N, (P, 8), generator=g), torch.randint(0, 2, (P, 8), generator=g),
                        torch.randint(0, 72, (P, 8), generator=g), torch.randint(0, 49, (P, 8), generator=g)],
                       2).to(torch.int32).cuda()
    o = binding.mref_proposal_device(sbj, ref, R, 1.0, RNG, RNG, prop, threshold=3.0e38, level2=True)
    torch.cuda.synchronize()
    assert o["n_level2"] == P and o["n_fallback"] == P
    assert torch.equal(o["best_ref"], bref)
    assert torch.equal(o["best_val"], bval)
    assert torch.equal(o["best_peak"], bpeak)
