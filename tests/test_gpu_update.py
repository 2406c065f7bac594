"""Reference update on the GPU (isac_update_references_device) against the oracle."""
import numpy as np
import pytest

from isac_amd import binding, synth
from oracle import isac_oracle as O

pytestmark = pytest.mark.gpu


def _peaks_from_search(ds, R, rng):
    import torch
    sbj = torch.from_numpy(ds.sbj).cuda()
    ref = torch.from_numpy(ds.ref).cuda()
    peak, d = binding.mref_align_device(sbj, ref, R, 1.0, rng, rng)
    return sbj, peak, d


def _check_against_oracle(sbj, peak, cls, N, fl, ff):
    import torch
    got = binding.update_references_device(sbj, peak, torch.from_numpy(cls).cuda(), N, fl, ff)
    refs, params, members, avg = (t.cpu().numpy() for t in got)
    w_refs, w_params, w_members, w_avg = O.update_references(sbj.cpu().numpy(), peak.cpu().numpy(),
                                                             cls, N, fl, ff)
    np.testing.assert_array_equal(members, w_members)
    np.testing.assert_allclose(avg, w_avg, atol=1e-5)
    np.testing.assert_allclose(params, w_params, atol=1e-4)
    scale = max(np.abs(w_refs).max(), 1e-30)
    assert np.abs(refs - w_refs).max() <= 2e-5 * scale
    return refs


@pytest.mark.parametrize("P,N,D,R,rng", [(300, 6, 64, 29, 2.0), (257, 5, 76, 35, 3.0),
                                         (120, 4, 63, 28, 2.0)])
def test_update_matches_oracle(gpu, P, N, D, R, rng):
    ds = synth.make_dataset(P, N, D, D / 2 - 4, 0.2, rng, False)
    sbj, peak, d = _peaks_from_search(ds, R, rng)
    cls = np.random.default_rng(P).integers(0, N, P).astype(np.int32)
    _check_against_oracle(sbj, peak, cls, N, 0.25, 0.2)


def test_update_with_an_empty_class(gpu):
    ds = synth.make_dataset(90, 4, 64, 28.0, 0.2, 2.0, True)
    sbj, peak, d = _peaks_from_search(ds, 29, 2.0)
    cls = np.where(ds.ref_idx == 2, 0, ds.ref_idx).astype(np.int32)   # class 2 empty
    refs = _check_against_oracle(sbj, peak, cls, 4, 0.3, 0.1)
    assert np.all(refs[2] == 0.0)


def test_update_rejects_bad_class_ids(gpu):
    import torch
    ds = synth.make_dataset(40, 3, 64, 28.0, 0.2, 2.0, True)
    sbj, peak, d = _peaks_from_search(ds, 29, 2.0)
    cls = torch.zeros(40, dtype=torch.int32, device="cuda")
    cls[7] = 3
    with pytest.raises(binding.GpuAlignError):
        binding.update_references_device(sbj, peak, cls, 3, 0.25, 0.2)
    binding.load().gpu_clear_error()


def _centred_back(refs, avg, ref, N):
    yy, xx = np.mgrid[0:64, 0:64] - 32
    inside = np.hypot(xx, yy) < 24
    out = []
    for k in range(N):
        back = O.filter_shift(refs[k:k + 1], 100.0, 0.01, [avg[k, 0]], [avg[k, 1]])[0]
        out.append(max(np.corrcoef(back[inside], ref[r][inside])[0, 1] for r in range(N)))
    return out


def test_update_recovers_the_references(gpu):
    """Noise-free particles in their generating classes with their exhaustive
    peaks: every class average, moved back by its centring shift, is its
    reference."""
    import torch
    P, N = 240, 4
    ds = synth.make_dataset(P, N, 64, 28.0, None, 2.0, True)
    sbj, peak, d = _peaks_from_search(ds, 29, 2.0)
    refs, params, members, avg = binding.update_references_device(
        sbj, peak, torch.from_numpy(ds.ref_idx.astype(np.int32)).cuda(), N, 100.0, 0.01)
    assert members.cpu().numpy().tolist() == np.bincount(ds.ref_idx, minlength=N).tolist()
    assert min(_centred_back(refs.cpu().numpy(), avg.cpu().numpy(), ds.ref, N)) > 0.999


def test_search_assign_update_chain(gpu):
    """One multi-reference iteration on the device: search, equal-size
    assignment (classes of P / N, so a few particles of over-full generating
    classes land elsewhere), update."""
    import torch
    P, N = 240, 4
    ds = synth.make_dataset(P, N, 64, 28.0, None, 2.0, True)
    sbj, peak, d = _peaks_from_search(ds, 29, 2.0)
    groups = binding.assign_groups_device(d)
    cls = binding.classes_from_assignment(groups.cpu().numpy(), N)
    assert sorted(np.asarray(groups.cpu().numpy()[:P]).tolist()) == list(range(P))
    refs, params, members, avg = binding.update_references_device(
        sbj, peak, torch.from_numpy(cls).cuda(), N, 100.0, 0.01)   # filter passes everything
    assert members.cpu().numpy().tolist() == [P // N] * N
    assert min(_centred_back(refs.cpu().numpy(), avg.cpu().numpy(), ds.ref, N)) > 0.99
