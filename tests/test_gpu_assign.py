"""Equal-size group assignment on the GPU against the oracle's greedy."""
import numpy as np
import pytest

from isac_amd import binding
from oracle import isac_oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N,P,kind", [(20, 2000, "normal"), (13, 997, "normal"), (7, 700, "ties"),
                                      (5, 123, "nan"), (100, 20000, "normal"), (1, 17, "normal"),
                                      (9, 9, "normal")])
def test_assign_groups_matches_greedy(gpu, N, P, kind):
    import torch
    rng = np.random.default_rng(N * 1000 + P)
    if kind == "ties":
        d = rng.integers(0, 5, (N, P)).astype(np.float32)
    else:
        d = rng.standard_normal((N, P)).astype(np.float32)
    if kind == "nan":
        d[rng.random((N, P)) < 0.1] = np.nan
    want = O.assign_groups(d)
    got = binding.assign_groups_device(torch.from_numpy(d).cuda()).cpu().numpy()
    np.testing.assert_array_equal(got, want)


def test_assign_groups_pitch_and_host_form(gpu):
    import torch
    rng = np.random.default_rng(5)
    d = rng.standard_normal((8, 301)).astype(np.float32)
    want = O.assign_groups(d[:, :290])
    got = binding.assign_groups_device(torch.from_numpy(d).cuda(), n_img=290).cpu().numpy()
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(binding.assign_groups(d[:, :290]), want)


def test_assign_groups_on_alignment_scores(gpu):
    """On real multi-reference scores: classes of equal size, stable."""
    import torch
    from isac_amd import synth
    sbj, ref, _ = synth.make_dataset_torch(600, 12, 64, 28.0, 0.1, 2.0, True)
    peak, d = binding.mref_align_device(sbj, ref, 29, 1.0, 2.0, 2.0)
    got = binding.assign_groups_device(d).cpu().numpy()
    np.testing.assert_array_equal(got, O.assign_groups(d.cpu().numpy()))
