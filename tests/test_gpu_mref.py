"""GPU parity of multi-reference alignment against the oracle (MI355X)."""
import numpy as np
import pytest

from isac_amd import binding, synth
from oracle import coracle
from parity import compare_mref

pytestmark = pytest.mark.gpu


def _run(ds, R, rng, step, acc, label, **kw):
    peak, d = binding.multi_ref_align(ds.sbj, ds.ref, R, step, rng, rng, acc_xy=acc, **kw)
    ref = coracle.mref(ds.sbj, ds.ref, acc, R, rng, rng, step)
    st = compare_mref(peak, d, ref, label)
    print(label, st)
    return peak, d, ref


def test_small_workload_subset(gpu):
    ds = synth.dataset_p1k_n8_d64(n_particles=96)
    acc = np.zeros((96, 2), np.float32)
    _run(ds, 29, 2.0, 1.0, acc, "64 px, 8 refs, 96 particles")


def test_headline_workload_subset(gpu):
    ds = synth.dataset_p20k_n100_d76(n_particles=64, n_refs=100)
    acc = np.random.default_rng(11).uniform(-1.5, 1.5, (64, 2)).astype(np.float32)
    _run(ds, 35, 3.0, 1.0, acc, "76 px, 100 refs, 64 particles")


@pytest.mark.parametrize("D,R,nref,rng", [(32, 13, 5, 1.0), (48, 20, 7, 4.0), (96, 44, 3, 2.0),
                                          (128, 60, 9, 3.0), (64, 30, 1, 0.0)])
def test_shapes(gpu, D, R, nref, rng):
    ds = synth.make_dataset(21, nref, D, 0.42 * D, 0.2, max(rng, 1.0), True, seed_refs=D)
    acc = np.random.default_rng(D).uniform(-2, 2, (21, 2)).astype(np.float32)
    _run(ds, R, rng, 1.0, acc, f"D{D}-R{R}-N{nref}-s{rng}")


def test_single_particle_and_pitch(gpu):
    ds = synth.make_dataset(1, 6, 64, 28.0, 0.1, 2, True)
    acc = np.zeros((1, 2), np.float32)
    peak, d = binding.multi_ref_align(ds.sbj, ds.ref, 29, 1.0, 2.0, 2.0, acc_xy=acc, d_pitch=5)
    ref = coracle.mref(ds.sbj, ds.ref, acc, 29, 2.0, 2.0, 1.0)
    compare_mref(peak, d, ref, "P1")
    assert np.all(d[:, 1:] == 0.0), "columns beyond sbj_num must stay untouched"


def test_device_entry_matches_abi(gpu):
    import torch
    ds = synth.dataset_p1k_n8_d64(n_particles=40)
    acc = np.random.default_rng(3).uniform(-1, 1, (40, 2)).astype(np.float32)
    peak_h, d_h = binding.multi_ref_align(ds.sbj, ds.ref, 29, 1.0, 2.0, 2.0, acc_xy=acc)
    s = torch.from_numpy(ds.sbj).cuda()
    r = torch.from_numpy(ds.ref).cuda()
    a = torch.from_numpy(acc).cuda()
    peak_d, d_d = binding.mref_align_device(s, r, 29, 1.0, 2.0, 2.0, acc_xy=a)
    torch.cuda.synchronize()
    assert np.array_equal(peak_d.cpu().numpy(), peak_h)
    assert np.array_equal(d_d.cpu().numpy(), d_h)


def test_abi_in_many_particle_batches(gpu):
    """A tiny memory request forces one-particle batches: each batch's images
    are gathered and copied on the copy stream just before its launches; the
    results equal the default call's (batches of a quarter of the particles)
    bit for bit."""
    ds = synth.dataset_p1k_n8_d64(n_particles=37)
    acc = np.random.default_rng(5).uniform(-1, 1, (37, 2)).astype(np.float32)
    peak1, d1 = binding.multi_ref_align(ds.sbj, ds.ref, 29, 1.0, 2.0, 2.0, acc_xy=acc)
    peak2, d2 = binding.multi_ref_align(ds.sbj, ds.ref, 29, 1.0, 2.0, 2.0, acc_xy=acc,
                                        mem_request=1e-9)
    assert np.array_equal(peak1, peak2)
    assert np.array_equal(d1, d2)


def test_profile_counts_the_search_launches(gpu):
    """isac_profile_enable / isac_profile_read, the HIP-event timing behind the
    benchmark's roofline: every particle batch's search launch and polar launch
    is counted and timed while profiling is on; enabling again clears the counts."""
    import ctypes
    import torch
    lib = binding.load()
    ds = synth.dataset_p1k_n8_d64(n_particles=40)
    s = torch.from_numpy(ds.sbj).cuda()
    r = torch.from_numpy(ds.ref).cuda()

    def read():
        n, ms = ctypes.c_int(-1), ctypes.c_double(-1.0)
        npol, mspol = ctypes.c_int(-1), ctypes.c_double(-1.0)
        lib.isac_profile_read(ctypes.byref(n), ctypes.byref(ms), ctypes.byref(npol), ctypes.byref(mspol))
        return n.value, ms.value, npol.value, mspol.value

    lib.isac_profile_enable(1)
    for _ in range(3):   # 40 particles: one batch per call
        binding.mref_align_device(s, r, 29, 1.0, 2.0, 2.0)
    torch.cuda.synchronize()
    n, ms, npol, mspol = read()
    assert (n, npol) == (3, 3) and ms > 0.0 and mspol > 0.0
    lib.isac_profile_enable(1)
    assert read() == (0, 0.0, 0, 0.0)
    lib.isac_profile_enable(0)
    binding.mref_align_device(s, r, 29, 1.0, 2.0, 2.0)
    torch.cuda.synchronize()
    assert read() == (0, 0.0, 0, 0.0)
    binding.check(lib)
