"""isac_mref_align_device_noalloc (the allocation-free form of
isac_mref_align_device) against isac_mref_align_device itself: bit-identical
peak_list and d on two calls in a row, the second smaller, the first one larger
than one particle batch of the static scratch array; no device memory taken by
either call; an oversized job is refused with an error before any launch."""
import numpy as np
import pytest

from isac_amd import binding, synth

pytestmark = pytest.mark.gpu


def _inputs(ds, n, seed):
    import torch
    acc = np.random.default_rng(seed).uniform(-1, 1, (n, 2)).astype(np.float32)
    return (torch.from_numpy(np.ascontiguousarray(ds.sbj[:n])).cuda(),
            torch.from_numpy(np.ascontiguousarray(ds.ref)).cuda(),
            torch.from_numpy(acc).cuda())


def _pair(s, r, a, ring, step, rx, ry):
    import torch
    P, N = s.shape[0], r.shape[0]
    peak0, d0 = binding.mref_align_device(s, r, ring, step, rx, ry, acc_xy=a)
    torch.cuda.synchronize()
    outs = []
    for _ in range(2):
        peak = torch.full((N, 4 * P), float("nan"), dtype=torch.float32, device=s.device)
        d = torch.full((N, P), float("nan"), dtype=torch.float32, device=s.device)
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        binding.mref_align_device_noalloc(s, r, ring, step, rx, ry, peak, d, acc_xy=a)
        torch.cuda.synchronize()
        free1 = torch.cuda.mem_get_info()[0]
        outs.append((peak.cpu().numpy(), d.cpu().numpy()))
    # the repeated call: device memory free before and after is unchanged
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
    return peak0.cpu().numpy(), d0.cpu().numpy(), outs[1][0], outs[1][1], free0 - free1


def test_noalloc_matches_the_allocating_entry_on_two_calls(gpu):
    # call 1: 76 px, 35 rings, 49 shifts, 100 references, 80 particles: the
    # scratch array holds about 35 particles' spectra, so three batches
    big = synth.dataset_p20k_n100_d76(n_particles=80, n_refs=100)
    s, r, a = _inputs(big, 80, 3)
    per = 49 * 35 * 1024
    assert 80 * per > binding.NOALLOC_SCRATCH_BYTES // 2   # several batches
    print("call 1: 80 x 100, 76 px, 35 rings, 49 shifts", flush=True)
    p0, d0, p1, d1, taken = _pair(s, r, a, 35, 1.0, 3.0, 3.0)
    assert np.array_equal(p0, p1) and np.array_equal(d0, d1)
    assert taken <= 0
    # call 2, smaller: 64 px, 29 rings, 25 shifts, 8 references, 40 particles
    small = synth.dataset_p1k_n8_d64(n_particles=40)
    s, r, a = _inputs(small, 40, 4)
    print("call 2: 40 x 8, 64 px, 29 rings, 25 shifts", flush=True)
    p0, d0, p1, d1, taken = _pair(s, r, a, 29, 1.0, 2.0, 2.0)
    assert np.array_equal(p0, p1) and np.array_equal(d0, d1)
    assert taken <= 0


def test_noalloc_refuses_a_job_beyond_its_scratch(gpu):
    import torch
    # 2000 references x 35 rings of spectra (70 MiB) exceed the 64 MiB array:
    # refused before any launch, outputs untouched
    s = torch.zeros((2, 76, 76), dtype=torch.float32, device=gpu)
    r = torch.zeros((2000, 76, 76), dtype=torch.float32, device=gpu)
    peak = torch.full((2000, 8), 7.0, dtype=torch.float32, device=gpu)
    d = torch.full((2000, 2), 7.0, dtype=torch.float32, device=gpu)
    lib = binding.load()
    lib.gpu_clear_error()
    with pytest.raises(binding.GpuAlignError, match="error 4"):
        binding.mref_align_device_noalloc(s, r, 35, 1.0, 3.0, 3.0, peak, d)
    assert lib.gpu_last_error() == 0   # reported once, then cleared
    torch.cuda.synchronize()
    assert bool((peak == 7.0).all()) and bool((d == 7.0).all())
