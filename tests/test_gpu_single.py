"""GPU parity of the pre-alignment and reference-free entry points (MI355X)."""
import numpy as np
import pytest

from isac_amd import binding, synth
from oracle import isac_oracle as O
from parity import circ_diff

pytestmark = pytest.mark.gpu


def single_ref_oracle(sbj, refs, ref_of, acc, R, shifts, D, tie_rel=1e-5):
    tab = O.single_ref_table(sbj, refs, ref_of, acc, R, shifts)
    prm = O.decode_single_ref(tab, acc, shifts, D, R)
    P, _, S, L = tab.shape
    flat = tab.reshape(P, 2 * S, L)
    best = prm["max"]
    slot = prm["mirror"] * S + prm["shift_idx"]
    runner = np.empty(P)
    for p in range(P):
        sm = flat[p].max(-1)
        sm[slot[p]] = -np.inf
        dist = np.abs((np.arange(L) - prm["t"][p] + L // 2) % L - L // 2)
        runner[p] = max(sm.max(), np.where(dist > 3, flat[p, slot[p]], -np.inf).max())
    prm["tie"] = (best - runner) <= tie_rel * np.abs(best)
    return prm


def assert_params_match(g, o, label):
    ok = ~o["tie"]
    assert np.array_equal(g["mirror"][ok], o["mirror"][ok]), f"{label}: mirror"
    assert np.allclose(g["shift_x"][ok], o["shift_x"][ok], atol=1e-4), f"{label}: shift_x"
    assert np.allclose(g["shift_y"][ok], o["shift_y"][ok], atol=1e-4), f"{label}: shift_y"
    da = circ_diff(g["angle"], o["angle"])[ok]
    assert da.max() <= 0.5, f"{label}: psi {da.max()}"
    print(label, "pairs", ok.size, "ties", int((~ok).sum()), "max psi err", float(da.max()))


@pytest.mark.parametrize("batch", [64, 16])
def test_pre_align_one_run(gpu, batch):
    D, R, rng = 64, 29, 2.0
    ds = synth.make_dataset(40, 1, D, 28.0, 0.5, 2, True, seed_refs=5)
    acc = np.random.default_rng(8).uniform(-1, 1, (40, 2)).astype(np.float32)
    g = binding.pre_align(ds.sbj, ds.ref[0], R, rng, 1.0, batch=batch, acc_xy=acc)
    shifts = O.shift_grid(rng, rng, 1.0)
    o = single_ref_oracle(ds.sbj, ds.ref, np.zeros(40, int), acc.astype(np.float64), R, shifts, D)
    assert_params_match(g, o, f"pre-align batch {batch}")


@pytest.mark.parametrize("D,R,rng,step", [(76, 35, 3.0, 1.0), (96, 44, 3.0, 1.0), (128, 60, 4.0, 1.0)])
def test_pre_align_shapes(gpu, D, R, rng, step):
    """Single-reference alignment at the larger boxes, ring counts and shift
    ranges (81 shifts at 128 px).  Half-pixel grids are covered by the tie
    tests: on them most pairs here are oracle ties (neighbouring shifts reach
    the same maximum within 1e-5), which this comparison leaves out."""
    P = 24
    ds = synth.make_dataset(P, 1, D, 0.42 * D, 0.5, rng, True, seed_refs=D + 3)
    acc = np.random.default_rng(D + 4).uniform(-1, 1, (P, 2)).astype(np.float32)
    g = binding.pre_align(ds.sbj, ds.ref[0], R, rng, step, batch=P, acc_xy=acc)
    shifts = O.shift_grid(rng, rng, step)
    o = single_ref_oracle(ds.sbj, ds.ref, np.zeros(P, int), acc.astype(np.float64), R, shifts, D)
    assert_params_match(g, o, f"pre-align D{D} R{R} rng {rng} step {step}")


def test_pre_align_two_runs_accumulate(gpu):
    D, R, rng = 64, 29, 2.0
    ds = synth.make_dataset(24, 1, D, 28.0, 1.0, 2, True, seed_refs=6)
    g = binding.pre_align(ds.sbj, ds.ref[0], R, rng, 1.0, batch=24, iterations=2)
    shifts = O.shift_grid(rng, rng, 1.0)
    acc = np.zeros((24, 2))
    for _ in range(2):
        o = single_ref_oracle(ds.sbj, ds.ref, np.zeros(24, int), acc, R, shifts, D)
        acc = np.stack([o["shift_x"], o["shift_y"]], 1)
    assert_params_match(g, o, "pre-align x2")


def _classes(n_cls, per, D, seed, snr):
    blobs = synth.make_blobs(n_cls, 0.4 * D, seed)
    rng = np.random.default_rng(seed)
    cid = np.repeat(np.arange(n_cls), per)
    P = cid.size
    sbj = synth.render_particles(blobs, D, cid, rng.uniform(0, 360, P), rng.uniform(-1, 1, (P, 2)),
                                 rng.integers(0, 2, P), snr, 0.4 * D, seed + 1)
    avg = np.stack([sbj[cid == k].mean(0) for k in range(n_cls)]).astype(np.float32)
    return sbj, cid, avg


def test_ref_free_iterations(gpu):
    D, R, rng = 48, 20, 1.0
    sbj, cid, avg = _classes(3, 10, D, 21, 2.0)
    g = binding.ref_free_align(sbj, cid, avg, R, rng, num_iter=2)
    shifts = O.shift_grid(rng, rng, 1.0)
    acc = np.zeros((sbj.shape[0], 2))
    refs = avg.astype(np.float64)
    for _ in range(2):
        o = single_ref_oracle(sbj, refs, cid, acc, R, shifts, D)
        prm, refs = O.ref_free_iteration(sbj, refs, cid, acc, R, shifts, D)
        acc = np.stack([prm["shift_x"], prm["shift_y"]], 1)
    assert_params_match(g, o, "ref-free x2")


@pytest.mark.parametrize("D,R,rng", [(96, 44, 2.0), (128, 60, 2.0)])
def test_ref_free_larger_boxes(gpu, D, R, rng):
    """Reference-free iterations (alignment, then transform and class averages on
    the device) at 96 and 128 px."""
    sbj, cid, avg = _classes(2, 8, D, D + 7, 2.0)
    g = binding.ref_free_align(sbj, cid, avg, R, rng, num_iter=2)
    shifts = O.shift_grid(rng, rng, 1.0)
    acc = np.zeros((sbj.shape[0], 2))
    refs = avg.astype(np.float64)
    for _ in range(2):
        o = single_ref_oracle(sbj, refs, cid, acc, R, shifts, D)
        prm, refs = O.ref_free_iteration(sbj, refs, cid, acc, R, shifts, D)
        acc = np.stack([prm["shift_x"], prm["shift_y"]], 1)
    assert_params_match(g, o, f"ref-free D{D} x2")


def test_ref_free_with_tangent_filter(gpu):
    D, R, rng = 48, 20, 1.0
    sbj, cid, avg = _classes(2, 12, D, 31, 3.0)
    n_iter, fh, ff = 3, 0.35, 0.2
    g = binding.ref_free_align(sbj, cid, avg, R, rng, num_iter=n_iter, cutoff_freq=fh, falloff=ff)
    shifts = O.shift_grid(rng, rng, 1.0)
    acc = np.zeros((sbj.shape[0], 2))
    refs = avg.astype(np.float64)
    for i in range(n_iter):
        refs = O.tangent_filter(refs, 0.1 + (fh - 0.1) * i / float(n_iter - 1), ff)
        o = single_ref_oracle(sbj, refs, cid, acc, R, shifts, D)
        prm, refs = O.ref_free_iteration(sbj, refs, cid, acc, R, shifts, D)
        acc = np.stack([prm["shift_x"], prm["shift_y"]], 1)
    assert_params_match(g, o, "ref-free + filter x3")


def test_size_checks(gpu):
    lib = binding.load()
    cfg = binding.make_config(1000, 1, 76, 35, 1.0, 4.0, 4.0, "pre_aln")
    assert lib.pre_align_size_check(1000, cfg, 0, 0.9, False) == 1
    assert lib.pre_align_size_check(1000, cfg, 0, 1e-9, False) == 0
    big = binding.make_config(5000, 100, 76, 35, 1.0, 1.0, 1.0, "multigroup_aln")
    assert lib.ref_free_alignment_2D_size_check(big, 0, 0.9, False) == 1
    assert lib.ref_free_alignment_2D_size_check(big, 0, 1e-9, False) == 0


def test_reset_shifts_keeps_count(gpu):
    lib = binding.load()
    cfg = binding.make_config(4, 1, 32, 10, 1.0, 2.0, 2.0, "pre_aln")
    lib.pre_align_init(4, cfg, 0)
    binding.check(lib)
    lib.reset_shifts(4.0, 2.0)          # 5x5 grid as before
    binding.check(lib)
    lib.reset_shifts(3.0, 1.0)          # 7x7: rejected
    assert lib.gpu_last_error() == 2
    lib.gpu_clear_error()
    lib.gpu_clear()


def test_oversized_shift_grid_is_rejected_promptly(gpu):
    # a range of 1e6 px in steps of 1 px (or a step far below the range's float
    # spacing) is beyond any grid the library takes: the calls return an error at
    # once instead of enumerating the grid
    import time
    lib = binding.load()
    t0 = time.perf_counter()
    lib.reset_shifts(1.0e6, 1.0)              # no context yet: a state error
    assert lib.gpu_last_error() != 0
    lib.gpu_clear_error()
    cfg = binding.make_config(4, 1, 32, 10, 1.0, 2.0, 2.0, "pre_aln")
    lib.pre_align_init(4, cfg, 0)
    binding.check(lib)
    lib.reset_shifts(1.0e6, 1.0)              # 2e6 x 2e6 shifts: not the context's count
    assert lib.gpu_last_error() == 2
    lib.gpu_clear_error()
    lib.reset_shifts(3.0e8, 1.0)              # a step below the float spacing of the range
    assert lib.gpu_last_error() == 2
    lib.gpu_clear_error()
    lib.gpu_clear()
    assert time.perf_counter() - t0 < 5.0
