"""Device pre-alignment loop (isac_pre_align_loop_device) against the oracle's
restatement of the driver loop (oracle.isac_oracle.pre_align_loop): shift steps
x iterations with the average, optional centring / low-pass and alignment of
every particle on the device.  Final parameters: mirror and accumulated integer
shifts exact, psi within 0.5 degrees, on particles whose last alignment is not a
near-tie; the final average to 1e-4 of its range."""
import numpy as np
import pytest

from isac_amd import binding, synth
from oracle import isac_oracle as O
from parity import circ_diff

pytestmark = pytest.mark.gpu


def _data(P, D, seed, snr):
    ds = synth.make_dataset(P, 1, D, 0.4 * D, snr, 2, True, seed_refs=seed, seed_particles=seed + 1,
                            seed_noise=seed + 2)
    # the driver subtracts each image's mean inside the mask first (:1401-1403)
    yy, xx = np.mgrid[:D, :D]
    mask = (xx - D // 2) ** 2 + (yy - D // 2) ** 2 <= (0.45 * D) ** 2
    sbj = ds.sbj - np.array([im[mask].mean() for im in ds.sbj])[:, None, None]
    return np.ascontiguousarray(sbj, dtype=np.float32)


def _near_tie(tab, prm, rel=1e-5):
    """Particles whose best (mirror, shift, psi) of the last alignment has a
    runner-up within rel (another slot, or the same slot more than 3 samples away)."""
    P, _, S, L = tab.shape
    flat = tab.reshape(P, 2 * S, L)
    slot = prm["mirror"] * S + prm["shift_idx"]
    tie = np.zeros(P, bool)
    for p in range(P):
        sm = flat[p].max(-1)
        sm[slot[p]] = -np.inf
        dist = np.abs((np.arange(L) - prm["t"][p] + L // 2) % L - L // 2)
        runner = max(sm.max(), np.where(dist > 3, flat[p, slot[p]], -np.inf).max())
        tie[p] = prm["max"][p] - runner <= rel * abs(prm["max"][p])
    return tie


@pytest.mark.parametrize("P,D,R,rng0,center,cutoff,falloff",
                         [(40, 48, 18, 2.0, 0, 0.0, 0.0), (40, 48, 18, 2.0, -1, 0.0, 0.0),
                          (40, 48, 18, 2.0, -1, 0.3, 0.1), (32, 76, 35, 3.0, -1, 0.3, 0.1)])
def test_pre_align_loop_matches_oracle(gpu, P, D, R, rng0, center, cutoff, falloff):
    import torch
    ranges, steps, max_iter = [rng0, 1.0], [1.0, 1.0], 3
    sbj = _data(P, D, 40 if D == 48 else D, 1.0)
    print(f"center {center} cutoff {cutoff}: {P} x {D} px, {R} rings, steps {ranges}, {max_iter} iterations",
          flush=True)
    g, avg_g = binding.pre_align_loop_device(torch.from_numpy(sbj).cuda(), R, ranges, steps,
                                             max_iter, center=center, cutoff=cutoff,
                                             falloff=falloff)
    torch.cuda.synchronize()
    o, avg_o, tab = O.pre_align_loop(sbj, R, ranges, steps, max_iter, center=center,
                                     cutoff=cutoff, falloff=falloff, return_table=True)
    ok = ~_near_tie(tab, o)
    print("near-ties", int((~ok).sum()), flush=True)
    assert np.array_equal(g["mirror"][ok], o["mirror"][ok])
    assert np.abs(g["shift_x"] - o["shift_x"])[ok].max() < 1e-4
    assert np.abs(g["shift_y"] - o["shift_y"])[ok].max() < 1e-4
    da = circ_diff(g["angle"], o["angle"])[ok]
    print("max psi err", float(da.max()), "avg err", float(np.abs(avg_g.cpu().numpy() - avg_o).max()),
          flush=True)
    assert da.max() <= 0.5
    span = float(avg_o.max() - avg_o.min())
    np.testing.assert_allclose(avg_g.cpu().numpy(), avg_o, atol=1e-4 * span)


def test_pre_align_loop_single_step_without_alignment(gpu):
    # one step, one iteration: only the average of the untransformed particles
    import torch
    sbj = _data(12, 32, 7, 2.0)
    g, avg = binding.pre_align_loop_device(torch.from_numpy(sbj).cuda(), 10, [1.0], [1.0], 1)
    torch.cuda.synchronize()
    assert np.all(g["shift_x"] == 0) and np.all(g["angle"] == 0) and np.all(g["mirror"] == 0)
    np.testing.assert_allclose(avg.cpu().numpy(), sbj.astype(np.float64).mean(0), atol=1e-5)
