"""What the trained proposal network gains over the method without it (the
exhaustive search), on a small problem: 20 references, 76 px, +-3 shifts."""
import time

import numpy as np
import pytest

from isac_amd import binding, learned, posenet, synth

pytestmark = pytest.mark.gpu

N, D, R, RNG = 20, 76, 35, 3.0


def _set(P, seed):
    sbj, ref, _ = synth.make_dataset_torch(P, N, D, 34.0, 0.1, RNG, True, seed_refs=77,
                                           seed_particles=seed, seed_noise=seed + 1)
    return sbj, ref


def test_trained_network_narrows_the_search(gpu):
    import torch
    tr, ref = _set(24000, 100)
    cal, _ = _set(4000, 200)
    te, _ = _set(4000, 300)
    lab, _, _ = learned.labels(tr, ref, R, RNG)
    net = posenet.PoseNet(N, 49, D=D, seed=0, **learned.NET)
    hist = net.train(tr, lab, epochs=12, **dict(learned.TRAIN, warmup=300))
    print("loss per epoch", [round(h, 3) for h in hist])
    assert hist[-1] < 0.6 * hist[0]
    _, dc, pc = learned.labels(cal, ref, R, RNG)
    pr, raw = learned.calibrate(net, cal, ref, dc, pc, R, RNG, target=0.998)
    d, peak = learned.exhaustive(te, ref, R, RNG)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    o = pr.run(te, ref)
    torch.cuda.synchronize()
    t_learned = time.perf_counter() - t0
    t0 = time.perf_counter()
    learned.exhaustive(te, ref, R, RNG)
    torch.cuda.synchronize()
    t_full = time.perf_counter() - t0
    agree = float(learned.agreement(o["best_ref"], o["best_peak"], d, peak).double().mean())
    f_tr = o["n_fallback"] / te.shape[0]
    # the same search with an untrained network and the same threshold
    un = learned.Proposer(posenet.PoseNet(N, 49, D=D, seed=1, **learned.NET), pr.threshold, R, RNG)
    ou = un.run(te, ref)
    torch.cuda.synchronize()
    f_un = ou["n_fallback"] / te.shape[0]
    agree_un = float(learned.agreement(ou["best_ref"], ou["best_peak"], d, peak).double().mean())
    print(f"narrowed-only agreement (cal) {raw:.4f}; threshold {pr.threshold:.4g}")
    print(f"trained: agreement {agree:.4f}, fallback {f_tr:.3f}, {t_learned * 1e3:.1f} ms; "
          f"untrained: agreement {agree_un:.4f}, fallback {f_un:.3f}; exhaustive {t_full * 1e3:.1f} ms")
    assert agree >= 0.995
    assert f_tr < 0.5 and f_un > f_tr + 0.3
