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


def test_config5_learned_form_against_fp64_oracle(gpu):
    """The learned form at its own workload (100 references, 76 px, +-3 px, SNR
    0.1, integer shifts; 200 000 training and 40 000 calibration particles, the
    report's recipe) scored on 3 000 held-out particles against the FP64 C
    oracle: best reference (either of the oracle's top two when they tie within
    1e-5), mirror and shift exact, psi within 0.5 degree, for >= 99.5 %."""
    import os
    import torch
    from oracle import coracle
    n_ref = 100

    def make(P, seed):
        s, r, _ = synth.make_dataset_torch(P, n_ref, D, 34.0, 0.1, RNG, True, seed_refs=1,
                                           seed_particles=seed, seed_noise=seed + 1)
        return s, r

    tr, ref = make(200000, 11)
    lab, _, _ = learned.labels(tr, ref, R, RNG)
    net = posenet.PoseNet(n_ref, 49, D=D, seed=0, **learned.NET)
    hist = net.train(tr, lab, epochs=learned.EPOCHS, **learned.TRAIN)
    print("loss per epoch", [round(h, 3) for h in hist])
    del tr, lab
    cal, _ = make(learned.N_CAL, 21)
    _, dc, pc = learned.labels(cal, ref, R, RNG)
    pr, raw = learned.calibrate(net, cal, ref, dc, pc, R, RNG, target=learned.TARGET)
    del cal, dc, pc
    k = 3000
    te, _ = make(k, 31)
    o = pr.run(te, ref)
    torch.cuda.synchronize()
    orc = coracle.mref(te.cpu().numpy(), ref.cpu().numpy(), np.zeros((k, 2), np.float32), R, RNG,
                       RNG, 1.0, nthreads=min(16, len(os.sched_getaffinity(0))), precision="f64")
    od = torch.from_numpy(orc["d"].astype(np.float32)).cuda()
    opk = torch.from_numpy(orc["peak_list"]).cuda()
    a = learned.agreement(o["best_ref"], o["best_peak"], od, opk).cpu().numpy()
    srt = np.sort(orc["d"], axis=0)
    ties = int(((srt[-1] - srt[-2]) <= 1e-5 * np.abs(srt[-1])).sum())
    print(f"held-out {k}: agreement with the FP64 oracle {a.mean():.4f} ({int((~a).sum())} misses), "
          f"oracle best-reference ties {ties}, fallback {o['n_fallback'] / k:.3f}, "
          f"threshold {pr.threshold:.4g}, confidence cut {pr.conf_min:.4g}")
    assert a.mean() >= 0.995
