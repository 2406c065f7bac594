"""Device stability test (isac_multi_align_stability_device) against the FP64
oracle (oracle/stability_oracle.py) and against closed forms.

The oracle fits the global transforms with scipy (Nelder-Mead then BFGS); the
device eliminates the translations in closed form and fits the rotations by a
monotone minorise-maximise iteration.  Both reach the same minimum; outputs are
FP32, so the checks allow 2e-4 px / 2e-3 degrees.  Members within 1e-4 px of the
pixel-error threshold may fall on either side and are not counted."""
import numpy as np
import pytest

from isac_amd import binding
from oracle import stability_oracle as SO

pytestmark = pytest.mark.gpu


def _class(rng, L, n, ang_sd, sh_sd, p_mflip, outliers=0):
    """L runs of one class: a true alignment per member, per-run noise, one global
    rigid transform (with mirror flip) per run, some members' mirror flipped in one run."""
    base = np.stack([rng.uniform(0, 360, n), rng.uniform(-3, 3, n), rng.uniform(-3, 3, n),
                     rng.integers(0, 2, n)], 1)
    runs = []
    for j in range(L):
        p = base.copy()
        p[:, 0] += rng.normal(0, ang_sd, n)
        p[:, 1:3] += rng.normal(0, sh_sd, (n, 2))
        if outliers:
            k = rng.choice(n, outliers, replace=False)
            p[k, 0] += rng.uniform(30, 90, outliers)
        a = rng.uniform(0, 360)
        t = rng.uniform(-1, 1, 2)
        f = int(rng.integers(0, 2)) if j < L - 1 else 0
        s = p[:, 1] + 1j * p[:, 2]
        if f:
            s = -np.conj(s)
            p[:, 0] = -p[:, 0]
        s = np.exp(1j * np.radians(a)) * s + (t[0] + 1j * t[1])
        p[:, 0] = (p[:, 0] + a) % 360
        p[:, 1], p[:, 2] = s.real, s.imag
        p[:, 3] = np.logical_xor(p[:, 3] != 0, f != 0)
        m = rng.random(n) < p_mflip
        p[m, 3] = 1 - p[m, 3]
        runs.append(p.reshape(-1).tolist())
    return runs


def _compare(dev, ref, err_thld, label):
    st_d, rate_d, err_d, ex_d = dev
    st_o, rate_o, err_o, ex_o = ref
    assert rate_d == pytest.approx(rate_o, abs=1e-6), label
    np.testing.assert_array_equal(ex_d["flips"], ex_o["flips"], err_msg=label)
    if err_o < 0:
        assert err_d == -1.0 and st_d == [], label
        return
    assert err_d == pytest.approx(err_o, rel=1e-5, abs=1e-6), label
    cons = ex_o["consistent"]
    np.testing.assert_allclose(ex_d["pix"][cons], ex_o["pix"][cons], atol=2e-4, rtol=1e-5,
                               err_msg=label)
    assert (ex_d["pix"][~cons] == -1.0).all(), label
    da = (ex_d["avg"][cons, 0] - ex_o["avg"][cons, 0] + 180.0) % 360.0 - 180.0
    assert np.abs(da).max() < 2e-3, label
    np.testing.assert_allclose(ex_d["avg"][cons, 1:3], ex_o["avg"][cons, 1:3], atol=2e-4,
                               err_msg=label)
    np.testing.assert_array_equal(ex_d["avg"][cons, 3], ex_o["avg"][cons, 3], err_msg=label)
    dg = (ex_d["G"][:, 0] - ex_o["G"][:, 0] + 180.0) % 360.0 - 180.0
    assert np.abs(dg).max() < 2e-3, label
    np.testing.assert_allclose(ex_d["G"][:, 1:3], ex_o["G"][:, 1:3], atol=2e-4, err_msg=label)
    # stable sets: same members away from the threshold, ordered by pixel error
    near = np.abs(ex_o["pix"] - err_thld) < 1e-4
    sd = {s[1] for s in st_d if not near[s[1]]}
    so = {s[1] for s in st_o if not near[s[1]]}
    assert sd == so, label
    pe = [s[0] for s in st_d]
    assert pe == sorted(pe), label


def test_random_classes_match_the_oracle(gpu):
    rng = np.random.default_rng(11)
    for L in (3, 4, 5):
        classes = [_class(rng, L, n, sd, 0.4, 0.05, out)
                   for n, sd, out in ((60, 2.0, 0), (120, 4.0, 3), (35, 6.0, 2))]
        thld = 1.5
        dev = binding.multi_align_stability_device(classes, 0.0, 1e4, thld, 60.0)
        for k, c in enumerate(classes):
            ref = SO.multi_align_stability(c, 0.0, 1e4, thld, 60.0)
            print(f"L{L} class {k}: rate {dev[k][1]:.4f} err dev {dev[k][2]:.6f} oracle {ref[2]:.6f} "
                  f"stable {len(dev[k][0])}/{len(ref[0])}", flush=True)
            _compare(dev[k], ref, thld, f"L{L} class {k}")


def test_identical_runs_give_zero_error(gpu):
    rng = np.random.default_rng(2)
    base = np.stack([rng.uniform(0, 360, 40), rng.uniform(-3, 3, 40), rng.uniform(-3, 3, 40),
                     rng.integers(0, 2, 40)], 1)
    cls = [base.reshape(-1).tolist()] * 4
    (st, rate, err, ex), = binding.multi_align_stability_device([cls], 0.0, 1e4, 0.5, 60.0)
    assert rate == 1.0 and err < 1e-3 and len(st) == 40
    assert np.abs(ex["G"][:, :3]).max() < 1e-4
    da = (ex["avg"][:, 0] - base[:, 0] + 180.0) % 360.0 - 180.0
    assert np.abs(da).max() < 1e-3 and np.allclose(ex["avg"][:, 1:3], base[:, 1:3], atol=1e-4)


def test_balanced_perturbation_has_the_analytic_error(gpu):
    # one run deviates by +-delta degrees and +-eps px, alternating: the fit stays
    # the identity and each member's squared error is
    # r^2 (1 - |(L - 1 + e^{i delta}) / L|) + eps^2 (L - 1) / L^2
    L, n, d, delta, eps = 4, 40, 60.0, 4.0, 0.6
    rng = np.random.default_rng(4)
    base = np.stack([rng.uniform(0, 360, n), np.zeros(n), np.zeros(n), rng.integers(0, 2, n)], 1)
    prm = np.stack([base.copy() for _ in range(L)])
    sgn = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    prm[0, :, 0] += sgn * delta
    prm[0, :, 1] += sgn * eps
    (st, rate, err, ex), = binding.multi_align_stability_device(
        [[p.reshape(-1).tolist() for p in prm]], 0.0, 1e4, 5.0, d)
    r = d / 2
    e2 = r * r * (1 - abs((L - 1) + np.exp(1j * np.radians(delta))) / L) + eps * eps * (L - 1) / L ** 2
    np.testing.assert_allclose(ex["pix"] ** 2, e2, rtol=2e-5)
    assert err == pytest.approx(np.sqrt(e2), rel=1e-5)


def test_rejections_and_mirror_counts(gpu):
    rng = np.random.default_rng(5)
    c = _class(rng, 3, 30, 1.0, 0.2, 0.0)
    prm = np.asarray(c).reshape(3, -1, 4)
    prm[1, :10, 3] = 1 - prm[1, :10, 3]          # 10 members flip their mirror in run 1
    c = [p.reshape(-1).tolist() for p in prm]
    # a mirror-consistency threshold above the rate rejects the class
    (st, rate, err, ex), = binding.multi_align_stability_device([c], 0.8, 1e4, 1.0, 60.0)
    ref = SO.multi_align_stability(c, 0.8, 1e4, 1.0, 60.0)
    assert rate == pytest.approx(ref[1]) and rate == pytest.approx(20 / 30)
    assert st == [] and err == -1.0 and (ex["pix"] == -1.0).all()
    # a class error above grp_err_thld leaves the stable set empty
    (st, rate, err, ex), = binding.multi_align_stability_device([c], 0.0, 1e-3, 1.0, 60.0)
    ref = SO.multi_align_stability(c, 0.0, 1e-3, 1.0, 60.0)
    assert st == [] == ref[0] and err == pytest.approx(ref[2], rel=1e-5)
    # an error threshold below every member's error leaves too few to fit
    (st, rate, err, ex), = binding.multi_align_stability_device([c], 0.0, 1e4, 1e-6, 60.0)
    ref = SO.multi_align_stability(c, 0.0, 1e4, 1e-6, 60.0)
    assert st == [] and err == -1.0 == ref[2]


def test_malformed_class_ranges_are_empty(gpu):
    # through the C ABI: ranges that run backwards, start below 0 or end past
    # n_total are empty classes (rate 0, error -1, nothing read or written)
    # while the valid classes around them are computed as usual
    import torch
    rng = np.random.default_rng(3)
    c0 = np.asarray(_class(rng, 3, 30, 1.0, 0.2, 0.0)).reshape(3, -1, 4)
    c1 = np.asarray(_class(rng, 3, 25, 1.0, 0.2, 0.0)).reshape(3, -1, 4)
    prm = np.ascontiguousarray(np.concatenate([c0, c1], axis=1))
    n_total = prm.shape[1]
    # classes 0 and 1 valid; 2: decreasing, 3: ends below 0, 4: starts below 0,
    # 5: ends past n_total, 6: starts past n_total
    off = np.array([0, 30, 55, 20, -5, 10, 90, 60], np.int32)
    n_class = len(off) - 1
    lib = binding.load()
    dev = torch.device("cuda")
    d_prm = torch.from_numpy(prm).to(dev)
    d_off = torch.from_numpy(off).to(dev)
    pix = torch.full((n_total,), 7.0, dtype=torch.float32, device=dev)
    avg = torch.zeros((n_total, 4), dtype=torch.float32, device=dev)
    stab = torch.zeros(n_total, dtype=torch.int32, device=dev)
    cls = torch.full((n_class, 2), 7.0, dtype=torch.float32, device=dev)
    glob = torch.zeros((n_class, 3, 4), dtype=torch.float32, device=dev)
    rc = lib.isac_multi_align_stability_device(d_prm.data_ptr(), n_total, d_off.data_ptr(), n_class, 3,
                                               0.0, 1e4, 1.0, 60.0, pix.data_ptr(), avg.data_ptr(),
                                               stab.data_ptr(), cls.data_ptr(), glob.data_ptr(),
                                               torch.cuda.current_stream().cuda_stream)
    assert rc == 0
    cls = cls.cpu().numpy()
    for c in (2, 3, 4, 5, 6):   # malformed: empty
        assert cls[c, 0] == 0.0 and cls[c, 1] == -1.0, (c, cls[c])
    ref = binding.multi_align_stability_device([c0.reshape(3, -1).tolist(), c1.reshape(3, -1).tolist()],
                                               0.0, 1e4, 1.0, 60.0)
    assert cls[0, 0] == pytest.approx(ref[0][1]) and cls[0, 1] == pytest.approx(ref[0][2])
    assert cls[1, 0] == pytest.approx(ref[1][1]) and cls[1, 1] == pytest.approx(ref[1][2])
    np.testing.assert_array_equal(pix.cpu().numpy(), np.concatenate([ref[0][3]["pix"], ref[1][3]["pix"]]))
