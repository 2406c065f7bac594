"""Pins of the stability-test oracle (oracle/stability_oracle.py) by closed forms:
identical runs, runs that differ by one global rigid transform (with mirror
flips), balanced per-particle perturbations with an analytic pixel error, and
mirror-consistency counting.  Parity with SPHIRE's own sp_pixel_error is
unpinned (the module is absent from the reference tree)."""
import numpy as np

from oracle import stability_oracle as SO


def _runs(base, Gs):
    """Run j's parameters such that compose(Gs[j], run_j) == base."""
    out = []
    for G in Gs:
        a, tx, ty, f = G
        # invert x -> R(a) F^f x + t:  theta = (-1)^f (theta' - a), s = F^f R(-a) (s' - t)
        th = base[:, 0] - a
        sx, sy = base[:, 1] - tx, base[:, 2] - ty
        ca, sa = np.cos(np.radians(-a)), np.sin(np.radians(-a))
        rx, ry = ca * sx - sa * sy, sa * sx + ca * sy
        if f:
            th, rx = -th, -rx
        out.append(np.stack([th % 360.0, rx, ry, np.logical_xor(base[:, 3] != 0, f != 0)], 1))
    return np.stack(out)


def _base(n, seed):
    rng = np.random.default_rng(seed)
    return np.stack([rng.uniform(0, 360, n), rng.uniform(-3, 3, n), rng.uniform(-3, 3, n),
                     rng.integers(0, 2, n)], 1).astype(np.float64)


def test_compose_inverts_the_run_generator():
    base = _base(20, 1)
    Gs = np.array([[37.0, 1.5, -0.5, 1], [-120.0, 0.25, 2.0, 0], [0, 0, 0, 0]])
    runs = _runs(base, Gs)
    for j in range(3):
        back = SO.compose(Gs[j], runs[j])
        d = (back[:, 0] - base[:, 0] + 180.0) % 360.0 - 180.0
        assert np.abs(d).max() < 1e-9 and np.allclose(back[:, 1:3], base[:, 1:3])
        assert np.array_equal(back[:, 3] != 0, base[:, 3] != 0)


def test_identical_runs_are_perfectly_stable():
    base = _base(30, 2)
    prm = np.stack([base] * 4)
    st, rate, err, ex = SO.multi_align_stability([p.reshape(-1) for p in prm], 0.0, 1e4, 0.5, 60.0)
    # a zero error up to rounding: r^2 (1 - |mean e^{i theta}|) is ~1e-13, its root ~1e-7
    assert rate == 1.0 and abs(err) < 1e-6
    assert len(st) == 30 and max(s[0] for s in st) < 1e-6
    for s in st:
        i = s[1]
        da = (s[2][0] - base[i, 0] + 180.0) % 360.0 - 180.0
        assert abs(da) < 1e-6 and np.allclose(s[2][1:3], base[i, 1:3], atol=1e-6)
        assert s[2][3] == base[i, 3]


def test_global_rigid_transforms_are_fitted_away():
    base = _base(40, 3)
    Gs = np.array([[25.0, 1.0, -2.0, 1], [-70.0, -0.5, 0.75, 0], [140.0, 2.0, 1.0, 1], [0, 0, 0, 0]])
    prm = _runs(base, Gs)
    st, rate, err, ex = SO.multi_align_stability([p.reshape(-1) for p in prm], 0.0, 1e4, 0.5, 60.0)
    assert rate == 1.0
    assert np.array_equal(ex["flips"], Gs[:, 3].astype(int))
    assert err < 1e-5 and len(st) == 40
    for j in range(3):
        da = (ex["G"][j, 0] - Gs[j, 0] + 180.0) % 360.0 - 180.0
        assert abs(da) < 1e-5 and np.allclose(ex["G"][j, 1:3], Gs[j, 1:3], atol=1e-5)


def test_balanced_perturbation_has_the_analytic_pixel_error():
    # run 0 deviates by +-delta in angle and +-eps in x, alternating over particles:
    # by symmetry the best global transforms stay the identity, and every particle's
    # squared pixel error is r^2 (1 - |(L - 1 + e^{i delta}) / L|) + eps^2 (L - 1) / L^2
    L, n, d = 4, 40, 60.0
    delta, eps = 4.0, 0.6
    base = _base(n, 4)
    base[:, 1:] = 0.0          # zero shifts: a global rotation leaves |s| unchanged
    prm = np.stack([base.copy() for _ in range(L)])
    sgn = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    prm[0, :, 0] += sgn * delta
    prm[0, :, 1] += sgn * eps
    st, rate, err, ex = SO.multi_align_stability([p.reshape(-1) for p in prm], 0.0, 1e4, 5.0, d)
    r = d / 2
    rho = abs((L - 1) + np.exp(1j * np.radians(delta))) / L
    e2 = r * r * (1 - rho) + eps * eps * (L - 1) / L ** 2
    assert np.allclose(ex["pix"] ** 2, e2, rtol=1e-6)
    assert abs(err - np.sqrt(e2)) < 1e-6
    # a threshold between the error and zero empties the stable set
    st2, _, _, _ = SO.multi_align_stability([p.reshape(-1) for p in prm], 0.0, 1e4, 0.5 * np.sqrt(e2), d)
    assert st2 == []


def test_mirror_consistency_counts_particles_flipped_in_one_run():
    base = _base(20, 5)
    # run 0 is globally mirrored (a pure mirror transform): still consistent
    prm = _runs(base, np.array([[0.0, 0.0, 0.0, 1], [0, 0, 0, 0], [0, 0, 0, 0]]))
    prm[1, :5, 3] = 1 - prm[1, :5, 3]          # 5 particles flip their mirror in run 1 only
    flips, ok = SO.mirror_flips(prm[..., 3])
    assert flips.tolist() == [1, 0, 0]
    assert ok.sum() == 15 and not ok[:5].any()
    st, rate, err, ex = SO.multi_align_stability([p.reshape(-1) for p in prm], 0.0, 1e4, 0.5, 60.0)
    assert rate == 0.75
    assert sorted(s[1] for s in st) == list(range(5, 20))
    # a threshold on the mirror-consistency rate rejects the class
    st3, rate3, err3, _ = SO.multi_align_stability([p.reshape(-1) for p in prm], 0.8, 1e4, 0.5, 60.0)
    assert st3 == [] and rate3 == 0.75 and err3 == -1.0


def test_outliers_are_purged_from_the_fit():
    # one particle with a wild angle in run 0 must not bias the global fit
    base = _base(50, 6)
    base[:, 3] = 0
    prm = np.stack([base.copy() for _ in range(3)])
    prm[0, 7, 0] += 90.0
    st, rate, err, ex = SO.multi_align_stability([p.reshape(-1) for p in prm], 0.0, 1e4, 1.0, 60.0)
    assert not ex["kept"][7] and ex["kept"].sum() == 49
    assert np.abs(ex["G"][:2, :3]).max() < 1e-6
    assert 7 not in [s[1] for s in st] and len(st) == 49
