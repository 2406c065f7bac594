"""Pins of the CPU oracle of the pose-proposal network (no GPU)."""
import numpy as np
import pytest

from oracle import posenet_oracle as N


def test_bf16_rounding_matches_torch():
    import torch
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(10000).astype(np.float32) * 10.0 ** rng.integers(-8, 8, 10000),
                        np.array([0.0, -0.0, 1.0, 65504.0, 3.0e38, -3.0e38, 1e-40], np.float32)])
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(N.bf16(x), want)


def test_conv_matches_direct_loop():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 5, 6, 3))
    w = rng.standard_normal((4, 27))
    b = rng.standard_normal(4)
    y = N.conv(x, w, b, 3)
    wk = w.reshape(4, 3, 3, 3)   # [co][ky][kx][ci]
    for i in range(2):
        for yy in range(5):
            for xx in range(6):
                for co in range(4):
                    s = b[co]
                    for ky in range(3):
                        for kx in range(3):
                            sy, sx = yy + ky - 1, xx + kx - 1
                            if 0 <= sy < 5 and 0 <= sx < 6:
                                s += wk[co, ky, kx] @ x[i, sy, sx]
                    assert abs(y[i, yy, xx, co] - s) < 1e-12


def test_col2im_is_the_adjoint_of_im2col():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((2, 7, 5, 3))
    c = rng.standard_normal((2, 7, 5, 9, 3))
    assert abs(np.sum(N.im2col(x, 3) * c) - np.sum(x * N.col2im(c, 3))) < 1e-10


def test_pool_first_max_and_backward_routing():
    y = np.zeros((1, 4, 4, 1))
    y[0, 0, 1, 0] = 2.0          # window (0,0): max at (0,1)
    y[0, 2, 2, 0] = y[0, 3, 3, 0] = 1.0   # window (1,1): tie -> first, (0,0) of the window
    a, arg = N.pool(y)
    assert a[0, 0, 0, 0] == 2.0 and arg[0, 0, 0, 0] == 1 and arg[0, 1, 1, 0] == 0
    da = np.ones((1, 2, 2, 1))
    dz = N.pool_bwd(y, arg, da)
    assert dz[0, 0, 1, 0] == 1.0 and dz[0, 2, 2, 0] == 1.0 and dz[0, 3, 3, 0] == 0.0
    assert dz.sum() == 2.0      # windows whose maximum is 0 pass nothing (ReLU)


@pytest.mark.parametrize("pool4,soft", [(True, {}),
                                        (False, dict(sigma_psi=0.8, sigma_shift=0.6, shift_ny=3))])
def test_backward_matches_finite_differences(pool4, soft):
    spec = N.NetSpec(D=20, channels=(4, 6, 8, 8), hidden=10, heads=(3, 4, 9), pool4=pool4)
    params = N.init_params(spec, seed=3)
    params = [p.astype(np.float64) for p in params]
    rng = np.random.default_rng(4)
    for p in params:
        p[:, -1] = rng.standard_normal(p.shape[0]) * 0.1
    img = rng.standard_normal((3, 20, 20)).astype(np.float32)
    lab = np.stack([rng.integers(0, 3, 3), rng.integers(0, 4, 3), rng.integers(0, 9, 3)], 1)

    def loss_of(ps):
        act = N.forward(spec, ps, img, quant=False)
        return N.xent(spec, act["logits"], lab, **soft)[0].mean()

    act = N.forward(spec, params, img, quant=False)
    _, grads = N.backward(spec, params, act, lab, quant=False, **soft)
    eps = 1e-6
    checked = 0
    for li, p in enumerate(params):
        flat = p.reshape(-1)
        for j in rng.choice(flat.size, size=min(12, flat.size), replace=False):
            if li == 5 and j // p.shape[1] >= spec.n_out:
                continue
            old = flat[j]
            flat[j] = old + eps
            lp = loss_of(params)
            flat[j] = old - eps
            lm = loss_of(params)
            flat[j] = old
            fd = (lp - lm) / (2 * eps)
            an = grads[li].reshape(-1)[j]
            assert abs(fd - an) <= 1e-6 + 1e-4 * abs(fd), (li, j, fd, an)
            checked += 1
    assert checked > 50


def test_adamw_matches_torch():
    import torch
    rng = np.random.default_rng(5)
    p0 = rng.standard_normal(100).astype(np.float32)
    tp = torch.nn.Parameter(torch.from_numpy(p0.copy()))
    opt = torch.optim.AdamW([tp], lr=1e-2, betas=(0.9, 0.999), eps=1e-8, weight_decay=0.01)
    p, m, v = p0.copy(), np.zeros(100, np.float32), np.zeros(100, np.float32)
    for t in range(1, 6):
        g = rng.standard_normal(100).astype(np.float32)
        tp.grad = torch.from_numpy(g.copy())
        opt.step()
        p, m, v = N.adamw(p, g, m, v, t, 1e-2, wd=0.01)
    np.testing.assert_allclose(p, tp.detach().numpy(), rtol=1e-5, atol=1e-6)


def test_proposals_brute_force_orders_by_summed_logits():
    rng = np.random.default_rng(6)
    heads = (5, 6, 4)
    z = rng.standard_normal((3, 15)).astype(np.float32)
    got = N.proposals(heads, z, k=8)
    for i in range(3):
        s = (z[i, :5, None, None] + z[i, None, 5:11, None]) + z[i, None, None, 11:15]
        order = np.argsort(-s, axis=None, kind="stable")[:8]
        r, pc, sh = np.unravel_index(order, s.shape)
        want = np.stack([r, pc // 3, pc % 3, sh], 1)
        np.testing.assert_array_equal(got[i], want)


def test_soft_targets_shape():
    spec = N.NetSpec(D=20, channels=(4, 6, 8, 8), hidden=10, heads=(3, 8, 9))
    t = N.soft_targets(spec, np.array([[0, 5, 4]]), sigma_psi=1.0, sigma_shift=1.0, shift_ny=3)
    assert np.allclose([x.sum() for x in t], 1.0)
    assert np.all(t[1][0, :4] == 0.0)                      # other mirror: no weight
    assert abs(t[1][0, 4] - t[1][0, 6]) < 1e-15            # circular neighbours of bin 1
    assert t[2][0].argmax() == 4 and abs(t[2][0, 1] - t[2][0, 3]) < 1e-15


def test_proposal_suppression_spreads_windows():
    """With suppression no two proposals of one reference and mirror lie within
    the suppression radius of each other; without it the top-k may."""
    rng = np.random.default_rng(11)
    z = rng.standard_normal((5, 3 + 24 + 9)).astype(np.float32)
    z[:, 0] += 5.0                                   # one dominant reference
    got = N.proposals((3, 24, 9), z, k=6, psi_r=2, shift_r=1, shift_ny=3)
    for p in got:
        for a in range(6):
            for b in range(a):
                if p[a, 0] == p[b, 0] and p[a, 1] == p[b, 1]:
                    db = abs(p[a, 2] - p[b, 2])
                    db = min(db, 12 - db)
                    near = abs(p[a, 3] // 3 - p[b, 3] // 3) <= 1 and abs(p[a, 3] % 3 - p[b, 3] % 3) <= 1
                    assert not (db <= 2 and near)


def test_joint_calibration_prefers_the_confidence_cut_when_it_helps():
    from isac_amd import learned
    rng = np.random.default_rng(3)
    n = 2000
    val = rng.normal(1.0, 0.1, n)
    conf = rng.uniform(0.5, 1.0, n)
    agree = np.ones(n, bool)
    bad = np.argsort(conf)[:40]            # the disagreeing particles have the lowest confidence
    agree[bad] = False
    val[bad] = rng.normal(1.05, 0.05, 40)  # ... and high values: the threshold alone fails
    cut, thr, fb = learned.calibrate_joint(val, agree, conf, target=0.997)
    assert cut > 0 and fb < 0.1
    keep = (conf >= cut) & (val >= thr)
    assert np.sum(keep & ~agree) <= int(0.003 * n)
    thr_only = learned.calibrate_threshold(val, agree, 0.997)
    assert np.mean(val < thr_only) > fb
