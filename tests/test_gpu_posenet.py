"""GPU kernels of the pose-proposal network against the CPU oracle."""
import numpy as np
import pytest

from oracle import posenet_oracle as N

pytestmark = pytest.mark.gpu


def _net(n_ref=100, n_shift=49, D=76, seed=0):
    from isac_amd import posenet
    return posenet.PoseNet(n_ref, n_shift, D=D, seed=seed)


def _spec(net):
    return N.NetSpec(D=net.D, channels=net.channels, hidden=net.hidden, heads=net.heads,
                     pool4=net.pool4)


def _perturbed_params(net, seed):
    rng = np.random.default_rng(seed)
    ps = net.params_numpy()
    for li, p in enumerate(ps):
        p[:, -1] = (rng.standard_normal(p.shape[0]) * 0.05).astype(np.float32)
        if li == 5:
            p[net.n_out:, :] = 0.0
    net.load_params(ps)
    return ps


def _images(n, D, seed):
    from isac_amd import synth
    ds = synth.make_dataset(n, 12, D, D / 2 - 4, 0.1, 3.0, False, seed_particles=seed)
    return ds.sbj


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def test_forward_matches_oracle(gpu):
    import torch
    net = _net()
    ps = _perturbed_params(net, 1)
    img = _images(6, 76, 2)
    logits = net.forward(torch.from_numpy(img).cuda()).cpu().numpy()
    act = N.forward(_spec(net), ps, img, quant=True)
    err = _rel(logits[:, :net.n_out], act["logits"][:, :net.n_out])
    print("forward logits rel err", err)
    assert err < 1e-2
    assert np.all(logits[:, net.n_out:] == 0.0)


@pytest.mark.parametrize("pool4,sp,ss", [(True, 0.0, 0.0), (False, 1.0, 0.7)])
def test_gradients_match_oracle(gpu, pool4, sp, ss):
    """Per-layer gradients (weights and biases) of the mean loss within 2e-2
    (relative, Frobenius) of the oracle with the same bf16 rounding points;
    one-hot and Gaussian targets, with and without the last pooling."""
    import torch
    from isac_amd import posenet
    net = posenet.PoseNet(100, 49, D=76, seed=0, pool4=pool4,
                          channels=(32, 64, 128, 128 if pool4 else 32))
    ps = _perturbed_params(net, 3)
    img = _images(8, 76, 4)
    rng = np.random.default_rng(5)
    lab = np.stack([rng.integers(0, 100, 8), rng.integers(0, 144, 8), rng.integers(0, 49, 8)], 1)
    net.forward(torch.from_numpy(img).cuda(), train=True)
    loss = net.backward(torch.from_numpy(lab.astype(np.int32)).cuda(), sp, ss).cpu().numpy()
    spec = _spec(net)
    act = N.forward(spec, ps, img, quant=True)
    oloss, og = N.backward(spec, ps, act, lab, quant=True, sigma_psi=sp, sigma_shift=ss, shift_ny=7)
    assert _rel(loss, oloss) < 1e-3, (loss, oloss)
    for li in range(6):
        g = net.grad(li).cpu().numpy()
        e = _rel(g, og[li])
        print(f"layer {li} grad rel err {e:.2e}")
        assert e < 2e-2, (li, e)


def test_adamw_step_matches_oracle(gpu):
    import torch
    net = _net(n_ref=10, n_shift=9, D=32)
    p0 = net.flat.cpu().numpy()
    rng = np.random.default_rng(7)
    g = rng.standard_normal(p0.size).astype(np.float32)
    net.gflat.copy_(torch.from_numpy(g))
    net.step(1e-3, wd=1e-4)
    p1, m1, v1 = N.adamw(p0, g, np.zeros_like(p0), np.zeros_like(p0), 1, 1e-3, wd=1e-4)
    np.testing.assert_allclose(net.flat.cpu().numpy(), p1, rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(net.m.cpu().numpy(), m1, rtol=1e-6, atol=1e-9)


@pytest.mark.parametrize("psi_r,min_refs", [(-1, 0), (1, 0), (2, 0), (2, 3), (-1, 9)])
def test_proposals_kernel_matches_brute_force(gpu, psi_r, min_refs):
    import torch
    from isac_amd import posenet
    lib = posenet._lib()
    rng = np.random.default_rng(8)
    heads = (10, 24, 9)
    n, ld = 37, 48
    z = rng.standard_normal((n, ld)).astype(np.float32)
    z[3, :10] = 0.5   # ties inside a head
    z[5, 10:34] = 0.25
    zt = torch.from_numpy(z).cuda()
    prop = torch.empty((n, 8, 4), dtype=torch.int32, device="cuda")
    conf = torch.empty(n, dtype=torch.float32, device="cuda")
    posenet._call(lib, "isac_nn_proposals", zt.data_ptr(), n, ld, 10, 12, 9, 8, psi_r, 1, 3,
                  min_refs, prop.data_ptr(), conf.data_ptr(), torch.cuda.current_stream().cuda_stream)
    want = N.proposals(heads, z, k=8, psi_r=psi_r, shift_r=1, shift_ny=3, min_refs=min_refs)
    np.testing.assert_array_equal(prop.cpu().numpy(), want)
    np.testing.assert_allclose(conf.cpu().numpy(), N.proposal_confidence(heads, z, want), rtol=1e-5)


@pytest.mark.parametrize("ks,H,cin,cout", [(3, 11, 32, 48), (1, 1, 64, 304), (3, 5, 16, 16)])
def test_conv_kernel_shapes(gpu, ks, H, cin, cout):
    """The MFMA convolution on partial tiles (Cout not a multiple of 64, pixel
    counts not a multiple of 256) against float64 on the bf16 inputs."""
    import torch
    from isac_amd import posenet
    lib = posenet._lib()
    rng = np.random.default_rng(ks * 100 + cout)
    n = 3
    x = N.bf16(rng.standard_normal((n, H, H, cin)))
    w = N.bf16(rng.standard_normal((cout, ks * ks * cin)) * 0.1)
    b = rng.standard_normal(cout).astype(np.float32)
    xt = torch.from_numpy(x.astype(np.float32)).cuda().to(torch.bfloat16)
    wt = torch.from_numpy(w.astype(np.float32)).cuda().to(torch.bfloat16)
    bt = torch.from_numpy(b).cuda()
    y = torch.empty((n, H, H, cout), dtype=torch.float32, device="cuda")
    posenet._call(lib, "isac_nn_conv_fwd", xt.data_ptr(), n, H, H, cin, wt.data_ptr(), bt.data_ptr(),
                  1, cout, ks, 1, 1, y.data_ptr(), torch.cuda.current_stream().cuda_stream)
    want = np.maximum(N.conv(x, w, b.astype(np.float64), ks), 0.0)
    np.testing.assert_allclose(y.cpu().numpy(), want, rtol=1e-5, atol=1e-4)


@pytest.mark.parametrize("pool4", [True, False])
def test_fused_inference_forward_is_bit_identical(gpu, pool4):
    """Inference fuses the 2x2 pools into the convolutions; the logits must be
    bit-identical to the training forward's (unfused kernels + pool)."""
    import torch
    from isac_amd import posenet
    net = posenet.PoseNet(100, 49, D=76, seed=2, pool4=pool4,
                          channels=(32, 64, 128, 128 if pool4 else 32))
    _perturbed_params(net, 6)
    img = torch.from_numpy(_images(37, 76, 8)).cuda()
    a = net.forward(img).clone()
    b = net.forward(img, train=True).clone()
    torch.cuda.synchronize()
    assert torch.equal(a, b)


@pytest.mark.parametrize("H,cin,cout", [(38, 32, 64), (19, 64, 128), (7, 32, 128), (11, 64, 32),
                                        (8, 32, 32), (13, 64, 64), (3, 64, 128), (76, 16, 64),
                                        (10, 16, 256), (9, 128, 128), (6, 16, 48)])
def test_pooled_conv_is_bit_identical_to_conv_then_pool(gpu, H, cin, cout):
    """The pooled convolution equals the convolution followed by the 2x2 max
    pool, bit for bit: the resident-weight form (Cin 32 / 64, Cout 32 / 64 /
    128, one or several bands per image, one to four wavefronts per output
    block), the banded form (Cin 16 / 128) and the unbanded kernel (Cout 48)."""
    import torch
    from isac_amd import posenet
    lib = posenet._lib()
    rng = np.random.default_rng(H * 1000 + cout)
    n = 5
    x = N.bf16(rng.standard_normal((n, H, H, cin)))
    w = N.bf16(rng.standard_normal((cout, 9 * cin)) * 0.1)
    b = rng.standard_normal(cout).astype(np.float32)
    xt = torch.from_numpy(x.astype(np.float32)).cuda().to(torch.bfloat16)
    wt = torch.from_numpy(w.astype(np.float32)).cuda().to(torch.bfloat16)
    bt = torch.from_numpy(b).cuda()
    st = torch.cuda.current_stream().cuda_stream
    y = torch.empty((n, H, H, cout), dtype=torch.bfloat16, device="cuda")
    posenet._call(lib, "isac_nn_conv_fwd", xt.data_ptr(), n, H, H, cin, wt.data_ptr(), bt.data_ptr(),
                  1, cout, 3, 1, 0, y.data_ptr(), st)
    want = torch.empty((n, H // 2, H // 2, cout), dtype=torch.bfloat16, device="cuda")
    posenet._call(lib, "isac_nn_maxpool2", y.data_ptr(), n, H, H, cout, want.data_ptr(), st)
    got = torch.full((n, H // 2, H // 2, cout), float("nan"), dtype=torch.bfloat16, device="cuda")
    posenet._call(lib, "isac_nn_conv_pool_fwd", xt.data_ptr(), n, H, H, cin, wt.data_ptr(),
                  bt.data_ptr(), 1, cout, got.data_ptr(), st)
    torch.cuda.synchronize()
    assert torch.equal(got.view(torch.int16), want.view(torch.int16))


@pytest.mark.parametrize("H,cin,cout,relu,bias", [(9, 128, 32, 1, True), (38, 32, 64, 1, True),
                                                  (19, 64, 32, 0, False), (19, 128, 64, 0, False),
                                                  (11, 32, 16, 1, True), (3, 128, 16, 1, True)])
def test_resident_conv_is_bit_identical(gpu, H, cin, cout, relu, bias):
    """The unpooled 3x3 convolution in its resident-weight form (Cin 32 / 64 /
    128, Cout 16 / 32 / 64, one or several bands) equals the unbanded MFMA
    kernel bit for bit (run at Cout 80, which has no resident form)."""
    import torch
    from isac_amd import posenet
    lib = posenet._lib()
    rng = np.random.default_rng(H * 7 + cin + cout)
    n = 3
    x = torch.from_numpy(rng.standard_normal((n, H, H, cin)).astype(np.float32)).cuda().to(torch.bfloat16)
    w = torch.from_numpy((rng.standard_normal((80, 9 * cin)) * 0.1).astype(np.float32)).cuda()
    w = w.to(torch.bfloat16)
    b = torch.from_numpy(rng.standard_normal(80).astype(np.float32)).cuda()
    st = torch.cuda.current_stream().cuda_stream
    bp = b.data_ptr() if bias else None
    got = torch.empty((n, H, H, cout), dtype=torch.bfloat16, device="cuda")
    posenet._call(lib, "isac_nn_conv_fwd", x.data_ptr(), n, H, H, cin, w.data_ptr(), bp, 1, cout, 3,
                  relu, 0, got.data_ptr(), st)
    # Cout 80 is not a resident shape: the unbanded kernel; its first Cout channels
    # see the same weights and bias
    ref = torch.empty((n, H, H, 80), dtype=torch.bfloat16, device="cuda")
    posenet._call(lib, "isac_nn_conv_fwd", x.data_ptr(), n, H, H, cin, w.data_ptr(), bp, 1, 80,
                  3, relu, 0, ref.data_ptr(), st)
    torch.cuda.synchronize()
    assert torch.equal(got.view(torch.int16), ref[..., :cout].contiguous().view(torch.int16))
