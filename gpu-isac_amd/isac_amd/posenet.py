"""Learned pose proposal: the network's host side (training and inference).

Every arithmetic step runs in the HIP kernels of csrc/posenet.hip through the
C ABI of include/isac_posenet.h; torch only allocates device memory and
provides streams and random permutations of the training set.  The reference
(ucdavis/GPU-ISAC) has no learned component; the exhaustive search it runs
(cuda/gpu_aln_noref.cu:326-554) produces the training labels and remains the
fallback of the narrowed search (isac_mref_proposal_device).

Network: 4 conv layers (3x3, 32/64/128/128 channels, ReLU, 2x2 max pool)
and a 2-layer MLP (384 hidden) with three heads: reference, psi class (mirror
x 5-degree bin, mirror-major) and shift index; about 1.1 M parameters at 76 px.
"""
from __future__ import annotations

import ctypes
import math
from typing import Optional

import numpy as np

from . import binding

PSI_BIN_DEG = 5.0


def _lib():
    lib = binding.load()
    if getattr(lib, "_posenet_bound", False):
        return lib
    vp, ci, cf, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_size_t
    sig = {
        "isac_nn_standardize": [vp, ci, ci, vp, vp],
        "isac_nn_conv1_fwd": [vp, ci, ci, ci, vp, ci, vp, vp],
        "isac_nn_conv1_pool_fwd": [vp, ci, ci, ci, vp, ci, vp, vp],
        "isac_nn_conv_pool_fwd": [vp, ci, ci, ci, ci, vp, vp, ci, ci, vp, vp],
        "isac_nn_conv_fwd": [vp, ci, ci, ci, ci, vp, vp, ci, ci, ci, ci, ci, vp, vp],
        "isac_nn_maxpool2": [vp, ci, ci, ci, ci, vp, vp],
        "isac_nn_maxpool2_bwd": [vp, vp, ci, ci, ci, ci, vp, vp],
        "isac_nn_relu_bwd": [vp, vp, sz, vp, vp],
        "isac_nn_conv_wgrad": [vp, vp, ci, ci, ci, ci, ci, ci, vp, vp, sz, vp],
        "isac_nn_softmax_xent": [vp, ci, ci, ci, ci, ci, vp, cf, cf, ci, vp, vp, vp],
        "isac_nn_adamw": [vp, vp, vp, vp, vp, sz, cf, cf, cf, cf, cf, ci, vp],
        "isac_nn_pack_weights": [vp, ci, ci, ci, vp, vp, vp],
        "isac_nn_proposals": [vp, ci, ci, ci, ci, ci, ci, ci, ci, ci, ci, vp, vp, vp],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ci
    lib._posenet_bound = True
    return lib


def _call(lib, name, *args):
    rc = getattr(lib, name)(*args)
    if rc:
        binding.check(lib)
        raise binding.GpuAlignError(f"{name} returned {rc}")


def _p(t):
    return None if t is None else t.data_ptr()


def he_init(layers, n_out, seed):
    """He-normal weights ([Cout][K+1], bias column zero); the last layer is
    scaled by 0.5 and its padding rows (beyond n_out) stay zero."""
    rng = np.random.default_rng(seed)
    out = []
    for li, (cin, cout, ks) in enumerate(layers):
        K = ks * ks * cin
        p = np.zeros((cout, K + 1), np.float32)
        p[:, :K] = rng.standard_normal((cout, K)) * np.sqrt(2.0 / K)
        if li == len(layers) - 1:
            p[n_out:, :] = 0.0
            p[:, :K] *= 0.5
        out.append(p)
    return out


class PoseNet:
    """The proposal network on cuda:0: fp32 master weights, bf16 copies for the
    kernels, AdamW state; activations of the last training forward are kept
    for the backward."""

    def __init__(self, n_ref: int, n_shift: int, D: int = 76, n_psi: int = 72,
                 channels=(32, 64, 128, 128), hidden: int = 384, seed: int = 0,
                 pool4: bool = True, device="cuda"):
        import torch
        if channels[0] != 32 or any(c % 16 for c in channels[1:]) or hidden % 16:
            raise ValueError("channels must be (32, 16k, 16k, 16k) and hidden a multiple of 16")
        self.lib = _lib()
        self.dev = torch.device(device)
        self.D, self.n_ref, self.n_psi, self.n_shift = D, n_ref, n_psi, n_shift
        self.heads = (n_ref, 2 * n_psi, n_shift)
        self.channels, self.hidden = tuple(channels), hidden
        self.sizes = []
        s = D
        for _ in range(4):
            self.sizes.append(s)
            s //= 2
        if s < 1:
            raise ValueError("image too small for four 2x2 poolings")
        self.pool4 = pool4
        if not pool4:
            s = self.sizes[3]
        self.feat = channels[3] * s * s
        self.n_out = sum(self.heads)
        self.ld = (self.n_out + 15) // 16 * 16
        c = channels
        self.layers = [(1, c[0], 3), (c[0], c[1], 3), (c[1], c[2], 3), (c[2], c[3], 3),
                       (self.feat, hidden, 1), (hidden, self.ld, 1)]
        sizes = [cout * (ks * ks * cin + 1) for cin, cout, ks in self.layers]
        self.offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        total = int(self.offsets[-1])
        f32 = torch.float32
        self.flat = torch.zeros(total, dtype=f32, device=self.dev)
        self.gflat = torch.zeros_like(self.flat)
        self.m = torch.zeros_like(self.flat)
        self.v = torch.zeros_like(self.flat)
        self.w, self.wflip = [], []
        for li, (cin, cout, ks) in enumerate(self.layers):
            self.w.append(torch.empty(cout * ks * ks * cin, dtype=torch.bfloat16, device=self.dev))
            self.wflip.append(None if li == 0 else
                              torch.empty(cout * ks * ks * cin, dtype=torch.bfloat16, device=self.dev))
        self.part = torch.empty(16 << 20, dtype=f32, device=self.dev)   # 64 MB wgrad scratch
        self.t = 0
        self.saved = None
        self.load_params(he_init(self.layers, self.n_out, seed))

    @property
    def n_params(self) -> int:
        return sum(cout * (ks * ks * cin + 1) for cin, cout, ks in self.layers[:-1]) + \
            self.n_out * (self.layers[-1][0] + 1)

    def master(self, li):
        cin, cout, ks = self.layers[li]
        return self.flat[self.offsets[li]:self.offsets[li + 1]].view(cout, ks * ks * cin + 1)

    def grad(self, li):
        cin, cout, ks = self.layers[li]
        return self.gflat[self.offsets[li]:self.offsets[li + 1]].view(cout, ks * ks * cin + 1)

    def load_params(self, params):
        import torch
        for li, p in enumerate(params):
            self.master(li).copy_(torch.as_tensor(np.asarray(p, np.float32)))
        self.pack()

    def params_numpy(self):
        return [self.master(li).cpu().numpy() for li in range(len(self.layers))]

    def pack(self):
        for li, (cin, cout, ks) in enumerate(self.layers):
            _call(self.lib, "isac_nn_pack_weights", _p(self.master(li)), cout, cin, ks,
                  _p(self.w[li]), _p(self.wflip[li]), self._st())

    def _st(self):
        import torch
        return torch.cuda.current_stream(self.dev).cuda_stream

    def _bias(self, li):
        cin, cout, ks = self.layers[li]
        K = ks * ks * cin
        return self.master(li).data_ptr() + 4 * K, K + 1

    # ------------------------------------------------------------------ forward
    def forward(self, img, train: bool = False):
        """Logits [n, ld] (fp32) of images [n, D, D] (fp32, contiguous, cuda)."""
        import torch
        n, D = img.shape[0], self.D
        assert img.shape[1:] == (D, D) and img.dtype == torch.float32 and img.is_contiguous()
        bf, st, lib = torch.bfloat16, self._st(), self.lib
        e = lambda *shape, dt=bf: torch.empty(shape, dtype=dt, device=self.dev)  # noqa: E731
        x0 = e(n, D, D)
        _call(lib, "isac_nn_standardize", _p(img), n, D, _p(x0), st)
        sv = {"x0": x0}
        if not train:   # pooled layers with the pool fused into the convolution
            a = e(n, D // 2, D // 2, self.channels[0])
            _call(lib, "isac_nn_conv1_pool_fwd", _p(x0), n, D, D, _p(self.master(0)),
                  self.channels[0], _p(a), st)
            for li in range(1, 4):
                H = self.sizes[li]
                cin, cout, _ = self.layers[li]
                b, bs = self._bias(li)
                if li == 3 and not self.pool4:
                    y = e(n, H, H, cout)
                    _call(lib, "isac_nn_conv_fwd", _p(a), n, H, H, cin, _p(self.w[li]), b, bs, cout,
                          3, 1, 0, _p(y), st)
                    a = y
                else:
                    an = e(n, H // 2, H // 2, cout)
                    _call(lib, "isac_nn_conv_pool_fwd", _p(a), n, H, H, cin, _p(self.w[li]), b, bs,
                          cout, _p(an), st)
                    a = an
            return self._head(a, n, sv, train)
        y = e(n, D, D, self.channels[0])
        _call(lib, "isac_nn_conv1_fwd", _p(x0), n, D, D, _p(self.master(0)), self.channels[0],
              _p(y), st)
        for li in range(4):
            H = self.sizes[li]
            cin, cout, _ = self.layers[li]
            if li > 0:
                y = e(n, H, H, cout)
                b, bs = self._bias(li)
                _call(lib, "isac_nn_conv_fwd", _p(a), n, H, H, cin, _p(self.w[li]), b, bs, cout, 3,
                      1, 0, _p(y), st)
            if li == 3 and not self.pool4:
                a = y
            else:
                a = e(n, H // 2, H // 2, cout)
                _call(lib, "isac_nn_maxpool2", _p(y), n, H, H, cout, _p(a), st)
            sv[f"y{li + 1}"], sv[f"a{li + 1}"] = y, a
        return self._head(a, n, sv, train)

    def _head(self, a, n, sv, train):
        import torch
        st, lib = self._st(), self.lib
        e = lambda *shape, dt=torch.bfloat16: torch.empty(shape, dtype=dt, device=self.dev)  # noqa: E731
        h = e(n, self.hidden)
        b, bs = self._bias(4)
        _call(lib, "isac_nn_conv_fwd", _p(a), n, 1, 1, self.feat, _p(self.w[4]), b, bs,
              self.hidden, 1, 1, 0, _p(h), st)
        logits = e(n, self.ld, dt=torch.float32)
        b, bs = self._bias(5)
        _call(lib, "isac_nn_conv_fwd", _p(h), n, 1, 1, self.hidden, _p(self.w[5]), b, bs,
              self.ld, 1, 0, 1, _p(logits), st)
        sv["h"], sv["logits"] = h, logits
        self.saved = sv if train else None
        return logits

    # ----------------------------------------------------------------- backward
    def backward(self, labels, sigma_psi: float = 0.0, sigma_shift: float = 0.0):
        """Gradients of the mean loss of the last training forward into the
        gradient buffer; returns the per-particle loss (device tensor).
        sigma_psi / sigma_shift > 0 use Gaussian targets over neighbouring psi
        bins / shifts (the shift grid is taken as square)."""
        import torch
        sv = self.saved
        if sv is None:
            raise RuntimeError("backward() needs forward(..., train=True) first")
        n = sv["logits"].shape[0]
        assert labels.shape == (n, 3) and labels.dtype == torch.int32 and labels.is_contiguous()
        bf, st, lib = torch.bfloat16, self._st(), self.lib
        e = lambda *shape: torch.empty(shape, dtype=bf, device=self.dev)  # noqa: E731
        loss = torch.empty(n, dtype=torch.float32, device=self.dev)
        dl = e(n, self.ld)
        h0, h1, h2 = self.heads
        ny = int(round(math.sqrt(h2)))
        _call(lib, "isac_nn_softmax_xent", _p(sv["logits"]), n, self.ld, h0, h1, h2, _p(labels),
              sigma_psi, sigma_shift, ny if ny * ny == h2 else 0, _p(loss), _p(dl), st)
        pb = self.part.numel() * 4
        _call(lib, "isac_nn_conv_wgrad", _p(sv["h"]), _p(dl), n, 1, 1, self.hidden, self.ld, 1,
              _p(self.grad(5)), _p(self.part), pb, st)
        dh = e(n, self.hidden)
        _call(lib, "isac_nn_conv_fwd", _p(dl), n, 1, 1, self.ld, _p(self.wflip[5]), None, 0,
              self.hidden, 1, 0, 0, _p(dh), st)
        dzh = e(n, self.hidden)
        _call(lib, "isac_nn_relu_bwd", _p(sv["h"]), _p(dh), n * self.hidden, _p(dzh), st)
        _call(lib, "isac_nn_conv_wgrad", _p(sv["a4"]), _p(dzh), n, 1, 1, self.feat, self.hidden,
              1, _p(self.grad(4)), _p(self.part), pb, st)
        da = e(n, self.feat)
        _call(lib, "isac_nn_conv_fwd", _p(dzh), n, 1, 1, self.hidden, _p(self.wflip[4]), None, 0,
              self.feat, 1, 0, 0, _p(da), st)
        for li in range(3, -1, -1):
            H = self.sizes[li]
            cin, cout, _ = self.layers[li]
            dz = e(n, H, H, cout)
            if li == 3 and not self.pool4:
                _call(lib, "isac_nn_relu_bwd", _p(sv["y4"]), _p(da), n * H * H * cout, _p(dz), st)
            else:
                _call(lib, "isac_nn_maxpool2_bwd", _p(sv[f"y{li + 1}"]), _p(da), n, H, H, cout,
                      _p(dz), st)
            xin = sv["x0"] if li == 0 else sv[f"a{li}"]
            _call(lib, "isac_nn_conv_wgrad", _p(xin), _p(dz), n, H, H, cin, cout, 3,
                  _p(self.grad(li)), _p(self.part), pb, st)
            if li > 0:
                da = e(n, H, H, cin)
                _call(lib, "isac_nn_conv_fwd", _p(dz), n, H, H, cout, _p(self.wflip[li]), None,
                      0, cin, 3, 0, 0, _p(da), st)
        return loss

    def step(self, lr: float, wd: float = 0.0, beta1: float = 0.9, beta2: float = 0.999,
             eps: float = 1e-8):
        """One AdamW step over every parameter, then the bf16 copies."""
        self.t += 1
        _call(self.lib, "isac_nn_adamw", _p(self.flat), _p(self.gflat), _p(self.m), _p(self.v),
              None, self.flat.numel(), lr, beta1, beta2, eps, wd, self.t, self._st())
        self.pack()

    # ---------------------------------------------------------------- training
    def train(self, imgs, labels, epochs: int, batch: int = 256, lr: float = 2e-3,
              wd: float = 1e-4, warmup: int = 200, seed: int = 0, sigma_psi: float = 0.0,
              sigma_shift: float = 0.0, log=None):
        """Mini-batch AdamW with linear warm-up and cosine decay over ``epochs``
        passes of (imgs [N, D, D], labels [N, 3] int32), both on the device."""
        import torch
        N = imgs.shape[0]
        nb = N // batch
        total = max(1, epochs * nb)
        g = torch.Generator(device=self.dev)
        g.manual_seed(seed)
        it = 0
        hist = []
        for ep in range(epochs):
            perm = torch.randperm(N, generator=g, device=self.dev)
            acc_loss = torch.zeros((), dtype=torch.float64, device=self.dev)
            for b in range(nb):
                idx = perm[b * batch:(b + 1) * batch]
                x = imgs.index_select(0, idx).contiguous()
                y = labels.index_select(0, idx).contiguous()
                self.forward(x, train=True)
                loss = self.backward(y, sigma_psi, sigma_shift)
                w = min(1.0, (it + 1) / warmup)
                cur = lr * w * 0.5 * (1.0 + math.cos(math.pi * it / total))
                self.step(cur, wd)
                acc_loss += loss.double().sum()
                it += 1
            mean = float(acc_loss) / max(1, nb * batch)
            hist.append(mean)
            if log:
                log(f"epoch {ep + 1}/{epochs}: mean loss {mean:.4f}")
        self.saved = None
        return hist

    # --------------------------------------------------------------- inference
    def propose(self, imgs, k: int = 8, chunk: int = 4096, psi_r: int = 2, shift_r: int = 1,
                min_refs: int = 3, with_conf: bool = False):
        """Top-k proposals [n, k, 4] int32 (ref, mirror, psi bin, shift index),
        spread by suppressing neighbours within psi_r bins and shift_r steps
        (psi_r < 0: plain top-k of the summed logits), the first min_refs taken
        from the min_refs best-ranked references.  with_conf: also the softmax
        mass of the reference head on the proposed references [n]."""
        import torch
        n = imgs.shape[0]
        ny = int(round(math.sqrt(self.n_shift)))
        if ny * ny != self.n_shift:
            psi_r, ny = -1, 1
        prop = torch.empty((n, k, 4), dtype=torch.int32, device=self.dev)
        conf = torch.empty(n, dtype=torch.float32, device=self.dev) if with_conf else None
        for i0 in range(0, n, chunk):
            i1 = min(n, i0 + chunk)
            logits = self.forward(imgs[i0:i1])
            _call(self.lib, "isac_nn_proposals", _p(logits), i1 - i0, self.ld, self.n_ref,
                  self.n_psi, self.n_shift, k, psi_r, shift_r, ny, min_refs, _p(prop[i0:i1]),
                  None if conf is None else _p(conf[i0:i1]), self._st())
        return (prop, conf) if with_conf else prop


def labels_from_search(d, peak_list, shift_rng: float, shift_step: float = 1.0,
                       n_psi: int = 72):
    """Training labels [P, 3] int32 (best reference, psi class, shift index) from
    an exhaustive multi-reference result (d [N, P], peak_list [N, 4P]) computed
    with zero accumulated shifts.  psi = 360 - angle (+180 when mirrored)."""
    import torch
    P = d.shape[1]
    best = torch.argmax(d, dim=0)
    pk = peak_list.view(d.shape[0], P, 4)[best, torch.arange(P, device=d.device)]
    mir = (pk[:, 3] > 0.5).long()
    psi = torch.remainder(360.0 - pk[:, 0].double() + 180.0 * mir.double(), 360.0)
    b = torch.clamp((psi / (360.0 / n_psi)).floor().long(), 0, n_psi - 1)
    ny = int(round(2 * shift_rng / shift_step)) + 1
    ix = torch.round((pk[:, 1].double() + shift_rng) / shift_step).long()
    iy = torch.round((pk[:, 2].double() + shift_rng) / shift_step).long()
    s = ix * ny + iy
    return torch.stack([best, mir * n_psi + b, s], 1).to(torch.int32).contiguous()
