"""CPU oracle of the pose-proposal network (test infrastructure only).

The reference (ucdavis/GPU-ISAC) has no learned component: its search is
exhaustive (cuda/gpu_aln_noref.cu:326-554).  This module states, in numpy, the
network whose forward, backward and optimiser kernels live in
gpu-isac_amd/csrc/posenet.hip, so that the GPU's activations and gradients can
be checked.  Only tests and benchmark checks import it.

Network (NHWC, ``quant=True`` rounds to bf16 where the GPU stores bf16):
  x0 = q(standardize(img))                       [n, D, D, 1]
  y_l = q(relu(conv3x3(a_{l-1}, q(W_l)) + b_l)),  a_l = maxpool2(y_l)   l = 1..4
  h = q(relu(flatten(a_4) q(W_5)^T + b_5))
  logits = h q(W_6)^T + b_6                       (fp32 on the GPU, not rounded)
  loss = sum over the three heads of softmax cross-entropy, mean over n.
Master parameters of a layer are [Cout][K+1] (weights in (tap, ci) order, then
the bias).  Backward: the gradient of the mean loss, passing through the bf16
roundings unchanged (straight-through), with the GPU's rounding of dlogits and
of every input gradient to bf16.

Parity pins: ``quant=False`` gradients are checked against central finite
differences in float64 (tests/test_posenet_oracle.py); the im2col convolution
against a direct loop; the col2im adjoint against <conv(x), y> = <x, conv^T(y)>.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


def bf16(x):
    """Round to bfloat16 (nearest even) and return float64 values."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, (u | 0x400000) & 0xFFFF0000, r)
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


@dataclass
class NetSpec:
    D: int = 76
    channels: tuple = (32, 64, 128, 128)
    hidden: int = 384
    heads: tuple = (100, 144, 49)
    pool4: bool = True                           # 2x2 max pool after the last conv layer
    sizes: list = field(default_factory=list)   # spatial size entering conv l (l = 1..4)

    def __post_init__(self):
        s, self.sizes = self.D, []
        for _ in range(4):
            self.sizes.append(s)
            s //= 2
        if not self.pool4:
            s = self.sizes[3]
        self.feat = self.channels[3] * s * s
        self.n_out = sum(self.heads)
        self.ld = (self.n_out + 15) // 16 * 16

    def layers(self):
        """(Cin, Cout, ks) of the six layers (fc2 padded to ld outputs)."""
        c = self.channels
        return [(1, c[0], 3), (c[0], c[1], 3), (c[1], c[2], 3), (c[2], c[3], 3),
                (self.feat, self.hidden, 1), (self.hidden, self.ld, 1)]


def init_params(spec: NetSpec, seed: int = 0):
    """He-normal weights, zero biases; fc2 rows beyond the heads are zero."""
    rng = np.random.default_rng(seed)
    out = []
    for li, (cin, cout, ks) in enumerate(spec.layers()):
        K = ks * ks * cin
        p = np.zeros((cout, K + 1), np.float32)
        p[:, :K] = rng.standard_normal((cout, K)) * np.sqrt(2.0 / K)
        if li == 5:
            p[spec.n_out:, :] = 0.0
            p[:, :K] *= 0.5
        out.append(p)
    return out


def standardize(img):
    """x0 = (img - mean) / (std + 1e-6) per image, float32 arithmetic as the GPU."""
    img = np.asarray(img, np.float32)
    n = img.shape[0]
    flat = img.reshape(n, -1).astype(np.float64)
    mu = flat.mean(1)
    sd = np.sqrt(np.maximum((flat * flat).mean(1) - mu * mu, 0.0))
    fm = mu.astype(np.float32)[:, None, None]
    fi = (1.0 / (sd + 1e-6)).astype(np.float32)[:, None, None]
    return ((img - fm) * fi).astype(np.float32)


def im2col(x, ks):
    """x [n, H, W, C] -> [n, H, W, ks*ks, C] (zero padding ks//2), tap = ky*ks + kx."""
    if ks == 1:
        return x[:, :, :, None, :]
    n, H, W, C = x.shape
    p = ks // 2
    xp = np.zeros((n, H + 2 * p, W + 2 * p, C), x.dtype)
    xp[:, p:p + H, p:p + W] = x
    return np.stack([xp[:, ky:ky + H, kx:kx + W] for ky in range(ks) for kx in range(ks)], axis=3)


def col2im(cols, ks):
    """Adjoint of im2col: cols [n, H, W, ks*ks, C] -> [n, H, W, C]."""
    if ks == 1:
        return cols[:, :, :, 0, :]
    n, H, W, T, C = cols.shape
    p = ks // 2
    xp = np.zeros((n, H + 2 * p, W + 2 * p, C), cols.dtype)
    t = 0
    for ky in range(ks):
        for kx in range(ks):
            xp[:, ky:ky + H, kx:kx + W] += cols[:, :, :, t]
            t += 1
    return xp[:, p:p + H, p:p + W]


def conv(x, w, b, ks):
    """x [n, H, W, Cin], w [Cout, ks*ks*Cin] -> x * w + b (float64)."""
    n, H, W, C = x.shape
    cols = im2col(x, ks).reshape(n * H * W, ks * ks * C)
    y = cols @ w.T
    if b is not None:
        y = y + b
    return y.reshape(n, H, W, -1)


def pool(y):
    """2x2 max pool (floor); also the first-max window index (order (0,0),(0,1),(1,0),(1,1))."""
    n, H, W, C = y.shape
    Ho, Wo = H // 2, W // 2
    c = y[:, :2 * Ho, :2 * Wo]
    cand = np.stack([c[:, 0::2, 0::2], c[:, 0::2, 1::2], c[:, 1::2, 0::2], c[:, 1::2, 1::2]], 0)
    arg = np.argmax(cand, axis=0)
    return np.take_along_axis(cand, arg[None], 0)[0], arg


def pool_bwd(y, arg, da):
    """Gradient through pool and the ReLU before it (first max, only if > 0)."""
    n, H, W, C = y.shape
    Ho, Wo = da.shape[1], da.shape[2]
    dz = np.zeros_like(y)
    for q, (oy, ox) in enumerate(((0, 0), (0, 1), (1, 0), (1, 1))):
        ys = y[:, oy:2 * Ho:2, ox:2 * Wo:2]
        sel = (arg == q) & (ys > 0)
        dz[:, oy:2 * Ho:2, ox:2 * Wo:2] = np.where(sel, da, 0.0)
    return dz


def forward(spec: NetSpec, params, img, quant=True):
    """All activations (float64) of the network for images [n, D, D]."""
    q = bf16 if quant else (lambda v: np.asarray(v, np.float64))
    n = img.shape[0]
    x0 = q(standardize(img)) if quant else standardize(img).astype(np.float64)
    act = {"x0": x0[..., None]}
    a = act["x0"]
    for l in range(4):
        cin, cout, ks = spec.layers()[l]
        P = params[l]
        y = q(np.maximum(conv(a, q(P[:, :-1]), P[:, -1].astype(np.float64), 3), 0.0))
        if l == 3 and not spec.pool4:
            a, arg = y, None
        else:
            a, arg = pool(y)
        act[f"y{l + 1}"], act[f"arg{l + 1}"], act[f"a{l + 1}"] = y, arg, a
    f = a.reshape(n, -1)
    act["f"] = f
    P5, P6 = params[4], params[5]
    h = q(np.maximum(f @ q(P5[:, :-1]).T + P5[:, -1], 0.0))
    act["h"] = h
    act["logits"] = h @ q(P6[:, :-1]).T + P6[:, -1]
    return act


def soft_targets(spec: NetSpec, labels, sigma_psi=0.0, sigma_shift=0.0, shift_ny=0):
    """Per-head target distributions: one-hot, or Gaussian over psi bins of the
    labelled mirror (circular, sigma in bins) and over the shift grid (x-major,
    shift_ny per row, sigma in steps)."""
    n = labels.shape[0]
    out = []
    for h, sz in enumerate(spec.heads):
        c = np.arange(sz)
        lab = labels[:, h][:, None]
        if h == 1 and sigma_psi > 0:
            npsi = sz // 2
            d = np.abs(c[None] % npsi - lab % npsi)
            d = np.minimum(d, npsi - d)
            t = np.where(c[None] // npsi == lab // npsi, np.exp(-0.5 * d * d / sigma_psi ** 2), 0.0)
        elif h == 2 and sigma_shift > 0:
            dx = c[None] // shift_ny - lab // shift_ny
            dy = c[None] % shift_ny - lab % shift_ny
            t = np.exp(-0.5 * (dx * dx + dy * dy) / sigma_shift ** 2)
        else:
            t = (c[None] == lab).astype(np.float64)
        out.append(t / t.sum(1, keepdims=True))
    return out


def xent(spec: NetSpec, logits, labels, sigma_psi=0.0, sigma_shift=0.0, shift_ny=0):
    """Per-particle summed head cross-entropies and d(mean loss)/d logits."""
    n = logits.shape[0]
    loss = np.zeros(n)
    g = np.zeros_like(logits)
    tg = soft_targets(spec, labels, sigma_psi, sigma_shift, shift_ny)
    o = 0
    for h, sz in enumerate(spec.heads):
        z = logits[:, o:o + sz]
        m = z.max(1, keepdims=True)
        lse = m[:, 0] + np.log(np.exp(z - m).sum(1))
        loss += (tg[h] * (lse[:, None] - z)).sum(1)
        pr = np.exp(z - lse[:, None])
        g[:, o:o + sz] = (pr - tg[h]) / n
        o += sz
    return loss, g


def backward(spec: NetSpec, params, act, labels, quant=True, **soft):
    """(per-particle loss, gradients [Cout][K+1] per layer); ``soft`` as xent."""
    q = bf16 if quant else (lambda v: np.asarray(v, np.float64))
    loss, g = xent(spec, act["logits"], labels, **soft)
    dl = q(g)
    grads = [None] * 6
    h, f = act["h"], act["f"]
    grads[5] = np.concatenate([dl.T @ h, dl.sum(0)[:, None]], 1)
    dh = q(dl @ q(params[5][:, :-1]))
    dzh = np.where(h > 0, dh, 0.0)
    grads[4] = np.concatenate([dzh.T @ f, dzh.sum(0)[:, None]], 1)
    da = q(dzh @ q(params[4][:, :-1])).reshape(act["a4"].shape)
    for l in range(3, -1, -1):
        cin, cout, ks = spec.layers()[l]
        y, arg = act[f"y{l + 1}"], act[f"arg{l + 1}"]
        dz = np.where(y > 0, da, 0.0) if arg is None else pool_bwd(y, arg, da)
        xin = act["x0"] if l == 0 else act[f"a{l}"]
        n, H, W, _ = xin.shape
        cols = im2col(xin, 3).reshape(n * H * W, 9 * cin)
        dzf = dz.reshape(n * H * W, cout)
        grads[l] = np.concatenate([dzf.T @ cols, dzf.sum(0)[:, None]], 1)
        if l > 0:
            dcols = (dzf @ q(params[l][:, :-1])).reshape(n, H, W, 9, cin)
            da = q(col2im(dcols, 3))
    return loss, grads


def adamw(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
    """One AdamW step (float32 state, as the GPU)."""
    p, g, m, v = (np.asarray(a, np.float32) for a in (p, g, m, v))
    m = np.float32(b1) * m + np.float32(1 - b1) * g
    v = np.float32(b2) * v + np.float32(1 - b2) * g * g
    c1 = np.float32(1 - b1 ** t)
    c2 = np.float32(1 - b2 ** t)
    p = p - np.float32(lr) * ((m / c1) / (np.sqrt(v / c2) + np.float32(eps)) + np.float32(wd) * p)
    return p, m, v


def proposals(spec_heads, logits, k=8, psi_r=-1, shift_r=1, shift_ny=1, pool=(8, 16, 16),
              min_refs=0):
    """Greedy top-k (ref, mirror, psi bin, shift) triples by summed head logits
    over the candidate pool (top pool[0] references x top pool[1] psi classes x
    top pool[2] shifts; ranks by value, ties to the lower index), best first,
    ties to the lower (ref, psi, shift) ranks.  psi_r >= 0 suppresses, after each
    choice, the candidates of the same reference and mirror within psi_r bins
    (circular) and shift_r steps in x and y.  Choice j < min_refs takes the best
    remaining candidate of reference rank j when there is one.  Brute force:
    small heads only."""
    n_ref, n_psi2, n_shift = spec_heads
    n_psi = n_psi2 // 2
    out = np.full((logits.shape[0], k, 4), 0, np.int64)
    for i, z in enumerate(np.asarray(logits, np.float32)):
        zr, zp, zs = z[:n_ref], z[n_ref:n_ref + n_psi2], z[n_ref + n_psi2:n_ref + n_psi2 + n_shift]
        rr = sorted(range(n_ref), key=lambda j: (-zr[j], j))[:pool[0]]
        rp = sorted(range(n_psi2), key=lambda j: (-zp[j], j))[:pool[1]]
        rs = sorted(range(n_shift), key=lambda j: (-zs[j], j))[:pool[2]]
        cand = []
        for a, ri in enumerate(rr):
            for b, pi in enumerate(rp):
                for c, si in enumerate(rs):
                    cand.append((-(np.float32(zr[ri]) + np.float32(zp[pi]) + np.float32(zs[si])),
                                 (a, b, c), ri, pi, si))
        cand.sort(key=lambda t: (t[0], t[1]))
        alive = [True] * len(cand)
        for j in range(k):
            pick = None
            if j < min_refs:
                pick = next((q for q in range(len(cand)) if alive[q] and cand[q][1][0] == j), None)
            if pick is None:
                pick = next((q for q in range(len(cand)) if alive[q]), None)
            if pick is None:
                out[i, j] = (-1, 0, 0, 0)
                continue
            _, _, ri, pi, si = cand[pick]
            out[i, j] = (ri, pi // n_psi, pi % n_psi, si)
            for q, (_, _, r2, p2, s2) in enumerate(cand):
                if not alive[q]:
                    continue
                if q == pick:
                    alive[q] = False
                elif psi_r >= 0 and r2 == ri and p2 // n_psi == pi // n_psi:
                    db = abs(p2 % n_psi - pi % n_psi)
                    db = min(db, n_psi - db)
                    if db <= psi_r and abs(s2 // shift_ny - si // shift_ny) <= shift_r and \
                            abs(s2 % shift_ny - si % shift_ny) <= shift_r:
                        alive[q] = False
    return out


def proposal_confidence(spec_heads, logits, prop):
    """Softmax mass of the reference head on each particle's distinct proposed
    references (entries with ref < 0 ignored)."""
    n_ref = spec_heads[0]
    out = np.zeros(len(prop))
    for i, z in enumerate(np.asarray(logits, np.float64)):
        zr = z[:n_ref]
        e = np.exp(zr - zr.max())
        refs = sorted(set(int(r) for r in np.asarray(prop[i])[:, 0] if r >= 0))
        out[i] = e[refs].sum() / e.sum()
    return out
