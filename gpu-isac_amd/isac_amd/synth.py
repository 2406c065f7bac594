"""Seeded synthetic particles for the 2D alignment path.

References are sums of 5-15 anisotropic Gaussian blobs inside a given radius.
A particle is the analytic blob sum evaluated at pixel centres after the
transform  x -> shift(v) . rotate(phi) . mirror  (about the image centre
D//2), plus i.i.d. Gaussian noise with variance var(clean inside radius)/SNR.

With the alignment path's conventions (angle = 360 - psi, +180 when
mirrored) a noiseless particle generated with (phi, v, m) is recovered as
angle = phi, shift = v, mirror = m.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class BlobSet:
    """Blob parameters of N references: per blob (ref index, cx, cy, s1, s2, theta, amp)."""
    ref: np.ndarray
    cx: np.ndarray
    cy: np.ndarray
    s1: np.ndarray
    s2: np.ndarray
    th: np.ndarray
    amp: np.ndarray
    n_ref: int


def make_blobs(n_ref: int, radius: float, seed: int) -> BlobSet:
    """Draw 5-15 anisotropic blobs per reference, centres within ``radius``."""
    rng = np.random.default_rng(seed)
    rows = []
    for r in range(n_ref):
        nb = int(rng.integers(5, 16))
        for _ in range(nb):
            s1 = rng.uniform(1.5, 0.22 * radius)
            s2 = rng.uniform(0.5, 1.0) * s1
            rc = np.sqrt(rng.uniform()) * max(radius - 2.0 * s1, 0.25 * radius)
            ac = rng.uniform(0, 2 * np.pi)
            rows.append((r, rc * np.cos(ac), rc * np.sin(ac), s1, s2,
                         rng.uniform(0, np.pi), rng.uniform(0.4, 1.6)))
    a = np.array(rows, dtype=np.float64)
    return BlobSet(a[:, 0].astype(np.int64), a[:, 1], a[:, 2], a[:, 3], a[:, 4],
                   a[:, 5], a[:, 6], n_ref)


def _eval(blobs: BlobSet, ref_idx: np.ndarray, X: np.ndarray, Y: np.ndarray) -> np.ndarray:
    """Blob sum of reference ref_idx[b] at coordinates X[b], Y[b] (relative to centre)."""
    out = np.zeros(X.shape, dtype=np.float64)
    ref_idx = np.asarray(ref_idx)
    for r in np.unique(ref_idx):
        bsel = np.flatnonzero(ref_idx == r)
        Xr, Yr = X[bsel], Y[bsel]
        acc = np.zeros(Xr.shape, dtype=np.float64)
        for j in np.flatnonzero(blobs.ref == r):
            ct, st = np.cos(blobs.th[j]), np.sin(blobs.th[j])
            dx = Xr - blobs.cx[j]
            dy = Yr - blobs.cy[j]
            u = dx * ct + dy * st
            v = -dx * st + dy * ct
            acc += blobs.amp[j] * np.exp(-0.5 * (u * u / blobs.s1[j] ** 2
                                                  + v * v / blobs.s2[j] ** 2))
        out[bsel] = acc
    return out


def render_refs(blobs: BlobSet, D: int) -> np.ndarray:
    """Clean references (N, D, D) float32 sampled at pixel centres about D//2."""
    c = D // 2
    yy, xx = np.meshgrid(np.arange(D) - c, np.arange(D) - c, indexing="ij")
    X = np.broadcast_to(xx, (blobs.n_ref, D, D)).astype(np.float64)
    Y = np.broadcast_to(yy, (blobs.n_ref, D, D)).astype(np.float64)
    return _eval(blobs, np.arange(blobs.n_ref), X, Y).astype(np.float32)


def render_particles(blobs: BlobSet, D: int, ref_idx, phi_deg, shift_xy, mirror,
                     snr, radius, seed) -> np.ndarray:
    """Transformed, noisy particles (P, D, D) float32.

    Inverse map of a particle pixel q (relative to D//2): q - v, rotate by -phi,
    then mirror x -> -x.  Noise sigma^2 = var(clean within radius)/snr, drawn
    from a generator seeded with ``seed`` (snr=None or inf -> no noise).
    """
    ref_idx = np.asarray(ref_idx)
    P = ref_idx.shape[0]
    c = D // 2
    yy, xx = np.meshgrid(np.arange(D) - c, np.arange(D) - c, indexing="ij")
    phi = np.radians(np.asarray(phi_deg, dtype=np.float64)).reshape(-1, 1, 1)
    sh = np.asarray(shift_xy, dtype=np.float64).reshape(-1, 2)
    qx = xx[None] - sh[:, 0].reshape(-1, 1, 1)
    qy = yy[None] - sh[:, 1].reshape(-1, 1, 1)
    cp, sp = np.cos(phi), np.sin(phi)
    rx = qx * cp + qy * sp
    ry = -qx * sp + qy * cp
    mir = np.asarray(mirror).reshape(-1, 1, 1)
    rx = np.where(mir == 1, -rx, rx)
    clean = _eval(blobs, ref_idx, rx, ry)
    if snr is None or not np.isfinite(snr):
        return clean.astype(np.float32)
    inside = (xx * xx + yy * yy) <= radius * radius
    var = clean[:, inside].var(axis=1)
    rng = np.random.default_rng(seed)
    snr_arr = np.broadcast_to(np.asarray(snr, dtype=np.float64), (P,))
    sigma = np.sqrt(var / snr_arr).reshape(-1, 1, 1)
    noise = rng.standard_normal(clean.shape) * sigma
    return (clean + noise).astype(np.float32)


@dataclass
class Dataset:
    sbj: np.ndarray        # (P, D, D) float32
    ref: np.ndarray        # (N, D, D) float32
    ref_idx: np.ndarray    # generating reference per particle
    phi: np.ndarray        # generating in-plane angle (deg)
    shift: np.ndarray      # generating shift (P, 2)
    mirror: np.ndarray     # generating mirror flag


def make_dataset(n_particles: int, n_refs: int, D: int, radius: float, snr,
                 max_shift: float, integer_shifts: bool, seed_refs: int = 1,
                 seed_particles: int = 2, seed_noise: int = 3) -> Dataset:
    """Seeded dataset (recommended seeds: references 1, particles 2, noise 3)."""
    blobs = make_blobs(n_refs, radius, seed_refs)
    ref = render_refs(blobs, D)
    rng = np.random.default_rng(seed_particles)
    ref_idx = rng.integers(0, n_refs, n_particles)
    phi = rng.uniform(0.0, 360.0, n_particles)
    if integer_shifts:
        m = int(round(max_shift))
        shift = rng.integers(-m, m + 1, (n_particles, 2)).astype(np.float64)
    else:
        shift = rng.uniform(-max_shift, max_shift, (n_particles, 2))
    mirror = rng.integers(0, 2, n_particles)
    sbj = np.empty((n_particles, D, D), dtype=np.float32)
    chunk = 512
    for p0 in range(0, n_particles, chunk):
        p1 = min(n_particles, p0 + chunk)
        sbj[p0:p1] = render_particles(blobs, D, ref_idx[p0:p1], phi[p0:p1], shift[p0:p1],
                                      mirror[p0:p1], snr, radius, seed_noise * 100003 + p0)
    return Dataset(sbj, ref, ref_idx, phi, shift, mirror)


def dataset_p1k_n8_d64(n_particles: int = 1000, snr=0.1) -> Dataset:
    """1,000 x 64^2 particles, 8 refs (r <= 28), int shifts +-2, SNR 0.1."""
    return make_dataset(n_particles, 8, 64, 28.0, snr, 2, True)


def dataset_p20k_n100_d76(n_particles: int = 20000, snr=0.05, n_refs: int = 100) -> Dataset:
    """20,000 x 76^2 particles, 100 refs (r <= 34), sub-pixel shifts +-3, SNR 0.05."""
    return make_dataset(n_particles, n_refs, 76, 34.0, snr, 3.0, False)


def make_dataset_torch(n_particles: int, n_refs: int, D: int, radius: float, snr,
                       max_shift: float, integer_shifts: bool, device="cuda",
                       seed_refs: int = 1, seed_particles: int = 2, seed_noise: int = 3,
                       chunk: int = 1024):
    """Same model and parameter draws as make_dataset, rendered with torch on
    ``device`` in float32 (fast for large particle counts; the noise stream
    differs from the numpy one).  Returns (sbj, ref) device tensors plus the
    generating parameters as numpy arrays."""
    import torch
    blobs = make_blobs(n_refs, radius, seed_refs)
    ref = render_refs(blobs, D)
    rng = np.random.default_rng(seed_particles)
    ref_idx = rng.integers(0, n_refs, n_particles)
    phi = rng.uniform(0.0, 360.0, n_particles)
    if integer_shifts:
        m = int(round(max_shift))
        shift = rng.integers(-m, m + 1, (n_particles, 2)).astype(np.float64)
    else:
        shift = rng.uniform(-max_shift, max_shift, (n_particles, 2))
    mirror = rng.integers(0, 2, n_particles)
    # padded blob table [n_refs, B, 7]: cx, cy, 1/s1^2, 1/s2^2, cos th, sin th, amp
    B = int(np.bincount(blobs.ref, minlength=n_refs).max())
    tab = np.zeros((n_refs, B, 7), np.float32)
    cnt = np.zeros(n_refs, np.int64)
    for j in range(blobs.ref.shape[0]):
        r = blobs.ref[j]
        tab[r, cnt[r]] = (blobs.cx[j], blobs.cy[j], 1.0 / blobs.s1[j] ** 2, 1.0 / blobs.s2[j] ** 2,
                          np.cos(blobs.th[j]), np.sin(blobs.th[j]), blobs.amp[j])
        cnt[r] += 1
    dev = torch.device(device)
    tab_t = torch.from_numpy(tab).to(dev)
    c = D // 2
    ax = torch.arange(D, device=dev, dtype=torch.float32) - c
    yy, xx = torch.meshgrid(ax, ax, indexing="ij")
    inside = (xx * xx + yy * yy) <= radius * radius
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed_noise)
    sbj = torch.empty((n_particles, D, D), dtype=torch.float32, device=dev)
    for p0 in range(0, n_particles, chunk):
        p1 = min(n_particles, p0 + chunk)
        ph = torch.from_numpy(np.radians(phi[p0:p1])).float().to(dev).view(-1, 1, 1)
        sh = torch.from_numpy(shift[p0:p1]).float().to(dev)
        qx = xx[None] - sh[:, 0].view(-1, 1, 1)
        qy = yy[None] - sh[:, 1].view(-1, 1, 1)
        cp, sp = torch.cos(ph), torch.sin(ph)
        rx = qx * cp + qy * sp
        ry = -qx * sp + qy * cp
        mir = torch.from_numpy(mirror[p0:p1]).to(dev).view(-1, 1, 1)
        rx = torch.where(mir == 1, -rx, rx)
        t = tab_t[torch.from_numpy(ref_idx[p0:p1]).to(dev)]          # (b, B, 7)
        dx = rx[:, None] - t[:, :, 0, None, None]
        dy = ry[:, None] - t[:, :, 1, None, None]
        u = dx * t[:, :, 4, None, None] + dy * t[:, :, 5, None, None]
        v = -dx * t[:, :, 5, None, None] + dy * t[:, :, 4, None, None]
        img = (t[:, :, 6, None, None] * torch.exp(-0.5 * (u * u * t[:, :, 2, None, None]
                                                          + v * v * t[:, :, 3, None, None]))).sum(1)
        if snr is not None and np.isfinite(snr):
            var = img[:, inside].var(dim=1, unbiased=False)
            sig = torch.sqrt(var / float(snr)).view(-1, 1, 1)
            img = img + sig * torch.randn(img.shape, generator=gen, device=dev)
        sbj[p0:p1] = img
    return sbj, torch.from_numpy(ref).to(dev), Dataset(None, ref, ref_idx, phi, shift, mirror)


BATCHED_BOXES = (64, 76, 96, 128)


def uneven_subset_shapes(n_subsets: int = 64, seed: int = 4):
    """Shapes of the batched uneven workload: per subset a particle count
    uniform in 500..5000, a box from {64, 76, 96, 128}, a reference count
    uniform in 20..200, a shift range of +-2..+-4 (step 1), rings 1..box/2-4 and
    SNR uniform in 0.05..0.2."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_subsets):
        D = int(rng.choice(BATCHED_BOXES))
        out.append({"P": int(rng.integers(500, 5001)), "D": D, "N": int(rng.integers(20, 201)),
                    "shift_rng": float(rng.integers(2, 5)), "ring_num": D // 2 - 4,
                    "snr": float(rng.uniform(0.05, 0.2)), "seed": 1000 + 3 * i})
    return out


def make_subsets_torch(shapes, device="cuda"):
    """Render every subset of ``shapes`` on the device: blob references inside
    radius box/2-4, particles with random psi, mirror and sub-pixel shifts within
    the subset's search range.  Returns dicts ready for binding.BatchedMref."""
    subs = []
    for sh in shapes:
        D = sh["D"]
        sbj, ref, _ = make_dataset_torch(sh["P"], sh["N"], D, float(D // 2 - 4), sh["snr"],
                                         sh["shift_rng"], False, device, seed_refs=sh["seed"],
                                         seed_particles=sh["seed"] + 1, seed_noise=sh["seed"] + 2)
        subs.append({"sbj": sbj, "ref": ref, "ring_num": sh["ring_num"],
                     "shift_rng": sh["shift_rng"], "acc_xy": None})
    return subs
