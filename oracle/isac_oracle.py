"""FP64 CPU oracle for the GPU-ISAC polar-ring 2D alignment path.

TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library imports, links or
calls this module; only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may use it, and only as the checker.

This is a numpy restatement, in double precision, of the semantics of the
reference CUDA library ``gpu_aln_pack.so``.  Each function cites the reference
source it restates (paths relative to the ucdavis/GPU-ISAC tree):

  polar_template          cuda/gpu_aln_common.cu:39-62
  shift_grid              cuda/gpu_aln_common.cu:64-86, cuda/gpu_aln_noref.cu:2380-2382
  bilinear_clamped        cuda/gpu_aln_noref.cu:1283-1287, :2314-2339 (texture fetch,
                          linear filter, clamp addressing, +0.5 texel-centre offset)
  polar_resample          cuda/gpu_aln_noref.cu:1257-1288
  normalize_polar         cuda/gpu_aln_noref.cu:1294-1351
  ring_spectra            cuda/gpu_aln_noref.cu:1885, :2121-2125 (R2C, unnormalised)
  ccf_spectra             cuda/gpu_aln_noref.cu:1483-1537 (and :1356-1481)
  psi_ccf                 cuda/gpu_aln_noref.cu:2393, :2453-2461 (C2R, unnormalised)
  parabola_delta          cuda/gpu_aln_noref.cu:1777-1787, :2535-2582
  multi_ref_align         cuda/gpu_aln_noref.cu:326-554, :1667-1794
  decode_single_ref       cuda/gpu_aln_noref.cu:2463-2520, :1624-1665
  transform_images        cuda/gpu_aln_noref.cu:1539-1590
  class_averages          cuda/gpu_aln_noref.cu:1592-1622
  tangent_filter          cuda/gpu_aln_noref.cu:1007-1037, :2203-2266
  driver_post_rotation    bin/sp_isac2_gpu.py:1016-1022
  filter_shift            the tanh low-pass of :1020-1035 in absolute frequency, with
                          the Fourier shift EMAN2's fshift applies (bin/sp_isac2_gpu.py:1221)
  update_references       bin/sp_isac2_gpu.py:1146-1232 with transform_images

Deliberate, documented divergences from the (unrunnable) CUDA path:
  * bilinear weights are exact (CUDA's texture unit uses 8-bit fractions);
  * clamp-to-edge is per image in both axes (the reference's stacked texture
    lets y over/underflow read the neighbouring particle);
  * the pre-align / reference-free row argmax scans only the L psi samples of
    each slot (the reference also scans two padding floats whose content
    after an in-place cuFFT C2R is unspecified).

Pinning: every function here is checked in ``tests/test_oracle.py`` against
closed forms, brute force on tiny inputs, library cross-checks (numpy FFT vs
O(L^2) DFT) or invariants (rotation equivariance, mirror, normalisation
moments).  See that file for which check pins which function.
"""
from __future__ import annotations

import numpy as np

RING_LEN = 256


# --------------------------------------------------------------------------- grid

def polar_template(ring_num: int, ring_len: int = RING_LEN):
    """Sampling offsets (x, y) = (i+1)(cos 2pi j/L, sin 2pi j/L), i<R, j<L.

    cuda/gpu_aln_common.cu:53-59.  Returns two (R, L) float64 arrays.
    """
    theta = 2.0 * np.pi * np.arange(ring_len) / ring_len
    rho = np.arange(1, ring_num + 1, dtype=np.float64)[:, None]
    return rho * np.cos(theta)[None, :], rho * np.sin(theta)[None, :]


def shift_grid(rng_x: float, rng_y: float, step: float):
    """Shift list in the reference's x-major order.

    cuda/gpu_aln_common.cu:78-83: ``for sx=-X; sx<=X; sx+=h: for sy=-Y; sy<=Y;
    sy+=h`` with float32 loop variables.  Returns an (S, 2) float32 array.
    """
    out = []
    sx = np.float32(-rng_x)
    while sx <= np.float32(rng_x):
        sy = np.float32(-rng_y)
        while sy <= np.float32(rng_y):
            out.append((sx, sy))
            sy = np.float32(sy + np.float32(step))
        sx = np.float32(sx + np.float32(step))
    return np.array(out, dtype=np.float32).reshape(-1, 2)


def shift_count(rng_x: float, rng_y: float, step: float) -> int:
    """S = (2X/h+1)(2Y/h+1)  (cuda/gpu_aln_noref.cu:2382)."""
    return int((2 * (rng_x / step) + 1) * (2 * (rng_y / step) + 1))


# ------------------------------------------------------------------ sampling

def bilinear_clamped(img: np.ndarray, x: np.ndarray, y: np.ndarray) -> np.ndarray:
    """Bilinear value of ``img[y, x]`` (row-major, pixel i centred at i).

    Restates an unnormalised-coordinate linear-filter texture fetch with clamp
    addressing (cuda/gpu_aln_noref.cu:2331-2335): the reference passes
    coordinate c+0.5 and the texture unit samples at (c+0.5)-0.5 = c, taking
    taps floor(c) and floor(c)+1, each clamped to [0, D-1].
    ``img`` may be (D, D) or (B, D, D) with x, y broadcasting to (B, ...).
    """
    h, w = img.shape[-2], img.shape[-1]
    x0 = np.floor(x)
    y0 = np.floor(y)
    fx = x - x0
    fy = y - y0
    x0 = x0.astype(np.int64)
    y0 = y0.astype(np.int64)
    xa = np.clip(x0, 0, w - 1)
    xb = np.clip(x0 + 1, 0, w - 1)
    ya = np.clip(y0, 0, h - 1)
    yb = np.clip(y0 + 1, 0, h - 1)
    if img.ndim == 2:
        f = img
        v00, v01 = f[ya, xa], f[ya, xb]
        v10, v11 = f[yb, xa], f[yb, xb]
    else:
        bidx = np.arange(img.shape[0]).reshape((-1,) + (1,) * (x.ndim - 1))
        v00, v01 = img[bidx, ya, xa], img[bidx, ya, xb]
        v10, v11 = img[bidx, yb, xa], img[bidx, yb, xb]
    return ((1 - fy) * ((1 - fx) * v00 + fx * v01)
            + fy * ((1 - fx) * v10 + fx * v11))


def polar_resample(imgs: np.ndarray, cx: np.ndarray, cy: np.ndarray,
                   ring_num: int, ring_len: int = RING_LEN) -> np.ndarray:
    """Polar samples of each image about (cx, cy) (pixel-index coordinates).

    cuda/gpu_aln_noref.cu:1262-1288: centre = img_dim/2 (integer division) +
    int grid shift + accumulated shift; sample at centre + template offset.
    ``imgs`` (B, D, D), ``cx``/``cy`` (B,) -> (B, R, L) float64.
    """
    ox, oy = polar_template(ring_num, ring_len)
    cx = np.asarray(cx, dtype=np.float64).reshape(-1, 1, 1)
    cy = np.asarray(cy, dtype=np.float64).reshape(-1, 1, 1)
    return bilinear_clamped(imgs.astype(np.float64), cx + ox[None], cy + oy[None])


def normalize_polar(p: np.ndarray) -> np.ndarray:
    """Radius-weighted mean/std normalisation of polar images (..., R, L).

    cuda/gpu_aln_noref.cu:1294-1351: A_i = sum_j p_ij, Q_i = sum_j p_ij^2,
    mu = sum (i+1) A_i / (L sum(i+1)), sd = sqrt|sum (i+1) Q_i/(L sum(i+1)) - mu^2|,
    p' = (p - mu)/(sd + 1e-7).  (The 2*pi factors of the source cancel.)
    """
    R, L = p.shape[-2], p.shape[-1]
    w = np.arange(1, R + 1, dtype=np.float64)
    A = p.sum(-1)
    Q = (p * p).sum(-1)
    nn = w.sum() * L
    mu = (A * w).sum(-1) / nn
    var = (Q * w).sum(-1) / nn - mu * mu
    sd = np.sqrt(np.abs(var))
    return (p - mu[..., None, None]) / (sd[..., None, None] + 1e-7)


# ------------------------------------------------------------------- spectra

def ring_spectra(p: np.ndarray) -> np.ndarray:
    """Unnormalised forward real FFT of every ring: P[k] = sum_j p_j e^{-2pi i jk/L}.

    cuda/gpu_aln_noref.cu:1885 (cufftPlan1d R2C), :2123.  (..., L) -> (..., L/2+1).
    """
    return np.fft.rfft(p, axis=-1)


def ccf_spectra(ref_spec: np.ndarray, sbj_spec: np.ndarray):
    """Ring-weighted cross-spectra for both mirrors.

    cuda/gpu_aln_noref.cu:1508-1536:
      C0[k] = sum_i (i+1) R_i[k] conj(S_i[k])
      C1[k] = sum_i (i+1) conj(R_i[k] S_i[k])
    Broadcasts over leading axes; ring axis is -2, frequency axis -1.
    """
    R = ref_spec.shape[-2]
    w = np.arange(1, R + 1, dtype=np.float64)[:, None]
    c0 = (w * ref_spec * np.conj(sbj_spec)).sum(-2)
    c1 = (w * np.conj(ref_spec * sbj_spec)).sum(-2)
    return c0, c1


def psi_ccf(c: np.ndarray, ring_len: int = RING_LEN) -> np.ndarray:
    """Unnormalised C2R over angular frequency (cufft C2R, :2393, :2459).

    c[t] = Re C[0] + Re C[L/2] (-1)^t + 2 sum_{k=1}^{L/2-1} Re(C[k] e^{+2pi i kt/L}).
    """
    return np.fft.irfft(c, n=ring_len, axis=-1) * ring_len


# ------------------------------------------------------------------ decoding

def parabola_delta(x7: np.ndarray) -> np.ndarray:
    """7-point parabola offset (sparx prb1d form used by the reference).

    cuda/gpu_aln_noref.cu:1781-1787 and :2571-2581:
      c2 = 49x0+6x1-21x2-32x3-27x4-6x5+31x6 ; c3 = 5x0-3x2-4x3-3x4+5x6
      delta = c2/(2 c3) - 4  if c3 != 0 else 0      (not clamped, quirk Q5)
    ``x7`` (..., 7) float64 -> (...,).
    """
    x = x7
    c2 = (49.0 * x[..., 0] + 6.0 * x[..., 1] - 21.0 * x[..., 2] - 32.0 * x[..., 3]
          - 27.0 * x[..., 4] - 6.0 * x[..., 5] + 31.0 * x[..., 6])
    c3 = (5.0 * x[..., 0] - 3.0 * x[..., 2] - 4.0 * x[..., 3] - 3.0 * x[..., 4]
          + 5.0 * x[..., 6])
    with np.errstate(divide="ignore", invalid="ignore"):
        d = np.where(c3 != 0.0, c2 / (2.0 * np.where(c3 != 0.0, c3, 1.0)) - 4.0, 0.0)
    return d


def neighbours7(c_slot: np.ndarray, t: np.ndarray) -> np.ndarray:
    """x_q = c[(t+q) mod L], q=-3..3 (unsigned wrap, :1777-1779, :2566)."""
    L = c_slot.shape[-1]
    idx = (t[..., None] + np.arange(-3, 4)) % L
    return np.take_along_axis(c_slot, idx, axis=-1)


def angle_multiref(t, delta, mirror, ring_len=RING_LEN):
    """Multi-ref angle decode (cuda/gpu_aln_noref.cu:1784-1791)."""
    step = 360.0 / ring_len
    ang = step * t + step * delta
    ang = 360.0 - ang
    ang = np.where(mirror == 1, ang + 180.0, ang)
    ang = np.where(ang >= 360.0, ang - 360.0, ang)
    return ang


def angle_single_ref(t, delta, mirror, ring_len=RING_LEN):
    """Pre-align / ref-free angle decode (cuda/gpu_aln_noref.cu:2513-2518).

    The >= 360 wrap is applied only inside the mirror branch.
    """
    step = 360.0 / ring_len
    ang = 360.0 - (step * t + step * delta)
    ang = np.where(mirror == 1, ang + 180.0, ang)
    ang = np.where((mirror == 1) & (ang >= 360.0), ang - 360.0, ang)
    return ang


# ----------------------------------------------------------- per-particle CCF

def particle_spectra(sbj: np.ndarray, acc_xy: np.ndarray, shifts: np.ndarray,
                     ring_num: int, ring_len: int = RING_LEN) -> np.ndarray:
    """Normalised ring spectra of every particle at every grid shift.

    Centre = D//2 + int(shift) + accumulated shift (:1264-1265; grid shift is
    truncated to int at the kernel argument, quirk Q1).
    sbj (P, D, D), acc_xy (P, 2), shifts (S, 2) -> (P, S, R, L/2+1) complex128.
    """
    P, D = sbj.shape[0], sbj.shape[-1]
    c = D // 2
    S = shifts.shape[0]
    si = np.trunc(shifts.astype(np.float64)).astype(np.int64)
    out = np.empty((P, S, ring_num, ring_len // 2 + 1), dtype=np.complex128)
    for s in range(S):
        cx = c + si[s, 0] + acc_xy[:, 0].astype(np.float64)
        cy = c + si[s, 1] + acc_xy[:, 1].astype(np.float64)
        pol = polar_resample(sbj, cx, cy, ring_num, ring_len)
        out[:, s] = ring_spectra(normalize_polar(pol))
    return out


def reference_spectra(ref: np.ndarray, ring_num: int, ring_len: int = RING_LEN):
    """Spectra of the references: centre D//2, no shift (:493, :2074)."""
    N, D = ref.shape[0], ref.shape[-1]
    c = D // 2
    pol = polar_resample(ref, np.full(N, c, np.float64), np.full(N, c, np.float64),
                         ring_num, ring_len)
    return ring_spectra(normalize_polar(pol))


def ccf_table(sspec: np.ndarray, rspec: np.ndarray, ring_len: int = RING_LEN):
    """Psi-domain CCF for all pairs: (P, S, Rr, K) x (N, Rr, K) -> (P, N, 2, S, L).

    Mirror-major order matches the reference table (:2427-2433) restricted to
    one (particle, reference) pair: offset m*S*(L+2) + s*(L+2) + t.
    """
    R = rspec.shape[-2]
    w = np.arange(1, R + 1, dtype=np.float64)[:, None]
    rw = rspec * w  # (N, R, K)
    # C0 = sum_i w R conj(S) ; C1 = sum_i w conj(R S)
    c0 = np.einsum("nik,psik->pnsk", rw, np.conj(sspec), optimize=True)
    c1 = np.conj(np.einsum("nik,psik->pnsk", rw, sspec, optimize=True))
    c = np.stack([c0, c1], axis=2)  # (P, N, 2, S, K)
    return psi_ccf(c, ring_len)


# ---------------------------------------------------------- multi-reference

def multi_ref_align(sbj, ref, acc_xy, ring_num, shift_rng_x, shift_rng_y,
                    shift_step, ring_len=RING_LEN, chunk=16, tie_rel=1e-5):
    """Multi-reference alignment of every particle against every reference.

    Restates multi_ref_align (cuda/gpu_aln_noref.cu:326-554) with the wizard
    decode (:1667-1794):
      * scan (m, s, t<L) in table order with strict '>' from -FLT_MAX, so the
        first maximum wins; d = the raw maximum (quirk Q15);
      * 7-point parabola around t* in the winning slot, angle decode S10;
      * peak = {angle, acc_x + shift_x[s*], acc_y + shift_y[s*], m} (:536-539).

    Returns a dict of arrays indexed [n, p]: d, angle, sx, sy, mirror, shift_idx,
    t, delta, runner (best value of a distinct alignment: another (m, s) slot,
    or |t - t*| > 3 circularly in the same slot), tie (runner within tie_rel).
    """
    sbj = np.asarray(sbj)
    ref = np.asarray(ref)
    acc_xy = np.asarray(acc_xy, dtype=np.float64).reshape(-1, 2)
    P, N = sbj.shape[0], ref.shape[0]
    shifts = shift_grid(shift_rng_x, shift_rng_y, shift_step)
    S = shifts.shape[0]
    L = ring_len
    rspec = reference_spectra(ref, ring_num, ring_len)
    keys = ("d", "angle", "sx", "sy", "mirror", "shift_idx", "t", "delta",
            "runner", "tie")
    res = {k: np.zeros((N, P), dtype=np.float64) for k in keys}
    for p0 in range(0, P, chunk):
        p1 = min(P, p0 + chunk)
        sspec = particle_spectra(sbj[p0:p1], acc_xy[p0:p1], shifts, ring_num, ring_len)
        tab = ccf_table(sspec, rspec, ring_len)            # (b, N, 2, S, L)
        b = p1 - p0
        flat = tab.reshape(b, N, 2 * S * L)
        idx = np.argmax(flat, axis=-1)                     # first max wins
        best = np.take_along_axis(flat, idx[..., None], -1)[..., 0]
        m = idx // (S * L)
        s = (idx // L) % S
        t = idx % L
        slot = tab.reshape(b, N, 2 * S, L)
        win = np.take_along_axis(slot, (m * S + s)[..., None, None], axis=2)[:, :, 0]
        x7 = neighbours7(win, t)
        delta = parabola_delta(x7)
        ang = angle_multiref(t, delta, m, L)
        # runner-up over distinct alignments
        slot_max = slot.max(-1)                            # (b, N, 2S)
        np.put_along_axis(slot_max, (m * S + s)[..., None], -np.inf, axis=-1)
        other = slot_max.max(-1)
        tt = np.arange(L)
        dist = np.abs(((tt[None, None, :] - t[..., None]) + L // 2) % L - L // 2)
        same = np.where(dist > 3, win, -np.inf).max(-1)
        runner = np.maximum(other, same)
        sl = np.s_[:, p0:p1]
        res["d"][sl] = best.T
        res["angle"][sl] = ang.T
        res["sx"][sl] = (acc_xy[p0:p1, 0][:, None] + shifts[s, 0]).T
        res["sy"][sl] = (acc_xy[p0:p1, 1][:, None] + shifts[s, 1]).T
        res["mirror"][sl] = m.T
        res["shift_idx"][sl] = s.T
        res["t"][sl] = t.T
        res["delta"][sl] = delta.T
        res["runner"][sl] = runner.T
        res["tie"][sl] = ((best - runner) <= tie_rel * np.abs(best)).T
    return res


def peak_list_layout(res, P_local: int):
    """Pack results into the driver's peak_list [N][4 P] (S11, :536-539)."""
    N = res["d"].shape[0]
    pk = np.zeros((N, 4 * P_local), dtype=np.float64)
    pk[:, 0::4] = res["angle"]
    pk[:, 1::4] = res["sx"]
    pk[:, 2::4] = res["sy"]
    pk[:, 3::4] = res["mirror"]
    return pk


def driver_post_rotation(angle, sx, sy):
    """Shift conversion after the call (bin/sp_isac2_gpu.py:1016-1022):
    (sx, sy) <- (-sx cos a - sy sin a, sx sin a - sy cos a), a in radians."""
    a = np.radians(angle)
    return -sx * np.cos(a) - sy * np.sin(a), sx * np.sin(a) - sy * np.cos(a)


# --------------------------------------------- pre-align / reference-free

def single_ref_table(sbj, ref, ref_id, acc_xy, ring_num, shifts, ring_len=RING_LEN):
    """CCF of each particle with its own reference (cu_ccf_mult, :1356-1481).

    Returns (P, 2, S, L) psi-domain CCF values.
    """
    rspec = reference_spectra(ref, ring_num, ring_len)
    sspec = particle_spectra(sbj, acc_xy, shifts, ring_num, ring_len)
    R = ring_num
    w = np.arange(1, R + 1, dtype=np.float64)[:, None]
    rs = rspec[np.asarray(ref_id)] * w                     # (P, R, K)
    c0 = np.einsum("pik,psik->psk", rs, np.conj(sspec))
    c1 = np.conj(np.einsum("pik,psik->psk", rs, sspec))
    return psi_ccf(np.stack([c0, c1], axis=1), ring_len)


def decode_single_ref(tab, acc_xy, shifts, img_dim, ring_num, ring_len=RING_LEN):
    """Row argmax + parameter update for pre-align / reference-free.

    cuda/gpu_aln_noref.cu:1638-1664 (argmax from value 0 / index 0, strict '>',
    first max wins) and :2486-2519 (mirror, accumulate grid shift, clamp to
    +-(D - R - 2), angle with the FP64 parabola; quirk Q4, Q16).
    tab (P, 2, S, L).  Returns dict with updated shift_x/shift_y, angle, mirror.
    """
    P, _, S, L = tab.shape
    flat = tab.reshape(P, -1)
    idx = np.argmax(flat, axis=-1)
    mx = flat[np.arange(P), idx]
    idx = np.where(mx > 0.0, idx, 0)          # init value 0.0 / index 0
    m = idx // (S * L)
    s = (idx // L) % S
    t = idx % L
    win = tab.reshape(P, 2 * S, L)[np.arange(P), m * S + s]
    delta = parabola_delta(neighbours7(win, t))
    lim = float(img_dim - ring_num - 2)
    acc = np.asarray(acc_xy, dtype=np.float64).reshape(-1, 2)
    sx = np.clip(acc[:, 0] + shifts[s, 0], -lim, lim)
    sy = np.clip(acc[:, 1] + shifts[s, 1], -lim, lim)
    ang = angle_single_ref(t, delta, m, L)
    return {"shift_x": sx, "shift_y": sy, "angle": ang, "mirror": m,
            "shift_idx": s, "t": t, "delta": delta, "max": np.where(mx > 0, mx, 0.0)}


def transform_images(imgs, angle, shift_x, shift_y, mirror):
    """Apply found parameters to images (cu_transform_batch, :1558-1588).

    Output pixel (x, y): xs = D - x if mirrored else x; rotate by +angle about
    (D//2, D//2); add the accumulated shift; bilinear (clamped) fetch.
    """
    imgs = np.asarray(imgs, dtype=np.float64)
    P, D = imgs.shape[0], imgs.shape[-1]
    c = float(D // 2)
    yy, xx = np.meshgrid(np.arange(D, dtype=np.float64), np.arange(D, dtype=np.float64),
                         indexing="ij")
    mir = np.asarray(mirror).reshape(-1, 1, 1)
    xs = np.where(mir == 1, D - xx[None], xx[None]) - c
    ys = yy[None] - c
    a = np.radians(np.asarray(angle, dtype=np.float64)).reshape(-1, 1, 1)
    ca, sa = np.cos(a), np.sin(a)
    xr = xs * ca - ys * sa + c + np.asarray(shift_x, np.float64).reshape(-1, 1, 1)
    yr = xs * sa + ys * ca + c + np.asarray(shift_y, np.float64).reshape(-1, 1, 1)
    return bilinear_clamped(imgs, xr, yr)


def class_averages(imgs, cid_idx):
    """Per-class mean over imgs[cid_idx[k]:cid_idx[k+1]] (:1605-1621)."""
    return np.stack([imgs[cid_idx[k]:cid_idx[k + 1]].mean(0)
                     for k in range(len(cid_idx) - 1)])


def tangent_filter(refs, cutoff, falloff):
    """Hyperbolic-tangent low-pass in 2-D Fourier space (:1020-1035, :2232-2262).

    H = 0.5 (tanh(c (d + f)) - tanh(c (d - f))), c = pi / (2 falloff f),
    x = kx / (D/2 + 1) * 0.5 over the R2C half plane kx = 0..D/2,
    y = min(ky, D - ky) / (D/2) * 0.5, d = sqrt(x^2 + y^2).
    Both cuFFT transforms are unnormalised and the reference divides by D*D
    after the forward and again after the inverse transform for the FIRST
    image only (quirk Q7): image 0 comes back at gain 1/(D*D), every other
    image at gain D*D.  The polar normalisation downstream removes the gain.
    """
    refs = np.asarray(refs, dtype=np.float64)
    N, D = refs.shape[0], refs.shape[-1]
    F = np.fft.rfft2(refs)                                   # (N, D, D//2+1)
    kx = np.arange(D // 2 + 1, dtype=np.float64)
    ky = np.arange(D, dtype=np.float64)
    x = kx / float(D // 2 + 1) * 0.5
    half = D / 2.0
    y = np.where(ky < half, ky, D - ky) / half * 0.5
    d = np.sqrt(x[None, :] ** 2 + y[:, None] ** 2)
    cc = np.pi / (2.0 * falloff * cutoff)
    H = 0.5 * (np.tanh(cc * (d + cutoff)) - np.tanh(cc * (d - cutoff)))
    out = np.fft.irfft2(F * H[None], s=(D, D)) * (D * D)     # unnormalised C2R
    gain = np.full(N, 1.0)
    gain[0] = 1.0 / float(D * D) ** 2
    return out * gain[:, None, None]


def ref_free_iteration(sbj, refs, cid, acc_xy, ring_num, shifts, img_dim,
                       ring_len=RING_LEN):
    """One ref_free_alignment_2D() call (cuda/gpu_aln_noref.cu:964-996).

    Returns (params dict, new references).
    """
    tab = single_ref_table(sbj, refs, cid, acc_xy, ring_num, shifts, ring_len)
    prm = decode_single_ref(tab, acc_xy, shifts, img_dim, ring_num, ring_len)
    tr = transform_images(sbj, prm["angle"], prm["shift_x"], prm["shift_y"], prm["mirror"])
    cid = np.asarray(cid)
    cid_idx = list(np.flatnonzero(np.r_[True, cid[1:] != cid[:-1]])) + [len(cid)]
    return prm, class_averages(tr, cid_idx)


# ------------------------------------------------------------ reference update
# Device reference update (isac_update_references_device): the host step of
# bin/sp_isac2_gpu.py:1146-1232, with the reference-free path's image transform
# (transform_images) in place of EMAN2's rot_shift2D, an absolute-frequency tanh
# low-pass for filt_tanl and a Fourier phase ramp for fshift.  EMAN2 is absent,
# so agreement with its own routines is unpinned; the pins are closed forms
# (integer shifts are circular shifts, a cosine keeps its shape, the filter gain
# on a cosine) and the integration check that noise-free particles averaged with
# their exhaustive peaks give back their reference (tests/test_oracle.py).

def filter_shift(imgs, cutoff, falloff, dx, dy):
    """Per image n: multiply the 2-D spectrum by H(r) exp(-2 pi i (kx dx_n + ky
    dy_n) / D) and invert (numpy irfft2: the real part of the kx = 0 and kx = D/2
    columns), so the content moves by (dx_n, dy_n).
    H(r) = 0.5 (tanh(c (r + f)) - tanh(c (r - f))), c = pi / (2 falloff f)
    (H = 1 when cutoff <= 0),
    r = sqrt(kx^2 + min(ky, D - ky)^2) / D cycles per pixel; ky is signed
    (ky = D/2 taken as +D/2) in the phase ramp."""
    imgs = np.asarray(imgs, dtype=np.float64)
    N, D = imgs.shape[0], imgs.shape[-1]
    F = np.fft.rfft2(imgs)                                   # (N, D, D//2+1)
    kx = np.arange(D // 2 + 1, dtype=np.float64)
    ky = np.arange(D, dtype=np.float64)
    r = np.sqrt((kx[None, :] / D) ** 2 + (np.minimum(ky, D - ky)[:, None] / D) ** 2)
    if cutoff > 0:
        cc = np.pi / (2.0 * falloff * cutoff)
        H = 0.5 * (np.tanh(cc * (r + cutoff)) - np.tanh(cc * (r - cutoff)))
    else:                                                    # no low-pass: the ramp alone
        H = np.ones_like(r)
    kys = np.where(2 * ky > D, ky - D, ky)
    dx = np.asarray(dx, np.float64).reshape(-1, 1, 1)
    dy = np.asarray(dy, np.float64).reshape(-1, 1, 1)
    ramp = np.exp(-2j * np.pi * (kx[None, None, :] * dx + kys[None, :, None] * dy) / D)
    return np.fft.irfft2(F * H[None] * ramp, s=(D, D))


def pre_align_loop(sbj, ring_num, ranges, steps, max_iter, center=0, cutoff=0.0, falloff=0.0,
                   acc_xy=None, return_table=False):
    """The pre-alignment driver loop, bin/sp_isac_applications.py:1486-1693, with
    the device path's stand-ins for EMAN2 (isac_pre_align_loop_device): per shift
    step (reset_shifts, :1489) and iteration, the average of the particles
    transformed with their current parameters (transform_images; sum_oe,
    :1495), optionally moved by minus the mean shift after the driver's rotation
    (center == -1, :1547-1552; the sums of the previous alignment, :1664-1690,
    zero before the first) and low-passed (cutoff > 0), then each particle
    aligned to it (single_ref_table + decode_single_ref, the pre_align_run
    step); the last iteration builds the average only (:1563).
    Returns (params dict of the last alignment, final average), plus the CCF
    table of the last alignment with return_table."""
    sbj = np.asarray(sbj, dtype=np.float64)
    P, D = sbj.shape[0], sbj.shape[-1]
    acc = np.zeros((P, 2)) if acc_xy is None else np.asarray(acc_xy, np.float64).reshape(P, 2).copy()
    angle = np.zeros(P)
    mirror = np.zeros(P, dtype=int)
    total, it = len(ranges) * max_iter, 0
    prm = tab = None
    for rng_, step in zip(ranges, steps):
        shifts = shift_grid(rng_, rng_, step)
        for _ in range(max_iter):
            it += 1
            avg = transform_images(sbj, angle, acc[:, 0], acc[:, 1], mirror).mean(0)
            shift_avg = center == -1 and it > 1
            if shift_avg or cutoff > 0:
                dx = dy = 0.0
                if shift_avg:
                    px, py = driver_post_rotation(angle, acc[:, 0], acc[:, 1])
                    dx = -np.where(mirror != 0, -px, px).sum() / P
                    dy = -py.sum() / P
                avg = filter_shift(avg[None], cutoff, falloff, [dx], [dy])[0]
            if it != total:
                tab = single_ref_table(sbj, avg[None], np.zeros(P, int), acc, ring_num, shifts)
                prm = decode_single_ref(tab, acc, shifts, D, ring_num)
                acc = np.stack([prm["shift_x"], prm["shift_y"]], 1)
                angle, mirror = prm["angle"], prm["mirror"]
    return (prm, avg, tab) if return_table else (prm, avg)


def update_references(imgs, peak_list, cls, n_ref, cutoff, falloff):
    """(refs [N, D, D], params [P, 4], members [N], shift_avg [N, 2]).

    Particle i of class k = cls[i] has the peak (alpha, sx, sy, m) =
    peak_list[k, 4i:4i+4]; (sx', sy') = driver_post_rotation(alpha, sx, sy);
    class sums of (m ? -sx' : sx', sy') in ascending particle order, rounded to
    float32 (the driver's MPI_FLOAT reduction) and divided by the member count;
    params = (alpha, m ? sx' + avg_x : sx' - avg_x, sy' - avg_y, m)
    (bin/sp_isac2_gpu.py:1200-1208); refs = filter_shift(class mean of
    transform_images(imgs, alpha, sx, sy, m), cutoff, falloff, -avg).  A class
    without members gives zeros."""
    imgs = np.asarray(imgs, dtype=np.float64)
    P = imgs.shape[0]
    cls = np.asarray(cls, dtype=np.int64)
    pk = np.asarray(peak_list, dtype=np.float64).reshape(n_ref, P, 4)
    q = pk[cls, np.arange(P)]
    alpha, sx, sy, m = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    mir = m != 0
    sxp, syp = driver_post_rotation(alpha, sx, sy)
    members = np.bincount(cls, minlength=n_ref)
    sum_x = np.zeros(n_ref)
    sum_y = np.zeros(n_ref)
    for i in range(P):
        sum_x[cls[i]] += -sxp[i] if mir[i] else sxp[i]
        sum_y[cls[i]] += syp[i]
    with np.errstate(invalid="ignore", divide="ignore"):
        avg_x = np.where(members > 0, sum_x.astype(np.float32).astype(np.float64) / members, 0.0)
        avg_y = np.where(members > 0, sum_y.astype(np.float32).astype(np.float64) / members, 0.0)
    params = np.stack([alpha, np.where(mir, sxp + avg_x[cls], sxp - avg_x[cls]),
                       syp - avg_y[cls], m], axis=1)
    tr = transform_images(imgs, alpha, sx, sy, mir.astype(int))
    D = imgs.shape[-1]
    avg = np.zeros((n_ref, D, D))
    for k in range(n_ref):
        if members[k]:
            avg[k] = tr[cls == k].mean(0)
    refs = filter_shift(avg, cutoff, falloff, -avg_x, -avg_y)
    return refs, params, members, np.stack([avg_x, avg_y], axis=1)


# ------------------------------------------------------------ narrowed search
# No counterpart in the reference (its search is exhaustive).  Restates the
# narrowed search of isac_mref_proposal_device: a proposal (ref, mirror, psi
# bin b, shift index s0) admits the shifts within hw grid steps of s0 and the psi
# samples t with |t * 360/L - (b + 0.5) * 360/n_bins| <= psi_half (t taken
# modulo L) of that mirror; per particle the maximum over every admitted
# (ref, m, s, t), ties to the lower ref, then the table order (m, s, t).
# Pin: when a window holds the exhaustive optimum it returns multi_ref_align's
# result (tests/test_oracle.py).

def proposal_mask(prop_p, n_ref, nx, ny, n_bins, psi_half, hw, ring_len=RING_LEN):
    """Admitted samples of one particle's proposals: bool [N, 2, S, L]."""
    S, L = nx * ny, ring_len
    mask = np.zeros((n_ref, 2, S, L), bool)
    step = 360.0 / L
    bw = 360.0 / n_bins
    for ref, m, b, s0 in np.asarray(prop_p).reshape(-1, 4):
        if not (0 <= ref < n_ref and 0 <= m <= 1 and 0 <= b < n_bins and 0 <= s0 < S):
            continue
        c = (b + 0.5) * bw
        half = float(np.float32(psi_half))
        t0 = int(np.ceil((c - half) / step))
        t1 = int(np.floor((c + half) / step))
        ts = np.arange(t0, min(t1, t0 + L - 1) + 1) % L
        ix0, iy0 = divmod(int(s0), ny)
        for ix in range(max(0, ix0 - hw), min(nx - 1, ix0 + hw) + 1):
            for iy in range(max(0, iy0 - hw), min(ny - 1, iy0 + hw) + 1):
                mask[ref, m, ix * ny + iy, ts] = True
    return mask


def windowed_best(sbj, ref, acc_xy, ring_num, shift_rng_x, shift_rng_y, shift_step, prop,
                  n_bins=72, psi_half=10.0, hw=1, ring_len=RING_LEN):
    """Best (ref, value, angle, sx, sy, mirror, s, t) per particle over the
    proposals' windows; ref = -1 when nothing is admitted.  ``interior``: every
    neighbour of the maximum (psi t +- 1, shift +- 1 in x and y inside the grid)
    is admitted for its reference and mirror."""
    sbj = np.asarray(sbj)
    acc_xy = np.asarray(acc_xy, dtype=np.float64).reshape(-1, 2)
    P, N = sbj.shape[0], ref.shape[0]
    shifts = shift_grid(shift_rng_x, shift_rng_y, shift_step)
    S, L = shifts.shape[0], ring_len
    ny = int(np.sum(shifts[:, 0] == shifts[0, 0]))
    nx = S // ny
    rspec = reference_spectra(np.asarray(ref), ring_num, ring_len)
    out = {k: np.zeros(P) for k in ("d", "angle", "sx", "sy", "mirror", "shift_idx", "t")}
    out["ref"] = np.full(P, -1, np.int64)
    out["interior"] = np.zeros(P, bool)
    for p in range(P):
        mask = proposal_mask(prop[p], N, nx, ny, n_bins, psi_half, hw, ring_len)
        refs = np.nonzero(mask.reshape(N, -1).any(1))[0]
        if refs.size == 0:
            out["d"][p] = -np.inf
            continue
        sspec = particle_spectra(sbj[p:p + 1], acc_xy[p:p + 1], shifts, ring_num, ring_len)
        tab = ccf_table(sspec, rspec[refs], ring_len)[0]       # (n, 2, S, L)
        vals = np.where(mask[refs], tab, -np.inf).reshape(len(refs), -1)
        idx = np.argmax(vals, axis=1)                           # first max in (m, s, t) order
        best = vals[np.arange(len(refs)), idx]
        j = int(np.argmax(best))                                # ties: lower reference
        m, s, t = idx[j] // (S * L), (idx[j] // L) % S, idx[j] % L
        val = best[j]

        def interior(mk, m, s, t):
            ix, iy = divmod(int(s), ny)
            nb = [(s, (t + 1) % L), (s, (t - 1) % L)]
            nb += [(s + dx * ny + dy, t) for dx, dy in ((-1, 0), (1, 0), (0, -1), (0, 1))
                   if 0 <= ix + dx < nx and 0 <= iy + dy < ny]
            return all(mk[m, s2, t2] for s2, t2 in nb)

        inside = interior(mask[refs[j]], m, s, t)
        x7 = neighbours7(tab[j, m, s][None], np.array([t]))
        delta = parabola_delta(x7)
        out["ref"][p] = refs[j]
        out["d"][p] = val
        out["angle"][p] = angle_multiref(np.array([t]), delta, np.array([m]), L)[0]
        out["sx"][p] = acc_xy[p, 0] + shifts[s, 0]
        out["sy"][p] = acc_xy[p, 1] + shifts[s, 1]
        out["mirror"][p], out["shift_idx"][p], out["t"][p] = m, s, t
        out["interior"][p] = inside
    return out


# ------------------------------------------------------- group assignment
# Restates the contract of the step after multi_ref_align (bin/sp_isac2_gpu.py:
# 1119-1144, where EMAN2's Util.assign_groups -- absent here -- does it): pairs in
# descending score order fill classes of cap = n_img // n_ref.  The tie order and
# the class given to the leftover particles are this library's choice (EMAN2's
# are unknown: parity with it is unpinned); the pins are the stability property
# and hand-worked cases (tests/test_oracle.py).

def assign_groups(d, n_img=None):
    """Greedy equal-size assignment; returns the n_img + 1 int layout: class k's
    particles ascending at [k*cap, (k+1)*cap), the leftover particles ascending,
    the class of the highest leftover pair (0 if none).  NaN = -inf."""
    d = np.asarray(d, np.float64)
    if n_img is not None:
        d = d[:, :n_img]
    N, P = d.shape
    v = np.where(np.isnan(d), -np.inf, d)
    cap = P // N
    n_idx, p_idx = np.meshgrid(np.arange(N), np.arange(P), indexing="ij")
    order = np.lexsort((p_idx.ravel(), n_idx.ravel(), -v.ravel()))
    cls = np.full(P, -1)
    cnt = np.zeros(N, np.int64)
    placed = 0
    for k in order:
        n, p = divmod(int(k), P)
        if cls[p] < 0 and cnt[n] < cap:
            cls[p] = n
            cnt[n] += 1
            placed += 1
            if placed == N * cap:
                break
    out = []
    for n in range(N):
        out += sorted(np.nonzero(cls == n)[0].tolist())
    rest = sorted(np.nonzero(cls < 0)[0].tolist())
    out += rest
    if rest:
        cand = [(-v[n, p], n, p) for p in rest for n in range(N)]
        out.append(min(cand)[1])
    else:
        out.append(0)
    return np.array(out, np.int64)
