"""FP64 CPU oracle of the multi-alignment stability test.

TEST INFRASTRUCTURE ONLY.  Only ``tests/`` import this module, and only as the
checker of the library's stability kernel (``isac_multi_align_stability_device``).

What it restates.  After the GPU class refinement, the driver calls, per class
k (bin/sp_isac2_gpu.py:1481-1491):

    stable_set, mirror_consistent_rate, err = sp_pixel_error.multi_align_stability(
        aln_param[k], 0.0, 10000.0, thld_err, False, last_ring * 2)

with ``aln_param[k]`` = one flat list [angle, sx, sy, mirror] * n per repeated
reference-free run (L = stab_ali runs, :1381-1393), and reads
``stable_set[i] = [pixel_error, member index, [angle, sx, sy, mirror]]``
(:1493-1521).  ``sp_pixel_error`` is SPHIRE code that the reference does not
ship (version unpinned), so this restates the published method it implements
(the multi-alignment pixel error of ISAC, Yang et al., Structure 20 (2012)
237-247, from SPARX): a particle's L alignments are compared at the points of a
circle of diameter d after every run has been brought into the last run's frame
by one global rigid transform; the transforms minimise the mean pixel error.

Definitions used here (unpinned against SPHIRE's own conventions):
  * a 2-D alignment (theta, s, m) maps x -> R(theta) F^m x + s, F = diag(-1, 1);
  * a global transform G = (alpha, t, f) composes as G o T:
        theta' = alpha + (-1)^f theta,  m' = m xor f,  s' = R(alpha) F^f s + t;
  * squared pixel error of particle i over its L aligned transforms:
        e_i = r^2 (1 - |mean_j exp(i theta'_ij)|) + var_j(s'x_ij) + var_j(s'y_ij),
    r = d / 2, var = population variance (SPARX's multi-alignment pixel error);
  * mirror stability: per-run flips f_j (f_{L-1} = 0) that maximise the number of
    particles whose flipped mirror flags agree in every run (first maximum in
    flip-pattern order); those particles are the mirror-consistent set;
  * global fit: from the mean differences to the last run, drop particles whose
    pixel error is above 3 thld_err, minimise the mean pixel error of the rest
    over (alpha_j, t_j), j < L - 1;
  * stable set: mirror-consistent particles with pixel error <= thld_err after
    the fit, sorted by pixel error (then index), each with its average
    alignment in the last run's frame (angle = arg of the mean unit vector, shift
    = mean shift, mirror = the last run's flag); err = sqrt of the mean squared
    pixel error of the mirror-consistent particles after the fit.
"""
from __future__ import annotations

import itertools

import numpy as np


def compose(G, prm):
    """G = (alpha_deg, tx, ty, flip); prm [..., 4] = (theta_deg, sx, sy, m)."""
    a = np.radians(G[0])
    f = int(G[3])
    th = prm[..., 0]
    sx, sy = prm[..., 1], prm[..., 2]
    if f:
        sx = -sx
        th = -th
    ca, sa = np.cos(a), np.sin(a)
    out = np.empty_like(prm, dtype=np.float64)
    out[..., 0] = G[0] + th
    out[..., 1] = ca * sx - sa * sy + G[1]
    out[..., 2] = sa * sx + ca * sy + G[2]
    out[..., 3] = np.logical_xor(prm[..., 3] != 0, f != 0)
    return out


def pixel_errors(aligned, r):
    """aligned [L, n, 4] (already in one frame) -> squared pixel error [n]."""
    th = np.radians(aligned[..., 0])
    rho = np.abs(np.exp(1j * th).mean(axis=0))
    return r * r * (1.0 - rho) + aligned[..., 1].var(axis=0) + aligned[..., 2].var(axis=0)


def apply_global(prm, Gs):
    """prm [L, n, 4]; Gs [L, 4] -> aligned [L, n, 4]."""
    return np.stack([compose(Gs[j], prm[j]) for j in range(prm.shape[0])])


def mirror_flips(mirrors):
    """mirrors [L, n] (0/1) -> (flips [L], consistent mask [n])."""
    L, n = mirrors.shape
    m = mirrors.astype(bool)
    best, best_f = -1, None
    for pat in itertools.product((0, 1), repeat=L - 1):
        f = np.array(list(pat) + [0], bool)
        ok = np.all((m ^ f[:, None]) == (m[-1] ^ f[-1])[None, :], axis=0)
        c = int(ok.sum())
        if c > best:
            best, best_f = c, f
    ok = np.all((m ^ best_f[:, None]) == m[-1][None, :], axis=0)
    return best_f.astype(int), ok


def initial_global(prm, flips):
    """Mean differences of each run to the last one (prm [L, n, 4], mirror-
    consistent particles): alpha_j = arg mean exp(i (theta_last - theta'_j)),
    t_j = mean(s_last - R(alpha_j) F^f_j s_j).  Returns Gs [L, 4]."""
    L = prm.shape[0]
    Gs = np.zeros((L, 4))
    for j in range(L - 1):
        f = flips[j]
        th = -prm[j, :, 0] if f else prm[j, :, 0]
        a = np.degrees(np.angle(np.exp(1j * np.radians(prm[-1, :, 0] - th)).mean()))
        moved = compose((a, 0.0, 0.0, f), prm[j])
        Gs[j] = (a, (prm[-1, :, 1] - moved[:, 1]).mean(), (prm[-1, :, 2] - moved[:, 2]).mean(), f)
    return Gs


def fit_global(prm, flips, r, G0):
    """Minimise the mean pixel error over (alpha_j, tx_j, ty_j), j < L - 1, with
    scipy's Nelder-Mead from G0 (then polished with BFGS).  Returns Gs [L, 4]."""
    from scipy.optimize import minimize
    L = prm.shape[0]

    def unpack(x):
        Gs = np.zeros((L, 4))
        Gs[:L - 1, :3] = x.reshape(L - 1, 3)
        Gs[:, 3] = flips
        return Gs

    def f(x):
        return pixel_errors(apply_global(prm, unpack(x)), r).mean()

    x0 = G0[:L - 1, :3].reshape(-1)
    res = minimize(f, x0, method="Nelder-Mead",
                   options={"xatol": 1e-10, "fatol": 1e-14, "maxiter": 40000, "maxfev": 80000})
    res = minimize(f, res.x, method="BFGS", options={"gtol": 1e-12})
    return unpack(res.x)


def multi_align_stability(ali_params, mir_stab_thld, grp_err_thld, err_thld, d):
    """ali_params: L flat lists [angle, sx, sy, mirror] * n (the driver's layout).
    Returns (stable_set, mirror_consistent_rate, err, extra) where stable_set is
    a list of [pixel_error, member index, [angle, sx, sy, mirror]] and extra holds
    the per-particle pixel errors, consistency mask and the fitted transforms."""
    prm = np.asarray(ali_params, np.float64).reshape(len(ali_params), -1, 4)
    L, n, _ = prm.shape
    r = d / 2.0
    flips, ok = mirror_flips(prm[..., 3] != 0)
    rate = ok.sum() / float(n)
    extra = {"flips": flips, "consistent": ok, "pix": np.full(n, -1.0), "avg": np.zeros((n, 4))}
    if rate <= mir_stab_thld or ok.sum() < 2:
        return [], rate, -1.0, extra
    idx = np.nonzero(ok)[0]
    sub = prm[:, idx]
    G0 = initial_global(sub, flips)
    e0 = pixel_errors(apply_global(sub, G0), r)
    keep = np.sqrt(np.maximum(e0, 0.0)) <= 3.0 * err_thld
    if keep.sum() < 2:
        return [], rate, -1.0, extra
    Gs = fit_global(sub[:, keep], flips, r, G0)
    al = apply_global(sub, Gs)
    e = np.maximum(pixel_errors(al, r), 0.0)
    err = float(np.sqrt(e.mean()))
    extra.update(G=Gs, G0=G0, kept=keep)
    extra["pix"][idx] = np.sqrt(e)
    th = np.radians(al[..., 0])
    ang = np.degrees(np.angle(np.exp(1j * th).mean(axis=0))) % 360.0
    avg = np.stack([ang, al[..., 1].mean(0), al[..., 2].mean(0), al[-1, :, 3]], 1)
    extra["avg"][idx] = avg
    if err > grp_err_thld:
        return [], rate, err, extra
    stable = [[float(np.sqrt(e[k])), int(idx[k]), avg[k].tolist()]
              for k in range(idx.size) if np.sqrt(e[k]) <= err_thld]
    stable.sort(key=lambda s: (s[0], s[1]))
    return stable, rate, err, extra
