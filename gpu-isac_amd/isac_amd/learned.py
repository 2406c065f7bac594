"""Learned-proposal form of the multi-reference search.

Pipeline: the exhaustive FP32 search (isac_mref_align_device) labels a
training set; PoseNet is trained on those labels with its own kernels; a
threshold on the narrowed search's best CCF is calibrated on held-out
particles so that the particles below it, sent to the exhaustive search,
bring the agreement with the exhaustive search to a target; inference is
PoseNet proposals -> isac_mref_proposal_device.  No counterpart in the
reference, whose search is always exhaustive (cuda/gpu_aln_noref.cu:326-554).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import binding, posenet

TIE_REL = 1e-5

# Network and training settings of the learned form: the last conv layer keeps
# its 9x9 map (128 -> 32 channels, no pooling) so the shift head sees position;
# Gaussian targets over neighbouring psi bins (sigma 1 bin) and shifts (0.7 step).
NET = dict(channels=(32, 64, 128, 32), hidden=384, pool4=False)
TRAIN = dict(batch=256, lr=2e-3, wd=1e-4, warmup=1000, sigma_psi=1.0, sigma_shift=0.7)
# Training length and calibration of the learned form.  Four network seeds
# (24 epochs, 40 000 calibration particles, target 0.997; 100 000 held out): 3.07-3.71x
# the exhaustive FP32 search at 99.66-99.76 % agreement with the FP64 oracle
# (10 000-particle sample).  With 10 000 calibration particles the confidence cut
# rests on a few dozen particles and the fallback share varied from 13 to 32 %
# between seeds at the same target.
EPOCHS = 24
N_CAL = 40000
# Agreement with the exhaustive search that the fallback threshold is calibrated
# for, on the calibration set.
TARGET = 0.997


def exhaustive(sbj, ref, ring_num, shift_rng, shift_step=1.0, chunk=20000):
    """Exhaustive FP32 search in particle chunks: (d [N, P], peak_list [N, 4P])."""
    import torch
    P, N = sbj.shape[0], ref.shape[0]
    d = torch.empty((N, P), dtype=torch.float32, device=sbj.device)
    peak = torch.empty((N, P, 4), dtype=torch.float32, device=sbj.device)
    for p0 in range(0, P, chunk):
        p1 = min(P, p0 + chunk)
        pk, dd = binding.mref_align_device(sbj[p0:p1], ref, ring_num, shift_step, shift_rng,
                                           shift_rng)
        d[:, p0:p1] = dd
        peak[:, p0:p1] = pk.view(N, p1 - p0, 4)
    return d, peak.view(N, 4 * P)


def labels(sbj, ref, ring_num, shift_rng, shift_step=1.0, n_psi=72):
    d, peak = exhaustive(sbj, ref, ring_num, shift_rng, shift_step)
    return posenet.labels_from_search(d, peak, shift_rng, shift_step, n_psi), d, peak


def agreement(best_ref, best_peak, d, peak, psi_tol=0.5, tie_rel=TIE_REL):
    """Per particle: does (best_ref, best_peak) match the exhaustive search?
    Same reference (either of the top two when they tie within tie_rel), same
    mirror and shift, psi within psi_tol degrees of that reference's angle."""
    import torch
    N, P = d.shape
    cols = torch.arange(P, device=d.device)
    top = torch.topk(d, 2, dim=0) if N > 1 else None
    r1 = torch.argmax(d, dim=0)
    ok_ref = best_ref.long() == r1
    if top is not None:
        v1, v2 = top.values[0], top.values[1]
        tie = (v1 - v2) <= tie_rel * v1.abs()
        ok_ref = ok_ref | (tie & (best_ref.long() == top.indices[1]))
    rr = torch.where(ok_ref, best_ref.long().clamp(0, N - 1), r1)
    pk = peak.view(N, P, 4)[rr, cols]
    dpsi = torch.remainder(best_peak[:, 0] - pk[:, 0] + 180.0, 360.0) - 180.0
    same = ok_ref & (best_peak[:, 3] == pk[:, 3]) & ((best_peak[:, 1] - pk[:, 1]).abs() < 1e-4) & \
        ((best_peak[:, 2] - pk[:, 2]).abs() < 1e-4) & (dpsi.abs() <= psi_tol)
    return same


def calibrate_threshold(val, agree, target: float = 0.997):
    """Smallest threshold t such that, with the particles below t sent to the
    exhaustive search (which then agree), at most (1 - target) of the
    calibration particles disagree."""
    val = np.asarray(val, np.float64)
    bad = np.sort(val[~np.asarray(agree, bool)])[::-1]
    allowed = int(math.floor((1.0 - target) * val.size))
    if bad.size <= allowed:
        return -3.0e38
    return float(np.nextafter(np.float32(bad[allowed]), np.float32(np.inf)))


def calibrate_joint(val, agree, conf, target: float = 0.997, grid: int = 41):
    """Confidence cut and CCF threshold together: for each candidate cut (a
    quantile of conf, up to the 40 % quantile), the particles below it go to
    the exhaustive search and the threshold is calibrate_threshold's over the
    rest (with the allowance of the whole set); the pair with the fewest
    fallbacks wins.  Returns (conf_min, threshold, expected fallback fraction)."""
    val = np.asarray(val, np.float64)
    agree = np.asarray(agree, bool)
    conf = np.asarray(conf, np.float64)
    n = val.size
    allowed = int(math.floor((1.0 - target) * n))
    best = (0.0, -3.0e38, 2.0)
    for q in np.linspace(0.0, 0.4, grid):
        cut = 0.0 if q == 0.0 else float(np.quantile(conf, q))
        keep = conf >= cut
        bad = np.sort(val[keep & ~agree])[::-1]
        thr = -3.0e38 if bad.size <= allowed else \
            float(np.nextafter(np.float32(bad[allowed]), np.float32(np.inf)))
        fb = float(np.mean(~keep | (val < thr)))
        if fb < best[2]:
            best = (cut, thr, fb)
    return best


@dataclass
class Proposer:
    net: posenet.PoseNet
    threshold: float
    ring_num: int
    shift_rng: float
    shift_step: float = 1.0
    k: int = 8
    psi_half: float = 10.0
    shift_half: int = 1
    level2: bool = True
    conf_min: float = 0.0
    # exhaustive fallback on MFMA with exact FP32 re-scoring of the 8 best references
    fallback_form: str = "fp16"

    def run(self, sbj, ref, acc_xy=None):
        prop, conf = self.net.propose(sbj, self.k, with_conf=True)
        return binding.mref_proposal_device(sbj, ref, self.ring_num, self.shift_step,
                                            self.shift_rng, self.shift_rng, prop, self.threshold,
                                            n_psi_bins=self.net.n_psi, psi_half=self.psi_half,
                                            shift_half=self.shift_half, level2=self.level2,
                                            conf=conf, conf_min=self.conf_min, acc_xy=acc_xy,
                                            fallback_form=self.fallback_form)


def calibrate(net, sbj, ref, d, peak, ring_num, shift_rng, target=0.997, use_conf=True, **kw):
    """Proposer whose threshold (and, with use_conf, confidence cut) is
    calibrated on (sbj, exhaustive d / peak).  Returns (proposer, agreement of
    the uncalibrated search)."""
    pr = Proposer(net, -3.0e38, ring_num, shift_rng, **kw)
    prop, conf = net.propose(sbj, pr.k, with_conf=True)
    o = binding.mref_proposal_device(sbj, ref, ring_num, pr.shift_step, shift_rng, shift_rng, prop,
                                     -3.0e38, n_psi_bins=net.n_psi, psi_half=pr.psi_half,
                                     shift_half=pr.shift_half, level2=pr.level2)
    agree = agreement(o["best_ref"], o["best_peak"], d, peak).cpu().numpy()
    val = o["best_val"].cpu().numpy()
    if use_conf:
        pr.conf_min, pr.threshold, _ = calibrate_joint(val, agree, conf.cpu().numpy(), target)
    else:
        pr.threshold = calibrate_threshold(val, agree, target)
    return pr, float(agree.mean())
