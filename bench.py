"""Benchmark of the multi-reference polar alignment hot path on one MI355X.

Workload (default "p20k_n100_d76"): 20,000 float32 particles of 76x76 px rendered from
100 seeded blob references (radius <= 34 px) with random psi, mirror and
sub-pixel shifts in +-3 px at SNR 0.05; search over rings 1..35, shifts +-3
step 1 (49 shifts), 256 psi samples and both mirrors.  A step is one
multi-reference alignment of every particle against every reference with the
images already resident in HBM (isac_mref_align_device); the metric is
particle x reference alignments per second.

Also reported:
  roofline      the fused contraction + psi-IFFT kernel (k_ccf_mref): algorithmic
                flops per launch (pairs x (8 S R K_f + 2 S 2.5 L log2 L)) over its
                average launch time, timed with HIP events on its stream
  cpu_baseline  the same algorithm's C restatement (FP32 build, OpenMP) on the
                host cores, on a time-bounded sample of the same particles; its
                FP64 build checks the GPU parameters on that sample
  host_abi      the same search behind the reference's host-pointer ABI
                (multi_ref_align: host images in, PCIe both ways), median of 5
  mfma_forms    the FP16 / BF16 matrix-core forms and the FP16 form with the exact
                FP32 best per particle, each with its own roofline: the MFMA flops
                the ring kernel (k_mref_rt) issues per step over its HIP-event
                time against the 2.5 PFLOP/s dense peak, and the step time against
                the algorithmic bound max(contraction / 2.5 PFLOP/s, (psi IFFTs +
                ring FFTs) / 157.3 TFLOP/s, image bytes / 8 TB/s)
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(ROOT, "gpu-isac_amd"), ROOT):
    if _p not in sys.path:
        sys.path.insert(0, _p)

WORKLOADS = {
    # name: (P, N, D, radius, snr, max_shift, integer_shifts, R, shift_rng)
    "p20k_n100_d76": (20000, 100, 76, 34.0, 0.05, 3.0, False, 35, 3.0),
    "p1k_n8_d64": (1000, 8, 64, 28.0, 0.1, 2.0, True, 29, 2.0),
}
PEAK_FP32_TFLOPS = 157.3     # MI355X published FP32 peak
PEAK_F16_TFLOPS = 2500.0     # MI355X published dense FP16 / BF16 MFMA peak
PEAK_HBM_TBS = 8.0           # MI355X HBM3E
RING_LEN = 256


def flops_per_pair(S, R, L=RING_LEN):
    K = L // 2 + 1
    return 8.0 * S * R * K + 2.0 * S * 2.5 * L * math.log2(L)


def mfma_issued_flops(P, N, S, R):
    """MFMA flops the reference-resident search kernel (k_mref_rt, rings <= 48)
    issues for P particles x N references: per workgroup (16 particles x 8
    references) and 16-row block, 8 wavefronts x 16 slots x 2 tiles of 16 x 16
    contraction steps (16x16x32 over rings 0..31, 16x16x16 over a further 16-ring
    step or the 8-ring tail), and per valid (row, reference) pair 4 16x16x32 psi
    steps.  Padding rows / references / rings count as issued."""
    full16, rem = R // 16, R % 16
    tail = 0 if rem == 0 else (8 if rem <= 8 else 16)
    nks = full16 + (1 if tail == 16 else 0)
    f32, f16 = 16 * 16 * 32 * 2, 16 * 16 * 16 * 2
    per_tile = (f32 if nks >= 2 else 0) + (f16 if nks % 2 else 0) + (f16 if tail == 8 else 0)
    groups, nb8, s_pad = -(-P // 16), -(-N // 8), max(S, 16)
    contraction = groups * nb8 * s_pad * 8 * 16 * 2 * per_tile
    psi = P * S * nb8 * 8 * 4 * f32
    return float(contraction + psi)


def algorithmic_bound_ms(P, N, S, R, L=RING_LEN, D=76):
    """max(F_contr / P_mfma, (F_ifft + F_fwd) / P_fp32, B / BW) per step: the ring
    contraction at the MFMA peak, the psi inverse FFTs and the ring FFTs at the
    FP32 peak, the image bytes (read once) at the HBM peak."""
    K = L // 2 + 1
    pairs = P * N
    f_contr = 8.0 * S * R * K
    f_ifft = 2.0 * S * 2.5 * L * math.log2(L)
    f_fwd = (S * R * 2.5 * L * math.log2(L)) / N       # per pair, amortised over references
    b = 4.0 * D * D / N + 20.0
    t = (pairs * f_contr / (PEAK_F16_TFLOPS * 1e12), pairs * (f_ifft + f_fwd) / (PEAK_FP32_TFLOPS * 1e12),
         pairs * b / (PEAK_HBM_TBS * 1e12))
    return max(t) * 1e3, {"contraction_mfma_ms": t[0] * 1e3, "ffts_fp32_ms": t[1] * 1e3,
                          "hbm_ms": t[2] * 1e3}


def wilson(k, n, z=1.96):
    """95 % Wilson score interval of a proportion k / n."""
    if n <= 0:
        return [0.0, 1.0]
    p = k / n
    den = 1.0 + z * z / n
    c = (p + z * z / (2 * n)) / den
    h = z * math.sqrt(p * (1.0 - p) / n + z * z / (4.0 * n * n)) / den
    return [max(0.0, c - h), min(1.0, c + h)]


def cpu_baseline(sbj_host, ref_host, acc, R, rng, gpu_peak, gpu_d, budget_s, cores,
                 min_agree=1000):
    import numpy as np
    from oracle import coracle
    P = sbj_host.shape[0]
    N = ref_host.shape[0]
    chunk = 16
    done = 0
    t0 = time.perf_counter()
    while done < P and time.perf_counter() - t0 < budget_s:
        n = min(chunk, P - done)
        coracle.mref(sbj_host[done:done + n], ref_host, acc[done:done + n], R, rng, rng, 1.0,
                     nthreads=cores, precision="f32")
        done += n
    el = time.perf_counter() - t0
    out = {"value": done * N / el, "unit": "pairs/s", "cores": cores, "kind": "port",
           "sample": f"first {done} particles x {N} refs of the same workload "
                     f"({el:.1f} s, C restatement, FP32, OpenMP)"}
    # parameter agreement of the GPU run with the FP64 oracle: every particle the
    # CPU leg processed (at least min_agree), every reference
    m = min(max(done, min_agree), P)
    t0 = time.perf_counter()
    ref = coracle.mref(sbj_host[:m], ref_host, acc[:m], R, rng, rng, 1.0, nthreads=cores,
                       precision="f64")
    t_or = time.perf_counter() - t0
    out["agreement_vs_fp64"] = agreement_vs_oracle(gpu_peak[:, :4 * m], gpu_d[:, :m], ref)
    out["agreement_vs_fp64"]["oracle_s"] = round(t_or, 2)
    return out


def agreement_vs_oracle(pk, gd, ref):
    """Parameter agreement of the library's peak_list pk [N, 4m] and d gd [N, m]
    with the FP64 oracle's records for the same m particles: every pair, ties
    (best two distinct alignments within 1e-5 relative) counted apart, and the
    best reference per particle."""
    import numpy as np
    N, m = gd.shape
    tie = ref["tie"]
    same = (pk[:, 3::4] == ref["mirror"]) & (np.abs(pk[:, 1::4] - ref["sx"]) < 1e-4) & \
           (np.abs(pk[:, 2::4] - ref["sy"]) < 1e-4)
    dpsi = np.abs((pk[:, 0::4] - ref["angle"] + 180.0) % 360.0 - 180.0)
    drel = np.abs(gd - ref["d"]) / np.abs(ref["d"])
    full = same & (dpsi <= 0.5)
    nt = ~tie
    k_nt, n_nt = int(full[nt].sum()), int(nt.sum())
    # the best reference per particle (what the classification uses)
    cols = np.arange(m)
    ob = np.argmax(ref["d"], axis=0)
    srt = np.sort(ref["d"], axis=0)
    btie = (srt[-1] - srt[-2]) <= 1e-5 * np.abs(srt[-1]) if N > 1 else np.zeros(m, bool)
    gb = np.argmax(gd, axis=0)
    best_ok = (gb == ob) & full[ob, cols]
    return {
        "particles": int(m), "pairs": int(m * N),
        "ties": int(tie.sum()),
        "pairs_param_agreement_non_tie": k_nt / max(n_nt, 1),
        "pairs_param_agreement_non_tie_ci95": wilson(k_nt, n_nt),
        "shift_mirror_exact": float(same.mean()),
        "shift_mirror_exact_non_tie": float(same[nt].mean()) if nt.any() else 1.0,
        "psi_within_0.5deg": float((dpsi <= 0.5).mean()),
        "psi_max_err_deg_non_tie": float(dpsi[nt].max()) if nt.any() else 0.0,
        "ccf_max_rel_err": float(drel.max()),
        "ccf_rel_err_p50_p99_p9999": [float(np.percentile(drel, q)) for q in (50, 99, 99.99)],
        # the relative error is largest where a pair's maximum is itself near zero:
        # the error against the particle's best value, and that pair's value
        "ccf_max_err_vs_particle_best": float((np.abs(gd - ref["d"]) / np.abs(ref["d"]).max(axis=0)).max()),
        "ccf_max_rel_err_pair_value_vs_particle_best":
            float(np.abs(ref["d"]).flat[int(np.argmax(drel))] / np.abs(ref["d"]).max(axis=0)[int(np.argmax(drel)) % m]),
        "best_ref_agreement_non_tie": float(best_ok[~btie].mean()) if (~btie).any() else 1.0,
        "best_ref_ties": int(btie.sum())}


def host_abi_rate(sbj_h, ref_h, R, rng, runs=5):
    """The reference's ABI end to end: multi_ref_align with host image pointers
    (gather into pinned memory, H2D, polar spectra, search, D2H of peak_list and
    d).  Median of `runs` warm calls after one untimed call, one init for all."""
    import ctypes
    import numpy as np
    from isac_amd import binding
    lib = binding.load()
    P, D, N = sbj_h.shape[0], sbj_h.shape[-1], ref_h.shape[0]
    peak = np.zeros((N, 4 * P), np.float32)
    d = np.zeros((N, P), np.float32)
    cfg = binding.make_config(P, N, D, R, 1.0, rng, rng, "mref_aln")   # retained until gpu_clear
    prm = lib.multi_ref_align_init(ctypes.byref(cfg), ctypes.c_uint(0))
    binding.check(lib)
    try:
        for i in range(P):
            prm[i].shift_x = 0.0
            prm[i].shift_y = 0.0
        ps = binding.get_c_ptr_array([sbj_h[i] for i in range(P)])
        pr = binding.get_c_ptr_array([ref_h[i] for i in range(N)])
        times = []
        for _ in range(runs + 1):
            t0 = time.perf_counter()
            lib.multi_ref_align(ps, pr, peak.ctypes.data_as(binding.float_ptr),
                                d.ctypes.data_as(binding.float_ptr), ctypes.c_uint(P),
                                ctypes.c_float(0.9))
            times.append(time.perf_counter() - t0)
        binding.check(lib)
    finally:
        lib.gpu_clear()
    med = float(np.median(times[1:]))
    return {"value": P * N / med, "unit": "pairs/s", "median_s": med, "runs": runs,
            "host_bytes_in": int(sbj_h.nbytes + ref_h.nbytes), "peak_list_d_bytes_out": int(peak.nbytes + d.nbytes)}


def precision_report(n_particles=1000):
    """Precision report on the small workload (1000 x 64^2, 8 references, rings 1..29, +-2 shifts)
    with the ring contraction in FP16 and in BF16 on MFMA, the top-8 candidates
    re-scored in FP32 and near-ties in FP64, against the FP64 oracle."""
    import numpy as np
    import torch
    from isac_amd import binding, synth
    from oracle import coracle
    ds = synth.dataset_p1k_n8_d64(n_particles=n_particles)
    P, N = ds.sbj.shape[0], ds.ref.shape[0]
    s = torch.from_numpy(ds.sbj).cuda()
    r = torch.from_numpy(ds.ref).cuda()
    t0 = time.perf_counter()
    ref = coracle.mref(ds.sbj, ds.ref, np.zeros((P, 2), np.float32), 29, 2.0, 2.0, 1.0,
                       nthreads=min(16, len(os.sched_getaffinity(0))))
    t_oracle = time.perf_counter() - t0
    cols = np.arange(P)
    order = np.argsort(-ref["d"], axis=0, kind="stable")
    ob = order[0]
    vb, v2 = ref["d"][ob, cols], ref["d"][order[1], cols]
    otie = (vb - v2) <= 1e-5 * np.abs(vb)
    out = {"report": "precision (FP16/BF16 contraction, 1000 particles x 8 refs, 64 px)", "particles": P, "refs": N,
           "oracle": f"FP64 C restatement, {t_oracle:.1f} s", "oracle_best_ref_ties": int(otie.sum())}
    for prec in ("fp16", "bf16"):
        o = binding.mref_best_mixed_device(s, r, 29, 1.0, 2.0, 2.0, prec)
        torch.cuda.synchronize()
        o = {k: v.cpu().numpy() for k, v in o.items()}
        d = o["d"].astype(np.float64)
        err = np.abs(d - ref["d"])
        rel = err / np.abs(ref["d"])
        pk = o["peak_list"]
        tie = np.asarray(ref["tie"], bool)
        pair_same = (pk[:, 3::4] == ref["mirror"]) & (np.abs(pk[:, 1::4] - ref["sx"]) < 1e-4) & \
                    (np.abs(pk[:, 2::4] - ref["sy"]) < 1e-4)
        dpsi_pair = np.abs((pk[:, 0::4] - ref["angle"] + 180) % 360 - 180)
        br = o["best_ref"]
        bp = o["best_peak"]
        same_ref = br == ob
        best_same = same_ref & (bp[:, 3] == ref["mirror"][ob, cols]) & \
            (np.abs(bp[:, 1] - ref["sx"][ob, cols]) < 1e-4) & (np.abs(bp[:, 2] - ref["sy"][ob, cols]) < 1e-4)
        dpsi = np.abs((bp[:, 0] - ref["angle"][ob, cols] + 180) % 360 - 180)
        ok = ~otie
        out[prec] = {
            "approx_all_pairs": {
                "ccf_max_abs_err": float(err.max()), "ccf_max_rel_err": float(rel.max()),
                "ccf_median_rel_err": float(np.median(rel)),
                "ref_shift_mirror_exact_frac": float(pair_same[~tie].mean()),
                "psi_err_deg_p50_p99_max": [float(np.percentile(dpsi_pair[pair_same], q)) for q in (50, 99, 100)]},
            "best_per_particle_after_rescoring": {
                "ref_shift_mirror_exact_frac": float(best_same[ok].mean()),
                "psi_within_0.5deg_frac": float((dpsi[ok & best_same] <= 0.5).mean()),
                "psi_err_deg_p50_p99_max": [float(np.percentile(dpsi[ok & best_same], q)) for q in (50, 99, 100)],
                "near_ties_flagged": int(o["tie"].sum())}}
    return out


def traffic_per_launch(pairs_per_launch):
    """HBM bytes per k_ccf_mref launch from the committed rocprofv3 --pmc summary
    (FETCH_SIZE + WRITE_SIZE per dispatch, each scaled by the factor a 1 GiB copy
    gives), as bytes per pair times the pairs of this launch; None without it."""
    path = os.path.join(ROOT, "profiles", "k_ccf_mref_traffic_p20k_n100_d76.json")
    try:
        with open(path) as f:
            t = json.load(f)
        return t["hbm_bytes_per_pair"] * pairs_per_launch
    except (OSError, KeyError, ValueError):
        return None


def batched_report(steps=3, warmup=1, n_subsets=64):
    """Config 4: 64 uneven subsets (500..5000 particles, box 64/76/96/128, 20..200
    references, +-2..+-4 shifts, rings to box/2-4) in one captured graph around
    the persistent work-queue kernel, against one launch sequence per subset."""
    import numpy as np
    import torch
    from isac_amd import binding, synth
    shapes = synth.uneven_subset_shapes(n_subsets)
    subs = synth.make_subsets_torch(shapes)
    torch.cuda.synchronize()
    pairs = sum(sh["P"] * sh["N"] for sh in shapes)
    bm = binding.BatchedMref(subs)
    out = {"report": "batched uneven subsets (64 subsets)", "subsets": n_subsets,
           "particles": sum(sh["P"] for sh in shapes), "pairs": pairs,
           "boxes": {str(b): sum(1 for sh in shapes if sh["D"] == b) for b in synth.BATCHED_BOXES}}
    t0 = time.perf_counter()
    bm.run(binding.BATCH_GRAPH)   # plan, tune, capture, first launch
    torch.cuda.synchronize()
    out["graph_first_call_s"] = time.perf_counter() - t0
    out["graph_plan"] = bm.report()
    ref_out = [(p.clone(), d.clone()) for p, d in bm.outputs]

    def timed(mode):
        for _ in range(warmup):
            bm.run(mode)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            bm.run(mode)
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / steps

    t_graph = timed(binding.BATCH_GRAPH)
    assert bm.report()["replayed"]
    t_each = timed(binding.BATCH_PER_SUBSET)
    same = all(torch.equal(a, p) and torch.equal(b, d)
               for (a, b), (p, d) in zip(ref_out, bm.outputs))
    out.update({
        "graph_s_per_batch": t_graph, "graph_pairs_per_s": pairs / t_graph,
        "per_subset_s_per_batch": t_each, "per_subset_pairs_per_s": pairs / t_each,
        "speedup_graph_vs_per_subset": t_each / t_graph,
        "outputs_identical": bool(same), "steps": steps, "warmup": warmup})
    return out


def agreement_report(workload="p20k_n100_d76"):
    """Parameter agreement of the FP32 search with the FP64 oracle over the whole
    headline workload: every particle and reference (the bench line checks a
    sample).  The oracle is the C restatement in FP64 on the host cores."""
    import numpy as np
    import torch
    from isac_amd import binding, synth
    from oracle import coracle
    P, N, D, radius, snr, max_shift, int_shifts, R, rng = WORKLOADS[workload]
    sbj, ref, _ = synth.make_dataset_torch(P, N, D, radius, snr, max_shift, int_shifts,
                                           torch.device("cuda:0"))
    acc = torch.zeros((P, 2), dtype=torch.float32, device="cuda:0")
    peak, d = binding.mref_align_device(sbj, ref, R, 1.0, rng, rng, acc_xy=acc)
    torch.cuda.synchronize()
    cores = min(16, len(os.sched_getaffinity(0)))
    t0 = time.perf_counter()
    o = coracle.mref(sbj.cpu().numpy(), ref.cpu().numpy(), np.zeros((P, 2), np.float32), R, rng, rng,
                     1.0, nthreads=cores, precision="f64")
    t_or = time.perf_counter() - t0
    out = {"report": f"FP32 search vs FP64 oracle, whole workload {workload}", "particles": P, "refs": N,
           "oracle": f"FP64 C restatement, {cores} threads, {t_or:.1f} s"}
    out.update(agreement_vs_oracle(peak.cpu().numpy(), d.cpu().numpy(), o))
    # the MFMA forms with the exact best per particle: best reference, shift,
    # mirror and psi (within 0.5 degree) against the oracle's best, on the
    # particles whose best is neither a reference tie nor a pair tie
    cols = np.arange(P)
    ob = np.argmax(o["d"], axis=0)
    srt = np.sort(o["d"], axis=0)
    ok = ~(((srt[-1] - srt[-2]) <= 1e-5 * np.abs(srt[-1])) | np.asarray(o["tie"], bool)[ob, cols])
    for prec in ("fp16", "bf16"):
        b = binding.mref_best_mixed_device(sbj, ref, R, 1.0, rng, rng, prec, acc_xy=acc)
        torch.cuda.synchronize()
        br, bp = b["best_ref"].cpu().numpy(), b["best_peak"].cpu().numpy()
        dpsi = np.abs((bp[:, 0] - o["angle"][ob, cols] + 180.0) % 360.0 - 180.0)
        same = (br == ob) & (bp[:, 3] == o["mirror"][ob, cols]) & \
            (np.abs(bp[:, 1] - o["sx"][ob, cols]) < 1e-4) & (np.abs(bp[:, 2] - o["sy"][ob, cols]) < 1e-4) & \
            (dpsi <= 0.5)
        k, n = int(same[ok].sum()), int(ok.sum())
        out[f"{prec}_best_exact"] = {"particles_non_tie": n, "best_agreement_non_tie": k / max(n, 1),
                                     "best_agreement_non_tie_ci95": wilson(k, n),
                                     "flagged_near_ties": int(b["tie"].sum())}
    return out


def iteration_report(steps=5, warmup=1, fl=0.25, ff=0.2):
    """One multi-reference generation on the device at the headline size (20 000
    particles x 100 references, 76 px): the search (isac_mref_align_device), the
    equal-size assignment from d (isac_assign_groups_device) and the reference
    update (isac_update_references_device), each timed over `steps` calls."""
    import numpy as np
    import torch
    from isac_amd import binding, synth
    P, N, D, radius, snr, max_shift, int_shifts, R, rng = WORKLOADS["p20k_n100_d76"]
    sbj, ref, _ = synth.make_dataset_torch(P, N, D, radius, snr, max_shift, int_shifts, "cuda")
    torch.cuda.synchronize()

    def timed(fn):
        for _ in range(warmup):
            out = fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            out = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / steps, out

    t_search, (peak, d) = timed(lambda: binding.mref_align_device(sbj, ref, R, 1.0, rng, rng))
    t_assign, groups = timed(lambda: binding.assign_groups_device(d))
    cls = torch.from_numpy(binding.classes_from_assignment(groups.cpu().numpy(), N)).cuda()
    t_update, (refs, params, members, avg) = timed(
        lambda: binding.update_references_device(sbj, peak, cls, N, fl, ff))
    m = members.cpu().numpy()
    return {"report": "one multi-reference generation on the device (20 000 particles x 100 refs, 76 px)",
            "particles": P, "refs": N, "img_dim": D, "steps": steps, "warmup": warmup,
            "search_s": t_search, "assign_s": t_assign, "update_s": t_update,
            "generation_s": t_search + t_assign + t_update,
            "class_size_min": int(m.min()), "class_size_max": int(m.max()),
            "filter": {"cutoff": fl, "falloff": ff}}


def learned_report(n_train=200000, n_cal=None, n_test=100000, epochs=None, oracle_sample=10000,
                   target=None, snr=0.1, integer_shifts=True, log=None):
    """Config 5: the pose-proposal network trained on exhaustive-search labels of
    200,000 particles drawn like the small workload's (blob references, uniform psi and mirror,
    integer shifts, SNR 0.1) at 76 px, 100 references, +-3 px (49 shifts);
    threshold calibrated on 40,000 more; evaluated on 100,000 held-out
    particles: agreement with the exhaustive FP32 search and with the FP64
    oracle (sample), fallback rate, end-to-end time against the plain FP32
    search and the MFMA form, and the same search with an untrained network."""
    import numpy as np
    import torch
    from isac_amd import binding, learned, posenet, synth
    from oracle import coracle
    log = log or (lambda m: print(m, file=sys.stderr, flush=True))
    target = learned.TARGET if target is None else target
    n_cal = learned.N_CAL if n_cal is None else n_cal
    epochs = learned.EPOCHS if epochs is None else epochs
    N, D, R, rng = 100, 76, 35, 3.0

    def make(P, seed):
        s, r, _ = synth.make_dataset_torch(P, N, D, 34.0, snr, rng, integer_shifts, seed_refs=1,
                                           seed_particles=seed, seed_noise=seed + 1)
        return s, r

    out = {"report": "learned proposal (100 000 held-out particles, 76 px)", "refs": N, "box": D, "shifts": 49,
           "snr": snr, "integer_shifts": integer_shifts, "train": n_train,
           "calibration": n_cal, "held_out": n_test, "epochs": epochs}
    tr, ref = make(n_train, 11)
    t0 = time.perf_counter()
    lab, _, _ = learned.labels(tr, ref, R, rng)
    torch.cuda.synchronize()
    out["label_s"] = time.perf_counter() - t0
    net = posenet.PoseNet(N, 49, D=D, seed=0, **learned.NET)
    out["params"] = net.n_params
    t0 = time.perf_counter()
    hist = net.train(tr, lab, epochs=epochs, log=log, **learned.TRAIN)
    torch.cuda.synchronize()
    out["train_s"] = time.perf_counter() - t0
    out["train_loss_per_epoch"] = [round(h, 4) for h in hist]
    del tr, lab
    cal, _ = make(n_cal, 21)
    _, dc, pc = learned.labels(cal, ref, R, rng)
    pr, raw = learned.calibrate(net, cal, ref, dc, pc, R, rng, target=target)
    out["calibration_set"] = {"narrowed_only_agreement": raw, "threshold": pr.threshold,
                              "confidence_cut": pr.conf_min, "target": target}
    del cal, dc, pc
    te, _ = make(n_test, 31)
    d, peak = learned.exhaustive(te, ref, R, rng)

    def timed(fn, reps=3):
        # one warm-up, then the median of reps timed runs (a host stall in one run
        # does not move it)
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            t = time.perf_counter()
            r = fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t)
        return sorted(ts)[len(ts) // 2], r

    t_plain, _ = timed(lambda: learned.exhaustive(te, ref, R, rng, chunk=n_test))
    t_mfma, _ = timed(lambda: binding.mref_best_mixed_device(te, ref, R, 1.0, rng, rng, "fp16"))
    t_learn, o = timed(lambda: pr.run(te, ref))
    t_cnn, prop = timed(lambda: net.propose(te, 8))
    t_narrow, on = timed(lambda: binding.mref_proposal_device(te, ref, R, 1.0, rng, rng, prop,
                                                              threshold=-3.0e38,
                                                              edge_fallback=False))
    oe = binding.mref_proposal_device(te, ref, R, 1.0, rng, rng, prop, threshold=-3.0e38)
    agree = learned.agreement(o["best_ref"], o["best_peak"], d, peak)
    agree_n = learned.agreement(on["best_ref"], on["best_peak"], d, peak)
    fb_conf = int((o["fallback"] == 2).sum())
    un = learned.Proposer(posenet.PoseNet(N, 49, D=D, seed=1, **learned.NET), pr.threshold, R, rng)
    t_un, ou = timed(lambda: un.run(te, ref), reps=1)
    agree_un = learned.agreement(ou["best_ref"], ou["best_peak"], d, peak)
    out["held_out_set"] = {
        "agreement_with_exhaustive_fp32": float(agree.double().mean()),
        "narrowed_only_agreement": float(agree_n.double().mean()),
        "fallback_fraction": o["n_fallback"] / n_test,
        "fallback_by_confidence_fraction": fb_conf / n_test,
        "level2_fraction": o["n_level2"] / n_test,
        "edge_fallback_fraction": oe["n_fallback"] / n_test,
        "plain_fp32_s": t_plain, "mfma_fp16_best_s": t_mfma, "learned_s": t_learn,
        "cnn_proposals_s": t_cnn, "narrowed_without_fallback_s": t_narrow,
        "speedup_vs_plain_fp32": t_plain / t_learn, "speedup_vs_mfma": t_mfma / t_learn,
        "learned_particles_per_s": n_test / t_learn,
        "plain_fp32_particles_per_s": n_test / t_plain}
    out["without_training"] = {"agreement": float(agree_un.double().mean()),
                               "fallback_fraction": ou["n_fallback"] / n_test, "time_s": t_un,
                               "speedup_vs_plain_fp32": t_plain / t_un}
    # FP64 oracle on a sample of the held-out particles
    k = oracle_sample
    s_h, r_h = te[:k].cpu().numpy(), ref.cpu().numpy()
    t0 = time.perf_counter()
    orc = coracle.mref(s_h, r_h, np.zeros((k, 2), np.float32), R, rng, rng, 1.0,
                       nthreads=min(16, len(os.sched_getaffinity(0))), precision="f64")
    out["oracle_s"] = time.perf_counter() - t0
    od = torch.from_numpy(orc["d"].astype(np.float32)).cuda()
    opk = torch.from_numpy(orc["peak_list"]).cuda()
    a_l = learned.agreement(o["best_ref"][:k], o["best_peak"][:k], od, opk)
    bref, _, bpk = binding.best_of_exhaustive(d[:, :k].contiguous(),
                                              peak.view(N, -1, 4)[:, :k].reshape(N, 4 * k))
    a_p = learned.agreement(bref, bpk, od, opk)
    k_l, k_p = int(a_l.sum()), int(a_p.sum())
    out["fp64_oracle_sample"] = {"particles": k, "learned_agreement": k_l / k,
                                 "learned_agreement_ci95": wilson(k_l, k),
                                 "plain_fp32_agreement": k_p / k,
                                 "plain_fp32_agreement_ci95": wilson(k_p, k)}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="p20k_n100_d76", choices=sorted(WORKLOADS))
    ap.add_argument("--particles", type=int, default=None, help="override particle count")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-forms", action="store_true", help="skip the mixed-precision forms")
    ap.add_argument("--no-reports", action="store_true",
                    help="skip the compact precision / batched / learned summaries")
    ap.add_argument("--no-host-abi", action="store_true",
                    help="skip the host-pointer ABI timing (multi_ref_align, PCIe included)")
    ap.add_argument("--probe-form", choices=["fp32", "fp16", "bf16", "fp16_best"], default="fp32",
                    help="the form --probe runs: the FP32 search (default), the FP16 / BF16 MFMA "
                         "search, or the FP16 MFMA search with FP32 re-scoring of the best")
    ap.add_argument("--probe", action="store_true",
                    help="one step of the workload plus a 1 GiB device copy, nothing printed: the "
                         "target of rocprofv3 --pmc runs (the copy calibrates the byte counters)")
    ap.add_argument("--report", choices=["precision", "batched", "learned", "iteration", "agreement"],
                    default=None,
                    help="precision: FP16/BF16 accuracy report; batched: uneven-subset "
                         "throughput, graph vs one launch per subset; learned: proposal "
                         "search; iteration: search + assignment + reference update, 20 000 x 100; "
                         "agreement: the FP32 search against the FP64 oracle on the whole workload")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (--gpus {args.gpus} requested); "
              "the path shards over particles (references replicated) with one gather of d and "
              "peak_list at the end", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from isac_amd import binding, synth

    if args.report == "precision":
        print(json.dumps(precision_report()))
        return
    if args.report == "learned":
        print(json.dumps(learned_report()))
        return
    if args.report == "iteration":
        print(json.dumps(iteration_report()))
        return
    if args.report == "agreement":
        print(json.dumps(agreement_report(args.workload)))
        return
    if args.report == "batched":
        print(json.dumps(batched_report(steps=args.steps if args.steps != 10 else 3,
                                        warmup=min(args.warmup, 1))))
        return

    P, N, D, radius, snr, max_shift, int_shifts, R, rng = WORKLOADS[args.workload]
    if args.particles:
        P = args.particles
    S = int((2 * rng + 1) ** 2)
    dev = torch.device("cuda:0")
    sbj, ref, meta = synth.make_dataset_torch(P, N, D, radius, snr, max_shift, int_shifts, dev)
    acc = torch.zeros((P, 2), dtype=torch.float32, device=dev)
    peak = torch.empty((N, 4 * P), dtype=torch.float32, device=dev)
    dval = torch.empty((N, P), dtype=torch.float32, device=dev)
    peak2 = torch.empty_like(peak)
    dval2 = torch.empty_like(dval)
    lib = binding.load()
    torch.cuda.synchronize()

    def step():
        binding.mref_align_device(sbj, ref, R, 1.0, rng, rng, acc_xy=acc, out=(peak, dval))

    if args.probe:
        if args.probe_form == "fp32":
            step()
        elif args.probe_form == "fp16_best":
            binding.mref_best_mixed_device(sbj, ref, R, 1.0, rng, rng, "fp16", acc_xy=acc)
        else:
            binding.mref_align_mixed_device(sbj, ref, R, 1.0, rng, rng, args.probe_form,
                                            acc_xy=acc, out=(peak2, dval2))
        src = torch.ones(1 << 28, dtype=torch.float32, device=dev)   # 1 GiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        return

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    lib.isac_profile_enable(1)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    import ctypes
    n_ccf, ms_ccf = ctypes.c_int(0), ctypes.c_double(0.0)
    n_pol, ms_pol = ctypes.c_int(0), ctypes.c_double(0.0)
    lib.isac_profile_read(ctypes.byref(n_ccf), ctypes.byref(ms_ccf), ctypes.byref(n_pol),
                          ctypes.byref(ms_pol))
    lib.isac_profile_enable(0)
    binding.check(lib)

    ms_step = el / args.steps * 1e3
    value = P * N / (el / args.steps)
    launches = max(n_ccf.value, 1)
    avg_launch_ms = ms_ccf.value / launches
    pairs_per_launch = P * N * args.steps / launches
    fl = flops_per_pair(S, R) * pairs_per_launch
    achieved = fl / (avg_launch_ms * 1e-3) / 1e12
    res = {
        "metric": "particle x reference alignments/sec (76 px, 49 shifts, 256 psi, 2 mirrors)",
        "value": value, "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded blob references, rendered particles + Gaussian noise)",
        "config": {"workload": f"{args.workload}: {P} particles x {N} refs, {D}x{D} px, "
                               f"rings 1..{R}, shifts +-{int(rng)} step 1 ({S}), psi 256 x 2 mirrors",
                   "particles": P, "refs": N, "img_dim": D, "rings": R, "shifts": S},
        "value_protocol": "device-resident: images already in HBM when the timed region starts "
                          "(isac_mref_align_device); host_abi below is the same search behind "
                          "the reference's host-pointer ABI with PCIe both ways",
        "roofline": {"bound": "valu_fp32", "kernel": "k_ccf_mref (ring contraction + psi IFFT + search)",
                     "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": traffic_per_launch(pairs_per_launch),
                     "avg_launch_ms": avg_launch_ms, "launches": launches,
                     "flop_per_pair": flops_per_pair(S, R),
                     "polar_kernel_ms_per_step": ms_pol.value / args.steps},
        "cpu_baseline": None,
    }
    if not args.no_forms:
        # the mixed-precision (MFMA) forms on the same workload, same timing protocol
        forms = {}
        import ctypes
        bound_ms, bound_parts = algorithmic_bound_ms(P, N, S, R, D=D)
        for name, fn in (
                ("fp16", lambda: binding.mref_align_mixed_device(sbj, ref, R, 1.0, rng, rng, "fp16",
                                                                 acc_xy=acc, out=(peak2, dval2))),
                ("bf16", lambda: binding.mref_align_mixed_device(sbj, ref, R, 1.0, rng, rng, "bf16",
                                                                 acc_xy=acc, out=(peak2, dval2))),
                ("fp16_best_exact", lambda: binding.mref_best_mixed_device(sbj, ref, R, 1.0, rng, rng,
                                                                           "fp16", acc_xy=acc))):
            fn()
            torch.cuda.synchronize()
            k = max(2, args.steps // 2)
            lib.isac_profile_enable(1)
            t0 = time.perf_counter()
            for _ in range(k):
                fn()
            torch.cuda.synchronize()
            dt = (time.perf_counter() - t0) / k
            n_g, ms_g = ctypes.c_int(0), ctypes.c_double(0.0)
            n_p, ms_p = ctypes.c_int(0), ctypes.c_double(0.0)
            lib.isac_profile_read(ctypes.byref(n_g), ctypes.byref(ms_g), ctypes.byref(n_p),
                                  ctypes.byref(ms_p))
            lib.isac_profile_enable(0)
            forms[name] = {"value": P * N / dt, "ms_per_step": dt * 1e3, "speedup_vs_fp32": ms_step / (dt * 1e3),
                           "dtype": "fp16" if name != "bf16" else "bf16"}
            if R <= 48 and n_g.value > 0:
                # the search kernel's issued MFMA rate, its launches timed with HIP events
                g_ms_step = ms_g.value / k
                tflops = mfma_issued_flops(P, N, S, R) / (g_ms_step * 1e-3) / 1e12
                forms[name]["roofline"] = {
                    "bound": "mfma", "kernel": "k_mref_rt (16-bit ring contraction + FP16 psi DFT16 x DFT16 + search)",
                    "achieved": tflops, "peak": PEAK_F16_TFLOPS, "unit": "TFLOP/s",
                    "frac": tflops / PEAK_F16_TFLOPS, "kernel_ms_per_step": g_ms_step,
                    "issued_flop_per_step": mfma_issued_flops(P, N, S, R)}
                # per stage: the ring kernel (contraction + psi + search) and the rest of
                # the step (16-bit operands from the images, decode, and for the exact-best
                # form the FP32 spectra and re-scoring), beside the bound's terms
                forms[name]["stages_ms_per_step"] = {
                    "ring_kernel": g_ms_step, "operands_decode_rescoring": dt * 1e3 - g_ms_step}
            forms[name]["algorithmic_bound_ms_per_step"] = bound_ms
            forms[name]["frac_of_algorithmic_bound"] = bound_ms / (dt * 1e3)
        forms["algorithmic_bound_parts_ms"] = bound_parts
        res["mfma_forms"] = forms
    if not args.no_host_abi:
        # not `value`: the same search behind the reference's host-pointer ABI,
        # PCIe transfers and the host-side gather included
        res["host_abi"] = host_abi_rate(sbj.cpu().numpy(), ref.cpu().numpy(), R, rng)
    if not args.no_cpu_baseline:
        cores = min(16, len(os.sched_getaffinity(0)))
        m = min(P, 4096)
        sbj_h = sbj[:m].cpu().numpy()
        res["cpu_baseline"] = cpu_baseline(sbj_h, ref.cpu().numpy(), np.zeros((m, 2), np.float32),
                                           R, rng, peak.cpu().numpy(), dval.cpu().numpy(),
                                           args.cpu_seconds, cores)
        res["cpu_baseline"]["gpu_over_cpu"] = value / res["cpu_baseline"]["value"]
    if not args.no_reports:
        # compact summaries of the other workloads, outside the timed region
        del sbj, ref, acc, peak, dval, peak2, dval2
        torch.cuda.empty_cache()
        res["reports"] = compact_reports(log=lambda m: print(m, file=sys.stderr, flush=True))
    print(json.dumps(res))


def compact_reports(log):
    """Precision (1 000 x 8 refs, 64 px), uneven-subset batch (64 subsets) and
    learned proposal (100 000 held-out particles, 100 refs, 76 px) reports,
    reduced to each report's headline figures."""
    import torch
    out = {}
    t0 = time.perf_counter()
    pr = precision_report()
    out["precision_p1k_n8_d64"] = {
        prec: {"approx_ccf_max_rel_err": pr[prec]["approx_all_pairs"]["ccf_max_rel_err"],
               "approx_ref_shift_mirror_exact": pr[prec]["approx_all_pairs"]["ref_shift_mirror_exact_frac"],
               "best_after_rescoring_exact": pr[prec]["best_per_particle_after_rescoring"]["ref_shift_mirror_exact_frac"],
               "best_after_rescoring_psi_within_0.5deg": pr[prec]["best_per_particle_after_rescoring"]["psi_within_0.5deg_frac"],
               "near_ties_flagged": pr[prec]["best_per_particle_after_rescoring"]["near_ties_flagged"]}
        for prec in ("fp16", "bf16")}
    out["precision_p1k_n8_d64"]["s"] = round(time.perf_counter() - t0, 1)
    log(f"precision report done ({out['precision_p1k_n8_d64']['s']} s)")
    torch.cuda.empty_cache()
    t0 = time.perf_counter()
    br = batched_report(steps=3, warmup=1)
    out["batched_64_uneven_subsets"] = {
        k: br[k] for k in ("particles", "pairs", "graph_pairs_per_s", "per_subset_pairs_per_s",
                           "speedup_graph_vs_per_subset", "outputs_identical", "graph_first_call_s")}
    out["batched_64_uneven_subsets"]["s"] = round(time.perf_counter() - t0, 1)
    log(f"batched report done ({out['batched_64_uneven_subsets']['s']} s)")
    torch.cuda.empty_cache()
    t0 = time.perf_counter()
    lr = learned_report(log=log)
    h = lr["held_out_set"]
    out["learned_p100k_n100_d76"] = {
        "speedup_vs_plain_fp32": h["speedup_vs_plain_fp32"],
        "learned_s": h["learned_s"], "plain_fp32_s": h["plain_fp32_s"],
        "agreement_with_exhaustive_fp32": h["agreement_with_exhaustive_fp32"],
        "fallback_fraction": h["fallback_fraction"],
        "fp64_oracle_sample": lr["fp64_oracle_sample"],
        "train_s": lr["train_s"],
        "without_training_speedup": lr["without_training"]["speedup_vs_plain_fp32"]}
    out["learned_p100k_n100_d76"]["s"] = round(time.perf_counter() - t0, 1)
    log(f"learned report done ({out['learned_p100k_n100_d76']['s']} s)")
    return out


if __name__ == "__main__":
    main()
