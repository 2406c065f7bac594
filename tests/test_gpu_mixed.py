"""Mixed-precision (FP16/BF16 MFMA) multi-reference form against FP32 and the FP64 oracle."""
import numpy as np
import pytest

from isac_amd import binding, synth
from oracle import coracle
from parity import circ_diff

pytestmark = pytest.mark.gpu


def _stats(peak, d, ref):
    tie = np.asarray(ref["tie"], bool)
    same = (peak[:, 3::4] == ref["mirror"]) & (np.abs(peak[:, 1::4] - ref["sx"]) < 1e-4) & \
           (np.abs(peak[:, 2::4] - ref["sy"]) < 1e-4)
    dpsi = circ_diff(peak[:, 0::4], ref["angle"])
    rel = np.abs(d - ref["d"]) / np.abs(ref["d"])
    ok = same & (dpsi <= 0.5)
    return {"pairs": int(ok.size), "agree": float(ok.mean()), "agree_non_tie": float(ok[~tie].mean()),
            "d_rel_max": float(rel.max()), "d_rel_median": float(np.median(rel)),
            "psi_err_p99": float(np.percentile(dpsi[same], 99)) if same.any() else 0.0}


@pytest.mark.parametrize("precision,min_agree,max_rel", [("fp16", 0.97, 2e-3), ("bf16", 0.85, 2e-2)])
def test_mixed_vs_oracle_p1k_n8_d64(gpu, precision, min_agree, max_rel):
    import torch
    ds = synth.dataset_p1k_n8_d64(n_particles=128)
    acc = np.random.default_rng(1).uniform(-1, 1, (128, 2)).astype(np.float32)
    s = torch.from_numpy(ds.sbj).cuda()
    r = torch.from_numpy(ds.ref).cuda()
    a = torch.from_numpy(acc).cuda()
    peak, d = binding.mref_align_mixed_device(s, r, 29, 1.0, 2.0, 2.0, precision, acc_xy=a)
    torch.cuda.synchronize()
    ref = coracle.mref(ds.sbj, ds.ref, acc, 29, 2.0, 2.0, 1.0)
    st = _stats(peak.cpu().numpy(), d.cpu().numpy(), ref)
    print(precision, "1000 x 8, 64 px", st)
    assert st["agree"] >= min_agree and st["d_rel_max"] <= max_rel


@pytest.mark.parametrize("precision", ["fp16", "bf16"])
def test_mixed_vs_oracle_p20k_n100_d76(gpu, precision):
    import torch
    ds = synth.dataset_p20k_n100_d76(n_particles=48, n_refs=100)
    s = torch.from_numpy(ds.sbj).cuda()
    r = torch.from_numpy(ds.ref).cuda()
    peak, d = binding.mref_align_mixed_device(s, r, 35, 1.0, 3.0, 3.0, precision)
    torch.cuda.synchronize()
    ref = coracle.mref(ds.sbj, ds.ref, np.zeros((48, 2), np.float32), 35, 3.0, 3.0, 1.0)
    st = _stats(peak.cpu().numpy(), d.cpu().numpy(), ref)
    print(precision, "100 refs, 76 px", st)
    assert st["d_rel_max"] <= (5e-3 if precision == "fp16" else 5e-2)
