"""Shared parity checks: HIP results against the oracle's (tests only)."""
import numpy as np

PSI_TOL_DEG = 0.5      # psi within 0.5 degree
D_RTOL = 1e-4          # FP32 CCF maximum vs FP64 oracle, relative


def circ_diff(a, b):
    return np.abs((np.asarray(a, np.float64) - np.asarray(b, np.float64) + 180.0) % 360.0 - 180.0)


def compare_mref(peak, d, ref, label=""):
    """peak [N, 4P], d [N, >=P] from the library; ref = oracle dict ([N, P]).

    Pairs whose best two distinct alignments are within 1e-5 relative (oracle
    'tie') must still agree on d, and are reported; every other pair must
    match mirror and shift exactly and psi within 0.5 degree.
    """
    P = ref["d"].shape[1]
    tie = np.asarray(ref["tie"], bool)
    ok = ~tie
    mir = peak[:, 3::4]
    sx = peak[:, 1::4]
    sy = peak[:, 2::4]
    ang = peak[:, 0::4]
    dd = d[:, :P]
    rel = np.abs(dd - ref["d"]) / np.maximum(np.abs(ref["d"]), 1e-30)
    assert rel.max() <= D_RTOL, f"{label}: CCF max rel err {rel.max():.3e}"
    mm = (mir != ref["mirror"]) | (np.abs(sx - ref["sx"]) > 1e-4) | (np.abs(sy - ref["sy"]) > 1e-4)
    bad = mm & ok
    assert not bad.any(), f"{label}: {bad.sum()} non-tie pairs differ in mirror/shift, e.g. {np.argwhere(bad)[:5]}"
    da = circ_diff(ang, ref["angle"])
    badang = (da > PSI_TOL_DEG) & ok
    assert not badang.any(), f"{label}: {badang.sum()} pairs with psi error > 0.5 deg (max {da[ok].max():.3f})"
    return {"pairs": int(ok.size), "ties": int(tie.sum()), "tie_mismatch": int((mm & tie).sum()),
            "max_psi_err": float(da[ok].max()) if ok.any() else 0.0, "max_d_rel": float(rel.max())}
