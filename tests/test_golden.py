"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py).

CPU: the oracle and its C restatement still reproduce the fixtures.
GPU: the HIP library reproduces them through the C ABI.
"""
import os

import numpy as np
import pytest

from isac_amd import binding
from oracle import coracle
from oracle import isac_oracle as O
from parity import circ_diff, compare_mref

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return dict(np.load(os.path.join(GOLD, name), allow_pickle=False))


def test_oracle_reproduces_mref_fixture():
    g = load("golden_mref.npz")
    r = O.multi_ref_align(g["sbj"], g["ref"], g["acc"], int(g["ring_num"]), float(g["shift_rng"]),
                          float(g["shift_rng"]), float(g["shift_step"]))
    for k in ("mirror", "shift_idx", "t"):
        assert np.array_equal(r[k], g[k])
    assert np.allclose(r["d"], g["d"], rtol=1e-12)
    assert np.allclose(r["angle"], g["angle"], atol=1e-9)


def test_c_oracle_reproduces_mref_fixture():
    g = load("golden_mref.npz")
    r = coracle.mref(g["sbj"], g["ref"], g["acc"], int(g["ring_num"]), float(g["shift_rng"]),
                     float(g["shift_rng"]), float(g["shift_step"]))
    for k in ("mirror", "shift_idx", "t"):
        assert np.array_equal(np.asarray(r[k]), g[k])
    assert np.allclose(r["d"], g["d"], rtol=1e-6)
    assert circ_diff(r["angle"], g["angle"]).max() < 1e-4


def test_oracle_reproduces_prealign_fixture():
    g = load("golden_prealign.npz")
    sh = O.shift_grid(float(g["shift_rng"]), float(g["shift_rng"]), 1.0)
    tab = O.single_ref_table(g["sbj"], g["ref"], np.zeros(len(g["sbj"]), int),
                             g["acc"].astype(np.float64), int(g["ring_num"]), sh)
    p = O.decode_single_ref(tab, g["acc"].astype(np.float64), sh, g["sbj"].shape[-1], int(g["ring_num"]))
    for k in ("mirror", "shift_idx", "t"):
        assert np.array_equal(p[k], g[k])
    assert np.allclose(p["angle"], g["angle"], atol=1e-9)


def test_oracle_reproduces_reffree_fixture():
    g = load("golden_reffree.npz")
    sh = O.shift_grid(float(g["shift_rng"]), float(g["shift_rng"]), 1.0)
    refs = g["avg"].astype(np.float64)
    acc = np.zeros((len(g["sbj"]), 2))
    n_iter, fh, ff = int(g["n_iter"]), float(g["cutoff"]), float(g["falloff"])
    for i in range(n_iter):
        refs = O.tangent_filter(refs, 0.1 + (fh - 0.1) * i / float(n_iter - 1), ff)
        prm, refs = O.ref_free_iteration(g["sbj"], refs, g["cid"], acc, int(g["ring_num"]), sh,
                                         g["sbj"].shape[-1])
        acc = np.stack([prm["shift_x"], prm["shift_y"]], 1)
    assert np.array_equal(prm["mirror"], g["mirror"])
    assert np.allclose(prm["angle"], g["angle"], atol=1e-9)
    assert np.allclose(refs, g["refs_out"], rtol=1e-9, atol=1e-12)


@pytest.mark.gpu
def test_gpu_mref_fixture(gpu):
    g = load("golden_mref.npz")
    peak, d = binding.multi_ref_align(g["sbj"], g["ref"], int(g["ring_num"]), float(g["shift_step"]),
                                      float(g["shift_rng"]), float(g["shift_rng"]), acc_xy=g["acc"])
    compare_mref(peak, d, g, "golden mref")


@pytest.mark.gpu
def test_gpu_prealign_fixture(gpu):
    g = load("golden_prealign.npz")
    p = binding.pre_align(g["sbj"], g["ref"][0], int(g["ring_num"]), float(g["shift_rng"]), 1.0,
                          batch=len(g["sbj"]), acc_xy=g["acc"])
    assert np.array_equal(p["mirror"], g["mirror"])
    assert np.allclose(p["shift_x"], g["shift_x"], atol=1e-5)
    assert np.allclose(p["shift_y"], g["shift_y"], atol=1e-5)
    assert circ_diff(p["angle"], g["angle"]).max() <= 0.5


@pytest.mark.gpu
def test_gpu_reffree_fixture(gpu):
    g = load("golden_reffree.npz")
    p = binding.ref_free_align(g["sbj"], g["cid"], g["avg"], int(g["ring_num"]), float(g["shift_rng"]),
                               int(g["n_iter"]), float(g["cutoff"]), float(g["falloff"]))
    assert np.array_equal(p["mirror"], g["mirror"])
    assert np.allclose(p["shift_x"], g["shift_x"], atol=1e-5)
    assert circ_diff(p["angle"], g["angle"]).max() <= 0.5
