"""Host helper that turns assign_groups' layout into a class per particle."""
import numpy as np

from isac_amd import binding
from oracle import isac_oracle as O


def test_classes_from_assignment_follows_the_driver():
    # bin/sp_isac2_gpu.py:1122-1135: blocks of cap per class, leftovers to the last id
    rng = np.random.default_rng(3)
    d = rng.standard_normal((3, 11))
    out = O.assign_groups(d)
    cls = binding.classes_from_assignment(out, 3)
    cap = 11 // 3
    for k in range(3):
        assert sorted(np.flatnonzero(cls == k).tolist()) == sorted(
            out[k * cap:(k + 1) * cap].tolist() + (out[9:11].tolist() if out[11] == k else []))
