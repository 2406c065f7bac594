"""Host helper that turns assign_groups' layout into a class per particle."""
import numpy as np
import pytest

from isac_amd import binding
from oracle import isac_oracle as O


def _driver_belongsto(id_list_long, numref, nima):
    # the driver's own loops, bin/sp_isac2_gpu.py:1120-1133, over plain lists
    id_list = [[] for _ in range(numref)]
    maxasi = nima // numref
    for i in range(maxasi * numref):
        id_list[i // maxasi].append(id_list_long[i])
    for i in range(nima % maxasi):
        id_list[id_list_long[-1]].append(id_list_long[maxasi * numref + i])
    belongsto = [0] * nima
    for k in range(numref):
        for i in id_list[k]:
            belongsto[i] = k
    return belongsto


@pytest.mark.parametrize("n_ref,n_img", [(3, 11), (20, 30), (7, 7), (4, 19), (100, 250)])
def test_classes_from_assignment_follows_the_driver(n_ref, n_img):
    rng = np.random.default_rng(n_ref * 1000 + n_img)
    d = rng.standard_normal((n_ref, n_img))
    out = O.assign_groups(d)
    cls = binding.classes_from_assignment(out, n_ref)
    assert cls.tolist() == _driver_belongsto([int(v) for v in out], n_ref, n_img)


def test_more_classes_than_capacity_leaves_class_zero():
    # 20 classes x 30 particles: cap = 1, 30 % 1 = 0 leftovers placed, so the
    # 10 particles beyond the first 20 entries keep the default class 0
    d = np.random.default_rng(1).standard_normal((20, 30))
    out = O.assign_groups(d)
    cls = binding.classes_from_assignment(out, 20)
    unplaced = out[20:30]
    assert np.all(cls[unplaced] == 0)
    assert sorted(cls[out[:20]].tolist()) == list(range(20))


def test_too_few_particles_raise():
    with pytest.raises(ValueError):
        binding.classes_from_assignment(np.zeros(4, np.int32), 5)
