"""Pins the dynamic scene model (oracle/dynamic.py) against hand-worked
sequences of tg's create/destroy bookkeeping (tgvk_raytracer.c:687-712,
805-868, 994-1067) and against its invariants under random sequences."""
from __future__ import annotations

import numpy as np

from oracle.dynamic import DynamicScene
from tests import scenes as S
import tg_amd as T


def luts(n=2):
    return np.zeros(256 * n, T.LUT_DTYPE)


def test_hand_worked_sequence():
    m = DynamicScene(4, 20, luts())
    eye, t = S.identity_rotation(), (0.0, 0.0, 0.0)
    a = m.create((2, 1, 1), t, eye, 0)
    b = m.create((1, 1, 3), t, eye, 1)
    c = m.create((2, 2, 1), t, eye, 0)
    assert (a, b, c) == (0, 1, 2)  # the stacks start as [cap-1, ..., 0]
    assert list(m.cluster_pointers[:9]) == list(range(9))
    assert [int(m.objects[o]["first_cluster_pointer"]) for o in (a, b, c)] == [0, 2, 5]
    m.destroy(b)
    # b's clusters 2, 3, 4 go back in pointer order; c's pointers shift down
    assert m.free_clusters[-3:] == [2, 3, 4]
    assert list(m.cluster_pointers[:6]) == [0, 1, 5, 6, 7, 8] and m.n_cluster_pointers == 6
    assert int(m.objects[c]["first_cluster_pointer"]) == 2 and not m.live(b)
    d = m.create((1, 1, 2), t, eye, 1)
    assert d == 1  # LIFO: the freed object index comes back first
    assert list(m.cluster_pointers[:8]) == [0, 1, 5, 6, 7, 8, 4, 3]
    assert int(m.objects[d]["first_cluster_pointer"]) == 6
    assert m.c2o[4] == d and m.c2o[3] == d and m.c2o[2] == b  # index 2 is free; its stale entry stays
    m.destroy(a)
    assert list(m.cluster_pointers[:6]) == [5, 6, 7, 8, 4, 3]
    assert int(m.objects[c]["first_cluster_pointer"]) == 0 and int(m.objects[d]["first_cluster_pointer"]) == 4
    m.check_invariants()


def test_random_sequences_keep_the_invariants():
    rng = np.random.default_rng(3)
    for trial in range(5):
        m = DynamicScene(10, 200, luts())
        alive = []
        for step in range(80):
            free_cl = len(m.free_clusters)
            if alive and (rng.random() < 0.4 or not m.free_objects or free_cl < 27):
                o = alive.pop(int(rng.integers(len(alive))))
                m.destroy(o)
            else:
                n = tuple(int(v) for v in rng.integers(1, 4, 3))
                if int(np.prod(n)) > free_cl or not m.free_objects:
                    continue
                alive.append(m.create(n, rng.uniform(-50, 50, 3), S.random_rotation(rng), int(rng.integers(2))))
            m.check_invariants()
