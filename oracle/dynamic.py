"""Dynamic scene model — TEST INFRASTRUCTURE ONLY.

A plain-Python restatement of tg's object bookkeeping, the checker of
tgx_scene_reserve / tgx_object_create / tgx_object_destroy
(tg_amd/csrc/tgx_dynamic.hip). It follows the reference line by line:

  - capacities and free stacks, tgvk_raytracer.c:687-712: the stacks hold
    [capacity - 1, ..., 0] and are popped from the end;
  - tg_raytracer_create_object, tgvk_raytracer.c:805-868: pop the object
    index, append the pointers, pop one cluster index per pointer in
    relative-pointer order, write both maps;
  - tg_raytracer_destroy_object, tgvk_raytracer.c:994-1067: push the
    cluster indices back in pointer order, walk the later objects in pointer
    order (each found through the cluster map of its first pointer) lowering
    their first pointer, shift the later pointers down, push the object
    index back.

Only the record fields and the voxel arrays differ from tg, as in the
library: the caller supplies rotation, LUT index, bits and materials.
"""
from __future__ import annotations

import numpy as np

import tg_amd as T


class DynamicScene:
    def __init__(self, object_capacity: int, cluster_capacity: int, luts: np.ndarray):
        self.objects = np.zeros(object_capacity, T.OBJECT_DTYPE)
        self.free_objects = [object_capacity - i - 1 for i in range(object_capacity)]
        self.cluster_pointers = np.zeros(cluster_capacity, np.uint32)
        self.n_cluster_pointers = 0
        self.free_clusters = [cluster_capacity - i - 1 for i in range(cluster_capacity)]
        self.c2o = np.zeros(cluster_capacity, np.uint32)
        self.bits = np.zeros(16 * cluster_capacity, np.uint32)
        self.materials = np.zeros(512 * cluster_capacity, np.uint8)
        self.luts = np.array(luts, dtype=T.LUT_DTYPE, copy=True)

    @staticmethod
    def count(ob) -> int:
        n = ob["n_cluster_pointers_per_dim"]
        return int(n[0]) * int(n[1]) * int(n[2])

    def live(self, o: int) -> bool:  # tg_object_is_initialized (tgvk_raytracer.c:1069-1077)
        return bool(np.all(self.objects[o]["n_cluster_pointers_per_dim"] != 0))

    def create(self, n_per_dim, translation, rotation, lut_idx: int, bits=None, materials=None) -> int:
        o = self.free_objects.pop()
        ob = self.objects[o]
        ob["n_cluster_pointers_per_dim"] = n_per_dim
        n = self.count(ob)
        ob["first_cluster_pointer"] = self.n_cluster_pointers
        first = self.n_cluster_pointers
        self.n_cluster_pointers += n
        ob["translation"] = translation
        ob["rotation"] = rotation
        ob["lut_idx"] = lut_idx
        for rel in range(n):
            cidx = self.free_clusters.pop()
            self.cluster_pointers[first + rel] = cidx
            self.c2o[cidx] = o
            self.bits[16 * cidx:16 * cidx + 16] = 0 if bits is None else bits[16 * rel:16 * rel + 16]
            self.materials[512 * cidx:512 * cidx + 512] = 0 if materials is None else materials[512 * rel:512 * rel + 512]
        return o

    def destroy(self, o: int) -> None:
        ob = self.objects[o]
        n = self.count(ob)
        first = int(ob["first_cluster_pointer"])
        end = first + n
        for p in range(first, end):
            self.free_clusters.append(int(self.cluster_pointers[p]))
        if end < self.n_cluster_pointers:
            p = end
            while p < self.n_cluster_pointers:
                other = int(self.c2o[self.cluster_pointers[p]])
                self.objects[other]["first_cluster_pointer"] -= n
                p += self.count(self.objects[other])
            for p in range(end, self.n_cluster_pointers):
                self.cluster_pointers[p - n] = self.cluster_pointers[p]
        self.n_cluster_pointers -= n
        self.free_objects.append(o)
        self.objects[o] = np.zeros(1, T.OBJECT_DTYPE)[0]

    def scene(self) -> T.Scene:
        """The arrays as tgx_scene_export writes them (live pointers only)."""
        return T.Scene(self.objects.copy(), self.cluster_pointers[:self.n_cluster_pointers].copy(), self.c2o.copy(),
                       self.bits.copy(), self.materials.copy(), self.luts.copy())

    def check_invariants(self) -> None:
        """Live objects tile [0, n_cluster_pointers) with their pointer
        ranges; each pointer maps to a cluster of its own object; the free
        cluster stack holds exactly the indices no pointer uses."""
        cover = np.zeros(self.n_cluster_pointers, np.int32)
        for o in range(self.objects.size):
            if not self.live(o):
                assert o in self.free_objects
                continue
            ob = self.objects[o]
            f, n = int(ob["first_cluster_pointer"]), self.count(ob)
            assert f + n <= self.n_cluster_pointers
            cover[f:f + n] += 1
            for p in range(f, f + n):
                assert self.c2o[self.cluster_pointers[p]] == o
        assert np.all(cover == 1)
        used = set(int(c) for c in self.cluster_pointers[:self.n_cluster_pointers])
        assert len(used) == self.n_cluster_pointers
        assert used.isdisjoint(self.free_clusters)
        assert len(used) + len(self.free_clusters) == self.c2o.size
