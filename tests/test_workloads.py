"""Host-side workload generators: the batched workload's scenes are drawn as
specified (log-uniform triangle counts, the resolution set, 1-4 spp) and are
reproducible from the seed."""
import math

from rtg import scenes

RESOLUTIONS = ((16, 16), (40, 24), (64, 48), (33, 17))


def test_batch_workload_shapes():
    ws = scenes.batch_workload(n_scenes=6, seed=7, tri_range=(1_000, 20_000), resolutions=RESOLUTIONS)
    assert len(ws) == 6
    for s in ws:
        assert (s.width, s.height) in RESOLUTIONS
        assert 1 <= s.spp <= 4
        assert 500 <= s.instanced_triangles <= 40_000


def test_batch_scene_hits_its_target():
    for target in (1_000, 5_000, 60_000, 200_000):
        s = scenes.batch_scene(11, target, 32, 32, 1)
        assert abs(math.log(s.instanced_triangles / target)) < math.log(1.6), (target, s.instanced_triangles)


def test_batch_workload_is_reproducible():
    a = scenes.batch_workload(n_scenes=3, seed=9, tri_range=(1_000, 8_000), resolutions=RESOLUTIONS)
    b = scenes.batch_workload(n_scenes=3, seed=9, tri_range=(1_000, 8_000), resolutions=RESOLUTIONS)
    for x, y in zip(a, b):
        assert (x.width, x.height, x.spp, x.instanced_triangles) == (y.width, y.height, y.spp, y.instanced_triangles)
        xa, ya = x.to_arrays(), y.to_arrays()
        assert xa.keys() == ya.keys()
        for k in xa:
            assert (xa[k] == ya[k]).all(), k
