"""bench.py's own arithmetic and its contract checks, on the CPU: the
algorithmic bytes of a traversal launch from the device counters, the
measured-traffic lookup, and the refusal of more than one GPU."""
import json
import os
import subprocess
import sys

from conftest import REPO

sys.path.insert(0, REPO)
_queues = os.environ.get("GPU_MAX_HW_QUEUES")
import bench  # noqa: E402  (imports no GPU code; it raises GPU_MAX_HW_QUEUES, restored below)

if _queues is None:
    os.environ.pop("GPU_MAX_HW_QUEUES", None)
else:
    os.environ["GPU_MAX_HW_QUEUES"] = _queues


def counters(**kw):
    names = ["primary_rays", "bounce_rays", "shadow_rays", "node_visits", "triangle_tests", "instance_visits",
             "anyhit_calls", "occ_node_visits", "occ_triangle_tests", "occ_instance_visits", "occ_anyhit_calls",
             "side_shadow_rays", "side_occ_node_visits", "side_occ_triangle_tests", "side_occ_instance_visits",
             "side_occ_anyhit_calls", "side_node_visits", "side_triangle_tests", "side_instance_visits",
             "side_anyhit_calls", "spill_pushes", "occ_spill_pushes", "side_spill_pushes", "side_occ_spill_pushes",
             "textured_hits", "side_textured_hits", "side_bounce_rays"]
    c = dict.fromkeys(names, 0)
    c.update(kw)
    return c


def test_algorithmic_bytes_of_the_frame_stream_launches():
    # ray 32 B + hit 24 B / flag 4 B, 128 B per node, 48 B per triangle,
    # 112 B per instance entered; the bounce stream's share is left out
    c = counters(primary_rays=10, bounce_rays=3, shadow_rays=40, node_visits=200, triangle_tests=30,
                 instance_visits=20, occ_node_visits=500, occ_triangle_tests=60, occ_instance_visits=50,
                 side_shadow_rays=8, side_occ_node_visits=100, side_occ_triangle_tests=10,
                 side_occ_instance_visits=5, side_node_visits=70, side_triangle_tests=6, side_instance_visits=4)
    closest = 10 * 56 + (200 - 70) * 128 + (30 - 6) * 48 + (20 - 4) * 112
    occlusion = (40 - 8) * 36 + (500 - 100) * 128 + (60 - 10) * 48 + (50 - 5) * 112
    assert bench.algorithmic_bytes(c, "closest") == closest
    assert bench.algorithmic_bytes(c, "occlusion") == occlusion
    # without a bounce stream the bounce rays are traced by the frame's launches
    c = counters(primary_rays=10, bounce_rays=3, node_visits=200)
    assert bench.algorithmic_bytes(c, "closest") == 13 * 56 + 200 * 128


def test_shade_bytes_of_the_level_0_launch():
    # per ray 60 B in; per miss 16 B out; per textured hit 292 B; per shadow
    # ray 44 B; per continuation 204 B; the bounce stream's share left out
    c = counters(primary_rays=100, bounce_rays=12, shadow_rays=280, textured_hits=70, side_textured_hits=5,
                 side_shadow_rays=20, side_bounce_rays=2, side_node_visits=1)
    tex, shadow, cont = 65, 260, 10
    misses = 100 - tex - cont
    assert bench.shade_bytes(c) == 100 * 60 + misses * 16 + tex * 292 + shadow * 44 + cont * 204


def test_measured_figures_read_the_committed_summary():
    with open(os.path.join(REPO, "profiles", "hbm_traffic_trace_kernels.json")) as f:
        d = json.load(f)
    v = d["per_launch_bytes"]["k_trace_occlusion"]
    traffic, hit, req = bench.measured("k_trace_occlusion")
    assert traffic == v["fetch"] + v["write"]
    assert 0.0 < hit < 1.0 and req == d["l2_requests_per_launch"]["k_trace_occlusion"]
    assert bench.measured("no_such_kernel") == (None, None, None)
    r = bench.kernel_roofline("k_trace_occlusion", 8e9, 2.0, 5)
    assert r["achieved"] == 4000.0 and r["frac"] == 0.5
    assert r["dram_gb_s"] == round(traffic / 2e-3 / 1e9, 1)
    assert r["dram_frac"] == round(r["dram_gb_s"] / 8000.0, 4)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]
