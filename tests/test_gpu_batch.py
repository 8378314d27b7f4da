"""Batched rendering (mt_batch_*), both launch modes: every job's image equals the single-job
render with the same scene, camera and settings, bit for bit, whatever the batch's tiling and
iteration size, and jobs submitted while the batch runs complete after their submission.  The
persistent mode serves a batch with one kernel launch (a new one only after the queue stayed
empty for its idle time), the graph mode with a captured graph per iteration."""
import time

import numpy as np
import pytest

from helpers import load_scene, small_settings
from mitrace_amd import abi, scenegen

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scenes(gpu):
    from mitrace_amd.tracer import GpuScene
    names = ("cornell_gen", "sphere_grid_coarse", "material_spheres", "light_creep")
    return {n: GpuScene(load_scene(n)) for n in names}


def _jobs():
    """uneven jobs: sizes, sample counts, depths, modes and outputs"""
    return [
        ("cornell_gen", small_settings(width=37, height=23, samples=16, bounces=8, seed=1)),
        ("sphere_grid_coarse", small_settings(width=64, height=36, samples=8, bounces=5, seed=2,
                                              hdri_rotation=30.0, hdri_rotation_apply=True)),
        ("material_spheres", small_settings(width=48, height=48, samples=12, bounces=6, seed=3,
                                            firefly_enable=True)),
        ("cornell_gen", small_settings(width=20, height=30, samples=32, bounces=12, seed=4,
                                       mode=abi.MT_MODE_FIXED, rr_depth=3,
                                       output=abi.MT_OUTPUT_LINEAR)),
        ("light_creep", small_settings(width=40, height=40, samples=4, bounces=3, seed=5)),
        ("sphere_grid_coarse", small_settings(width=17, height=9, samples=64, bounces=10, seed=6)),
    ]


def _check_launches(b, mode):
    from mitrace_amd.batch import PERSISTENT
    s = b.stats()
    if mode == PERSISTENT:
        assert s["launches"] == 1 and s["graphs_captured"] == 0, s
    else:
        assert s["launches"] == 0 and s["graphs_captured"] >= 1, s
    return s


@pytest.mark.parametrize("mode", [1, 2], ids=["persistent", "graphs"])
@pytest.mark.parametrize("options", [
    {"paths_per_iteration": 1 << 14, "tile_pixels": 16, "iteration_tiles": 64},
    {"paths_per_iteration": 1 << 20, "tile_pixels": 1024, "iteration_tiles": 4096,
     "trace_blocks_per_cu": 4},
    {"paths_per_iteration": 1 << 16, "tile_pixels": 256, "streams": 1},
    {"paths_per_iteration": 1 << 15, "tile_pixels": 64, "streams": 16},
])
def test_batch_images_equal_single_job_renders(scenes, options, mode):
    from mitrace_amd._lib import MiTraceError
    from mitrace_amd.batch import GRAPHS, Batch
    print(f"batch mode {mode} options {options}", flush=True)
    if mode == GRAPHS and options.get("streams", 3) > 8:
        with pytest.raises(MiTraceError) as e:  # the graph mode has at most 8 streams
            Batch(dict(options, mode=mode))
        assert e.value.code == abi.MT_E_UNSUPPORTED
        return
    b = Batch(dict(options, mode=mode))
    ids = []
    for name, st in _jobs():
        ids.append(b.submit(scenes[name], st))
    b.start()
    b.finish()
    for (name, st), j in zip(_jobs(), ids):
        ref, _, _ = scenes[name].render(st)
        np.testing.assert_array_equal(b.image(j), ref, err_msg=f"{name} {st}")
        info = b.info(j)
        assert info["done"] and info["complete_ms"] >= info["submit_ms"]
        assert info["samples"] == st.width * st.height * st.samples
    s = _check_launches(b, mode)
    assert s["jobs"] == len(ids) and s["scenes"] == 4 and s["iterations"] >= 4
    b.close()


@pytest.mark.parametrize("mode", [1, 2], ids=["persistent", "graphs"])
def test_jobs_submitted_while_running(scenes, mode):
    from mitrace_amd.batch import Batch
    print(f"batch mode {mode}", flush=True)
    b = Batch({"paths_per_iteration": 1 << 16, "tile_pixels": 64, "mode": mode})
    b.start()
    ids = []
    for k, (name, st) in enumerate(_jobs()):
        ids.append(b.submit(scenes[name], st))
        time.sleep(0.01 * k)
    b.finish()
    for (name, st), j in zip(_jobs(), ids):
        ref, _, _ = scenes[name].render(st)
        np.testing.assert_array_equal(b.image(j), ref)
        info = b.info(j)
        assert info["done"] and info["complete_ms"] >= info["submit_ms"]
    _check_launches(b, mode)
    b.close()


def test_persistent_launch_ends_when_idle_and_restarts(scenes):
    """after the queue stays empty for the idle time (1 s) the launch ends; a later submission
    starts a second launch, which continues the same queue"""
    from mitrace_amd.batch import Batch
    b = Batch({"paths_per_iteration": 1 << 16, "tile_pixels": 64})
    b.start()
    jobs = _jobs()[:3]
    ids = [b.submit(scenes[jobs[0][0]], jobs[0][1])]
    t0 = time.perf_counter()
    while not b.info(ids[0])["done"]:
        assert time.perf_counter() - t0 < 60
        time.sleep(0.01)
    time.sleep(2.0)
    for name, st in jobs[1:]:
        ids.append(b.submit(scenes[name], st))
    b.finish()
    for (name, st), j in zip(jobs, ids):
        ref, _, _ = scenes[name].render(st)
        np.testing.assert_array_equal(b.image(j), ref)
    assert b.stats()["launches"] == 2
    b.close()


def test_persistent_launch_aborts_on_its_stall_limit(scenes, monkeypatch):
    """the persistent kernel's barrier wait is bounded: one tile for eight lane groups leaves
    seven groups whose first block polls the queue for tiles while the group's other blocks wait
    at its barrier; held open for 200 ms, they wait beyond a 1 ms limit, set the abort word,
    every block returns (the pollers once the queue closes), and the batch reports the error
    instead of hanging; a batch created afterwards with the normal limit renders correctly"""
    import time
    from mitrace_amd._lib import MiTraceError
    from mitrace_amd.batch import Batch
    monkeypatch.setenv("MT_BATCH_STALL_MS", "1")
    b = Batch({"paths_per_iteration": 1 << 20, "tile_pixels": 16384, "streams": 8})
    monkeypatch.setenv("MT_BATCH_STALL_MS", "20000")
    b.submit(scenes["cornell_gen"], small_settings(width=128, height=128, samples=64, bounces=8, seed=9))
    b.start()
    time.sleep(0.2)
    with pytest.raises(MiTraceError) as e:
        b.finish()
    assert e.value.code == abi.MT_E_HIP and "aborted" in str(e.value)
    b.close()
    name, st = _jobs()[0]
    b = Batch({"paths_per_iteration": 1 << 16})
    j = b.submit(scenes[name], st)
    b.finish()
    ref, _, _ = scenes[name].render(st)
    np.testing.assert_array_equal(b.image(j), ref)
    b.close()


def test_batch_rejects_jobs_beyond_its_limits(scenes):
    import ctypes as C
    from mitrace_amd.batch import Batch
    b = Batch({"paths_per_iteration": 1 << 10})
    st = small_settings(samples=2048)
    g = scenes["cornell_gen"]
    rc = b.L.mt_batch_submit(b.h, g.h, C.byref(g.scene.camera.to_desc()), C.byref(st.to_desc()),
                             None, None)
    assert rc == abi.MT_E_INVALID
    b.finish()  # an empty batch finishes
    b.close()


def test_autotune_picks_from_measured_candidates(scenes):
    from mitrace_amd.batch import autotune
    calib = [(scenes["cornell_gen"], small_settings(width=64, height=64, samples=16, seed=1), None),
             (scenes["sphere_grid_coarse"], small_settings(width=96, height=54, samples=16, seed=2),
              None)]
    best, trials = autotune(calib)
    assert len(trials) >= 8
    assert best["paths_per_iteration"] in (1 << 20, 1 << 21, 1 << 22, 1 << 23, 1 << 24)
    assert best["tile_pixels"] in (256, 1024, 4096, 16384)
    assert best in [t["options"] for t in trials]
    # each trial times its second round (the first, at 1 spp, captures the graphs)
    top = max(t["msamples_per_s"] for t in trials)
    chosen = max(t["msamples_per_s"] for t in trials if t["options"] == best)
    assert chosen == top
