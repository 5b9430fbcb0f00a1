"""Launch tuning of the particle contact pass: while a handle or batch tunes,
its substeps run different pooling limits; every limit must give the same
contacts, so a tuned run equals a run at the fixed default, bit for bit."""
import numpy as np
import pytest

import scenarios_dynamic as sd

pytestmark = pytest.mark.gpu

CANDIDATES = {0, 8, 16, 32, 64}
EVERY = 16   # the first step with dynamic contacts and every 16th after it are timed


@pytest.fixture
def modes(nb, hip):
    prev_graph, prev_tune = nb.set_graph_mode(hip, True), nb.set_autotune(hip, True)
    yield
    nb.set_graph_mode(hip, prev_graph)
    nb.set_autotune(hip, prev_tune)


def run_batch(nb, hip, scenes, tune, frames):
    nb.set_autotune(hip, tune)
    a, b, scene = sd.two_cloth_scene(nb, hip, scenes, nb.GlobalDynamicScene.CROSS_LOWER_LAYERS)
    batch = nb.SolverBatch([a, b])
    out, limits = [], []
    for _ in range(frames):
        sd.refresh_scene(scene, [a, b], device_gather=True)
        batch.step(6, 4, synchronize=True)
        out.append((a.download_positions(), b.download_positions()))
        limits.append(batch.tuned_pool_limit())
    stats = batch.graph_stats()
    batch.close()
    return out, limits, stats


def test_tuned_batch_equals_fixed_limit(nb, hip, scenes, modes):
    frames = 20
    fixed, fixed_limits, fixed_stats = run_batch(nb, hip, scenes, False, frames)
    tuned, limits, stats = run_batch(nb, hip, scenes, True, frames)
    for f, ((af, bf), (at, bt)) in enumerate(zip(fixed, tuned)):
        assert np.array_equal(af, at) and np.array_equal(bf, bt), f
    assert fixed_limits == [-1] * frames and fixed_stats == (1, frames - 1)
    # a timed step's choice is read at the start of the next step
    assert limits[0] == -1 and all(v in CANDIDATES for v in limits[1:])
    assert len(set(limits[1:EVERY + 1])) == 1 and len(set(limits[EVERY + 1:])) == 1
    # the timed steps (0 and 16) launch directly; the others replay a graph,
    # captured again when the second timed step changed the limit (read at
    # the start of step 17)
    captures, replays = stats
    assert captures == (2 if limits[EVERY + 1] != limits[1] else 1)
    assert captures + replays == frames - 2


def test_tuned_single_handles_equal_fixed_limit(nb, hip, scenes, modes):
    runs = []
    for tune in (False, True):
        nb.set_autotune(hip, tune)
        a, b, scene = sd.two_cloth_scene(nb, hip, scenes, nb.GlobalDynamicScene.CROSS_LOWER_LAYERS)
        frames = []
        for _ in range(6):
            sd.refresh_scene(scene, [a, b], device_gather=True)
            a.step(6, 4, diagnostics=True)
            b.step(6, 4, diagnostics=True)
            d = b.diagnostics()
            frames.append((a.download_positions(), b.download_positions(), d.dynamic_particle_contacts))
        runs.append((frames, [hip.ssbl_ext_tuned_pool_limit(s.handle, None) for s in (a, b)]))
    (fixed, lf), (tuned, lt) = runs
    for f, (x, y) in enumerate(zip(fixed, tuned)):
        assert np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1]) and x[2] == y[2], f
    assert lf == [-1, -1] and all(v in CANDIDATES for v in lt)
