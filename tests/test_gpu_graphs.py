"""Graph replay of the substep loop: a step replayed from its captured HIP
graph must equal the same step launched directly, bit for bit, on every
replay, with the program's own eager work (parameter uploads, scene gathers
and hashes, downloads, synchronisations) between the replays."""
import numpy as np
import pytest

import scenarios_dynamic as sd

pytestmark = pytest.mark.gpu


@pytest.fixture
def graph_mode(nb, hip):
    # launch tuning keeps its timed steps out of graphs; these tests count
    # captures and replays without it
    prev, prev_tune = nb.set_graph_mode(hip, True), nb.set_autotune(hip, False)
    yield lambda on: nb.set_graph_mode(hip, on)
    nb.set_graph_mode(hip, prev)
    nb.set_autotune(hip, prev_tune)


def expected_captures(keys):
    """A handle or batch keeps the graphs of up to four launch sequences: a
    step captures when its key is new."""
    assert len(set(keys)) <= 4
    return len(set(keys))


def test_solo_replays_equal_direct_launches(nb, hip, scenes, graph_mode):
    sc = scenes.config1(frames=12)
    direct = nb.NativeXpbdSolver(sc.cloths[0], sc.options[0], sc.static_triangles, lib=hip)
    graphed = nb.NativeXpbdSolver(sc.cloths[0], sc.options[0], sc.static_triangles, lib=hip)
    # a change of iterations and the diagnostics steps change the launch
    # sequence, so they are captured anew
    schedule = [(sc.substeps, sc.iterations, f == 3) for f in range(5)] + [(4, 3, False)] * 3
    schedule += [(sc.substeps, sc.iterations, f == 2) for f in range(4)]
    for f, (sub, it, diag) in enumerate(schedule):
        graph_mode(False)
        direct.step(sub, it, diagnostics=diag, synchronize=True)
        graph_mode(True)
        graphed.step(sub, it, diagnostics=diag, synchronize=True)
        assert np.array_equal(direct.download_positions(), graphed.download_positions()), f
        if diag:
            dd, dg = direct.diagnostics(), graphed.diagnostics()
            for name in ("abi41_soft_contact_count", "abi41_exact_impulse_contact_count", "finite_flag"):
                assert getattr(dd, name) == getattr(dg, name), (f, name)
    assert nb.graph_stats(hip, handle=direct.handle) == (0, 0)
    captures = expected_captures(schedule)
    assert nb.graph_stats(hip, handle=graphed.handle) == (captures, len(schedule) - captures)


def run_two_cloth_batch(nb, hip, scenes, graphed, frames):
    a, b, scene = sd.two_cloth_scene(nb, hip, scenes, nb.GlobalDynamicScene.CROSS_LOWER_LAYERS)
    batch = nb.SolverBatch([a, b])
    nb.set_graph_mode(hip, graphed)
    out = []
    for _ in range(frames):
        sd.refresh_scene(scene, [a, b], device_gather=True)
        batch.step(6, 4, synchronize=True)
        out.append((a.download_positions(), b.download_positions()))
    stats = batch.graph_stats()
    batch.close()
    return out, stats


def test_batch_with_global_scene_replays_equal_direct(nb, hip, scenes, graph_mode):
    """The batched frame of the cross-cloth workload: scene gather and hashes
    run eagerly before each replay of the substep graph (dynamic contact
    passes with the fused closing steps and next integration)."""
    frames = 16
    direct, sd_ = run_two_cloth_batch(nb, hip, scenes, False, frames)
    graphed, sg = run_two_cloth_batch(nb, hip, scenes, True, frames)
    assert sd_ == (0, 0)
    assert sg == (1, frames - 1)
    for f, ((ad, bd), (ag, bg)) in enumerate(zip(direct, graphed)):
        assert np.array_equal(ad, ag) and np.array_equal(bd, bg), f
    # the cloths met: B rests on A, so the contact passes did work
    assert not np.array_equal(graphed[0][1], graphed[-1][1])


def test_batch_of_mixed_cloths_replays_equal_direct(nb, hip, scenes, graph_mode):
    sc = scenes.config3(seed=2, count=6, frames=6)
    sc.static_triangles = np.empty((0, 9), np.float32)
    runs = []
    for graphed in (False, True):
        graph_mode(graphed)
        members = [nb.NativeXpbdSolver(c, o, sc.static_triangles, lib=hip) for c, o in zip(sc.cloths, sc.options)]
        batch = nb.SolverBatch(members)
        frames = []
        for _ in range(sc.frames):
            batch.step(sc.substeps, sc.iterations, synchronize=True)
            frames.append([m.download_positions() for m in members])
        runs.append((frames, batch.graph_stats()))
        batch.close()
    (direct, sd_), (graphed, sg) = runs
    assert sd_ == (0, 0) and sg == (1, sc.frames - 1)
    for f, (xd, xg) in enumerate(zip(direct, graphed)):
        for k, (pd, pg) in enumerate(zip(xd, xg)):
            assert np.array_equal(pd, pg), (f, k)


def _replay_vs_direct(nb, hip, graph_mode, make, frames, step):
    """Builds the handles twice, runs `frames` frames with graph replay off
    and on, and returns the per-frame positions and the graph statistics."""
    runs = []
    for graphed in (False, True):
        graph_mode(graphed)
        solvers, owner = make()
        out = []
        for _ in range(frames):
            step(solvers, owner)
            nb.synchronize(hip)
            out.append([s.download_positions() for s in solvers])
        stats = owner.graph_stats() if isinstance(owner, nb.SolverBatch) else nb.graph_stats(hip, handle=owner.handle)
        runs.append((out, stats))
        if isinstance(owner, nb.SolverBatch):
            owner.close()
    return runs


@pytest.mark.parametrize("mode", [1, 2])
def test_self_collision_steps_replay_equal_direct(nb, hip, scenes, graph_mode, mode):
    """Strict (2) and fast (1) self collision: hashes, candidate tables and
    responses captured with the substep loop."""
    def make():
        cloth = scenes.grid_cloth(24, 0.5, 0.15, pin_corners=False, seed=5)
        opts = scenes.SolverOptions(use_ground=True, self_collision=True, self_collision_mode=mode, cloth_thickness=0.005,
                                    self_collision_distance=0.006)
        s = nb.NativeXpbdSolver(cloth, opts, None, lib=hip)
        return [s], s
    frames = 10
    (direct, sd_), (graphed, sg) = _replay_vs_direct(nb, hip, graph_mode, make, frames,
                                                     lambda ss, o: o.step(4, 4, diagnostics=False, synchronize=False))
    assert sd_ == (0, 0) and sg == (1, frames - 1)
    for f, (a, b) in enumerate(zip(direct, graphed)):
        assert np.array_equal(a[0], b[0]), f


def pcg_batch_cloths(scenes):
    """Hard-stretch cloths of different sizes: strength 1.0 (polish and caps)
    coloured and uncoloured, strength 0.4, and one cloth on the direct
    stretch, on a ground."""
    out = []
    for side, strength, colored, seed in ((20, 1.0, True, 1), (12, 1.0, False, 2), (28, 0.4, True, 3), (16, 0.0, True, 4)):
        cloth = scenes.grid_cloth(side, 0.03 * (side - 1), 0.6, seed=seed)
        if not colored:
            cloth.edge_color_offsets = np.empty(0, np.int32)
        opts = scenes.SolverOptions(use_ground=True, stretch_optimization_enabled=strength > 0.0,
                                    stretch_optimization_strength=strength, cloth_thickness=0.005)
        out.append((cloth, opts))
    return out


def test_pcg_batch_replays_equal_direct(nb, hip, scenes, graph_mode):
    def make():
        solvers = [nb.NativeXpbdSolver(c, o, None, lib=hip) for c, o in pcg_batch_cloths(scenes)]
        return solvers, nb.SolverBatch(solvers)
    frames = 6
    (direct, sd_), (graphed, sg) = _replay_vs_direct(nb, hip, graph_mode, make, frames, lambda ss, b: b.step(5, 3))
    assert sd_ == (0, 0) and sg == (1, frames - 1)
    for f, (xd, xg) in enumerate(zip(direct, graphed)):
        for k, (pd, pg) in enumerate(zip(xd, xg)):
            assert np.array_equal(pd, pg), (f, k)


def test_many_handles_first_captures_back_to_back(nb, hip, scenes, graph_mode):
    """The per-handle loop of the 64-cloth workload (SDF, ground, pins): every
    handle captures its first graph while the previous handles' graphs still
    run on the library stream (no synchronisation between steps)."""
    sc = scenes.config3(seed=0, cross_cloth=False)
    def make():
        solvers = [nb.NativeXpbdSolver(c, o, sc.static_triangles, lib=hip) for c, o in zip(sc.cloths, sc.options)]
        return solvers, solvers[-1]
    def step(solvers, _):
        for s in solvers:
            s.step(sc.substeps, sc.iterations, diagnostics=False, synchronize=False)
    frames = 3
    (direct, _), (graphed, sg) = _replay_vs_direct(nb, hip, graph_mode, make, frames, step)
    assert sg == (1, frames - 1)
    for f, (xd, xg) in enumerate(zip(direct, graphed)):
        for k, (pd, pg) in enumerate(zip(xd, xg)):
            assert np.array_equal(pd, pg), (f, k)
