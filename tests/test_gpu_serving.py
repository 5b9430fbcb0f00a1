"""Serving on the GPU: sessions served in a continuously batched mix (PCG
stretch solve on, wind, a moving sphere) give exactly the result of the same
session stepped alone, and that equals the oracle's result for the session
(tolerance 0: batched stepping is bit-identical to single stepping, and the
HIP path reproduces the oracle's FP32 operations); the admission estimate
bounds the device memory a handle really holds."""
from dataclasses import replace

import numpy as np
import pytest

from test_serving import small_requests

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def serving():
    from ssbl_xpbd_amd import serving as s
    return s


def test_mix_equals_alone_on_gpu(serving, scenes, hip, oracle):
    reqs = small_requests(serving, scenes, sides=[40, 100, 24, 64, 32, 80], frames=[8, 5, 12, 6, 10, 4],
                          arrivals=[0, 0, 1, 2, 2, 5])
    est = sorted(serving.estimate_device_bytes(r.cloth, r.options) for r in reqs)
    budget = sum(est[-3:])        # the largest sessions cannot all be resident at once
    srv = serving.ClothServer(hbm_budget_bytes=budget, substeps=10, iterations=8, lib=hip)
    rep = srv.run(reqs, clock="tick")
    print("served:", rep.summary())
    assert rep.unfinished == 0 and len(rep.results) == len(reqs)
    assert rep.peak_device_bytes <= budget
    assert rep.peak_sessions >= 3                   # sessions overlapped in one batch
    for r in reqs:
        assert r.options.stretch_optimization_enabled      # PCG on, as served sessions run
        alone = serving.run_alone(r, 10, 8, lib=hip)
        ref = serving.run_alone(r, 10, 8, lib=oracle)
        got = rep.results[r.sid].positions
        assert np.isfinite(got).all()
        diff = float(np.max(np.abs(got - alone)))
        dref = float(np.max(np.abs(alone - ref)))
        print(f"session {r.sid}: {r.vertex_count} vertices, {r.frames} frames, max |mix - alone| = {diff:.3e}, "
              f"max |alone - oracle| = {dref:.3e}")
        assert diff == 0.0
        assert dref == 0.0


def test_device_bytes_within_estimate(serving, scenes, hip, nb):
    cloth = serving.session_cloth(60, seed=1)
    base = scenes.SolverOptions(use_ground=True, collision_margin=0.005, cloth_thickness=0.005, static_sdf_max_resolution=64)
    variants = {
        "plain": base,
        "self": replace(base, self_collision=True, self_collision_distance=0.005),
        "pcg": replace(base, stretch_optimization_enabled=True, stretch_optimization_strength=1.0),
    }
    box = scenes.box_triangles((0.0, 0.0, 0.3), 0.3)
    for name, opts in variants.items():
        s = nb.NativeXpbdSolver(cloth, opts, box, lib=hip)
        held = int(hip.ssbl_ext_device_bytes(s.handle))
        est = serving.estimate_device_bytes(cloth, opts)
        s.close()
        print(f"{name}: held {held} B, estimate {est} B ({est / max(held, 1):.2f}x)")
        assert 0 < held <= est
