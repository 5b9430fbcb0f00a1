"""A batch stepped as two concurrent parts (two streams, the default for
batches of 64 tiles or more) equals the same batch stepped as one table, bit
for bit, with the diagnostic counters of every member equal too: the
64-cloth workload's stacks 0-5 (24 cloths, four collision layers, the 128^3
box SDF, the cross-cloth scene gathered on the device every frame), with
graph replay on and off.  Tolerance 0."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

COUNTERS = ("fast_triangle_pair_contacts", "dynamic_particle_contacts", "dynamic_triangle_candidate_count",
            "dynamic_particle_candidate_count", "dynamic_triangle_aabb_reject_count", "static_sdf_contact_count",
            "abi41_exact_impulse_contact_count")


@pytest.mark.parametrize("graphs", [True, False])
def test_two_part_batch_equals_one_table(nb, hip, scenes, graphs):
    from test_gpu_configs import _make
    sc = scenes.config3(seed=0)
    picked = list(range(24))
    assert sum((len(sc.cloths[k].positions_world) + 255) // 256 for k in picked) >= 64   # large enough to split
    previous_graphs = nb.set_graph_mode(hip, graphs)
    try:
        split, cross_s = _make(nb, hip, sc, picked)
        one, cross_o = _make(nb, hip, sc, picked)
        nb.set_batch_parts(hip, 2)
        batch_split = nb.SolverBatch(split)
        nb.set_batch_parts(hip, 1)
        batch_one = nb.SolverBatch(one)
        nb.set_batch_parts(hip, 2)
        for f in range(20):
            cross_s.refresh(device=True)
            batch_split.step(sc.substeps, sc.iterations, diagnostics=True)
            cross_o.refresh(device=True)
            batch_one.step(sc.substeps, sc.iterations, diagnostics=True)
            for k, (a, b) in enumerate(zip(split, one)):
                da, db = a.diagnostics(), b.diagnostics()
                for name in COUNTERS:
                    assert getattr(da, name) == getattr(db, name), (f, picked[k], name)
            if (f + 1) % 5 == 0:
                for k, (a, b) in enumerate(zip(split, one)):
                    for x, y in zip(a.download_state(), b.download_state()):
                        assert np.array_equal(x, y), ("two parts vs one table", f, picked[k])
        print(f"24 cloths, 20 frames, graphs {graphs}: two concurrent parts equal one table", flush=True)
        batch_split.close()
        batch_one.close()
    finally:
        nb.set_graph_mode(hip, previous_graphs)
        nb.set_batch_parts(hip, 2)
