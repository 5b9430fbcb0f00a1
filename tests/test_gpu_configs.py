"""GPU parity on the benchmark workloads themselves, against the CPU oracle.

* The 64-cloth uneven batch (10 substeps x 8 iterations, ground, the 128^3
  box SDF built from the box triangles, cross-cloth collision over four
  layers gathered on the device every frame): the cloths of stacks 0, 1
  and 10 (12 cloths, 77k vertices; stack 10 stands over the box and lands
  on it) for 32 frames, past the layers' first contact (about frame 8) and
  the box contact (about frame 18).  The batched HIP form, the per-handle
  HIP form and the oracle are compared bit for bit: positions every 8
  frames, the dynamic-contact and SDF counters (the batch's read back with
  its step) every frame.  Then all 64
  cloths for 12 frames: batched equals per-handle bit for bit, all finite.
* The 10k-vertex strict self-collision cloth (100x100, 20 substeps x 5
  iterations, sphere, box SDF): frames where the oracle stays within a few
  seconds each; self-collision hash cells, candidate tables and positions
  bit-exact after every frame.

Tolerance: 0 (the HIP path reproduces the oracle's FP32 operations).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STACKS = (0, 1, 10)
DYN_COUNTERS = ("fast_triangle_pair_contacts", "dynamic_particle_contacts", "dynamic_triangle_candidate_count",
                "dynamic_particle_candidate_count", "dynamic_triangle_aabb_reject_count", "abi41_exact_impulse_contact_count",
                "static_sdf_contact_count", "global_dynamic_hash_overflow")


def _stack_cloths(scenes, stacks):
    sc = scenes.config3(seed=0)
    picked = [4 * s + k for s in stacks for k in range(4)]
    return sc, picked


def _make(nb, lib, sc, picked):
    from ssbl_xpbd_amd import cross_cloth
    solvers = [nb.NativeXpbdSolver(sc.cloths[k], sc.options[k], sc.static_triangles, lib=lib) for k in picked]
    cross = cross_cloth.CrossClothScene(solvers, [sc.layers[k] for k in picked], contact_distance=sc.contact_distance, lib=lib)
    return solvers, cross


def test_config3_stacks_vs_oracle(nb, hip, oracle, scenes):
    sc, picked = _stack_cloths(scenes, STACKS)
    frames = 32
    bat, cross_b = _make(nb, hip, sc, picked)
    one, cross_o = _make(nb, hip, sc, picked)
    ref, cross_r = _make(nb, oracle, sc, picked)
    batch = nb.SolverBatch(bat)
    totals = dict.fromkeys(DYN_COUNTERS, 0)
    print(f"config 3 stacks {STACKS}: {len(picked)} cloths, {sum(s.vertex_count for s in ref)} vertices", flush=True)
    for f in range(frames):
        cross_b.refresh(device=True)
        batch.step(sc.substeps, sc.iterations, diagnostics=True)     # counters of every member read back
        cross_o.refresh(device=True)
        cross_r.refresh(device=True)
        for b, g, r in zip(bat, one, ref):
            g.step(sc.substeps, sc.iterations, diagnostics=True, synchronize=False)
            r.step(sc.substeps, sc.iterations, diagnostics=True, synchronize=True)
            db, dg, dr = b.diagnostics(), g.diagnostics(), r.diagnostics()
            for k in DYN_COUNTERS:
                assert getattr(dg, k) == getattr(dr, k), (f, k, getattr(dg, k), getattr(dr, k))
                assert getattr(db, k) == getattr(dr, k), ("batched", f, k, getattr(db, k), getattr(dr, k))
                totals[k] += getattr(dr, k)
        if (f + 1) % 8 == 0:
            nb.synchronize(hip)
            for c, (b, g, r) in enumerate(zip(bat, one, ref)):
                pb, pg, pr = b.download_state(), g.download_state(), r.download_state()
                for x, y, z in zip(pb, pg, pr):
                    assert np.array_equal(y, z), ("per-handle HIP vs oracle", f, picked[c])
                    assert np.array_equal(x, y), ("batched vs per-handle", f, picked[c])
            print(f"frame {f + 1}: bit-exact; counters so far {totals}", flush=True)
    assert totals["fast_triangle_pair_contacts"] + totals["dynamic_particle_contacts"] > 0
    assert totals["static_sdf_contact_count"] > 0
    batch.close()


def test_config3_all_64_batched_equals_per_handle(nb, hip, scenes):
    sc = scenes.config3(seed=0)
    picked = list(range(len(sc.cloths)))
    bat, cross_b = _make(nb, hip, sc, picked)
    one, cross_o = _make(nb, hip, sc, picked)
    batch = nb.SolverBatch(bat)
    contacts = 0
    for f in range(12):
        cross_b.refresh(device=True)
        batch.step(sc.substeps, sc.iterations)
        cross_o.refresh(device=True)
        for g in one:
            g.step(sc.substeps, sc.iterations, diagnostics=f == 11, synchronize=False)
    nb.synchronize(hip)
    for c, (b, g) in enumerate(zip(bat, one)):
        for x, y in zip(b.download_state(), g.download_state()):
            assert np.isfinite(x).all(), c
            assert np.array_equal(x, y), ("batched vs per-handle", c)
        d = g.diagnostics()
        contacts += d.fast_triangle_pair_contacts + d.dynamic_particle_contacts
    print(f"64 cloths, 12 frames: batched == per-handle, {contacts} cross-cloth contacts in the last frame")
    assert contacts > 0
    batch.close()


def test_config2_full_size_vs_oracle(nb, hip, oracle, scenes):
    sc = scenes.config2(seed=0)
    frames = 3
    g = nb.NativeXpbdSolver(sc.cloths[0], sc.options[0], sc.static_triangles, lib=hip)
    r = nb.NativeXpbdSolver(sc.cloths[0], sc.options[0], sc.static_triangles, lib=oracle)
    sg, og = g.static_sdf()
    sr, orr = r.static_sdf()
    assert sg.shape == sr.shape and max(sg.shape) == 128
    assert np.array_equal(og, orr)
    print(f"config 2 SDF {sg.shape}: max |hip - oracle| = {np.max(np.abs(sg - sr)):.3e}", flush=True)
    assert np.max(np.abs(sg - sr)) <= 1.0e-6
    for f in range(frames):
        g.step(sc.substeps, sc.iterations, diagnostics=True, synchronize=True)
        r.step(sc.substeps, sc.iterations, diagnostics=True, synchronize=True)
        assert np.array_equal(g.self_hash_cells(), r.self_hash_cells()), f
        cg, ig = g.self_neighbors()
        cr, ir = r.self_neighbors()
        assert np.array_equal(cg, cr), f
        for v in np.flatnonzero(cr):
            assert np.array_equal(ig[v, :cr[v]], ir[v, :cr[v]]), (f, v)
        for x, y in zip(g.download_state(), r.download_state()):
            assert np.array_equal(x, y), f
        dg, dr = g.diagnostics(), r.diagnostics()
        for k in ("self_candidate_count", "abi41_soft_contact_count", "abi41_exact_impulse_contact_count",
                  "abi41_edge_edge_contact_count", "static_sdf_contact_count"):
            assert getattr(dg, k) == getattr(dr, k), (f, k)
        print(f"config 2 frame {f + 1}: bit-exact, {int(cr.sum())} candidate pairs", flush=True)
