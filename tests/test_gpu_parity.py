"""GPU parity: the HIP library against the CPU oracle, through the C ABI.

Tolerance (stated by the project's accuracy target): positions within 1e-4
relative or 1e-5 m.  The vertex-local path is built so that it reproduces the
oracle's FP32 operation order; the printed max difference shows how close it
is (0.0 means bit-exact).
"""
from dataclasses import replace

import numpy as np
import pytest

from parity_util import make_pair, rel_close, run_and_compare, state_diff

pytestmark = pytest.mark.gpu
TOL = 1.0e-5


def test_config1_plain_120_frames(nb, hip, oracle, scenes):
    sc = scenes.config1()
    g, r = make_pair(nb, hip, oracle, sc.cloths[0], sc.options[0], sc.static_triangles)
    worst = run_and_compare(g, r, sc.frames, sc.substeps, sc.iterations, every=10)
    print(f"config1 32x32 120 frames: max |hip - oracle| = {worst:.3e}")
    assert worst <= TOL
    pg, pr = g.download_positions(), r.download_positions()
    assert rel_close(pg, pr)
    dg, dr = g.diagnostics(), r.diagnostics()
    assert dg.finite_flag == 1


def test_config1_diagnostic_counts_match(nb, hip, oracle, scenes):
    sc = scenes.config1(frames=3)
    g, r = make_pair(nb, hip, oracle, sc.cloths[0], sc.options[0], sc.static_triangles)
    run_and_compare(g, r, 3, sc.substeps, sc.iterations, diagnostics=True)
    dg, dr = g.diagnostics(), r.diagnostics()
    for name in ("abi41_bending_wing_count", "abi41_bending_guarded", "abi41_hard_projection_fallbacks",
                 "external_friction_corrections"):
        assert getattr(dg, name) == getattr(dr, name), name


def test_soft_pins_sphere_and_wall(nb, hip, oracle, scenes):
    cloth = scenes.grid_cloth(24, 1.0, 1.2, seed=3)
    cloth.pin_weights = np.array([0.5, 0.3], np.float32)
    opts = scenes.SolverOptions(use_ground=True, use_sphere=True, sphere_center=(0.1, -0.1, 0.6), sphere_radius=0.3,
                                use_wall=True, wall_origin=(0.45, 0.0, 0.0), wall_normal=(-1.0, 0.0, 0.0),
                                contact_friction=0.4, contact_tangent_damping=0.2, damping=0.99)
    g, r = make_pair(nb, hip, oracle, cloth, opts, None)
    worst = run_and_compare(g, r, 40, 8, 6, every=5)
    print(f"pins/sphere/wall: max |hip - oracle| = {worst:.3e}")
    assert worst <= TOL


def test_static_sdf_grid_and_contacts(nb, hip, oracle, scenes):
    cloth = scenes.grid_cloth(20, 0.8, 0.9, pin_corners=False, seed=5)
    box = scenes.box_triangles((0.0, 0.0, 0.3), 0.3)
    opts = scenes.SolverOptions(use_ground=True, static_sdf_max_resolution=48, collision_margin=0.01)
    g, r = make_pair(nb, hip, oracle, cloth, opts, box)
    sg, og = g.static_sdf()
    sr, orr = r.static_sdf()
    assert sg.shape == sr.shape and np.array_equal(og, orr)
    print(f"sdf {sg.shape}: max |hip - oracle| = {np.max(np.abs(sg - sr)):.3e}")
    assert np.max(np.abs(sg - sr)) <= 1.0e-6
    worst = run_and_compare(g, r, 60, 6, 4, diagnostics=True, every=10)
    print(f"sdf contacts: max |hip - oracle| = {worst:.3e}")
    assert worst <= TOL
    assert g.diagnostics().static_sdf_contact_count == r.diagnostics().static_sdf_contact_count


def test_pressure_and_wind(nb, hip, oracle, scenes):
    pos, tris = scenes.icosphere(2, 0.3)
    pos[:, 2] += 0.6
    cloth = scenes.build_cloth(pos, tris, np.zeros(len(pos), bool), jitter=1.0e-4, seed=9)
    opts = scenes.SolverOptions(use_ground=True, use_volume_pressure=True, pressure_strength=1.2, damping=0.98)
    g, r = make_pair(nb, hip, oracle, cloth, opts, None)
    wind = dict(type=1, strength=8.0, direction=(1.0, 0.2, 0.0), axis=(0.0, 0.0, 1.0), apply_to_location=1, noise=0.3, seed=4)
    g.update_force_fields([wind])
    r.update_force_fields([wind])
    worst = run_and_compare(g, r, 30, 5, 5, every=5)
    print(f"pressure+wind: max |hip - oracle| = {worst:.3e}")
    assert worst <= TOL


def test_all_force_field_types(nb, hip, oracle, scenes):
    cloth = scenes.grid_cloth(20, 0.8, 1.0, seed=13)
    opts = scenes.SolverOptions(use_ground=True)
    g, r = make_pair(nb, hip, oracle, cloth, opts, None)
    base = dict(apply_to_location=1, axis=(0.0, 0.0, 1.0))
    fields = [
        dict(base, type=2, strength=-3.0, origin=(0.0, 0.0, 0.2), use_max_distance=1, distance_max=2.0, use_min_distance=1,
             distance_min=0.1, falloff_power=1.5),
        dict(base, type=3, strength=2.0, origin=(0.1, 0.0, 0.0), flow=0.5),
        dict(base, type=4, strength=1.5, size=0.3, origin=(0.0, 0.2, 0.0)),
        dict(base, type=5, strength=0.5, origin=(0.3, 0.3, 0.5)),
        dict(base, type=6, strength=1.0, rest_length=0.5, harmonic_damping=0.2, origin=(0.0, 0.0, 1.0)),
        dict(base, type=7, strength=0.2, size=0.2, origin=(0.0, 0.0, 0.5)),
        dict(base, type=8, strength=0.7, direction=(0.0, 1.0, 0.0)),
        dict(base, type=9, strength=0.3, linear_drag=0.2, quadratic_drag=0.1),
        dict(base, type=10, strength=0.8, noise=1.5, seed=7, use_radial_max=1, radial_max=3.0, use_2d_force=1),
    ]
    g.update_force_fields(fields)
    r.update_force_fields(fields)
    worst = run_and_compare(g, r, 20, 4, 4, every=5)
    print(f"force fields: max |hip - oracle| = {worst:.3e}")
    assert worst <= TOL


def test_lra_tethers(nb, hip, oracle, scenes):
    cloth = scenes.grid_cloth(16, 0.6, 1.0, seed=11)
    pins = cloth.pin_indices
    dyn = np.setdiff1d(np.arange(len(cloth.positions_world)), pins)
    anchor = pins[np.argmin(np.linalg.norm(cloth.positions_world[pins][None, :, :] - cloth.positions_world[dyn][:, None, :], axis=2), axis=1)]
    cloth.lra_edges = np.stack([anchor, dyn], 1).astype(np.int32)
    cloth.lra_rest_lengths = (np.linalg.norm(cloth.positions_world[dyn] - cloth.positions_world[anchor], axis=1) * 1.02).astype(np.float32)
    opts = scenes.SolverOptions(use_ground=True, lra_compliance=1.0e-6)
    g, r = make_pair(nb, hip, oracle, cloth, opts, None)
    worst = run_and_compare(g, r, 20, 5, 4, diagnostics=True, every=5)
    print(f"lra: max |hip - oracle| = {worst:.3e}")
    assert worst <= TOL
    assert g.diagnostics().abi41_lra_tack_count == r.diagnostics().abi41_lra_tack_count


def test_batch_equals_single(nb, hip, oracle, scenes):
    sc = scenes.config3(seed=2, count=6, frames=4)
    sc.static_triangles = np.empty((0, 9), np.float32)
    singles = [nb.NativeXpbdSolver(c, o, None, lib=hip) for c, o in zip(sc.cloths, sc.options)]
    members = [nb.NativeXpbdSolver(c, o, None, lib=hip) for c, o in zip(sc.cloths, sc.options)]
    batch = nb.SolverBatch(members)
    for _ in range(sc.frames):
        for s in singles:
            s.step(sc.substeps, sc.iterations, diagnostics=False, synchronize=False)
        batch.step(sc.substeps, sc.iterations)
    nb.synchronize(hip)
    for s, m in zip(singles, members):
        assert np.array_equal(s.download_positions(), m.download_positions())
    ref = nb.NativeXpbdSolver(sc.cloths[0], sc.options[0], None, lib=oracle)
    for _ in range(sc.frames):
        ref.step(sc.substeps, sc.iterations)
    assert state_diff(members[0], ref) <= TOL


def test_batch_set_members_equals_fresh_batch(nb, hip, scenes):
    """A batch pointed at new members (a serving loop's changing set of
    stepped sessions) steps them exactly as a batch created for them and as
    each handle alone: first cloths 0-3, then 2-5 (two new, one PCG), then 0 and 5."""
    sc = scenes.config3(seed=4, count=6, frames=3, cross_cloth=False)
    sc.options[3] = replace(sc.options[3], stretch_optimization_enabled=True, stretch_optimization_strength=1.0)
    sc.options[5] = replace(sc.options[5], stretch_optimization_enabled=True, stretch_optimization_strength=0.4)
    make = lambda: [nb.NativeXpbdSolver(c, o, sc.static_triangles, lib=hip) for c, o in zip(sc.cloths, sc.options)]  # noqa: E731
    a, b, c = make(), make(), make()
    moving = nb.SolverBatch([a[k] for k in (0, 1, 2, 3)])
    for members in ((0, 1, 2, 3), (2, 3, 4, 5), (0, 5)):
        moving.set_members([a[k] for k in members])
        fresh = nb.SolverBatch([b[k] for k in members])
        for _ in range(2):
            moving.step(sc.substeps, sc.iterations)
            fresh.step(sc.substeps, sc.iterations)
            for k in members:
                c[k].step(sc.substeps, sc.iterations, diagnostics=False, synchronize=False)
        fresh.close()
        nb.synchronize(hip)
        for k in members:
            pa, pb, pc = a[k].download_positions(), b[k].download_positions(), c[k].download_positions()
            assert np.array_equal(pa, pb) and np.array_equal(pb, pc), (members, k)
    moving.close()


def test_batch_pass_timing_needs_a_step_and_keeps_the_state(nb, hip, scenes):
    """The pass timing refuses a batch that has not been stepped since it was
    built or re-targeted (its device view table is filled by a step); after
    a step it leaves every member's state as it was."""
    sc = scenes.config3(seed=1, count=4, frames=2, cross_cloth=False)
    members = [nb.NativeXpbdSolver(c, o, sc.static_triangles, lib=hip) for c, o in zip(sc.cloths, sc.options)]
    batch = nb.SolverBatch(members)
    with pytest.raises(nb.NativeSolverError, match="stepping it once"):
        batch.time_pass(8)
    batch.step(sc.substeps, sc.iterations)
    nb.synchronize(hip)
    before = [m.download_state() for m in members]
    assert batch.time_pass(8) > 0.0
    for m, b in zip(members, before):
        for x, y in zip(m.download_state(), b):
            assert np.array_equal(x, y)
    batch.set_members(members[:2])
    with pytest.raises(nb.NativeSolverError, match="stepping it once"):
        batch.time_pass(8)
    batch.close()
