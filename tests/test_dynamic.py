"""Dynamic colliders in the oracle: closed forms for one triangle contact and
one particle contact (both on the brute-force and the hashed query path),
hash queries equal to brute-force scans, and the global scene's source and
layer filter (A41:3984-3997)."""
import ctypes

import numpy as np
import pytest

import scenarios_dynamic as sd


@pytest.mark.parametrize("quads", [1, 8])          # 2 triangles: scan; 128: hashed
def test_plane_contact_closed_form(oracle, quads):
    (p, q, v), d = sd.plane_contact(oracle, quads)
    assert np.allclose(p[:, 2], 0.5 * (0.01 - 0.004), atol=1.0e-8)
    assert np.allclose(p[:, :2], [[-0.2, -0.2], [0.2, -0.2], [0.2, 0.2], [-0.2, 0.2]], atol=0)
    assert np.all(v == 0.0)
    assert d.fast_triangle_pair_contacts == 4 and d.abi41_exact_impulse_contact_count == 4
    assert d.dynamic_triangle_count == 2 * quads * quads
    assert d.dynamic_triangle_candidate_count >= 4


@pytest.mark.parametrize("left_first", [True, False])
@pytest.mark.parametrize("count", [4, 64])
def test_valley_tie_goes_to_first_triangle(oracle, count, left_first):
    """Equal depths in two mirrored faces: the first triangle of the set wins."""
    (p, _, _), d = sd.valley_tie(oracle, count, left_first)
    sign = 1.0 if left_first else -1.0
    assert np.all(np.sign(p[:, 0]) == sign), p[:, 0]
    # half the depth m - h / sqrt 2 along (1, 0, 1) / sqrt 2 (no gravity here)
    push = 0.5 * (0.01 - 0.008 / np.sqrt(2.0)) / np.sqrt(2.0)
    assert np.allclose(p[:, 0], sign * push, atol=1.0e-6)
    assert d.fast_triangle_pair_contacts == count


@pytest.mark.parametrize("far", [0, 200])         # 1 particle: scan; 201: hashed
def test_particle_contact_closed_form(oracle, far):
    (p, q, v), d = sd.particle_contact(oracle, far_particles=far)
    expected = np.float32(0.5) * (np.float32(0.008) + np.float32(0.005) - np.float32(0.01))
    assert abs(float(p[0, 2]) - float(expected)) < 1.0e-8
    assert d.dynamic_particle_contacts == 1 and d.dynamic_particle_candidate_count == 1


COUNTERS = ("fast_triangle_pair_contacts", "dynamic_particle_contacts", "dynamic_triangle_candidate_count",
            "dynamic_particle_candidate_count", "dynamic_triangle_aabb_reject_count", "abi41_exact_impulse_contact_count")


def run_cloth(nb, lib, scenes, frames=30, **kw):
    """Positions per frame and the dynamic counters summed over the frames."""
    s = sd.cloth_over_colliders(nb, lib, scenes, **kw)
    out, tot = [], dict.fromkeys(COUNTERS, 0)
    for _ in range(frames):
        s.step(6, 4, diagnostics=True, synchronize=True)
        out.append(s.download_positions())
        d = s.diagnostics()
        for k in COUNTERS:
            tot[k] += getattr(d, k)
    s.close()
    return out, tot


def test_hash_queries_equal_brute_force(nb, oracle, scenes):
    hashed, dh = run_cloth(nb, oracle, scenes)
    oracle.ssbl_oracle_dynamic_hash.argtypes = [ctypes.c_int]
    oracle.ssbl_oracle_dynamic_hash(0)
    try:
        scanned, ds = run_cloth(nb, oracle, scenes)
    finally:
        oracle.ssbl_oracle_dynamic_hash(1)
    assert dh["fast_triangle_pair_contacts"] > 0 and dh["dynamic_particle_contacts"] > 0
    for a, b in zip(hashed, scanned):
        assert np.array_equal(a, b)
    assert dh["fast_triangle_pair_contacts"] == ds["fast_triangle_pair_contacts"]
    assert dh["dynamic_particle_contacts"] == ds["dynamic_particle_contacts"]
    # the hash visits fewer triangles than the scan
    assert dh["dynamic_triangle_candidate_count"] < ds["dynamic_triangle_candidate_count"]


def test_scene_layers_and_sources(nb, oracle, scenes):
    a, b, scene = sd.two_cloth_scene(nb, oracle, scenes, nb.GlobalDynamicScene.CROSS_LOWER_LAYERS)
    seen_a = seen_b = contacts_b = 0
    for _ in range(30):
        sd.refresh_scene(scene, [a, b], device_gather=False)
        a.step(6, 4, diagnostics=True)
        b.step(6, 4, diagnostics=True)
        da, db = a.diagnostics(), b.diagnostics()
        seen_a += da.dynamic_triangle_candidate_count + da.dynamic_particle_candidate_count
        seen_b += db.dynamic_triangle_candidate_count + db.dynamic_particle_candidate_count
        contacts_b += db.dynamic_particle_contacts + db.fast_triangle_pair_contacts
    # lower layers: B (layer 1) sees A (layer 0); A sees nothing, and nobody sees itself
    assert seen_a == 0 and seen_b > 0 and contacts_b > 0
    assert db.global_dynamic_particle_count == 16 * 16 + 10 * 10
    assert db.global_dynamic_triangle_count == 2 * 15 * 15 + 2 * 9 * 9


def test_scene_gather_equals_host_packing(nb, oracle, scenes):
    res = []
    for gather in (False, True):
        a, b, scene = sd.two_cloth_scene(nb, oracle, scenes, nb.GlobalDynamicScene.CROSS_ALL)
        for _ in range(6):
            sd.refresh_scene(scene, [a, b], device_gather=gather)
            a.step(4, 3, diagnostics=False)
            b.step(4, 3, diagnostics=False)
        res.append((a.download_positions(), b.download_positions()))
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])


def test_given_sdf_grid_equals_built_grid(nb, oracle, scenes):
    """A grid handed over with ssbl_ext_set_static_sdf behaves as the same grid built from triangles."""
    box = scenes.box_triangles((0.0, 0.0, 0.3), 0.3)
    cloth = scenes.grid_cloth(16, 0.5, 0.8, pin_corners=False, seed=3)
    opts = scenes.SolverOptions(collision_margin=0.01, cloth_thickness=0.005, static_sdf_max_resolution=64)
    built = nb.NativeXpbdSolver(cloth, opts, box, lib=oracle)
    values, ov = built.static_sdf()
    given = nb.NativeXpbdSolver(cloth, opts, None, lib=oracle)
    given.set_static_sdf(values, ov[:3], float(ov[3]))
    contacts = 0
    for _ in range(25):
        built.step(6, 4, diagnostics=True)
        given.step(6, 4, diagnostics=True)
        contacts += built.diagnostics().static_sdf_contact_count
        assert np.array_equal(built.download_positions(), given.download_positions())
    assert contacts > 0
    d = given.diagnostics()
    assert d.static_sdf_grid_x == values.shape[2] and d.static_sdf_voxel_count == values.size


def test_analytic_box_grid_keeps_cloth_outside(nb, oracle, scenes):
    grid = sd.box_sdf_grid(n=48)
    out, d = sd.cloth_on_sdf_grid(nb, oracle, scenes, grid, frames=25)
    p = out[-1].astype(np.float64)
    q = np.abs(p - np.array([0.0, 0.0, 0.3])) - 0.3
    dist = np.linalg.norm(np.maximum(q, 0.0), axis=1) + np.minimum(q.max(axis=1), 0.0)
    assert d.static_sdf_contact_count > 0
    # without the SDF the cloth would fall through to the ground (30 cm inside);
    # corners of a 4 cm-voxel trilinear grid leave at most a few mm
    assert dist.min() > -0.01


# --- signed (indexed) dynamic triangles
@pytest.mark.parametrize("height", [0.002, -0.002, -0.004, 0.0045, 0.02])
@pytest.mark.parametrize("quads", [1, 40])
def test_signed_plane_closed_form(oracle, height, quads):
    """One quad (two triangles, brute-force query) or 40 x 40 quads of 1 cm
    (3,200 triangles, hash query over the 27 or 125 cells around the vertex,
    whose normals are shortened by the reference's 1e-7 clamp)."""
    half = 1.0 if quads == 1 else 0.2
    r = sd.Raw(oracle, [[0.01, 0.02, height]], [1.0], collision_margin=sd.SIGNED_MARGIN, cloth_thickness=sd.SIGNED_THICKNESS,
               gravity=(0.0, 0.0, 0.0))
    verts, idx = sd.indexed_plane(0.0, half, quads)
    sd.upload_indexed(oracle, r, verts, idx)
    r.step(1, 1)
    p, q, v = r.state()
    d = r.diag()
    side = 2.0 * half / quads
    expect = sd.signed_plane_expected(height, cross_len=side * side)
    assert abs(float(p[0, 2]) - expect) < 2e-7, (p[0], expect)
    assert abs(p[0, 0] - 0.01) < 1e-7 and abs(p[0, 1] - 0.02) < 1e-7
    touched = height < sd.SIGNED_MARGIN
    assert (d.fast_triangle_pair_contacts > 0) == touched
    assert d.dynamic_triangle_count == 2 * quads * quads


def test_signed_collider_velocity(oracle):
    """A plane that rose by 2 mm since its last upload moves at 2 mm / (1/60 s):
    the vertex it pushes leaves with that normal speed (A41:1034-1066)."""
    (p, q, v), d, r = sd.signed_plane_contact(oracle, 0.002, 1, moved=0.002)
    assert abs(float(p[0, 2]) - sd.signed_plane_expected(0.002)) < 2e-7
    assert abs(float(v[0, 2]) - 0.002 * 60.0) < 1e-4 and abs(v[0, 0]) < 1e-7 and abs(v[0, 1]) < 1e-7
    (p2, q2, v2), _, _ = sd.signed_plane_contact(oracle, 0.002, 1)
    assert abs(float(v2[0, 2])) < 1e-6


def test_soup_upload_turns_signed_off(oracle):
    """A soup upload after an indexed one uses the unsigned contact (half the
    depth, A41:4310-4322) again."""
    r = sd.Raw(oracle, [[0.01, 0.02, 0.002]], [1.0], collision_margin=sd.SIGNED_MARGIN, cloth_thickness=sd.SIGNED_THICKNESS,
               gravity=(0.0, 0.0, 0.0))
    verts, idx = sd.indexed_plane(0.0, 1.0, 1)
    sd.upload_indexed(oracle, r, verts, idx)
    soup = sd.plane_soup(0.0, 1.0, 1)
    r.check(oracle.ssbl_update_dynamic_triangles(r.h, soup.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), len(soup)))
    r.step(1, 1)
    p = r.state()[0]
    assert abs(float(p[0, 2]) - (0.002 + 0.5 * (sd.SIGNED_MARGIN - 0.002))) < 2e-7


def test_cloth_over_indexed_collider_runs(nb, oracle, scenes):
    s, out = sd.cloth_over_indexed_collider(nb, oracle, scenes, frames=12)
    d = s.diagnostics()
    assert d.fast_triangle_pair_contacts > 0 and d.dynamic_particle_contacts >= 0
    pos = out[-1][0]
    assert np.all(np.isfinite(pos))
    # nothing ends more than 1 mm below the dome's surface under it
    x, y = pos[:, 0], pos[:, 1]
    inside = (np.abs(x) < 0.6) & (np.abs(y) < 0.6)
    dome = 0.05 + 0.004 * 11 + 0.04 * np.exp(-4.0 * (x ** 2 + y ** 2))
    assert np.min(pos[inside, 2] - dome[inside]) > -0.002


def test_bridge_frame_inputs_signatures(nb, oracle, scenes):
    """NativeXpbdSolver.update_frame_inputs (native_backend.py:1064-1279): an
    unchanged soup or particle set is not sent again; indexed triangles with
    the same topology keep the previous centres (collider velocity), a new
    topology signature drops them."""
    def vertex_solver():
        cloth = scenes.build_cloth(np.array([[0.01, 0.02, 0.002]], np.float32), np.zeros((0, 3), np.int32), np.zeros(1, bool))
        opts = scenes.SolverOptions(collision_margin=sd.SIGNED_MARGIN, cloth_thickness=sd.SIGNED_THICKNESS, gravity=(0.0, 0.0, 0.0))
        return nb.NativeXpbdSolver(cloth, opts, None, lib=oracle)

    s = vertex_solver()
    soup = sd.plane_soup(-0.5, 1.0, 1)
    assert s.update_frame_inputs(dynamic_triangles=soup.reshape(-1, 3, 3), update_dynamic=True)
    assert not s.update_frame_inputs(dynamic_triangles=soup.reshape(-1, 3, 3), update_dynamic=True)
    parts = {"positions": np.array([[5.0, 5.0, 5.0]], np.float32), "radii": np.array([0.01], np.float32)}
    assert s.update_frame_inputs(dynamic_particles=parts, update_dynamic_particles=True)
    assert not s.update_frame_inputs(dynamic_particles=parts, update_dynamic_particles=True)

    for signature, expect_v in ((None, 0.002 * 60.0), ("changed", 0.0)):
        s = vertex_solver()
        v0, idx = sd.indexed_plane(-0.002, 1.0, 1)
        v1, _ = sd.indexed_plane(0.0, 1.0, 1)
        assert s.update_frame_inputs(dynamic_triangle_vertices=v0, dynamic_triangle_indices=idx, update_dynamic=True,
                                     dynamic_triangle_topology_signature=None if signature is None else "a")
        assert s.update_frame_inputs(dynamic_triangle_vertices=v1, dynamic_triangle_indices=idx, update_dynamic=True,
                                     dynamic_triangle_topology_signature=None if signature is None else "b")
        s.step(1, 1)
        p, q, v = s.download_state()
        assert abs(float(p[0, 2]) - sd.signed_plane_expected(0.002)) < 2e-7
        assert abs(float(v[0, 2]) - expect_v) < 1e-4
