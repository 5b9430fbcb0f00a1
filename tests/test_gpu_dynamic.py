"""Dynamic colliders on the GPU against the oracle: closed forms, a cloth
falling onto a triangle soup and a particle cloud (hashed queries), and two
cloths colliding through a global dynamic scene -- refreshed from host
copies or gathered on the device, stepped alone or batched.  Tolerance 0:
the HIP path reproduces the oracle's FP32 operations."""
import numpy as np
import pytest

import scenarios_dynamic as sd
from test_dynamic import COUNTERS, run_cloth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("quads", [1, 8])
def test_plane_contact_closed_form_gpu(hip, oracle, quads):
    (p, q, v), d = sd.plane_contact(hip, quads)
    (po, qo, vo), do = sd.plane_contact(oracle, quads)
    assert np.array_equal(p, po) and np.array_equal(q, qo) and np.array_equal(v, vo)
    assert np.allclose(p[:, 2], 0.5 * (0.01 - 0.004), atol=1.0e-8)
    for k in ("fast_triangle_pair_contacts", "abi41_exact_impulse_contact_count", "dynamic_triangle_candidate_count",
              "dynamic_triangle_aabb_reject_count", "dynamic_triangle_count", "dynamic_triangle_max_bucket_occupancy"):
        assert getattr(d, k) == getattr(do, k), k


@pytest.mark.parametrize("left_first", [True, False])
@pytest.mark.parametrize("count", [4, 64])          # 4: one round of pooled items; 64: several rounds
def test_valley_tie_gpu(hip, oracle, count, left_first):
    """Unsigned buckets are stored unordered on the GPU: the depth tie between
    two mirrored faces must still go to the set's first triangle."""
    (p, q, v), d = sd.valley_tie(hip, count, left_first)
    (po, qo, vo), do = sd.valley_tie(oracle, count, left_first)
    assert np.array_equal(p, po) and np.array_equal(q, qo) and np.array_equal(v, vo)
    assert np.all(np.sign(p[:, 0]) == (1.0 if left_first else -1.0))
    for k in ("fast_triangle_pair_contacts", "dynamic_triangle_candidate_count", "dynamic_triangle_aabb_reject_count"):
        assert getattr(d, k) == getattr(do, k), k


@pytest.mark.parametrize("far", [0, 200])
def test_particle_contact_closed_form_gpu(hip, oracle, far):
    (p, _, _), d = sd.particle_contact(hip, far_particles=far)
    (po, _, _), do = sd.particle_contact(oracle, far_particles=far)
    assert np.array_equal(p, po)
    assert d.dynamic_particle_contacts == do.dynamic_particle_contacts == 1


def test_cloth_over_soup_and_particles_gpu(nb, hip, oracle, scenes):
    g, dg = run_cloth(nb, hip, scenes)
    r, dr = run_cloth(nb, oracle, scenes)
    worst = max(float(np.max(np.abs(a - b))) for a, b in zip(g, r))
    print(f"cloth over dynamic colliders, 30 frames: max |hip - oracle| = {worst:.3e}; counters {dg}")
    assert worst == 0.0
    assert dg == dr
    assert dg["fast_triangle_pair_contacts"] > 0 and dg["dynamic_particle_contacts"] > 0


def run_mixed(nb, lib, scenes, *, batched, frames=12):
    """Cloth A over its own dynamic triangles and particles beside cloth B,
    which has no dynamic colliders, stepped as one batch or one by one."""
    a = sd.cloth_over_colliders(nb, lib, scenes)
    opts = scenes.SolverOptions(collision_margin=0.01, cloth_thickness=0.005)
    b = nb.NativeXpbdSolver(scenes.grid_cloth(12, 0.6, 1.2, pin_corners=True, seed=7), opts, None, lib=lib)
    batch = nb.SolverBatch([a, b]) if batched else None
    out = []
    for _ in range(frames):
        if batch is not None:
            batch.step(6, 4, synchronize=True)
        else:
            a.step(6, 4, synchronize=True)
            b.step(6, 4, synchronize=True)
        out.append((a.download_positions(), b.download_positions()))
    if batch is not None:
        batch.close()
    a.close()
    b.close()
    return out


def test_mixed_batch_with_and_without_colliders_gpu(nb, hip, oracle, scenes):
    """In a batch where only some cloths have dynamic colliders, the contact
    launches also close the substep (pin polish, velocity, next integration)
    for the cloths without any: batched equals alone equals the oracle."""
    batched = run_mixed(nb, hip, scenes, batched=True)
    alone = run_mixed(nb, hip, scenes, batched=False)
    ref = run_mixed(nb, oracle, scenes, batched=False)
    for (ab, bb), (aa, ba), (ao, bo) in zip(batched, alone, ref):
        assert np.array_equal(ab, aa) and np.array_equal(bb, ba)
        assert np.array_equal(aa, ao) and np.array_equal(ba, bo)
    # cloth B moved under gravity, which the fused integration applies
    assert not np.array_equal(batched[0][1], batched[-1][1])


def run_scene(nb, lib, scenes, *, gather, batched, frames=24):
    a, b, scene = sd.two_cloth_scene(nb, lib, scenes, nb.GlobalDynamicScene.CROSS_LOWER_LAYERS)
    batch = nb.SolverBatch([a, b]) if batched else None
    out, contacts = [], 0
    for _ in range(frames):
        sd.refresh_scene(scene, [a, b], device_gather=gather)
        if batch is not None:
            batch.step(6, 4, synchronize=True)
        else:
            a.step(6, 4, diagnostics=True)
            b.step(6, 4, diagnostics=True)
            d = b.diagnostics()
            contacts += d.dynamic_particle_contacts + d.fast_triangle_pair_contacts
        out.append((a.download_positions(), b.download_positions()))
    if batch is not None:
        batch.close()
    return out, contacts


def test_two_cloth_scene_gpu(nb, hip, oracle, scenes):
    ref, cr = run_scene(nb, oracle, scenes, gather=False, batched=False)
    host, ch = run_scene(nb, hip, scenes, gather=False, batched=False)
    dev, cd = run_scene(nb, hip, scenes, gather=True, batched=False)
    bat, _ = run_scene(nb, hip, scenes, gather=True, batched=True)
    assert cr > 0 and ch == cr and cd == cr
    for f, (x, y, z, w) in enumerate(zip(ref, host, dev, bat)):
        for k in range(2):
            assert np.array_equal(x[k], y[k]), ("host-packed scene vs oracle", f, k)
            assert np.array_equal(y[k], z[k]), ("device gather vs host packing", f, k)
            assert np.array_equal(z[k], w[k]), ("batched vs single", f, k)


def test_given_128_box_grid_gpu(nb, hip, oracle, scenes):
    """The 128^3 FP32 box SDF over a 2 m cube handed over as a grid: HIP equals the oracle."""
    grid = sd.box_sdf_grid(n=128)
    g, dg = sd.cloth_on_sdf_grid(nb, hip, scenes, grid)
    r, dr = sd.cloth_on_sdf_grid(nb, oracle, scenes, grid)
    worst = max(float(np.max(np.abs(a - b))) for a, b in zip(g, r))
    print(f"cloth on a given 128^3 box SDF: max |hip - oracle| = {worst:.3e}")
    assert worst == 0.0
    assert dg.static_sdf_contact_count == dr.static_sdf_contact_count > 0


# --- signed (indexed) dynamic triangles
SIGNED_COUNTERS = ("fast_triangle_pair_contacts", "abi41_exact_impulse_contact_count", "dynamic_triangle_candidate_count",
                   "dynamic_triangle_aabb_reject_count", "dynamic_triangle_count", "dynamic_particle_candidate_count",
                   "dynamic_particle_contacts", "external_friction_corrections")


@pytest.mark.parametrize("height", [0.002, -0.004])
@pytest.mark.parametrize("quads", [1, 40])
def test_signed_plane_closed_form_gpu(hip, oracle, height, quads):
    half = 1.0 if quads == 1 else 0.2
    out = []
    for lib in (hip, oracle):
        r = sd.Raw(lib, [[0.01, 0.02, height]], [1.0], collision_margin=sd.SIGNED_MARGIN, cloth_thickness=sd.SIGNED_THICKNESS,
                   gravity=(0.0, 0.0, 0.0))
        verts, idx = sd.indexed_plane(0.0, half, quads)
        sd.upload_indexed(lib, r, verts, idx)
        r.step(1, 1)
        out.append((r.state(), r.diag()))
    (pg, qg, vg), dg = out[0]
    (po, qo, vo), do = out[1]
    side = 2.0 * half / quads
    assert abs(float(pg[0, 2]) - sd.signed_plane_expected(height, cross_len=side * side)) < 2e-7
    assert np.array_equal(pg, po) and np.array_equal(qg, qo) and np.array_equal(vg, vo)
    for k in SIGNED_COUNTERS:
        assert getattr(dg, k) == getattr(do, k), k


def test_signed_collider_velocity_gpu(hip, oracle):
    (pg, qg, vg), dg, _ = sd.signed_plane_contact(hip, 0.002, 1, moved=0.002)
    (po, qo, vo), do, _ = sd.signed_plane_contact(oracle, 0.002, 1, moved=0.002)
    assert np.array_equal(pg, po) and np.array_equal(vg, vo) and np.array_equal(qg, qo)
    assert abs(float(vg[0, 2]) - 0.002 * 60.0) < 1e-4


@pytest.mark.parametrize("batched", [False, True])
def test_cloth_over_indexed_collider_gpu(nb, hip, oracle, scenes, batched):
    """A cloth on a rising indexed dome plus moving particles, 16 frames of 4
    substeps x 3 iterations: every frame re-uploads the dome (collider
    velocity from its centres), signed passes with local averaging and the
    inside-only passes; stepped alone or through a batch."""
    sg, og = sd.cloth_over_indexed_collider(nb, hip, scenes, frames=16, batched=batched)
    so, oo = sd.cloth_over_indexed_collider(nb, oracle, scenes, frames=16, batched=batched)
    worst = max(float(np.max(np.abs(a[k] - b[k]))) for a, b in zip(og, oo) for k in range(3))
    print(f"cloth over indexed dome (batched={batched}): max |hip - oracle| = {worst:.3e}")
    assert worst == 0.0
    if not batched:
        dg, do = sg.diagnostics(), so.diagnostics()
        for k in SIGNED_COUNTERS:
            assert getattr(dg, k) == getattr(do, k), k
        assert dg.fast_triangle_pair_contacts > 0
