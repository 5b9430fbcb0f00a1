"""Self collision on the GPU against the oracle: hash cells and candidate
pairs bit-exact, positions bit-exact (or within 1e-5 m)."""
import numpy as np
import pytest

import scenarios as S
from parity_util import make_pair, run_and_compare

pytestmark = pytest.mark.gpu
TOL = 1.0e-5


@pytest.mark.parametrize("scene", [S.self_vv, S.self_vt, S.self_ee])
def test_reference_smoke_scenes_strict(hip, oracle, scene):
    a0, a1, dg, rg = scene(hip)
    b0, b1, dr, rr = scene(oracle)
    assert a1 > a0 + 1e-4
    pg, pr = rg.state(), rr.state()
    diff = max(float(np.max(np.abs(x - y))) for x, y in zip(pg, pr))
    print(f"{scene.__name__}: max |hip - oracle| = {diff:.3e}")
    assert diff <= TOL
    for name in ("abi41_soft_contact_count", "abi41_exact_impulse_contact_count", "abi41_edge_edge_contact_count",
                 "fast_soft_repulsion_candidates"):
        assert getattr(dg, name) == getattr(dr, name), name
    assert dg.abi41_max_smoothed_delta == dr.abi41_max_smoothed_delta


def test_folded_cloth_cells_and_pairs_bit_exact(nb, hip, oracle, scenes):
    cloth, opts = S.folded_cloth(scenes)
    g, r = make_pair(nb, hip, oracle, cloth, opts, None)
    for frame in range(3):
        g.step(1, 1, diagnostics=True)
        r.step(1, 1, diagnostics=True)
        assert np.array_equal(g.self_hash_cells(), r.self_hash_cells())
        cg, ig = g.self_neighbors()
        cr, ir = r.self_neighbors()
        assert np.array_equal(cg, cr)
        mask = np.arange(32)[None, :] < cg[:, None]
        assert np.array_equal(ig[mask], ir[mask])
        pairs = int(cg.sum())
        pg, pr = g.download_state(), r.download_state()
        diff = max(float(np.max(np.abs(x - y))) for x, y in zip(pg, pr))
        print(f"folded frame {frame}: {pairs} candidate pairs, max |hip - oracle| = {diff:.3e}")
        assert pairs > 100 and diff <= TOL


@pytest.mark.parametrize("n", [8, 40])
def test_crumpling_cloth(nb, hip, oracle, scenes, n):
    # n = 8: 64 vertices, below the 128-vertex hash threshold (brute-force table)
    cloth, opts, frames = S.crumple(scenes, n=n, frames=20)
    g, r = make_pair(nb, hip, oracle, cloth, opts, None)
    worst = run_and_compare(g, r, frames, 5, 3, diagnostics=True, every=5)
    dg, dr = g.diagnostics(), r.diagnostics()
    print(f"crumple {n}x{n}: max |hip - oracle| = {worst:.3e}, candidates {dg.self_candidate_count}")
    assert worst <= TOL
    assert dg.self_candidate_count == dr.self_candidate_count


# --- fast mode (self_collision_mode 1)
FAST_FIELDS = ("abi41_soft_contact_count", "abi41_exact_impulse_contact_count", "abi41_edge_edge_contact_count",
               "fast_soft_repulsion_candidates", "self_filter_seen", "self_filter_accepted_vv", "self_filter_accepted_vt",
               "self_filter_accepted_ee", "self_filter_rejected_rest", "self_filter_rejected_duplicate",
               "self_filter_rejected_ownership", "self_filter_cache_hits", "self_filter_cache_misses", "self_cluster_count",
               "self_cluster_owned_contacts", "fast_overlap_island_candidates", "fast_overlap_island_clusters",
               "fast_overlap_island_vertex_refs", "fast_overlap_island_applied_vertices", "fast_overlap_island_guarded",
               "fast_overlap_island_max_delta")


def assert_fast_counters_equal(dg, dr):
    for name in FAST_FIELDS:
        assert getattr(dg, name) == getattr(dr, name), (name, getattr(dg, name), getattr(dr, name))


@pytest.mark.parametrize("scene", [S.self_vv, S.self_vt, S.self_ee])
@pytest.mark.parametrize("passes", [1, 3])
def test_reference_smoke_scenes_fast(hip, oracle, scene, passes):
    a0, a1, dg, rg = scene(hip, mode=1, passes=passes)
    b0, b1, dr, rr = scene(oracle, mode=1, passes=passes)
    assert a1 > a0 + 1e-4 and dg.fast_soft_repulsion_candidates > 0
    pg, pr = rg.state(), rr.state()
    diff = max(float(np.max(np.abs(x - y))) for x, y in zip(pg, pr))
    print(f"{scene.__name__} fast x{passes}: max |hip - oracle| = {diff:.3e}")
    assert diff <= TOL
    assert_fast_counters_equal(dg, dr)
    assert dg.abi41_max_smoothed_delta == dr.abi41_max_smoothed_delta


@pytest.mark.parametrize("same_cluster", [True, False])
def test_vv_cluster_ownership_fast(hip, oracle, same_cluster):
    start, out, dg = S.vv_clusters(hip, 1, same_cluster)
    _, ref, dr = S.vv_clusters(oracle, 1, same_cluster)
    assert np.allclose(out[0], S.vv_clusters_expected(start, 1, same_cluster), atol=2e-6)
    assert np.max(np.abs(out - ref)) <= TOL
    assert_fast_counters_equal(dg, dr)


@pytest.mark.parametrize("crossing,copies,drop", [(True, 3, 0.02), (True, 3, 0.009), (True, 2, 0.02), (False, 3, 0.02)])
def test_overlap_island_fast(hip, oracle, crossing, copies, drop):
    out, dg = S.overlap_island(hip, crossing, copies=copies, drop=drop)
    ref, dr = S.overlap_island(oracle, crossing, copies=copies, drop=drop)
    z, _ = S.overlap_island_expected(crossing, copies=copies, drop=drop)
    assert abs(float(out[0, 2]) - z) < 2e-6
    assert np.max(np.abs(out - ref)) <= TOL
    assert_fast_counters_equal(dg, dr)


@pytest.mark.parametrize("gap", [0.004, 0.0005])
def test_folded_cloth_fast(nb, hip, oracle, scenes, gap):
    # gap 0.0005 (under a quarter of the contact radius) forms overlap islands
    cloth, opts = S.folded_cloth(scenes, mode=1, passes=4, gap=gap)
    g, r = make_pair(nb, hip, oracle, cloth, opts, None)
    islands = 0
    for frame in range(3):
        g.step(4, 3, diagnostics=True)
        r.step(4, 3, diagnostics=True)
        dg, dr = g.diagnostics(), r.diagnostics()
        pg, pr = g.download_state(), r.download_state()
        diff = max(float(np.max(np.abs(x - y))) for x, y in zip(pg, pr))
        print(f"folded fast gap {gap} frame {frame}: accepted vv/vt/ee {dg.self_filter_accepted_vv}/"
              f"{dg.self_filter_accepted_vt}/{dg.self_filter_accepted_ee}, islands {dg.fast_overlap_island_applied_vertices}, "
              f"max |hip - oracle| = {diff:.3e}")
        assert diff <= TOL
        assert_fast_counters_equal(dg, dr)
        islands += dg.fast_overlap_island_applied_vertices
    if gap < 0.001:
        assert islands > 0


@pytest.mark.parametrize("n", [8, 40])
def test_crumpling_cloth_fast(nb, hip, oracle, scenes, n):
    cloth, opts, frames = S.crumple(scenes, n=n, frames=20, mode=1)
    g, r = make_pair(nb, hip, oracle, cloth, opts, None)
    worst = run_and_compare(g, r, frames, 5, 3, diagnostics=True, every=5)
    dg, dr = g.diagnostics(), r.diagnostics()
    print(f"crumple fast {n}x{n}: max |hip - oracle| = {worst:.3e}, seen {dg.self_filter_seen}")
    assert worst <= TOL
    assert_fast_counters_equal(dg, dr)
