"""Self collision on the oracle (CPU): the reference smoke scenes' criteria,
the candidate table against brute-force enumeration, and the fast mode's
filters and overlap islands against closed forms."""
import numpy as np
import pytest

import scenarios as S


def test_vertex_vertex_pushes_apart(oracle):
    g0, g1, d, _ = S.self_vv(oracle)
    assert g1 > g0 + 1e-4                                               # abi41_smoke.cpp:301-309
    assert d.abi41_soft_contact_count + d.abi41_exact_impulse_contact_count > 0 and d.fast_soft_repulsion_candidates > 0
    assert d.abi41_max_smoothed_delta > 0.0


def test_vertex_triangle_pushes_up(oracle):
    z0, z1, d, _ = S.self_vt(oracle)
    assert z1 > z0 + 1e-4
    assert d.abi41_soft_contact_count > 0 and d.abi41_max_smoothed_delta > 0.0


def test_edge_edge_pushes_up(oracle):
    z0, z1, d, _ = S.self_ee(oracle)
    assert z1 > z0 + 1e-4
    assert d.abi41_edge_edge_contact_count > 0 and d.abi41_max_smoothed_delta > 0.0


def test_hash_cells_and_table_match_brute_force(oracle, nb, scenes):
    """The hash build's cells are floor(p / cell), and the hashed candidate
    table lists exactly the pairs an O(N^2) enumeration with the same
    acceptance test finds (A41:4696-4717), in visiting order: 27 cells in
    z, y, x order, ascending vertex index within a cell."""
    import ctypes
    cloth, opts = S.folded_cloth(scenes)
    s = nb.NativeXpbdSolver(cloth, opts, None, lib=oracle)
    s.step(1, 1, diagnostics=False)
    n = s.vertex_count
    P, Q = np.empty((n, 3), np.float32), np.empty((n, 3), np.float32)
    fp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))  # noqa: E731
    oracle.ssbl_oracle_self_snapshot.argtypes = [ctypes.c_void_p] + [ctypes.POINTER(ctypes.c_float)] * 2 + [ctypes.c_int]
    assert oracle.ssbl_oracle_self_snapshot(s.handle, fp(P), fp(Q), P.size)
    cells = s.self_hash_cells()
    counts, idx = s.self_neighbors()
    r = np.float32(opts.self_collision_distance)
    cs = np.float32(max(r * np.float32(1.8), np.float32(1e-3)))
    assert np.array_equal(cells, np.floor(P / cs).astype(np.int32))
    onset = max(np.float32(r * np.float32(1.8)), np.float32(r + np.float32(1e-7)))
    skip = max(np.float32(onset * np.float32(1.25)), np.float32(r * np.float32(2.5)))
    rest = cloth.positions_world
    inv = cloth.inv_mass
    total = 0
    for v in range(n):
        got = list(idx[v, :counts[v]])
        if inv[v] <= 0:
            assert counts[v] == 0
            continue
        near = np.all(np.abs(cells - cells[v]) <= 1, axis=1)
        cand = np.flatnonzero(near)
        cand = cand[cand != v]
        rd = rest[cand].astype(np.float64) - rest[v]
        ok = np.sum(rd * rd, 1) >= float(skip) ** 2
        cd = P[cand].astype(np.float64) - P[v]
        pd = Q[cand].astype(np.float64) - Q[v]
        ok &= (np.sum(cd * cd, 1) < float(onset) ** 2) | (np.sum(pd * pd, 1) < float(onset) ** 2)
        expect = cand[ok]
        # visiting order: cell offset (dz, dy, dx) lexicographic, then index
        off = cells[expect] - cells[v]
        order = np.lexsort((expect, off[:, 0], off[:, 1], off[:, 2]))
        expect = list(expect[order][:32])
        assert got == expect, (v, got, expect)
        total += len(got)
    assert total > 100


# --- fast mode (self_collision_mode 1): contact filters and overlap islands
@pytest.mark.parametrize("scene", [S.self_vv, S.self_vt, S.self_ee])
@pytest.mark.parametrize("passes", [1, 3])
def test_reference_smoke_scenes_fast(oracle, scene, passes):
    """The reference smoke runs these scenes in fast mode (abi41_smoke.cpp:244,
    369, 513) with one pass; its pass criteria hold for one and three passes."""
    v0, v1, d, _ = scene(oracle, mode=1, passes=passes)
    assert v1 > v0 + 1e-4
    assert d.abi41_max_smoothed_delta > 0.0
    assert d.fast_soft_repulsion_candidates > 0
    assert d.self_filter_seen > 0
    assert d.self_cluster_count == d.self_cluster_owned_contacts
    assert d.self_filter_accepted_vv + d.self_filter_accepted_vt + d.self_filter_accepted_ee == d.self_cluster_owned_contacts
    # the table is built once per self substep and reused by the other passes
    assert d.self_filter_cache_misses == 600 and d.self_filter_cache_hits == 600 * (passes - 1)


@pytest.mark.parametrize("same_cluster", [True, False])
@pytest.mark.parametrize("mode", [1, 2])
def test_vv_cluster_ownership(oracle, mode, same_cluster):
    start, out, d = S.vv_clusters(oracle, mode, same_cluster)
    expect = S.vv_clusters_expected(start, mode, same_cluster)
    assert np.allclose(out[0], expect, atol=2e-6), (out[0], expect)
    assert np.array_equal(out[1:], start[1:])
    if mode == 1:
        kept = 1 if same_cluster else 2
        assert d.self_filter_seen == 2
        assert d.self_filter_accepted_vv == kept and d.abi41_exact_impulse_contact_count == kept
        assert d.self_filter_rejected_ownership == 2 - kept
        assert d.self_cluster_count == kept and d.fast_soft_repulsion_candidates == kept
    else:
        assert d.self_filter_seen == 0 and d.self_cluster_count == 0
        assert d.abi41_exact_impulse_contact_count == 2 and d.fast_soft_repulsion_candidates == 2


@pytest.mark.parametrize("crossing,copies,drop", [(True, 3, 0.02), (True, 3, 0.009), (True, 2, 0.02), (False, 3, 0.02)])
def test_overlap_island_closed_form(oracle, crossing, copies, drop):
    out, d = S.overlap_island(oracle, crossing, copies=copies, drop=drop)
    z, dl = S.overlap_island_expected(crossing, copies=copies, drop=drop)
    assert abs(float(out[0, 2]) - z) < 2e-6, (out[0], z)
    assert abs(out[0, 0]) < 1e-7 and abs(out[0, 1]) < 1e-7
    assert d.abi41_soft_contact_count == 0 and d.abi41_exact_impulse_contact_count == 0
    if crossing:
        assert d.fast_overlap_island_candidates == copies
        if copies >= 3:
            assert (d.fast_overlap_island_clusters, d.fast_overlap_island_applied_vertices) == (1, 1)
            assert d.fast_overlap_island_vertex_refs == copies and d.fast_overlap_island_guarded == 0
            assert abs(d.fast_overlap_island_max_delta - dl) <= 1.5e-6
            assert d.abi41_max_smoothed_delta == pytest.approx(dl, abs=1e-6)
        else:
            assert d.fast_overlap_island_guarded == 1 and d.fast_overlap_island_applied_vertices == 0
    else:
        assert d.fast_overlap_island_candidates == 0 and d.fast_overlap_island_applied_vertices == 0


def test_fast_mode_filter_bookkeeping(oracle, nb, scenes):
    """Counter identities of the fast filters on a folded cloth: one
    contact per cluster, every accepted contact counted once as a self
    candidate, each accepted vertex-vertex contact applied as an exact
    impulse, at most one response per accepted triangle or edge contact."""
    cloth, opts = S.folded_cloth(scenes, n=24, mode=1, passes=4)
    s = nb.NativeXpbdSolver(cloth, opts, None, lib=oracle)
    seen = owned = 0
    for f in range(3):
        s.step(4, 3, diagnostics=True)
        d = s.diagnostics()
        acc = d.self_filter_accepted_vv + d.self_filter_accepted_vt + d.self_filter_accepted_ee
        assert d.self_cluster_count == d.self_cluster_owned_contacts == acc
        assert d.fast_soft_repulsion_candidates == acc
        assert d.abi41_exact_impulse_contact_count == d.self_filter_accepted_vv
        assert d.abi41_soft_contact_count <= d.self_filter_accepted_vt
        assert d.abi41_edge_edge_contact_count <= d.self_filter_accepted_ee
        assert d.self_filter_cache_misses == s.vertex_count * 4
        assert d.self_filter_cache_hits == s.vertex_count * 4 * 3
        seen += d.self_filter_seen
        owned += acc
    assert owned > 0 and seen > owned
