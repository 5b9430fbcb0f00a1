"""Self collision on the oracle (CPU): the reference smoke scenes' criteria and
the candidate table against brute-force enumeration."""
import numpy as np

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
