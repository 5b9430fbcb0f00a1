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
