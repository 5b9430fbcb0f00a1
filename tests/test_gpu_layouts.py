"""GPU parity for the incidence layouts the benchmark meshes do not reach.

The stretch gather reads 6-byte incidence slots (16-bit neighbour offset)
when every neighbour of a cloth lies within 16383 indices, and the PCG
stores its blocks slot-major when no vertex has more than 8 incidences.
Every grid and sphere of the workloads takes those forms, so here:

* a grid whose vertex numbering is shuffled (neighbours far apart in index):
  the 8-byte slots;
* a wheel (a hub joined to 16 rim vertices, so 16 incidences at the hub):
  the incidences past the 8th slot come from the CSR list, and the PCG keeps
  its blocks in CSR order; with the caps coloured and uncoloured.

Each against the oracle, bit for bit, over enough frames for the cloth to
swing and stretch.
"""
import numpy as np
import pytest

from parity_util import make_pair, state_diff

pytestmark = pytest.mark.gpu


def _shuffled_grid(scenes, n, seed):
    pos, tris = scenes.grid_mesh(n, 1.0)
    pos[:, 1] += 1.0
    perm = np.random.default_rng(seed).permutation(n * n)   # new index of old vertex i
    new_pos = np.empty_like(pos)
    new_pos[perm] = pos
    pin = np.zeros(n * n, bool)
    pin[perm[n * (n - 1)]] = True
    pin[perm[n * n - 1]] = True
    return scenes.build_cloth(scenes.y_up_to_z_up(new_pos), perm[tris].astype(np.int32), pin, jitter=1.0e-4, seed=seed)


def _wheel(scenes, rim, seed, colored=True):
    ang = np.linspace(0.0, 2.0 * np.pi, rim, endpoint=False)
    pos = np.concatenate([[[0.0, 1.0, 0.0]], np.stack([0.3 * np.cos(ang), np.ones(rim), 0.3 * np.sin(ang)], 1)])
    tris = np.array([[0, 1 + i, 1 + (i + 1) % rim] for i in range(rim)], np.int32)
    pin = np.zeros(rim + 1, bool)
    pin[1] = pin[1 + rim // 2] = True
    cloth = scenes.build_cloth(scenes.y_up_to_z_up(pos), tris, pin, jitter=1.0e-4, seed=seed)
    if not colored:
        cloth.edge_color_offsets = np.empty(0, np.int32)
    return cloth


def _run(nb, hip, oracle, cloth, opts, frames, substeps, iterations):
    g, r = make_pair(nb, hip, oracle, cloth, opts, None)
    worst = 0.0
    for _ in range(frames):
        g.step(substeps, iterations, diagnostics=True, synchronize=True)
        r.step(substeps, iterations, diagnostics=True, synchronize=True)
        worst = max(worst, state_diff(g, r))
    moved = float(np.max(np.abs(g.download_positions() - cloth.positions_world)))
    return worst, moved, g.diagnostics(), r.diagnostics()


def test_far_neighbours_take_the_wide_slots(nb, hip, oracle, scenes):
    cloth = _shuffled_grid(scenes, 130, seed=3)
    e = np.asarray(cloth.edges).reshape(-1, 2)
    assert np.max(np.abs(e[:, 0] - e[:, 1])) > 16383   # some neighbours lie beyond a 16-bit offset
    opts = scenes.SolverOptions(use_ground=True)
    worst, moved, dg, dr = _run(nb, hip, oracle, cloth, opts, 12, 10, 8)
    print(f"shuffled 130x130 grid: max |hip - oracle| = {worst:.3e}, moved {moved:.3f} m")
    assert worst == 0.0 and moved > 0.01
    assert dg.abi41_hard_projection_fallbacks == dr.abi41_hard_projection_fallbacks


@pytest.mark.parametrize("pcg,colored", [(False, True), (True, True), (True, False)])
def test_wheel_hub_past_the_slots(nb, hip, oracle, scenes, pcg, colored):
    cloth = _wheel(scenes, 16, seed=5, colored=colored)
    deg = np.bincount(np.asarray(cloth.edges).ravel())
    assert deg.max() == 16   # the hub: more incidences than the 8 slots
    opts = scenes.SolverOptions(use_ground=True, stretch_optimization_enabled=pcg, stretch_optimization_strength=1.0,
                                cloth_thickness=0.005)
    worst, moved, dg, dr = _run(nb, hip, oracle, cloth, opts, 20, 6, 4)
    print(f"wheel pcg={pcg} colored={colored}: max |hip - oracle| = {worst:.3e}, moved {moved:.3f} m")
    assert worst == 0.0 and moved > 0.01
    if pcg:
        assert dg.abi41_pcg_iterations == dr.abi41_pcg_iterations and dg.abi41_pcg_csr_nnz == dr.abi41_pcg_csr_nnz
