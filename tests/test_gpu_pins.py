"""Pins as the vertex passes apply them, against the oracle: vertices with
several pins each (hard and soft, in pin order, with different targets), an
ignored pin (weight below 0.05), pinned vertices (inverse mass 0) with extra
pins, and the contacts between the pass's two pin projections (ground,
sphere and the box SDF), so that the projection after the contacts runs in
every pass: skipped for pinned vertices whose pins the pass already applied,
read again for the others.  Tolerance: positions within 1e-5 m (the printed
difference is 0 when bit-exact)."""
from dataclasses import replace

import numpy as np
import pytest

from parity_util import make_pair, run_and_compare

pytestmark = pytest.mark.gpu
TOL = 1.0e-5


def test_repeated_hard_and_soft_pins_with_contacts(nb, hip, oracle, scenes):
    n = 24
    cloth = scenes.grid_cloth(n, 1.0, 0.35, pin_corners=True, seed=11)
    rest = cloth.positions_world
    corners = [n * (n - 1), n * n - 1]                      # pinned (inverse mass 0)
    row = list(range(3, n - 3, 2))                          # free vertices of the first row
    idx = corners + corners + row + row[::2] + [n * (n // 2) + n // 2]
    targets = rest[idx].astype(np.float64)
    k = len(corners)
    targets[k:2 * k] += (0.02, 0.0, 0.01)                   # the corners' second pins, elsewhere
    targets[2 * k + len(row):2 * k + len(row) + len(row[::2])] += (0.0, 0.03, -0.02)
    weights = np.concatenate([np.ones(k), np.full(k, 0.9),             # hard, hard
                              np.where(np.arange(len(row)) % 2 == 0, 0.3, 0.8),   # soft and hard
                              np.full(len(row[::2]), 0.5),             # a second, soft pin
                              [0.01]]).astype(np.float32)              # ignored
    cloth = replace(cloth, pin_indices=np.asarray(idx, np.int32), pin_targets_world=targets.astype(np.float32),
                    pin_weights=weights)
    box = scenes.box_triangles((0.2, 0.1, 0.1), 0.12)
    opts = scenes.SolverOptions(use_ground=True, use_sphere=True, sphere_center=(-0.2, -0.1, 0.15), sphere_radius=0.15,
                                static_sdf_max_resolution=48, collision_margin=0.01, contact_friction=0.3)
    g, r = make_pair(nb, hip, oracle, cloth, opts, box)
    worst = run_and_compare(g, r, 30, 8, 6, every=3)
    print(f"repeated pins with contacts: max |hip - oracle| = {worst:.3e}", flush=True)
    assert worst <= TOL
    d = g.diagnostics()
    assert d.finite_flag == 1
