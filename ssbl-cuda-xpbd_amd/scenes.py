"""Cloth inputs for the ABI41 solver: constraint builders and the benchmark scenes.

The builders restate the reference's numpy input code (xpbd_core.py:715-881):
edges of a triangle list, "bend" pairs (the two vertices opposite each
interior edge), greedy graph colouring of distance constraints, and lumped
inverse masses.  The scene functions generate the synthetic cloths of the
benchmark configurations (grids and an icosphere), seeded with
numpy Generator(PCG64(seed)).

Scene coordinates are Y-up as the configurations are written; the solver's
ground plane is Z-up (ssbl_xpbd_cuda_abi41.cu:3520-3533), so every scene is
mapped with (x, y, z) -> (x, -z, y), a proper rotation that keeps winding.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

_EPS = 1.0e-7


# ------------------------------------------------------------- builders
def _pair_keys(lo: np.ndarray, hi: np.ndarray, n: int) -> np.ndarray:
    return lo.astype(np.int64) * n + hi.astype(np.int64)


def edge_constraints(triangles: np.ndarray, rest: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Unique undirected edges, sorted, with rest lengths (xpbd_core.py:715-727)."""
    tri = np.asarray(triangles, dtype=np.int64).reshape(-1, 3)
    if len(tri) == 0:
        return np.empty((0, 2), np.int32), np.empty(0, np.float32)
    e = np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]])
    e = np.sort(e, axis=1)
    n = int(e.max()) + 1
    keys = np.unique(_pair_keys(e[:, 0], e[:, 1], n))      # lexicographic (lo, hi) order
    e = np.stack([keys // n, keys % n], axis=1)
    d = rest[e[:, 1]].astype(np.float64) - rest[e[:, 0]].astype(np.float64)
    lengths = np.maximum(np.linalg.norm(d, axis=1), _EPS).astype(np.float32)
    return e.astype(np.int32), lengths


def bend_constraints(triangles: np.ndarray, rest: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Pairs of vertices opposite each interior edge (xpbd_core.py:730-747).

    Triangle edges are visited in order (triangle by triangle, edges ab, bc,
    ca).  The first triangle seen on an edge records its opposite vertex; a
    later triangle with a different opposite vertex forms the pair.  Computed
    by a stable sort on the edge key, so "first seen" is the first of each
    key's run.
    """
    tri = np.asarray(triangles, dtype=np.int64).reshape(-1, 3)
    if len(tri) == 0:
        return np.empty((0, 2), np.int32), np.empty(0, np.float32)
    x = tri[:, [0, 1, 2]].ravel()
    y = tri[:, [1, 2, 0]].ravel()
    opp = tri[:, [2, 0, 1]].ravel()
    n = int(tri.max()) + 1
    keys = _pair_keys(np.minimum(x, y), np.maximum(x, y), n)
    order = np.argsort(keys, kind="stable")
    k_sorted, o_sorted = keys[order], opp[order]
    run_start = np.ones(len(k_sorted), bool)
    run_start[1:] = k_sorted[1:] != k_sorted[:-1]
    first_opp = o_sorted[np.maximum.accumulate(np.where(run_start, np.arange(len(k_sorted)), 0))]
    later = ~run_start & (o_sorted != first_opp)
    a, b = first_opp[later], o_sorted[later]
    if len(a) == 0:
        return np.empty((0, 2), np.int32), np.empty(0, np.float32)
    pk = np.unique(_pair_keys(np.minimum(a, b), np.maximum(a, b), n))
    arr = np.stack([pk // n, pk % n], axis=1)
    d = rest[arr[:, 1]].astype(np.float64) - rest[arr[:, 0]].astype(np.float64)
    return arr.astype(np.int32), np.maximum(np.linalg.norm(d, axis=1), _EPS).astype(np.float32)


def color_constraints(pairs: np.ndarray, rest: np.ndarray, vertex_count: int):
    """Greedy colouring, then a stable sort by colour (xpbd_core.py:828-859).

    Each constraint takes the smallest colour not used by either endpoint so
    far.  Returns (sorted pairs, sorted rest lengths, colour offsets).
    """
    pairs = np.asarray(pairs, dtype=np.int32).reshape(-1, 2)
    if len(pairs) == 0:
        return pairs, np.asarray(rest, np.float32), np.asarray([0], np.int32)
    used = [0] * vertex_count  # bit set of colours used at each vertex
    colors = np.empty(len(pairs), dtype=np.int32)
    for k, (a, b) in enumerate(pairs.tolist()):
        blocked = (used[a] if 0 <= a < vertex_count else 0) | (used[b] if 0 <= b < vertex_count else 0)
        c = 0
        while blocked >> c & 1:
            c += 1
        colors[k] = c
        if 0 <= a < vertex_count:
            used[a] |= 1 << c
        if 0 <= b < vertex_count:
            used[b] |= 1 << c
    order = np.argsort(colors, kind="stable")
    counts = np.bincount(colors, minlength=int(colors.max()) + 1).astype(np.int32)
    offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    return (np.ascontiguousarray(pairs[order]), np.ascontiguousarray(np.asarray(rest, np.float32)[order]), offsets)


def vertex_inverse_mass(rest: np.ndarray, triangles: np.ndarray, density: float, pin_mask: np.ndarray) -> np.ndarray:
    """Lumped mass: a third of each triangle's area x density (xpbd_core.py:862-881)."""
    tri = np.asarray(triangles, dtype=np.int64).reshape(-1, 3)
    p = rest.astype(np.float64)
    area = 0.5 * np.linalg.norm(np.cross(p[tri[:, 1]] - p[tri[:, 0]], p[tri[:, 2]] - p[tri[:, 0]]), axis=1)
    share = np.maximum(area * max(density, _EPS), _EPS) / 3.0
    mass = np.zeros(len(rest), dtype=np.float64)
    for k in range(3):
        np.add.at(mass, tri[:, k], share)
    mass[mass <= 0.0] = 1.0
    inv = 1.0 / mass
    inv[np.asarray(pin_mask, bool)] = 0.0
    return inv.astype(np.float32)


def y_up_to_z_up(p: np.ndarray) -> np.ndarray:
    p = np.asarray(p, dtype=np.float64)
    return np.stack([p[:, 0], -p[:, 2], p[:, 1]], axis=1)


# --------------------------------------------------------------- meshes
def grid_mesh(n: int, size: float) -> tuple[np.ndarray, np.ndarray]:
    """n x n vertex grid in the Y-up XZ plane (y = 0), centred; one diagonal per quad."""
    s = np.linspace(-0.5 * size, 0.5 * size, n)
    x, z = np.meshgrid(s, s, indexing="xy")
    pos = np.stack([x.ravel(), np.zeros(n * n), z.ravel()], axis=1)
    r, c = np.meshgrid(np.arange(n - 1), np.arange(n - 1), indexing="ij")
    a = (r * n + c).ravel()
    b = a + 1
    cc = a + n
    d = a + n + 1
    tris = np.concatenate([np.stack([a, d, b], 1), np.stack([a, cc, d], 1)])
    return pos, tris.astype(np.int32)


def icosphere(subdivisions: int, radius: float) -> tuple[np.ndarray, np.ndarray]:
    """Loop-free icosahedron subdivision projected to a sphere (10*4^k + 2 vertices)."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
             (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
             (8, 6, 7), (9, 8, 1)]
    v = np.asarray(verts, dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.asarray(faces, dtype=np.int64)
    for _ in range(subdivisions):
        e = np.sort(np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]]), axis=1)
        uniq, inv = np.unique(e, axis=0, return_inverse=True)
        mids = v[uniq[:, 0]] + v[uniq[:, 1]]
        mids /= np.linalg.norm(mids, axis=1, keepdims=True)
        base = len(v)
        v = np.concatenate([v, mids])
        m = inv.reshape(3, -1).T + base  # midpoints of edges (01, 12, 20)
        a, b, c = f[:, 0], f[:, 1], f[:, 2]
        ab, bc, ca = m[:, 0], m[:, 1], m[:, 2]
        f = np.concatenate([np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1), np.stack([c, ca, bc], 1),
                            np.stack([ab, bc, ca], 1)])
    return v * radius, f.astype(np.int32)


def box_triangles(center, half) -> np.ndarray:
    """12 outward-wound triangles of an axis-aligned box (float32 soup, 9 per triangle)."""
    cx, cy, cz = center
    hx, hy, hz = (half, half, half) if np.isscalar(half) else half
    c = np.array([[cx + sx * hx, cy + sy * hy, cz + sz * hz] for sz in (-1, 1) for sy in (-1, 1) for sx in (-1, 1)])
    quads = [(0, 2, 3, 1), (4, 5, 7, 6), (0, 1, 5, 4), (2, 6, 7, 3), (0, 4, 6, 2), (1, 3, 7, 5)]
    tris = []
    for a, b, cc, d in quads:
        tris += [c[a], c[b], c[cc], c[a], c[cc], c[d]]
    return np.asarray(tris, dtype=np.float32).reshape(-1, 9)


# ---------------------------------------------------------------- cloth
@dataclass
class ClothBuildData:
    """Solver inputs of one cloth (fields of xpbd_core.py:87-107 that reach the ABI)."""
    positions_world: np.ndarray
    inv_mass: np.ndarray
    triangles: np.ndarray
    edges: np.ndarray
    edge_rest_lengths: np.ndarray
    edge_color_offsets: np.ndarray
    bends: np.ndarray
    bend_rest_lengths: np.ndarray
    bend_color_offsets: np.ndarray
    lra_edges: np.ndarray = field(default_factory=lambda: np.empty((0, 2), np.int32))
    lra_rest_lengths: np.ndarray = field(default_factory=lambda: np.empty(0, np.float32))
    lra_color_offsets: np.ndarray = field(default_factory=lambda: np.empty(0, np.int32))
    pin_indices: np.ndarray = field(default_factory=lambda: np.empty(0, np.int32))
    pin_targets_world: np.ndarray = field(default_factory=lambda: np.empty((0, 3), np.float32))
    pin_weights: np.ndarray = field(default_factory=lambda: np.empty(0, np.float32))
    start_positions: np.ndarray | None = None   # jittered start uploaded with ssbl_update_positions


@dataclass
class SolverOptions:
    """The SsblXpbdConfig fields a caller chooses (xpbd_core.py:45-84)."""
    dt: float = 1.0 / 60.0
    damping: float = 1.0
    gravity: tuple = (0.0, 0.0, -9.81)
    stretch_compliance: float = 0.0
    stretch_optimization_enabled: bool = False
    stretch_optimization_strength: float = 0.0
    bend_compliance: float = 1.0e-4
    lra_compliance: float = 0.0
    collision_margin: float = 0.005
    use_ground: bool = False
    ground_height: float = 0.0
    use_wall: bool = False
    wall_origin: tuple = (0.0, 0.0, 0.0)
    wall_normal: tuple = (0.0, 0.0, 1.0)
    use_sphere: bool = False
    sphere_center: tuple = (0.0, 0.0, 0.0)
    sphere_radius: float = 0.0
    self_collision: bool = False
    self_collision_mode: int = 2
    cloth_thickness: float = 0.005
    self_collision_distance: float = 0.0
    self_collision_interval: int = 1
    max_self_collision_neighbors: int = 32
    fast_self_collision_passes: int = 4
    use_volume_pressure: bool = False
    volume_compliance: float = 0.0
    pressure_strength: float = 0.0
    volume_target_scale: float = 1.0
    volume_solve_interval: int = 1
    self_probe_interval: int = 0
    self_surface_pair_interval: int = 0
    jitter_stabilizer_enabled: bool = False
    contact_friction: float = 0.3
    contact_tangent_damping: float = 0.0
    contact_compliance: float = 0.0
    static_sdf_voxel_size: float = 0.0
    static_sdf_band_voxels: int = 4
    static_sdf_max_resolution: int = 160
    rest_volume: float = 0.0


def build_cloth(rest_z_up: np.ndarray, triangles: np.ndarray, pin_mask: np.ndarray, *, density: float = 1.0,
                pin_weights: np.ndarray | None = None, jitter: float = 0.0, seed: int = 0) -> ClothBuildData:
    rest = np.ascontiguousarray(rest_z_up, dtype=np.float32)
    n = len(rest)
    tris = np.ascontiguousarray(triangles, dtype=np.int32)
    edges, edge_rest = edge_constraints(tris, rest)
    edges, edge_rest, edge_off = color_constraints(edges, edge_rest, n)
    bends, bend_rest = bend_constraints(tris, rest)
    bends, bend_rest, bend_off = color_constraints(bends, bend_rest, n)
    pin_mask = np.asarray(pin_mask, bool)
    inv_mass = vertex_inverse_mass(rest, tris, density, pin_mask)
    pins = np.flatnonzero(pin_mask).astype(np.int32)
    weights = np.ones(len(pins), np.float32) if pin_weights is None else np.asarray(pin_weights, np.float32)
    start = None
    if jitter > 0.0:
        rng = np.random.Generator(np.random.PCG64(seed))
        start = (rest.astype(np.float64) + rng.normal(0.0, jitter, size=rest.shape)).astype(np.float32)
    return ClothBuildData(
        positions_world=rest, inv_mass=inv_mass, triangles=tris, edges=edges, edge_rest_lengths=edge_rest,
        edge_color_offsets=edge_off, bends=bends, bend_rest_lengths=bend_rest, bend_color_offsets=bend_off,
        pin_indices=pins, pin_targets_world=np.ascontiguousarray(rest[pins]), pin_weights=weights, start_positions=start)


# --------------------------------------------------------------- scenes
@dataclass
class Scene:
    name: str
    cloths: list
    options: list
    static_triangles: np.ndarray
    substeps: int
    iterations: int
    frames: int
    layers: list | None = None          # collision layer per cloth when cross-cloth collision is on
    contact_distance: float = 0.0       # particle radius and scene margin of the cross-cloth scene

    @property
    def vertex_count(self) -> int:
        return sum(len(c.positions_world) for c in self.cloths)

    def vertex_substeps(self, frames: int | None = None) -> int:
        return self.vertex_count * self.substeps * (self.frames if frames is None else frames)


def grid_cloth(n: int, size: float, height: float, *, pin_corners: bool = True, seed: int = 0, jitter: float = 1.0e-4,
               offset=(0.0, 0.0)) -> ClothBuildData:
    pos, tris = grid_mesh(n, size)
    pos[:, 1] += height
    pos[:, 0] += offset[0]
    pos[:, 2] += offset[1]
    pin = np.zeros(n * n, bool)
    if pin_corners:
        pin[n * (n - 1)] = True
        pin[n * n - 1] = True
    return build_cloth(y_up_to_z_up(pos), tris, pin, jitter=jitter, seed=seed)


def config1(seed: int = 0, n: int = 32, frames: int = 120) -> Scene:
    """Plain XPBD: 32x32 grid, 1 m square at 1 m, two corners pinned, ground with friction 0.3."""
    cloth = grid_cloth(n, 1.0, 1.0, seed=seed)
    opts = SolverOptions(use_ground=True, ground_height=0.0, contact_friction=0.3, contact_tangent_damping=0.0,
                         stretch_compliance=0.0, bend_compliance=1.0e-4, collision_margin=0.005, cloth_thickness=0.005)
    return Scene("plain_32x32", [cloth], [opts], np.empty((0, 9), np.float32), 10, 10, frames)


def config2(seed: int = 0, n: int = 100, frames: int = 240) -> Scene:
    """10k cloth with strict self collision, dropped from 1.5 m on a sphere above a static box."""
    cloth = grid_cloth(n, 2.0, 1.5, pin_corners=False, seed=seed)
    box = box_triangles((0.0, 0.0, 0.3), 0.3)
    opts = SolverOptions(use_ground=True, use_sphere=True, sphere_center=(0.0, 0.0, 1.0), sphere_radius=0.4,
                         contact_friction=0.3, collision_margin=0.005, cloth_thickness=0.005, self_collision=True,
                         self_collision_mode=2, self_collision_distance=0.01 / 1.8, self_collision_interval=1,
                         fast_self_collision_passes=4, max_self_collision_neighbors=32, static_sdf_max_resolution=128)
    return Scene("selfcol_100x100", [cloth], [opts], box, 20, 5, frames)


def config3(seed: int = 0, count: int = 64, frames: int = 120, self_collision: bool = False,
            cross_cloth: bool = True) -> Scene:
    """64 uneven grid cloths (sides log-uniform in [16, 200]) over a shared ground and box.

    The cloths stand in stacks of four (collision layers 0-3, 10 cm apart in
    height) on a 4 m grid, so with cross-cloth collision on (lower layers
    only) each cloth falls onto the ones below it.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    sides = np.rint(np.exp(rng.uniform(math.log(16), math.log(200), size=count))).astype(int)
    cloths, opts, layers = [], [], []
    spacing = 0.02
    stacks = (count + 3) // 4
    cols = int(math.ceil(math.sqrt(stacks)))
    pitch = 200 * spacing + 0.5
    for k, side in enumerate(sides.tolist()):
        size = (side - 1) * spacing
        stack, layer = k // 4, k % 4
        ox = (stack % cols - cols / 2) * pitch
        oz = (stack // cols - cols / 2) * pitch
        pos, tris = grid_mesh(side, size)
        pos[:, 1] += 1.0 + 0.1 * layer
        pos[:, 0] += ox
        pos[:, 2] += oz
        pin = np.zeros(side * side, bool)
        segs = int(rng.integers(1, 4))
        for _ in range(segs):
            edge = int(rng.integers(0, 4))
            a, b = sorted(rng.integers(0, side, size=2).tolist())
            idx = np.arange(a, b + 1)
            if edge == 0:
                pin[idx] = True
            elif edge == 1:
                pin[(side - 1) * side + idx] = True
            elif edge == 2:
                pin[idx * side] = True
            else:
                pin[idx * side + side - 1] = True
        cloths.append(build_cloth(y_up_to_z_up(pos), tris, pin, jitter=1.0e-4, seed=seed * 1000 + k))
        opts.append(SolverOptions(use_ground=True, contact_friction=0.3, collision_margin=0.005, cloth_thickness=0.005,
                                  self_collision=self_collision, static_sdf_max_resolution=128))
        layers.append(layer)
    box = box_triangles((0.0, 0.0, 0.3), 0.3)
    return Scene("batched_64_uneven", cloths, opts, box, 10, 8, frames, layers if cross_cloth else None,
                 0.005 if cross_cloth else 0.0)


def config5(seed: int = 0, subdivisions: int = 8, towel_n: int = 1000, frames: int = 120) -> Scene:
    """Large closed-mesh stress case: an icosphere balloon (subdivision 8:
    655,362 vertices, 1,310,720 triangles, r = 1 m) with volume pressure 1.2,
    and a towel_n x towel_n grid towel (2 m square) twisted by its two end rows,
    which are pinned and turn in opposite directions at 1 rev/s about the
    towel's long axis (`towel_pin_targets`).  Strict self collision on both,
    hard-stretch PCG at strength 1.0 (polish and caps active), 20 substeps x 10
    iterations, seed-jittered starts."""
    pos, tris = icosphere(subdivisions, 1.0)
    pos = pos + np.array([3.0, 0.0, 1.5])
    balloon = build_cloth(pos, tris, np.zeros(len(pos), bool), jitter=1.0e-4, seed=seed * 1000 + 1)
    tpos, ttris = grid_mesh(towel_n, 2.0)
    tpos[:, 1] += 1.5
    rows = np.arange(towel_n * towel_n) // towel_n
    ends = (rows == 0) | (rows == towel_n - 1)
    towel = build_cloth(y_up_to_z_up(tpos), ttris, ends, jitter=1.0e-4, seed=seed * 1000 + 2)
    common = dict(collision_margin=0.005, cloth_thickness=0.005, self_collision=True, self_collision_mode=2,
                  self_collision_interval=1, fast_self_collision_passes=4, max_self_collision_neighbors=32,
                  stretch_optimization_enabled=True, stretch_optimization_strength=1.0)
    o_balloon = SolverOptions(use_volume_pressure=True, pressure_strength=1.2, **common)
    o_towel = SolverOptions(**common)
    return Scene("stress_balloon_towel", [balloon, towel], [o_balloon, o_towel], np.empty((0, 9), np.float32), 20, 10, frames)


def towel_pin_targets(towel: ClothBuildData, t: float) -> np.ndarray:
    """Targets of the towel's pinned end rows at time t: each row turns about
    the towel's long axis (through the two rows' centres) at 1 rev/s, the two
    rows in opposite directions."""
    rest = towel.positions_world[towel.pin_indices].astype(np.float64)
    centre = towel.positions_world.astype(np.float64).mean(axis=0)
    axis = np.array([0.0, 1.0, 0.0])   # rows run along x; the long axis is y (z-up world)
    sign = np.where(rest[:, 1] < centre[1], 1.0, -1.0)
    ang = 2.0 * math.pi * t * sign
    rel = rest - centre
    c, s_ = np.cos(ang), np.sin(ang)
    # rotation about the y axis
    x = c * rel[:, 0] + s_ * rel[:, 2]
    z = -s_ * rel[:, 0] + c * rel[:, 2]
    out = np.stack([x, rel[:, 1], z], 1) + centre
    return out.astype(np.float32)

