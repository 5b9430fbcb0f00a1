"""Small scenes with known answers, run on any ABI41 library.

Each function returns what its checks need; the checks live in the test
modules (oracle on the CPU, HIP library on the GPU) so both are held to the
same answers.  The expected values come from closed forms or from the
reference's own smoke-test thresholds (native/smoke/abi41_smoke.cpp).
"""
import numpy as np

from raw_abi import Raw

G = (0.0, 0.0, -9.81)


def free_fall(lib, substeps=7, damping=0.99):
    r = Raw(lib, [[0.0, 0.0, 2.0]], [1.0], gravity=G, damping=damping)
    r.step(substeps, 3)
    return r.state()


def free_fall_closed_form(substeps=7, damping=0.99):
    h = np.float64(np.float32(np.float32(1.0 / 60.0) / np.float32(substeps)))
    x, v = 2.0, 0.0
    for _ in range(substeps):
        v = damping * (v + G[2] * h)
        x = x + v * h
    return x, v


def hard_and_weighted_pins(lib):
    # native/smoke/abi41_smoke.cpp:1171-1269
    r = Raw(lib, [[0, 0, 0], [1, 0, 0], [2, 0, 0], [3, 0, 0]], [1, 1, 1, 1], gravity=(0.0, -9.8, 0.0))
    r.pins([0, 1, 2], [[0, 1, 0], [1, 1, 0], [2, 1, 0]], [1.0, 0.375, 0.04])
    r.step(1, 3)
    return r.positions()


def soft_pin(lib, weight, iterations):
    r = Raw(lib, [[0.0, 0.0, 0.0]], [1.0])
    r.pins([0], [[1.0, 0.0, 0.0]], [weight])
    r.step(1, iterations)
    return r.state()


def single_edge(lib, start, inv_mass=(1.0, 1.0), rest=1.0):
    r = Raw(lib, [[0, 0, 0], [rest, 0, 0]], list(inv_mass), edges=[[0, 1]], edge_rest=[rest])
    r.set_positions(start)
    r.step(1, 1)
    return r.positions()


def ground_friction(lib, friction, gx=2.0, substeps=4, frames=5):
    """A vertex resting on the ground under gravity with a tangential component."""
    r = Raw(lib, [[0.0, 0.0, 0.005]], [1.0], gravity=(gx, 0.0, -9.81), use_ground=1, ground_height=0.0,
            collision_margin=0.005, contact_friction=friction, contact_tangent_damping=0.0)
    for _ in range(frames):
        r.step(substeps, 2)
    return r.state(), substeps * frames


def sphere_push(lib):
    r = Raw(lib, [[0.1, 0.05, 0.2]], [1.0], use_sphere=1, sphere_center=(0.0, 0.0, 0.0), sphere_radius=0.5,
            collision_margin=0.01)
    r.step(1, 1)
    return r.positions()


def static_plane(lib):
    # native/smoke/abi41_smoke.cpp:71-143
    plane = np.array([[-1, -1, 0, 1, -1, 0, -1, 1, 0], [1, -1, 0, 1, 1, 0, -1, 1, 0]], np.float32)
    r = Raw(lib, [[0.0, 0.0, 0.02], [0.25, 0.0, 0.02]], [1.0, 0.0], static_triangles=plane, collision_margin=0.08,
            cloth_thickness=0.02, static_sdf_voxel_size=0.04, static_sdf_band_voxels=2, static_sdf_max_resolution=48)
    r.step(1, 1)
    return r.positions(), r.diag()


def box_sdf(lib, half=0.3, res=40):
    from ssbl_xpbd_amd import scenes
    box = scenes.box_triangles((0.0, 0.0, 0.0), half)
    r = Raw(lib, [[2.0, 2.0, 2.0]], [0.0], static_triangles=box, collision_margin=0.01, static_sdf_max_resolution=res)
    return r


def bend_pair(lib, compliance=0.0):
    # folded quad: 0-1 shared edge, 2 and 3 opposite
    pos = np.array([[0, 0, 0], [1, 0, 0], [0.5, 0.8, 0], [0.5, -0.8, 0]], np.float32)
    rest = np.linalg.norm(pos[2] - pos[3])
    folded = pos.copy()
    folded[2] = [0.5, 0.4, 0.6]
    folded[3] = [0.5, -0.4, 0.6]
    r = Raw(lib, pos, [0.0, 0.0, 1.0, 1.0], bends=[[2, 3]], bend_rest=[rest], bend_offsets=[0, 1],
            bend_compliance=compliance, cloth_thickness=0.02)
    r.set_positions(folded)
    r.step(1, 1)
    return folded, r.positions(), rest


# --- self collision: the reference smoke scenes (abi41_smoke.cpp:227-597),
# with the smoke's own pass criteria.  The smoke runs them in fast mode
# (self_collision_mode 1) with one pass; mode 2 is the strict mode.
SELF_CFG = dict(self_collision=1, cloth_thickness=0.10, collision_margin=0.0, max_self_collision_neighbors=32)


def _self_cfg(mode, passes):
    return dict(SELF_CFG, self_collision_mode=mode, fast_self_collision_passes=passes)


def self_vv(lib, mode=2, passes=4):
    n = 600
    rest = np.zeros((n, 3), np.float32)
    rest[:, 0] = np.arange(n)
    r = Raw(lib, rest, np.ones(n, np.float32), **_self_cfg(mode, passes))
    close = rest.copy()
    close[n - 1, 0] = 0.03
    r.set_positions(close)
    r.step(1, 6)
    out = r.positions()
    return float(np.linalg.norm(close[0] - close[n - 1])), float(np.linalg.norm(out[0] - out[n - 1])), r.diag(), r


def self_vt(lib, mode=2, passes=4):
    n, ntri = 600, 64
    rest = np.zeros((n, 3), np.float32)
    rest[0], rest[1], rest[2] = (-1, -1, 0), (1, -1, 0), (0, 1, 0)
    inv = np.ones(n, np.float32)
    inv[:3] = 0.0
    for i in range(3, n):
        rest[i] = (20.0 + i, 20.0 + (i % 7) * 0.25, (i % 5) * 0.20)
    rest[n - 1] = (5.0, 5.0, 1.0)
    tris = [[0, 1, 2]] + [[3 + t * 3, 4 + t * 3, 5 + t * 3] for t in range(1, ntri)]
    r = Raw(lib, rest, inv, triangles=tris, **_self_cfg(mode, passes))
    close = rest.copy()
    close[n - 1] = (0.0, -0.2, 0.03)
    r.set_positions(close)
    r.step(1, 6)
    out = r.positions()
    return float(close[n - 1, 2]), float(out[n - 1, 2]), r.diag(), r


def self_ee(lib, mode=2, passes=4):
    n, ne = 600, 64
    rest = np.zeros((n, 3), np.float32)
    rest[:, 0] = 200.0 + np.arange(n) * 3.0
    rest[:, 1] = 100.0 + (np.arange(n) % 11)
    rest[0], rest[1], rest[2], rest[3] = (-1, 0, 0), (1, 0, 0), (5, -1, 0), (5, 1, 0)
    inv = np.ones(n, np.float32)
    inv[:2] = 0.0
    edges = [[0, 1], [2, 3]]
    for e in range(2, ne):
        a = 4 + (e - 2) * 2
        rest[a] = (50.0 + e * 4.0, 30.0, 0.0)
        rest[a + 1] = (51.0 + e * 4.0, 30.0, 0.0)
        edges.append([a, a + 1])
    edges = np.asarray(edges, np.int32)
    lengths = np.linalg.norm(rest[edges[:, 0]] - rest[edges[:, 1]], axis=1).astype(np.float32)
    r = Raw(lib, rest, inv, edges=edges, edge_rest=lengths, **_self_cfg(mode, passes))
    close = rest.copy()
    close[2] = (0.0, -1.0, 0.03)
    close[3] = (0.0, 1.0, 0.03)
    r.set_positions(close)
    r.step(1, 6)
    out = r.positions()
    return float((close[2, 2] + close[3, 2]) / 2), float((out[2, 2] + out[3, 2]) / 2), r.diag(), r


def crumple(scenes, n=40, frames=25, seed=1, self_distance=0.01 / 1.8, mode=2, max_neighbors=32):
    """A grid dropped onto a small sphere so it folds over itself."""
    cloth = scenes.grid_cloth(n, 0.8, 0.5, pin_corners=False, seed=seed)
    opts = scenes.SolverOptions(use_ground=True, use_sphere=True, sphere_center=(0.0, 0.0, 0.2), sphere_radius=0.12,
                                contact_friction=0.3, cloth_thickness=0.005, self_collision=True, self_collision_mode=mode,
                                self_collision_distance=self_distance, fast_self_collision_passes=4,
                                max_self_collision_neighbors=max_neighbors)
    return cloth, opts, frames


def folded_cloth(scenes, n=30, size=0.6, gap=0.004, seed=2, mode=2, passes=1):
    """A grid whose upper half starts folded onto the lower half, `gap` above it."""
    cloth = scenes.grid_cloth(n, size, 0.3, pin_corners=False, seed=seed, jitter=0.0)
    p = cloth.positions_world.astype(np.float64).copy()
    h = size / (n - 1)
    # rows are along scene z (solver -y); fold rows >= n/2 back over the first half
    rows = np.arange(n * n) // n
    upper = rows >= n // 2
    mirror_row = (n - 1) - rows[upper]
    p[upper, 1] = -(mirror_row * h - size / 2)
    p[upper, 2] += gap
    rng = np.random.Generator(np.random.PCG64(seed))
    cloth.start_positions = (p + rng.normal(0, 1e-4, p.shape)).astype(np.float32)
    opts = scenes.SolverOptions(use_ground=True, contact_friction=0.3, cloth_thickness=0.005, self_collision=True,
                                self_collision_mode=mode, self_collision_distance=0.01 / 1.8, fast_self_collision_passes=passes,
                                max_self_collision_neighbors=32, gravity=(0.0, 0.0, 0.0))
    return cloth, opts


def vv_clusters(lib, mode, same_cluster):
    """One free vertex s with two pinned neighbours inside the contact
    thickness (0.1 m).  Each neighbour pushes s by 0.035 m (the 0.35 x
    thickness clamp) along their separation.  With same_cluster the two
    contact midpoints share a 0.25 m cluster cell and normal bucket, so the
    fast mode keeps only the first neighbour's push; otherwise (and in the
    strict mode) s moves by the mean of both pushes (A41:4816-4951, 793-829)."""
    start = np.array([[0.1, 0.1, 0.1], [0.15, 0.11, 0.1], [0.15, 0.09, 0.1] if same_cluster else [0.05, 0.1, 0.1]],
                     np.float32)
    # pinned vertices hold their rest positions; s rests far from both
    rest = start.copy()
    rest[0] = (5.0, 5.0, 5.0)
    inv = np.array([1.0, 0.0, 0.0], np.float32)
    r = Raw(lib, rest, inv, self_collision=1, self_collision_mode=mode, fast_self_collision_passes=1,
            self_collision_distance=0.1, cloth_thickness=0.1, collision_margin=0.0, max_self_collision_neighbors=32,
            gravity=(0.0, 0.0, 0.0))
    r.set_positions(start)
    r.step(1, 1)
    return start, r.positions(), r.diag()


def vv_clusters_expected(start, mode, same_cluster):
    s = start[0].astype(np.float64)
    pushes = []
    for o in (1, 2):
        d = s - start[o].astype(np.float64)
        pushes.append(d / np.linalg.norm(d) * 0.035)
    keep = pushes[:1] if (mode == 1 and same_cluster) else pushes
    return s + np.mean(keep, axis=0)


def overlap_island(lib, crossing, copies=3, r=0.03, drop=0.02, triangles=64):
    """A free vertex falls through `copies` coincident triangles (0.6 m wide,
    centroid at the origin, z = 0) in one substep of 1/60 s and ends `drop`
    lower, inside the contact radius r below them.  Gravity makes the fall
    exactly `drop` up to float rounding (x1 = x0 + g h^2, no damping).

    The triangles span more than 64 cells of the 1.8 r triangle hash, so each
    is hashed at its centroid cell only (A41:871-900): the overlap-island
    search over the 27 cells around the vertex finds them, the
    vertex-triangle query of the vertex's own cell (one cell lower) does not,
    and the vertex-vertex table is empty, so only the island moves the
    vertex.  With `crossing` it started above the triangles; otherwise below
    and deeper than r/4, which the island test skips (A41:5371-5374).  The
    remaining far-away triangles bring the count to the 64 that enable the
    triangle hash."""
    h = np.float32(np.float32(1.0) / np.float32(60.0))
    g = -drop / float(h) ** 2
    z0 = 0.004 if crossing else -0.008
    tri = [(-0.3, -0.2, 0.0), (0.3, -0.2, 0.0), (0.0, 0.4, 0.0)]
    pos = [(0.0, 0.0, z0)] + tri * copies
    tris = [[1 + 3 * k, 2 + 3 * k, 3 + 3 * k] for k in range(copies)]
    for t in range(copies, triangles):
        b = len(pos)
        pos += [(20.0 + 3 * t, 20.0, 0.0), (21.0 + 3 * t, 20.0, 0.0), (20.0 + 3 * t, 21.0, 0.0)]
        tris.append([b, b + 1, b + 2])
    pos = np.asarray(pos, np.float32)
    rest = pos.copy()
    rest[0] = (10.0, 10.0, 10.0)   # far from the triangles at rest: no rest-neighbour skip
    inv = np.zeros(len(pos), np.float32)
    inv[0] = 1.0
    raw = Raw(lib, rest, inv, triangles=tris, self_collision=1, self_collision_mode=1, fast_self_collision_passes=1,
              self_collision_distance=r, cloth_thickness=r, collision_margin=0.0, max_self_collision_neighbors=16,
              gravity=(0.0, 0.0, g))
    raw.set_positions(pos)
    raw.step(1, 1)
    return raw.positions(), raw.diag()


def overlap_island_expected(crossing, copies=3, r=0.03, drop=0.02):
    """Closed form of A41:5396-5417 for the scene above: every contact has
    depth r - d along -z; the vertex moves by min(0.5 x depth x share, 0.3 r)
    with share = w_v / (w_v + mean triangle weight at the contact point) = 1
    (pinned triangles weigh 0)."""
    z0 = 0.004 if crossing else -0.008
    z1 = z0 - drop
    if not crossing or copies < 3:
        return z1, 0.0
    depth = r - abs(z1)
    dl = min(depth * 0.5, 0.3 * r)
    return z1 - dl, dl


# --- hard-stretch PCG
def pcg_chain(lib, strength=0.4, start=None, inv_mass=(0.0, 1.0, 1.0), rests=(1.0, 1.0)):
    """A pinned-start chain, stretched, one substep of one iteration (no gravity)."""
    n = len(inv_mass)
    rest_pos = np.zeros((n, 3), np.float32)
    rest_pos[:, 0] = np.concatenate([[0.0], np.cumsum(rests)])
    edges = [[k, k + 1] for k in range(n - 1)]
    r = Raw(lib, rest_pos, list(inv_mass), edges=edges, edge_rest=list(rests), stretch_optimization_enabled=1,
            stretch_optimization_strength=strength, stretch_compliance=1.0e6)
    if start is not None:
        r.set_positions(start)
    r.step(1, 1)
    return r.state(), r.diag(), r


def pcg_exact_solution(start, inv_mass, rests, strength):
    """Dense FP64 solve of the block system I + sum k n n^T (A41:3085-3226)."""
    p = np.asarray(start, np.float64)
    n = len(p)
    s = min(max(strength * 2.5, 0.0), 1.0)
    k = max(s, 1e-3)
    A = np.eye(3 * n)
    b = np.zeros(3 * n)
    w = np.array(inv_mass, np.float64)
    for e in range(n - 1):
        i, j = e, e + 1
        d = p[j] - p[i]
        L = np.linalg.norm(d)
        nrm = d / L
        err = L - rests[e]
        limit = max(1e-5, min(max(rests[e], 0.0) * 0.5, 0.25))
        proj = max(-limit, min(err * s, limit))
        ws = w[i] + w[j]
        desired = nrm * proj / ws
        blk = k * np.outer(nrm, nrm)
        if w[i] > 0:
            b[3 * i:3 * i + 3] += desired * w[i]
            A[3 * i:3 * i + 3, 3 * i:3 * i + 3] += blk
        if w[j] > 0:
            b[3 * j:3 * j + 3] -= desired * w[j]
            A[3 * j:3 * j + 3, 3 * j:3 * j + 3] += blk
        if w[i] > 0 and w[j] > 0:
            A[3 * i:3 * i + 3, 3 * j:3 * j + 3] -= blk
            A[3 * j:3 * j + 3, 3 * i:3 * i + 3] -= blk
    free = np.repeat(w > 0, 3)
    x = np.zeros(3 * n)
    x[free] = np.linalg.solve(A[np.ix_(free, free)], b[free])
    return p + x.reshape(n, 3)


# --- the large closed-mesh stress case (balloon + twisted towel)
def run_stress(nb, lib, scenes, sc, frames, *, batched):
    """Steps the balloon and the towel `frames` frames, turning the towel's end
    rows before every frame; batched (one launch per phase for both) or each
    handle alone.  Returns both cloths' states and diagnostics."""
    solvers = [nb.NativeXpbdSolver(c, o, None, lib=lib) for c, o in zip(sc.cloths, sc.options)]
    towel = sc.cloths[1]
    batch = nb.SolverBatch(solvers) if batched else None
    dt = sc.options[1].dt
    for f in range(frames):
        print(f"stress frame {f} ({'batched' if batched else 'alone'}, {sc.vertex_count} vertices)", flush=True)
        solvers[1].update_pin_targets(towel.pin_indices, scenes.towel_pin_targets(towel, (f + 1) * dt), towel.pin_weights)
        if batched:
            batch.step(sc.substeps, sc.iterations, synchronize=True)
        else:
            for s in solvers:
                s.step(sc.substeps, sc.iterations, diagnostics=False)
    if batch is not None:
        batch.close()
    states = [s.download_state() for s in solvers]
    return solvers, states
