"""Dynamic collider scenes (triangle soups, particles, a global scene of
several cloths) shared by the oracle tests and the GPU parity tests."""
import ctypes

import numpy as np

from raw_abi import Raw
from ssbl_xpbd_amd import abi


def plane_soup(z, half=1.0, n=1):
    """An n x n grid of quads (2 n^2 triangles) covering [-half, half]^2 at height z."""
    s = np.linspace(-half, half, n + 1)
    tris = []
    for i in range(n):
        for j in range(n):
            a = (s[i], s[j], z)
            b = (s[i + 1], s[j], z)
            c = (s[i + 1], s[j + 1], z)
            d = (s[i], s[j + 1], z)
            tris.append([a, b, c])
            tris.append([a, c, d])
    return np.asarray(tris, np.float32).reshape(-1, 9)


def square(lib, margin=0.01):
    return Raw(lib, [[-0.2, -0.2, 0.0], [0.2, -0.2, 0.0], [0.2, 0.2, 0.0], [-0.2, 0.2, 0.0]], [1.0] * 4,
               collision_margin=margin, cloth_thickness=0.005)


def plane_contact(lib, quads_per_side, depth=0.004, margin=0.01):
    """Four free vertices at z = 0 over a dynamic plane at z = -depth: each is
    pushed up by half the penetration, margin - depth (A41:4310-4322)."""
    r = square(lib, margin)
    soup = plane_soup(-depth, 1.0, quads_per_side)
    r.check(lib.ssbl_update_dynamic_triangles(r.h, soup.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), len(soup)))
    r.step(1, 1)
    return r.state(), r.diag()


def valley_soup(left_first, fillers=64):
    """A V-shaped valley along y with its crease on x = 0: a left face rising
    to x = -0.1, z = 0.1 and its mirror image, two triangles each (the left
    ones first when `left_first`), then far-away filler triangles so that the
    set is hashed (64 triangles or more)."""
    def face(sx):
        a, b, c, d = (sx * 0.1, -0.3, 0.1), (0.0, -0.3, 0.0), (0.0, 0.3, 0.0), (sx * 0.1, 0.3, 0.1)
        return [[a, b, c], [a, c, d]]
    faces = face(-1.0) + face(1.0) if left_first else face(1.0) + face(-1.0)
    fill = [[(50.0 + k, 0.0, 0.0), (50.1 + k, 0.0, 0.0), (50.0 + k, 0.1, 0.0)] for k in range(fillers)]
    return np.asarray(faces + fill, np.float32).reshape(-1, 9)


def valley_tie(lib, count, left_first, height=0.008, margin=0.01):
    """`count` free vertices on the valley's mirror plane x = 0 at `height`.
    Each is exactly as deep in both faces (mirror images give the same
    distance bit for bit) with pushes of opposite x: the tie goes to the
    triangle first in the set, so every vertex moves to +x when the left face
    comes first (its push is along (1, 0, 1) / sqrt 2) and to -x otherwise.
    Four vertices give a wave one round of pooled (vertex, entry) items; 64
    give it several rounds, each requesting the next round's entries early."""
    ys = np.linspace(-0.2, 0.2, count)
    pts = [[0.0, float(y), height] for y in ys]
    r = Raw(lib, pts, [1.0] * count, collision_margin=margin, cloth_thickness=0.005)
    soup = valley_soup(left_first)
    r.check(lib.ssbl_update_dynamic_triangles(r.h, soup.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), len(soup)))
    r.step(1, 1)
    return r.state(), r.diag()


def particle_contact(lib, far_particles=0, distance=0.01, radius=0.008, margin=0.005):
    """A free vertex over one particle (plus far-away ones): pushed up by half of
    radius + margin - distance (A41:4504-4507)."""
    r = Raw(lib, [[0.0, 0.0, 0.0]], [1.0], collision_margin=margin, cloth_thickness=0.005)
    pos = [[0.0, 0.0, -distance]] + [[10.0 + 0.5 * k, 3.0, 0.0] for k in range(far_particles)]
    p = np.asarray(pos, np.float32)
    rad = np.full(len(p), radius, np.float32)
    fi = abi.FrameInputs()
    fi.update_dynamic_particles = 1
    fi.dynamic_particle_positions = p.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    fi.dynamic_particle_radii = rad.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    fi.dynamic_particle_count = len(p)
    r.check(lib.ssbl_update_frame_inputs(r.h, ctypes.byref(fi)))
    r.step(1, 1)
    return r.state(), r.diag()


def wavy_soup(n=24, half=0.8, z0=0.6, amp=0.08, seed=0):
    rng = np.random.Generator(np.random.PCG64(seed))
    s = np.linspace(-half, half, n + 1)
    x, y = np.meshgrid(s, s, indexing="ij")
    z = z0 + amp * np.sin(3.0 * x) * np.cos(2.0 * y) + rng.normal(0.0, 0.002, size=x.shape)
    tris = []
    for i in range(n):
        for j in range(n):
            a = (x[i, j], y[i, j], z[i, j])
            b = (x[i + 1, j], y[i + 1, j], z[i + 1, j])
            c = (x[i + 1, j + 1], y[i + 1, j + 1], z[i + 1, j + 1])
            d = (x[i, j + 1], y[i, j + 1], z[i, j + 1])
            tris.append([a, b, c])
            tris.append([a, c, d])
    return np.asarray(tris, np.float32).reshape(-1, 9)


def particle_cloud(count=400, seed=1):
    rng = np.random.Generator(np.random.PCG64(seed))
    p = np.stack([rng.uniform(-0.5, 0.5, count), rng.uniform(-0.5, 0.5, count), rng.uniform(0.45, 0.75, count)], 1)
    return p.astype(np.float32), rng.uniform(0.01, 0.03, count).astype(np.float32)


def cloth_over_colliders(nb, lib, scenes, *, triangles=True, particles=True, soup_n=24, n_particles=400):
    """A 20 x 20 cloth falling onto a wavy triangle soup and a particle cloud."""
    cloth = scenes.grid_cloth(20, 1.0, 1.0, pin_corners=False, seed=4)
    opts = scenes.SolverOptions(collision_margin=0.01, cloth_thickness=0.005)
    s = nb.NativeXpbdSolver(cloth, opts, None, lib=lib)
    if triangles:
        s.update_dynamic_triangles(wavy_soup(soup_n))
    if particles:
        p, r = particle_cloud(n_particles)
        s.update_dynamic_particles(p, r)
    return s


def two_cloth_scene(nb, lib, scenes, mode):
    """Cloth A (source 1, layer 0) below cloth B (source 2, layer 1); both
    attached to one global scene refreshed from their positions every frame."""
    pos, tris = scenes.grid_mesh(16, 0.8)
    pos[:, 1] += 0.5
    pin = np.zeros(16 * 16, bool)
    pin[[0, 15, 240, 255]] = True                      # a trampoline: four corners held
    cloth_a = scenes.build_cloth(scenes.y_up_to_z_up(pos), tris, pin, jitter=1.0e-4, seed=1)
    a = nb.NativeXpbdSolver(cloth_a, scenes.SolverOptions(collision_margin=0.01, cloth_thickness=0.005), None, lib=lib)
    b = nb.NativeXpbdSolver(scenes.grid_cloth(10, 0.3, 0.58, pin_corners=False, seed=2),
                            scenes.SolverOptions(collision_margin=0.01, cloth_thickness=0.005), None, lib=lib)
    scene = nb.GlobalDynamicScene(lib=lib)
    a.attach_global_dynamic_scene(scene, 1, 0, mode)
    b.attach_global_dynamic_scene(scene, 2, 1, mode)
    return a, b, scene


def refresh_scene(scene, solvers, device_gather):
    ids, layers, radii = [1, 2], [0, 1], [0.01, 0.01]
    if device_gather:
        scene.update_from_solvers(solvers, ids, layers, radii, 0.01, 0.01)
        return
    pos = [s.download_positions() for s in solvers]
    tris = _triangles_of(solvers, pos)
    scene.update(particle_positions=np.concatenate(pos), particle_radii=np.concatenate([np.full(len(p), r, np.float32) for p, r in zip(pos, radii)]),
                 particle_source_ids=np.concatenate([np.full(len(p), i, np.int32) for p, i in zip(pos, ids)]),
                 particle_collision_layers=np.concatenate([np.full(len(p), l, np.int32) for p, l in zip(pos, layers)]),
                 triangle_vertices=np.concatenate([t for t, _ in tris]),
                 triangle_source_ids=np.concatenate([np.full(len(t), i, np.int32) for (t, _), i in zip(tris, ids)]),
                 triangle_collision_layers=np.concatenate([np.full(len(t), l, np.int32) for (t, _), l in zip(tris, layers)]),
                 collision_margin=0.01, cloth_thickness=0.01)


def _triangles_of(solvers, pos):
    out = []
    for s, p in zip(solvers, pos):
        t = s.triangles
        out.append((p[t].reshape(-1, 9), t))
    return out


def box_sdf_grid(n=128, extent=2.0, center=(0.0, 0.0, 0.3), half=0.3):
    """An n^3 FP32 signed distance grid of an axis-aligned box over a cube of
    side `extent` centred on the box (values [z, y, x], negative inside)."""
    s = np.linspace(-extent / 2, extent / 2, n)
    voxel = float(s[1] - s[0])
    origin = (center[0] - extent / 2, center[1] - extent / 2, center[2] - extent / 2)
    z, y, x = np.meshgrid(s + center[2], s + center[1], s + center[0], indexing="ij")
    q = np.stack([np.abs(x - center[0]), np.abs(y - center[1]), np.abs(z - center[2])], -1) - half
    outside = np.linalg.norm(np.maximum(q, 0.0), axis=-1)
    inside = np.minimum(np.max(q, axis=-1), 0.0)
    return (outside + inside).astype(np.float32), origin, voxel


def cloth_on_sdf_grid(nb, lib, scenes, grid, frames=20):
    values, origin, voxel = grid
    cloth = scenes.grid_cloth(20, 0.8, 0.9, pin_corners=False, seed=6)
    opts = scenes.SolverOptions(collision_margin=0.01, cloth_thickness=0.005, use_ground=True)
    s = nb.NativeXpbdSolver(cloth, opts, None, lib=lib)
    s.set_static_sdf(values, origin, voxel)
    out = []
    for _ in range(frames):
        s.step(8, 4, diagnostics=True)
        out.append(s.download_positions())
    return out, s.diagnostics()


# --- signed (indexed) dynamic triangles, A41:8433-8590 and 4520-4540
def indexed_plane(z, half=1.0, n=1):
    """An n x n grid of quads at height z as indexed vertices (normals +z)."""
    s = np.linspace(-half, half, n + 1)
    x, y = np.meshgrid(s, s, indexing="ij")
    verts = np.stack([x.ravel(), y.ravel(), np.full(x.size, z)], 1).astype(np.float32)
    idx = []
    for i in range(n):
        for j in range(n):
            a, b, c, d = i * (n + 1) + j, (i + 1) * (n + 1) + j, (i + 1) * (n + 1) + j + 1, i * (n + 1) + j + 1
            idx += [[a, b, c], [a, c, d]]
    return verts, np.asarray(idx, np.int32)


def upload_indexed(lib, r, verts, idx, topology_changed=0):
    verts = np.ascontiguousarray(verts, np.float32)
    idx = np.ascontiguousarray(idx, np.int32)
    fi = abi.FrameInputs()
    fi.update_dynamic_indexed_triangles = 1
    fi.dynamic_triangle_vertices = verts.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    fi.dynamic_triangle_vertex_count = len(verts)
    fi.dynamic_triangle_indices = idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    fi.dynamic_indexed_triangle_count = len(idx)
    fi.dynamic_triangle_topology_changed = topology_changed
    r.check(lib.ssbl_update_frame_inputs(r.h, ctypes.byref(fi)))


SIGNED_MARGIN, SIGNED_THICKNESS = 0.005, 0.004   # contact margin max(0.005, 0.75 x 0.004) = 0.005


def signed_plane_contact(lib, height, quads_per_side, moved=None):
    """A free vertex at `height` over an indexed plane at z = 0 (no gravity),
    one substep of one iteration, so the pass after self collision and the
    six inside-only passes all run.  With `moved`, the plane was uploaded
    before at z = -moved, so it rises by `moved` over the frame."""
    r = Raw(lib, [[0.01, 0.02, height]], [1.0], collision_margin=SIGNED_MARGIN, cloth_thickness=SIGNED_THICKNESS,
            gravity=(0.0, 0.0, 0.0))
    if moved is not None:
        v0, idx = indexed_plane(-moved, 1.0, quads_per_side)
        upload_indexed(lib, r, v0, idx)
    verts, idx = indexed_plane(0.0, 1.0, quads_per_side)
    upload_indexed(lib, r, verts, idx)
    r.step(1, 1)
    return r.state(), r.diag(), r


def signed_plane_expected(height, cross_len=4.0):
    """Closed form of A41:995-1068 for the scene above (plane normal +z): the
    first pass pushes by a third of (margin - signed distance); the six
    inside-only passes each push by a third of (1 mm - signed distance) while
    the vertex is more than 1 mm below that.

    `cross_len` is |(b - a) x (c - a)| of the plane's triangles.  The
    reference divides the normal by sqrt(max(|n|^2, 1e-7)) (A41:4073-4075), so
    for triangles with |n| < sqrt(1e-7) both the signed distance and the push
    shrink by s = |n| / sqrt(1e-7)."""
    sc = min(1.0, cross_len / np.sqrt(np.float32(1.0e-7)))
    z = float(height)
    pen = SIGNED_MARGIN - sc * z
    if pen > 0.0:
        z += sc * pen / 3.0
    for _ in range(6):
        pen = 0.001 - sc * z
        if pen > 0.0:
            z += sc * pen / 3.0
    return z


def cloth_over_indexed_collider(nb, lib, scenes, frames=20, particles=True, batched=False):
    """A 24 x 24 cloth dropped onto a rising indexed dome (a bumpy grid) that
    is re-uploaded every frame, plus a few dynamic particles: the signed
    passes with collider velocity, local averaging, and inside-only passes."""
    cloth = scenes.grid_cloth(24, 0.8, 0.25, pin_corners=False, seed=3)
    opts = scenes.SolverOptions(collision_margin=SIGNED_MARGIN, cloth_thickness=SIGNED_THICKNESS, contact_friction=0.3)
    s = nb.NativeXpbdSolver(cloth, opts, None, lib=lib)
    n = 16
    g = np.linspace(-0.7, 0.7, n + 1)
    x, y = np.meshgrid(g, g, indexing="ij")
    idx = indexed_plane(0.0, 0.7, n)[1]
    p = np.stack([np.linspace(-0.2, 0.2, 5), np.zeros(5), np.full(5, 0.12)], 1).astype(np.float32)
    rad = np.full(5, 0.03, np.float32)
    out = []
    for f in range(frames):
        z = 0.05 + 0.004 * f + 0.04 * np.exp(-4.0 * (x ** 2 + y ** 2))
        verts = np.stack([x.ravel(), y.ravel(), z.ravel()], 1).astype(np.float32)
        fi = abi.FrameInputs()
        fi.update_dynamic_indexed_triangles = 1
        fi.dynamic_triangle_vertices = verts.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        fi.dynamic_triangle_vertex_count = len(verts)
        fi.dynamic_triangle_indices = idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        fi.dynamic_indexed_triangle_count = len(idx)
        if particles:
            pf = p + np.float32(0.002 * f) * np.array([1.0, 0.0, 0.0], np.float32)
            fi.update_dynamic_particles = 1
            fi.dynamic_particle_positions = pf.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
            fi.dynamic_particle_radii = rad.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
            fi.dynamic_particle_count = len(pf)
        if not lib.ssbl_update_frame_inputs(s.handle, ctypes.byref(fi)):
            raise nb.NativeSolverError(nb.last_error(lib))
        if batched:
            b = nb.SolverBatch([s])
            b.step(4, 3, synchronize=True)
            b.close()
        else:
            s.step(4, 3, diagnostics=True)
        out.append(s.download_state() if hasattr(s, "download_state") else s.download_positions())
    return s, out
