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
