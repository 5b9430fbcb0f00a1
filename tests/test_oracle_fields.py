"""Pins the oracle's force fields (A41:1263-1558) and overpressure
(A41:1883-1922, 7413-7584) against closed forms.  No GPU.

The observable is the one-substep velocity map of unconstrained vertices:
with damping 1 and no constraints, a step of length h gives
    v_new = v + (g + a_fields(x, v, n)) h
where a_fields is the capped sum of the fields' accelerations at the state
before the step (the velocity update re-measures (x_new - x) / h, which
returns v_new to within float rounding).  The expected accelerations below
are written from the physics each field type states -- wind pushes along its
direction in proportion to the facing of the surface, a force field has a
constant magnitude along the line from its origin, a charge falls off as 1/r^2,
a vortex turns about its axis, a harmonic spring pulls toward its rest
length with velocity damping, Lennard-Jones is r^-12 - r^-6 in units of its
radius with the ratio clamped at 6, a magnet is v x axis, drag opposes the
velocity -- and from the limits the reference states (5000 m/s^2 per field
and in total, the distance falloffs).  A mistake in a sign, an axis, a
power of r or a falloff fails one of them.
"""
import ctypes
import math

import numpy as np
import pytest

from raw_abi import Raw
from ssbl_xpbd_amd import abi

H = 0.5            # step length (s): dt = H with one substep
REL, ABS = 2.0e-5, 2.0e-6


def set_fields(r, fields):
    arr = (abi.ForceField * max(len(fields), 1))()
    for k, f in enumerate(fields):
        f = dict(f)
        f.setdefault("apply_to_location", 1)
        for key, value in f.items():
            cur = getattr(arr[k], key)
            setattr(arr[k], key, (ctypes.c_float * 3)(*value) if hasattr(cur, "__len__") else value)
    fi = abi.FrameInputs()
    fi.update_force_fields = 1
    fi.force_fields = ctypes.cast(arr, ctypes.POINTER(abi.ForceField))
    fi.force_field_count = len(fields)
    r.check(r.lib.ssbl_update_frame_inputs(r.h, ctypes.byref(fi)))


def solver(lib, positions, *, triangles=None, gravity=(0.0, 0.0, 0.0)):
    n = len(positions)
    return Raw(lib, positions, np.ones(n, np.float32), triangles=triangles, gravity=gravity, dt=H, damping=1.0)


def one_step(r, fields):
    """State before, and velocity after one substep with `fields` set."""
    set_fields(r, fields)
    p, _, v = r.state()
    r.step(1, 1, fetch=0)
    _, _, v1 = r.state()
    return p.astype(np.float64), v.astype(np.float64), v1.astype(np.float64)


def check(v_new, v_old, accel, gravity=(0.0, 0.0, 0.0)):
    expect = v_old + (np.asarray(gravity, np.float64) + accel) * H
    err = np.abs(v_new - expect)
    assert np.all(err <= ABS + REL * np.abs(expect)), (v_new, expect)


def cap(a, limit=5000.0):
    a = np.asarray(a, np.float64)
    n = np.linalg.norm(a, axis=-1, keepdims=True)
    return np.where(n > limit, a * (limit / np.maximum(n, 1e-30)), a)


def unit(x):
    x = np.asarray(x, np.float64)
    return x / np.linalg.norm(x)


# Random interior points of a cube, away from the fields' origins.
PTS = np.random.Generator(np.random.PCG64(5)).uniform(-1.0, 1.0, size=(9, 3)).astype(np.float32) + np.float32([0.0, 0.0, 2.0])


def moving_state(lib, gravity=(0.3, -0.7, 0.4)):
    """Vertices at PTS that have moved one substep under `gravity`, so that
    the velocity-dependent fields see a nonzero velocity."""
    r = solver(lib, PTS, gravity=gravity)
    r.step(1, 1, fetch=0)
    return r, gravity


# ------------------------------------------------------------------ wind
def flat_quad():
    pos = np.array([[0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], np.float32)
    tri = np.array([[0, 1, 2], [0, 2, 3]], np.int32)         # counter-clockwise seen from +z: normal +z
    return pos, tri


@pytest.mark.parametrize("direction,strength", [((0.6, 0.0, 0.8), 7.0), ((0.0, 0.0, 2.0), 3.0), ((1.0, 2.0, -0.5), 4.0)])
def test_wind_on_flat_quad(oracle, direction, strength):
    pos, tri = flat_quad()
    r = solver(oracle, pos, triangles=tri)
    p, v, v1 = one_step(r, [dict(type=1, strength=strength, direction=direction, noise=0.0)])
    d = unit(direction)
    facing = max(0.0, float(d @ np.array([0.0, 0.0, 1.0])))
    check(v1, v, np.tile(d * strength * facing, (4, 1)))
    if facing == 0.0:
        assert np.all(v1 == 0.0)                    # wind from behind the surface does nothing


def test_wind_needs_a_surface_normal(oracle):
    r = solver(oracle, PTS)                          # no triangles: no normal, no wind
    _, _, v1 = one_step(r, [dict(type=1, strength=5.0, direction=(0.0, 0.0, 1.0))])
    assert np.all(v1 == 0.0)


def test_wind_noise_bounded(oracle):
    pos, tri = flat_quad()
    r = solver(oracle, pos, triangles=tri)
    _, _, v1 = one_step(r, [dict(type=1, strength=4.0, direction=(0.0, 0.0, 1.0), noise=0.5, seed=3)])
    # magnitude scaled by max(0, 1 + noise * u) with u in [-1, 1): within [0.5, 1.5] x strength
    mag = v1[:, 2] / H
    assert np.all(mag >= 0.5 * 4.0 - 1e-5) and np.all(mag <= 1.5 * 4.0 + 1e-5)
    assert np.allclose(v1[:, :2], 0.0)


# -------------------------------------------------------- force / charge
@pytest.mark.parametrize("strength", [2.5, -4.0])
def test_force_has_constant_magnitude_along_the_radius(oracle, strength):
    origin = np.array([0.2, -0.1, 1.5])
    r = solver(oracle, PTS)
    p, v, v1 = one_step(r, [dict(type=2, strength=strength, origin=tuple(origin))])
    d = p - origin
    check(v1, v, strength * d / np.linalg.norm(d, axis=1, keepdims=True))


def test_charge_falls_off_as_inverse_square(oracle):
    origin = np.array([0.0, 0.0, 1.0])
    r = solver(oracle, PTS)
    p, v, v1 = one_step(r, [dict(type=5, strength=3.0, origin=tuple(origin))])
    d = p - origin
    dist = np.linalg.norm(d, axis=1, keepdims=True)
    assert np.all(dist > 0.1)                        # beyond the 0.01 m^2 floor of r^2
    check(v1, v, 3.0 * d / dist ** 3)


def test_force_linear_falloff_to_max_distance(oracle):
    origin = np.array([0.0, 0.0, 2.0])
    mn, mx = 0.3, 1.2
    r = solver(oracle, PTS)
    p, v, v1 = one_step(r, [dict(type=2, strength=2.0, origin=tuple(origin), use_max_distance=1, distance_min=mn, distance_max=mx)])
    d = p - origin
    dist = np.linalg.norm(d, axis=1, keepdims=True)
    att = np.where(dist >= mx, 0.0, np.where(dist > mn, np.clip((mx - dist) / (mx - mn), 0.0, 1.0), 1.0))
    assert (att == 0.0).any() and ((att > 0.0) & (att < 1.0)).any()
    check(v1, v, 2.0 * att * d / dist)


def test_force_power_falloff_beyond_min_distance(oracle):
    origin = np.array([0.0, 0.0, 2.0])
    mn, power = 0.25, 1.5
    r = solver(oracle, PTS)
    p, v, v1 = one_step(r, [dict(type=2, strength=2.0, origin=tuple(origin), use_min_distance=1, distance_min=mn,
                                 falloff_power=power)])
    d = p - origin
    dist = np.linalg.norm(d, axis=1, keepdims=True)
    att = np.where(dist > mn, (mn / dist) ** power, 1.0)
    check(v1, v, 2.0 * att * d / dist)


# ---------------------------------------------------------------- vortex
def test_vortex_turns_about_its_axis(oracle):
    origin = np.array([0.1, 0.2, 2.0])
    axis = unit((0.0, 0.3, 1.0))
    r = solver(oracle, PTS)
    p, v, v1 = one_step(r, [dict(type=3, strength=1.7, origin=tuple(origin), axis=tuple(axis))])
    d = p - origin
    radial = d - np.outer(d @ axis, axis)
    t = np.cross(axis, radial)
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    check(v1, v, 1.7 * t)
    a = (v1 - v) / H
    assert np.allclose(a @ axis, 0.0, atol=1e-5)                     # no push along the axis
    assert np.allclose(np.einsum("ij,ij->i", a, radial), 0.0, atol=1e-5)   # nor toward it


# -------------------------------------------------------------- harmonic
def test_harmonic_spring_and_damping(oracle):
    r, g = moving_state(oracle)
    origin = np.array([0.0, 0.0, 2.0])
    k, rest, c = 3.0, 0.6, 0.4
    p, v, v1 = one_step(r, [dict(type=6, strength=k, origin=tuple(origin), rest_length=rest, harmonic_damping=c, flow=5.0)])
    d = p - origin
    dist = np.linalg.norm(d, axis=1, keepdims=True)
    u = d / dist
    vr = np.einsum("ij,ij->i", v, u)[:, None]
    # flow does not apply to the harmonic type
    check(v1, v, u * (-k * (dist - rest) - vr * c * abs(k)), g)


# --------------------------------------------------------- Lennard-Jones
def test_lennard_jones_ratio_clamp(oracle):
    origin = np.array([0.0, 0.0, 2.0])
    radius = 0.2
    # near points (ratio > 6 before the clamp) and far points
    pts = np.array([[0.01, 0.0, 2.0], [0.0, -0.02, 2.0], [0.0, 0.0, 2.3], [0.5, 0.5, 2.0]], np.float32)
    r = solver(oracle, pts)
    s = 1.0e-7
    p, v, v1 = one_step(r, [dict(type=7, strength=s, origin=tuple(origin), size=radius)])
    d = p - origin
    dist = np.linalg.norm(d, axis=1, keepdims=True)
    ratio = np.minimum(radius / np.maximum(dist, 1e-3), 6.0)
    assert ratio[0] == 6.0 and ratio[1] == 6.0 and ratio[3] < 1.0
    check(v1, v, cap(d / dist * s * (ratio ** 12 - ratio ** 6)))


# ---------------------------------------------------------------- magnet
def test_magnet_is_velocity_cross_axis(oracle):
    r, g = moving_state(oracle)
    ax = unit((0.2, 1.0, -0.4))
    p, v, v1 = one_step(r, [dict(type=8, strength=0.9, direction=(0.2, 1.0, -0.4))])
    check(v1, v, 0.9 * np.cross(v, ax), g)


# ------------------------------------------------------------------ drag
def test_drag_linear_and_quadratic(oracle):
    r, g = moving_state(oracle)
    p, v, v1 = one_step(r, [dict(type=9, strength=0.5, linear_drag=0.3, quadratic_drag=0.2)])
    speed = np.linalg.norm(v, axis=1, keepdims=True)
    check(v1, v, -(0.3 + 0.2 * speed) * v, g)


def test_drag_falls_back_to_strength(oracle):
    r, g = moving_state(oracle)
    p, v, v1 = one_step(r, [dict(type=9, strength=-0.8, linear_drag=0.0, quadratic_drag=0.0)])
    check(v1, v, -0.8 * v, g)                       # |strength| as the linear coefficient


def test_flow_adds_velocity_damping(oracle):
    r, g = moving_state(oracle)
    origin = np.array([0.0, 0.0, 0.0])
    p, v, v1 = one_step(r, [dict(type=2, strength=1.5, origin=tuple(origin), flow=0.7)])
    d = p - origin
    check(v1, v, 1.5 * d / np.linalg.norm(d, axis=1, keepdims=True) - 0.7 * v, g)


# ------------------------------------------------------------ the caps
def test_acceleration_cap_per_field_and_total(oracle):
    origin = np.array([0.0, 0.0, 0.0])
    r = solver(oracle, PTS)
    p, v, v1 = one_step(r, [dict(type=2, strength=8000.0, origin=tuple(origin)),
                            dict(type=2, strength=3000.0, origin=(0.0, 0.0, 5.0))])
    d1 = p - origin
    d2 = p - np.array([0.0, 0.0, 5.0])
    a1 = cap(8000.0 * d1 / np.linalg.norm(d1, axis=1, keepdims=True))    # 8000 -> 5000 per field
    a2 = 3000.0 * d2 / np.linalg.norm(d2, axis=1, keepdims=True)
    total = cap(a1 + a2)
    check(v1, v, total)
    assert np.allclose(np.linalg.norm((v1 - v) / H, axis=1), np.minimum(np.linalg.norm(a1 + a2, axis=1), 5000.0), rtol=1e-5)


# ------------------------------------------------------------ turbulence
def test_turbulence_zero_on_lattice_points(oracle):
    # q = (x - origin) + axis is a lattice point for size 1: every octave's
    # Perlin value is 0, so the noise is 0.5 in each component and the push 0
    pts = np.array([[1.0, 2.0, 3.0], [-2.0, 0.0, 5.0], [4.0, -3.0, 1.0]], np.float32)
    r = solver(oracle, pts)
    p, v, v1 = one_step(r, [dict(type=4, strength=3.0, size=1.0, origin=(0.0, 0.0, 0.0), axis=(0.0, 0.0, 1.0))])
    assert np.all(v1 == 0.0)


def test_turbulence_and_texture_bounded(oracle):
    r = solver(oracle, PTS)
    p, v, v1 = one_step(r, [dict(type=4, strength=2.0, size=0.37, origin=(0.1, 0.2, 0.3))])
    a = (v1 - v) / H
    assert np.all(np.abs(a) <= 2.0 + 1e-5) and np.any(np.abs(a) > 1e-3)
    r = solver(oracle, PTS)
    p, v, v1 = one_step(r, [dict(type=10, strength=2.0, noise=1.3, seed=11)])
    a = (v1 - v) / H
    assert np.all(np.abs(a) <= 2.0 + 1e-5) and np.any(np.abs(a) > 1e-3)


def test_field_off_when_not_applied_or_zero(oracle):
    r = solver(oracle, PTS)
    _, _, v1 = one_step(r, [dict(type=2, strength=3.0, origin=(0.0, 0.0, 0.0), apply_to_location=0),
                            dict(type=2, strength=0.0, origin=(0.0, 0.0, 0.0))])
    assert np.all(v1 == 0.0)


# ----------------------------------------------------------- overpressure
def one_triangle(lib, pressure):
    """A right isosceles triangle (legs 1 m) in the z = 1 plane, normal +z,
    its three edges at rest length, no gravity, one substep of 1/60 s.  The
    edges are nearly free (compliance 1e3: a correction of ~1e-9 m) so that
    the trimmed vertices, which move different distances, are not pulled
    back by the stretch pass; they still set each vertex's rest scale."""
    pos = np.array([[0, 0, 1], [1, 0, 1], [0, 1, 1]], np.float32)
    edges = np.array([[0, 1], [1, 2], [0, 2]], np.int32)
    rest = np.array([1.0, math.sqrt(2.0), 1.0], np.float32)
    r = Raw(lib, pos, np.ones(3, np.float32), edges=edges, edge_rest=rest, triangles=np.array([[0, 1, 2]], np.int32),
            gravity=(0.0, 0.0, 0.0), damping=1.0, use_volume_pressure=1, pressure_strength=pressure, stretch_compliance=1.0e3)
    r.step(1, 1, fetch=0)
    return r.state()


def test_overpressure_single_triangle(oracle):
    h = 1.0 / 60.0
    # every vertex holds a third of the 0.5 m^2 area: a = p * 1792 * A/3 * w along the normal
    _, _, v = one_triangle(oracle, 1.2)
    a = 1.2 * 1792.0 * (0.5 / 3.0) * 1.0
    assert np.allclose(v[:, 2], a * h, rtol=1e-5) and np.allclose(v[:, :2], 0.0, atol=1e-6)
    # a large pressure is trimmed so that one substep moves a vertex at most
    # 0.85 x its mean incident rest edge length: v = 0.85 * rest_mean / h
    _, _, v = one_triangle(oracle, 20.0)
    rest_mean = np.array([1.0, (1.0 + math.sqrt(2.0)) / 2.0, (1.0 + math.sqrt(2.0)) / 2.0])
    assert 20.0 * 1792.0 * (0.5 / 3.0) * h * h > 0.85 * rest_mean.max()      # the trim applies to all three
    assert np.allclose(v[:, 2], 0.85 * rest_mean / h, rtol=1e-5)
