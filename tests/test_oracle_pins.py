"""Pins the CPU oracle against closed forms, invariants and the reference's own
smoke-test thresholds (native/smoke/abi41_smoke.cpp).  No GPU."""
import ctypes

import numpy as np
import pytest

import scenarios as S


@pytest.fixture(scope="module")
def o64(oracle):
    oracle.ssbl_oracle_set_precision.argtypes = [ctypes.c_int]
    yield oracle
    oracle.ssbl_oracle_set_precision(32)


def test_free_fall_matches_closed_form(oracle, o64):
    x, v = S.free_fall_closed_form()
    o64.ssbl_oracle_set_precision(64)
    p, _, vel = S.free_fall(o64)
    o64.ssbl_oracle_set_precision(32)
    assert abs(p[0, 2] - x) < 1e-6 and abs(vel[0, 2] - v) < 1e-6   # values pass through float32 on download
    p32, _, v32 = S.free_fall(oracle)
    # FP32: the velocity is re-measured from (pos - prev) / h, so one position ulp
    # (2.4e-7 m at 2 m) becomes ~1e-4 m/s
    assert abs(p32[0, 2] - x) < 2e-6 and abs(v32[0, 2] - v) < 2e-4
    assert p32[0, 0] == 0.0 and p32[0, 1] == 0.0


def test_pin_weights_reference_thresholds(oracle):
    out = S.hard_and_weighted_pins(oracle)
    hard_y, soft_y, low_y, free_y = out[0, 1], out[1, 1], out[2, 1], out[3, 1]
    assert abs(hard_y - 1.0) <= 1e-5                     # abi41_smoke.cpp:1238
    assert 0.2 < soft_y < 0.95                           # abi41_smoke.cpp:1242-1249
    assert abs(low_y - free_y) <= 1e-6                   # abi41_smoke.cpp:1251-1258


@pytest.mark.parametrize("weight,iterations", [(0.5, 3), (0.3, 5), (0.74, 1)])
def test_soft_pin_geometric_approach(oracle, weight, iterations):
    # relaxation r per projection; the step projects after integrate, once per
    # iteration and once more on the last iteration: (iterations + 2) times
    e = np.float32(1.0) / np.float32(iterations + 1)
    r = np.float32(1.0) - np.power(np.float32(1.0) - np.float32(weight) / np.float32(0.75), e, dtype=np.float32)
    p, _, v = S.soft_pin(oracle, weight, iterations)
    expect = 1.0 - (1.0 - float(r)) ** (iterations + 2)
    assert abs(p[0, 0] - expect) < 1e-6
    assert np.all(v == 0.0)


def test_hard_pin_is_exact(oracle):
    p, q, v = S.soft_pin(oracle, 0.9, 4)
    assert p[0, 0] == 1.0 and q[0, 0] == 1.0 and v[0, 0] == 0.0


def test_single_edge_relaxation_and_momentum(oracle):
    # unclamped: each pass removes 18 % of the constraint error, midpoint fixed
    out = S.single_edge(oracle, [[0, 0, 0], [1.1, 0, 0]])
    assert abs((out[1, 0] - out[0, 0]) - (1.1 - 0.18 * 0.1)) < 1e-6
    assert abs(out[0, 0] + out[1, 0] - 1.1) < 1e-6
    # clamped: a single correction is capped at min(0.1 * rest, 0.025) = 0.025
    out = S.single_edge(oracle, [[0, 0, 0], [2.0, 0, 0]])
    assert abs((out[1, 0] - out[0, 0]) - (2.0 - 0.05)) < 1e-6
    # unequal masses: sum of m * displacement is zero
    out = S.single_edge(oracle, [[0, 0, 0], [1.1, 0, 0]], inv_mass=(1.0, 3.0))
    d0, d1 = out[0, 0] - 0.0, out[1, 0] - 1.1
    assert abs(d0 / 1.0 + d1 / 3.0) < 1e-7 and d0 > 0 > d1
    # pinned end does not move
    out = S.single_edge(oracle, [[0, 0, 0], [1.1, 0, 0]], inv_mass=(0.0, 1.0))
    assert out[0, 0] == 0.0 and abs(out[1, 0] - (1.1 - 0.18 * 0.1)) < 1e-6


def test_ground_friction_turns_quadratic_drift_linear(oracle):
    (p0, _, _), n = S.ground_friction(oracle, 0.0)
    (p1, _, v1), _ = S.ground_friction(oracle, 1.0)
    h = np.float32(np.float32(1 / 60) / np.float32(4))
    assert abs(p0[0, 0] - 2.0 * h * h * n * (n + 1) / 2) < 1e-6     # frictionless: free tangential motion
    assert abs(p1[0, 0] - 2.0 * h * h * n) < 1e-6                  # static friction: only the integrate step moves it
    assert abs(p0[0, 2] - 0.005) < 1e-7 and abs(p1[0, 2] - 0.005) < 1e-7
    assert abs(v1[0, 0]) < 1e-6


def test_sphere_pushes_to_surface(oracle):
    p = S.sphere_push(oracle)
    assert abs(np.linalg.norm(p[0]) - 0.51) < 1e-6
    assert np.allclose(p[0] / np.linalg.norm(p[0]), np.array([0.1, 0.05, 0.2]) / np.linalg.norm([0.1, 0.05, 0.2]), atol=1e-6)


def test_static_plane_reference_thresholds(oracle):
    out, d = S.static_plane(oracle)
    assert out[0, 2] > 0.02 + 0.01                       # abi41_smoke.cpp:122
    assert abs(out[1, 2] - 0.02) <= 1e-5                 # abi41_smoke.cpp:127
    assert d.static_sdf_rebuild_count > 0 and d.static_sdf_voxel_count > 0
    assert min(d.static_sdf_grid_x, d.static_sdf_grid_y, d.static_sdf_grid_z) > 1
    assert d.static_sdf_contact_count > 0 and d.static_sdf_unsigned_fallback_count > 0


def box_distance(p, half):
    q = np.abs(p) - half
    outside = np.linalg.norm(np.maximum(q, 0.0), axis=-1)
    inside = np.minimum(np.max(q, axis=-1), 0.0)
    return outside + inside


def test_box_sdf_matches_analytic_distance(oracle, nb):
    r = S.box_sdf(oracle)
    dims = np.zeros(3, np.int32)
    ov = np.zeros(4, np.float32)
    fp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))  # noqa: E731
    assert oracle.ssbl_ext_static_sdf(r.h, dims.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), fp(ov), None, 0)
    vals = np.empty(int(np.prod(dims)), np.float32)
    assert oracle.ssbl_ext_static_sdf(r.h, dims.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), fp(ov), fp(vals), vals.size)
    grid = vals.reshape(dims[2], dims[1], dims[0])
    z, y, x = np.meshgrid(*[np.arange(d) for d in dims[::-1]], indexing="ij")
    pts = np.stack([ov[0] + x * ov[3], ov[1] + y * ov[3], ov[2] + z * ov[3]], -1).astype(np.float64)
    exact = box_distance(pts, 0.3)
    err = np.abs(np.abs(grid) - np.abs(exact))
    assert err.max() < 2e-6, err.max()
    # Sign by +x ray parity (A41:1725-1747) with the reference's 1e-5
    # barycentric slack: a ray that crosses a face diagonal is counted by both
    # triangles and flips the sign.  Here the +x face diagonal is y == z, so
    # the only mis-signed voxels are interior ones with y == z.
    wrong = (np.sign(grid) != np.sign(exact)) & (np.abs(exact) > 1e-6)
    assert wrong.any() and np.all(np.abs(pts[wrong][:, 1] - pts[wrong][:, 2]) < 1e-5)
    assert np.all(exact[wrong] < 0)


def test_bend_pair_once_per_frame(oracle):
    folded, out, rest = S.bend_pair(oracle)
    d0 = np.linalg.norm(folded[2] - folded[3])
    d1 = np.linalg.norm(out[2] - out[3])
    # error -0.8: correction 0.1*|c|/2 = 0.04 per end exceeds the 0.02 cap
    assert abs(d1 - (d0 + 0.04)) < 1e-6
    assert abs(d1 - rest) < abs(d0 - rest)               # abi41_smoke.cpp:1010-1086


def test_builder_counts_and_colouring(scenes):
    for n, (v, t, e, b) in {32: (1024, 1922, 2945, 2821), 100: (10000, 19602, 29601, 29205)}.items():
        c = scenes.grid_cloth(n, 1.0, 1.0)
        assert (len(c.positions_world), len(c.triangles), len(c.edges), len(c.bends)) == (v, t, e, b)
        assert len(c.edge_color_offsets) - 1 == 6 and len(c.bend_color_offsets) - 1 == 6
        for pairs, off in ((c.edges, c.edge_color_offsets), (c.bends, c.bend_color_offsets)):
            for k in range(len(off) - 1):
                seg = pairs[off[k]:off[k + 1]].ravel()
                assert len(np.unique(seg)) == len(seg)   # no vertex twice in a colour


def test_fp32_oracle_close_to_fp64_on_plain_config(o64, nb, scenes):
    """FP32 rounding against the same algorithm in FP64.  The hanging cloth is
    chaotic once it swings into the ground (the difference grows to ~0.15 m by
    frame 120), so agreement is checked where it is a statement about the
    arithmetic: the first frame."""
    sc = scenes.config1(frames=1)
    o64.ssbl_oracle_set_precision(64)
    d = nb.NativeXpbdSolver(sc.cloths[0], sc.options[0], sc.static_triangles, lib=o64)
    o64.ssbl_oracle_set_precision(32)
    f = nb.NativeXpbdSolver(sc.cloths[0], sc.options[0], sc.static_triangles, lib=o64)
    d.step(sc.substeps, sc.iterations, diagnostics=False)
    f.step(sc.substeps, sc.iterations, diagnostics=False)
    err = np.max(np.abs(d.download_positions().astype(np.float64) - f.download_positions()))
    print(f"FP32 vs FP64 oracle, plain 32x32, 1 frame: {err:.3e} m")
    assert err < 5e-5


def test_reference_smoke_binary_on_oracle():
    """The reference's own native smoke test, compiled unchanged against the
    oracle (oracle/Makefile target `ref`): every section passes, including
    the fast-mode self-collision scenes."""
    import subprocess
    from conftest import ROOT
    exe = ROOT / "oracle" / "_ref" / "abi41_smoke_oracle"
    if not exe.exists():
        pytest.skip("reference smoke binary not built (needs the reference tree at build time)")
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    for marker in ("HARD_STRETCH_OK", "OVERPRESSURE_OK", "TACK_BENDING_OK", "STATIC_SDF_OK", "PIN_WEIGHT_OK", "NATIVE_OK",
                   "SELF_COLLISION_OK", "SELF_VT_OK", "SELF_EE_OK"):
        assert f"SSBL_ABI41_{marker}" in out.stdout, out.stdout + out.stderr
    assert out.returncode == 0


def _chain(lib, order, mode):
    """Three unit-mass vertices on the x axis at 0, 1.1, 2.2 with two edges of
    rest length 1 in colours 0 and 1 (`order` lists the edges in colour
    order), no gravity, one substep of one iteration."""
    from raw_abi import Raw
    edges = np.array(order, np.int32)
    r = Raw(lib, [[0, 0, 0], [1.1, 0, 0], [2.2, 0, 0]], [1, 1, 1], edges=edges, edge_rest=[1.0, 1.0], edge_offsets=[0, 1, 2],
            gravity=(0.0, 0.0, 0.0))
    r.check(lib.ssbl_ext_set_stretch_mode(r.h, mode))
    r.step(1, 1)
    return r.positions()[:, 0].astype(np.float64)


def _edge(x, i, j):
    """One edge's XPBD projection in 1D (unit masses, compliance 0): each end
    moves 18 % of half the constraint error toward the rest length."""
    c = (x[j] - x[i]) - 1.0
    x[i] += 0.18 * c / 2.0
    x[j] -= 0.18 * c / 2.0


@pytest.mark.parametrize("order", [[[0, 1], [1, 2]], [[1, 2], [0, 1]]])
def test_colored_stretch_is_gauss_seidel_in_colour_order(oracle, order):
    x = np.array([0.0, 1.1, 2.2])
    for i, j in order:                  # colour 0, then colour 1, each seeing the other's update
        _edge(x, i, j)
    got = _chain(oracle, order, 1)
    assert np.allclose(got, x, atol=2e-7), (got, x)
    jac = np.array([0.0, 1.1, 2.2])
    jac[0] += 0.009; jac[2] -= 0.009    # Jacobi: both edges read the start, the middle vertex's pushes cancel
    assert np.allclose(_chain(oracle, order, 0), jac, atol=2e-7)
    assert abs(got[1] - 1.1) > 5e-4     # the two modes differ (GS: 1.10081, Jacobi: 1.1)


def test_colored_mode_requires_a_valid_colouring(oracle):
    from raw_abi import Raw
    r = Raw(oracle, [[0, 0, 0], [1, 0, 0], [2, 0, 0]], [1, 1, 1], edges=np.array([[0, 1], [1, 2]], np.int32),
            edge_rest=[1.0, 1.0], edge_offsets=[0, 2])            # vertex 1 twice in colour 0
    assert oracle.ssbl_ext_set_stretch_mode(r.h, 1) == 0
    assert b"valid edge colouring" in oracle.ssbl_last_error()
    r2 = Raw(oracle, [[0, 0, 0], [1, 0, 0]], [1, 1], edges=np.array([[0, 1]], np.int32), edge_rest=[1.0])   # no colours
    assert oracle.ssbl_ext_set_stretch_mode(r2.h, 1) == 0
    assert oracle.ssbl_ext_set_stretch_mode(r2.h, 0) == 1
