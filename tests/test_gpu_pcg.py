"""Hard-stretch PCG, polish and caps on the GPU against the oracle."""
import subprocess

import numpy as np
import pytest

import scenarios as S
from parity_util import make_pair, run_and_compare

pytestmark = pytest.mark.gpu
TOL = 1.0e-5


@pytest.mark.parametrize("strength", [0.4, 1.0])
@pytest.mark.parametrize("inv_mass,start", [
    ((0.0, 1.0), [[0, 0, 0], [2.5, 0.1, 0]]),
    ((1.0, 1.0), [[0, 0, 0], [2.5, 0.0, 0]]),
    ((0.0, 1.0, 1.0, 1.0), [[0, 0, 0], [1, 0, 0], [2, 0, 0], [5, 0, 0]]),
])
def test_chain_matches_oracle(hip, oracle, strength, inv_mass, start):
    rests = [1.0] * (len(inv_mass) - 1)
    (pg, qg, vg), dg, _ = S.pcg_chain(hip, strength=strength, start=start, inv_mass=inv_mass, rests=rests)
    (pr, qr, vr), dr, _ = S.pcg_chain(oracle, strength=strength, start=start, inv_mass=inv_mass, rests=rests)
    diff = max(np.max(np.abs(a - b)) for a, b in ((pg, pr), (qg, qr), (vg, vr)))
    print(f"chain n={len(inv_mass)} strength={strength}: max |hip - oracle| = {diff:.3e}")
    assert diff <= TOL
    assert dg.abi41_pcg_iterations == dr.abi41_pcg_iterations and dg.abi41_pcg_csr_nnz == dr.abi41_pcg_csr_nnz
    assert abs(dg.abi41_pcg_initial_residual - dr.abi41_pcg_initial_residual) <= 1e-6 * max(1.0, dr.abi41_pcg_initial_residual)


@pytest.mark.parametrize("strength", [0.5, 1.0])
def test_grid_cloth_with_pcg(nb, hip, oracle, scenes, strength):
    cloth = scenes.grid_cloth(20, 0.8, 1.0, seed=21)
    opts = scenes.SolverOptions(use_ground=True, stretch_optimization_enabled=True, stretch_optimization_strength=strength,
                                cloth_thickness=0.005)
    g, r = make_pair(nb, hip, oracle, cloth, opts, None)
    worst = run_and_compare(g, r, 20, 4, 3, diagnostics=True, every=5)
    dg, dr = g.diagnostics(), r.diagnostics()
    print(f"grid 20x20 PCG strength {strength}: max |hip - oracle| = {worst:.3e}, residual {dg.abi41_pcg_final_residual:.3e}")
    assert worst <= TOL
    assert dg.abi41_pcg_iterations == dr.abi41_pcg_iterations


SMOKE_MARKERS = ("HARD_STRETCH_OK", "OVERPRESSURE_OK", "TACK_BENDING_OK", "STATIC_SDF_OK", "PIN_WEIGHT_OK", "NATIVE_OK",
                 "SELF_COLLISION_OK", "SELF_VT_OK", "SELF_EE_OK")


def test_reference_smoke_binary_on_hip():
    """The reference's own native smoke test, compiled unchanged against the
    HIP library (oracle/Makefile target `ref`)."""
    from conftest import ROOT
    exe = ROOT / "oracle" / "_ref" / "abi41_smoke_hip"
    if not exe.exists():
        pytest.skip("reference smoke binary not built (needs the reference tree at build time)")
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(out.stdout[-2000:], out.stderr[-2000:])
    for marker in SMOKE_MARKERS:
        assert f"SSBL_ABI41_{marker}" in out.stdout, out.stdout
    assert out.returncode == 0


def test_stress_case_reduced_parity(nb, hip, oracle, scenes):
    """Balloon (subdivision 3, 642 vertices) + 24 x 24 twisted towel: strict
    self collision, PCG at strength 1, pressure, turning pins; three frames of
    20 x 10, bit-exact against the oracle, alone and batched."""
    sc = scenes.config5(subdivisions=3, towel_n=24, frames=3)
    _, hip_alone = S.run_stress(nb, hip, scenes, sc, 3, batched=False)
    _, hip_batch = S.run_stress(nb, hip, scenes, sc, 3, batched=True)
    _, ora = S.run_stress(nb, oracle, scenes, sc, 3, batched=False)
    worst = max(float(np.max(np.abs(a - b))) for sa, sb in zip(hip_alone, ora) for a, b in zip(sa, sb))
    print(f"stress case reduced, 3 frames: max |hip - oracle| = {worst:.3e}")
    assert worst == 0.0
    for sa, sb in zip(hip_alone, hip_batch):
        for a, b in zip(sa, sb):
            assert np.array_equal(a, b)


def test_stress_case_full_size(nb, hip, scenes):
    """The full stress case: a 655,362-vertex balloon and a 1,000,000-vertex
    towel, both with strict self collision and PCG, two frames of 20 x 10.
    The oracle cannot run this size, so the checks do not depend on it:
    batched equals alone bit for bit, the pinned rows sit exactly on their
    turning targets, and all state is finite.  (At this resolution pressure
    1.2 against the stretch caps doubles the balloon's radius within two
    frames; the reduced case above pins those dynamics against the oracle.)"""
    import time
    sc = scenes.config5(frames=2)
    t0 = time.perf_counter()
    solvers, batched = S.run_stress(nb, hip, scenes, sc, 2, batched=True)
    t_batch = time.perf_counter() - t0
    d = [s.diagnostics() for s in solvers]
    for s in solvers:
        s.close()
    t0 = time.perf_counter()
    solvers, alone = S.run_stress(nb, hip, scenes, sc, 2, batched=False)
    t_alone = time.perf_counter() - t0
    for s in solvers:
        s.close()
    print(f"stress case {sc.vertex_count} vertices: 2 frames batched {t_batch:.2f} s, alone {t_alone:.2f} s (incl. set-up)")
    for sa, sb in zip(alone, batched):
        for a, b in zip(sa, sb):
            assert np.array_equal(a, b)
    towel = sc.cloths[1]
    assert np.array_equal(alone[1][0][towel.pin_indices], scenes.towel_pin_targets(towel, 2 * sc.options[1].dt))
    assert all(np.all(np.isfinite(x)) for st in alone for x in st)


@pytest.mark.parametrize("cadence", [2, 4])
def test_pcg_substep_cadence_gpu(nb, hip, oracle, scenes, cadence):
    from test_pcg import _cadence_run
    g, rg = _cadence_run(nb, hip, scenes, cadence)
    r, rr = _cadence_run(nb, oracle, scenes, cadence)
    assert rg == rr
    for x, y in zip(g.download_state(), r.download_state()):
        assert np.array_equal(x, y)


def test_pcg_batch_equals_oracle(nb, hip, oracle, scenes):
    """Hard-stretch cloths of several sizes and strengths (coloured and
    uncoloured caps) and a direct-stretch cloth stepped as one batch: the
    solve is one launch per phase over all of them, the reductions one
    workgroup per cloth.  Each equals the oracle's run of it bit for bit."""
    from test_gpu_graphs import pcg_batch_cloths
    cloths = pcg_batch_cloths(scenes)
    members = [nb.NativeXpbdSolver(c, o, None, lib=hip) for c, o in cloths]
    refs = [nb.NativeXpbdSolver(c, o, None, lib=oracle) for c, o in cloths]
    batch = nb.SolverBatch(members)
    for f in range(8):
        batch.step(5, 3)
        for r in refs:
            r.step(5, 3, diagnostics=False, synchronize=True)
        nb.synchronize(hip)
        for k, (m, r) in enumerate(zip(members, refs)):
            for x, y in zip(m.download_state(), r.download_state()):
                assert np.array_equal(x, y), (f, k)
    batch.close()
