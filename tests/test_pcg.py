"""Hard-stretch PCG on the oracle: the 8-iteration block-Jacobi CG reaches the
dense FP64 solution of the same block system on systems small enough to
converge, and the reference smoke test's hard-stretch criteria hold."""
import numpy as np
import pytest

import scenarios as S


@pytest.mark.parametrize("inv_mass,start", [
    ((0.0, 1.0), [[0, 0, 0], [1.2, 0.1, 0]]),
    ((0.0, 1.0, 1.0), [[0, 0, 0], [1.1, 0.0, 0.0], [2.3, 0.2, 0.0]]),
    ((1.0, 0.5, 2.0), [[0, 0, 0], [1.05, 0.0, 0.1], [2.2, -0.1, 0.0]]),
])
def test_pcg_reaches_dense_solution(oracle, inv_mass, start):
    rests = [1.0] * (len(inv_mass) - 1)
    (p, _, _), d, _ = S.pcg_chain(oracle, strength=0.4, start=start, inv_mass=inv_mass, rests=rests)
    expect = S.pcg_exact_solution(start, inv_mass, rests, 0.4)
    assert np.max(np.abs(p - expect)) < 2e-6, (p, expect)
    assert d.abi41_pcg_iterations == 8 and d.abi41_pcg_csr_nnz == 2 * len(rests)
    assert d.abi41_pcg_final_residual < d.abi41_pcg_initial_residual


def test_reference_hard_stretch_criteria(oracle):
    # native/smoke/abi41_smoke.cpp:763-860 (strength 1.0: polish and caps active)
    start = [[0, 0, 0], [2.5, 0, 0]]
    (off, _, _), _, _ = (None, None, None), None, None
    import raw_abi
    r_off = raw_abi.Raw(oracle, [[0, 0, 0], [1, 0, 0]], [1, 1], edges=[[0, 1]], edge_rest=[1.0], stretch_compliance=1.0e6)
    r_off.set_positions(start)
    r_off.step(1, 1)
    off_d = np.linalg.norm(r_off.positions()[1] - r_off.positions()[0])
    (p_on, _, _), d, _ = S.pcg_chain(oracle, strength=1.0, start=start, inv_mass=(1.0, 1.0), rests=(1.0,))
    on_d = np.linalg.norm(p_on[1] - p_on[0])
    assert on_d < off_d - 0.05
    assert d.abi41_pcg_csr_nnz == 2 and d.abi41_pcg_texture_ready == 1 and d.abi41_pcg_iterations > 0
    assert 0 < d.abi41_pcg_final_residual < d.abi41_pcg_initial_residual and d.abi41_pcg_max_delta > 0
    (p_pin, _, _), _, _ = S.pcg_chain(oracle, strength=1.0, start=start, inv_mass=(0.0, 1.0), rests=(1.0,))
    assert abs(p_pin[0, 0]) <= 1e-5 and p_pin[1, 0] < 2.5 - 0.05


def test_reference_smoke_binary_on_oracle():
    """The reference's own native smoke test, compiled unchanged against the
    oracle; every case before the self-collision ones must pass (those use the
    fast self-collision mode)."""
    import subprocess
    from conftest import ROOT
    exe = ROOT / "oracle" / "_ref" / "abi41_smoke_oracle"
    if not exe.exists():
        pytest.skip("reference smoke binary not built (needs the reference tree)")
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120).stdout
    for marker in ("HARD_STRETCH_OK", "OVERPRESSURE_OK", "TACK_BENDING_OK", "STATIC_SDF_OK", "PIN_WEIGHT_OK", "NATIVE_OK"):
        assert f"SSBL_ABI41_{marker}" in out, out


def test_stress_case_reduced_on_oracle(nb, oracle, scenes):
    """The balloon + twisted towel scene at small size: the towel's pinned rows
    sit on their turning targets, batched and alone agree, everything stays
    finite and the balloon keeps its size (pressure against the hard-stretch
    caps makes it pulse by about 15 %)."""
    sc = scenes.config5(subdivisions=2, towel_n=12, frames=2)
    solvers, alone = S.run_stress(nb, oracle, scenes, sc, 2, batched=False)
    _, batched = S.run_stress(nb, oracle, scenes, sc, 2, batched=True)
    for a, b in zip(alone, batched):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
    towel = sc.cloths[1]
    target = scenes.towel_pin_targets(towel, 2 * sc.options[1].dt)
    assert np.array_equal(alone[1][0][towel.pin_indices], target)
    assert all(np.all(np.isfinite(x)) for st in alone for x in st)
    # the balloon keeps its size to within the pressure/stretch oscillation
    balloon = alone[0][0]
    r = np.linalg.norm(balloon - balloon.mean(axis=0), axis=1).mean()
    assert 0.7 < r < 1.3


def _cadence_run(nb, lib, scenes, cadence, frames=3, substeps=6):
    """A 16x16 cloth on the hard-stretch solve, created with
    SSBL_ABI41_PCG_SUBSTEP_CADENCE set (read when the handle is created)."""
    import os
    old = os.environ.get("SSBL_ABI41_PCG_SUBSTEP_CADENCE")
    os.environ["SSBL_ABI41_PCG_SUBSTEP_CADENCE"] = str(cadence)
    try:
        cloth = scenes.grid_cloth(16, 0.6, 1.0, seed=8)
        opts = scenes.SolverOptions(use_ground=True, stretch_optimization_enabled=True, stretch_optimization_strength=1.0)
        s = nb.NativeXpbdSolver(cloth, opts, None, lib=lib)
    finally:
        if old is None:
            del os.environ["SSBL_ABI41_PCG_SUBSTEP_CADENCE"]
        else:
            os.environ["SSBL_ABI41_PCG_SUBSTEP_CADENCE"] = old
    runs = []
    for _ in range(frames):
        s.step(substeps, 3, diagnostics=True, synchronize=True)
        runs.append(s.diagnostics().abi41_pcg_iterations // 8)
    return s, runs


@pytest.mark.parametrize("cadence,solves", [(1, 6), (2, 3), (4, 2), (8, 1), (0, 6), (-3, 6)])
def test_pcg_substep_cadence(nb, oracle, scenes, cadence, solves):
    # A41:284-292, 9158-9177: the solve runs on substeps s with (s + 1) % c == 0
    # and always on the last; c = min(env, substeps), unset or <= 0 means 1
    s, runs = _cadence_run(nb, oracle, scenes, cadence)
    assert runs == [solves] * len(runs)
    if cadence == 2:
        base, _ = _cadence_run(nb, oracle, scenes, 1)
        assert not np.array_equal(s.download_positions(), base.download_positions())
