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
