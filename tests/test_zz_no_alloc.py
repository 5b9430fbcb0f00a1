"""The allocation-free step (ssbl_ext_step_solver_ex_noalloc) against
ssbl_step_solver_ex on identical handles, through the C ABI.

The hard-stretch PCG solve (strength 1.0: polish and caps too) of the
variant takes its scratch from the library's static array.  Two calls in a
row -- a 120x120 cloth, then a 40x40 one (a smaller input reusing the array
the first left behind), then the large one again -- each equal the existing
entry point's result bit for bit (tolerance 0), diagnostics included.  A cloth
whose solve needs more scratch than the array holds is rejected with an error
and keeps stepping through the existing entry point.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

COUNTERS = ("abi41_pcg_iterations", "abi41_pcg_csr_nnz", "abi41_pcg_guarded", "abi41_hard_projection_fallbacks",
            "abi41_soft_contact_count", "external_friction_corrections")
FLOATS = ("abi41_pcg_initial_residual", "abi41_pcg_final_residual", "abi41_pcg_max_delta")


def _pair(nb, hip, scenes, side, seed):
    cloth = scenes.grid_cloth(side, 0.02 * (side - 1), 1.0, seed=seed)
    opts = scenes.SolverOptions(use_ground=True, stretch_optimization_enabled=True, stretch_optimization_strength=1.0,
                                collision_margin=0.005, cloth_thickness=0.005)
    return (nb.NativeXpbdSolver(cloth, opts, None, lib=hip), nb.NativeXpbdSolver(cloth, opts, None, lib=hip))


def _noalloc(hip, solver, substeps, iterations, fetch):
    ok = hip.ssbl_ext_step_solver_ex_noalloc(solver.handle, substeps, iterations, int(fetch), 1)
    assert ok == 1, hip.ssbl_last_error()


def _compare(ref, got, label):
    for x, y in zip(ref.download_state(), got.download_state()):
        assert np.isfinite(x).all(), label
        assert np.array_equal(x, y), label
    dr, dg = ref.diagnostics(), got.diagnostics()
    for k in COUNTERS + FLOATS:
        assert getattr(dr, k) == getattr(dg, k), (label, k, getattr(dr, k), getattr(dg, k))


def test_noalloc_step_equals_step_on_two_calls(nb, hip, scenes):
    hip.ssbl_ext_step_solver_ex_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    big_ref, big = _pair(nb, hip, scenes, 120, 1)
    small_ref, small = _pair(nb, hip, scenes, 40, 2)
    for call, (ref, got, label) in enumerate([(big_ref, big, "120x120"), (small_ref, small, "40x40"), (big_ref, big, "120x120 again")]):
        for frame in range(3):
            fetch = frame == 2
            ref.step(6, 4, diagnostics=fetch, synchronize=True)
            _noalloc(hip, got, 6, 4, fetch)
        _compare(ref, got, label)
        print(f"call {call + 1} ({label}): noalloc == step, pcg iterations {got.diagnostics().abi41_pcg_iterations}", flush=True)


def test_noalloc_rejects_a_solve_beyond_the_scratch(nb, hip, scenes):
    cloth = scenes.grid_cloth(700, 7.0, 1.0, seed=3)
    opts = scenes.SolverOptions(use_ground=True, stretch_optimization_enabled=True, stretch_optimization_strength=0.4)
    s = nb.NativeXpbdSolver(cloth, opts, None, lib=hip)
    assert hip.ssbl_ext_step_solver_ex_noalloc(s.handle, 1, 1, 0, 1) == 0
    msg = hip.ssbl_last_error().decode()
    print("rejected:", msg)
    assert "scratch" in msg
    s.step(1, 1, diagnostics=True, synchronize=True)          # the existing entry point still steps it
    assert np.isfinite(s.download_positions()).all()
