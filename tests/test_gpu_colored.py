"""Coloured Gauss-Seidel stretch (ssbl_ext_set_stretch_mode 1) on the GPU
against the oracle's coloured sweep: one launch per edge colour over every
cloth of the step, in place.  Tolerance 0 (a valid colouring makes a colour's
updates independent, so the launch equals the oracle's sequential sweep)."""
import numpy as np
import pytest

from parity_util import make_pair, run_and_compare
from test_oracle_pins import _chain

pytestmark = pytest.mark.gpu


def test_chain_closed_form_gpu(hip, oracle):
    for order in ([[0, 1], [1, 2]], [[1, 2], [0, 1]]):
        for mode in (0, 1):
            assert np.array_equal(_chain(hip, order, mode), _chain(oracle, order, mode)), (order, mode)


def test_invalid_colouring_refused_like_the_oracle(hip, oracle):
    from raw_abi import Raw
    msgs = []
    for lib in (hip, oracle):
        r = Raw(lib, [[0, 0, 0], [1, 0, 0], [2, 0, 0]], [1, 1, 1], edges=np.array([[0, 1], [1, 2]], np.int32),
                edge_rest=[1.0, 1.0], edge_offsets=[0, 2])
        assert lib.ssbl_ext_set_stretch_mode(r.h, 1) == 0
        msgs.append(lib.ssbl_last_error())
    assert msgs[0] == msgs[1]


def test_config1_colored_120_frames(nb, hip, oracle, scenes):
    sc = scenes.config1()
    g, r = make_pair(nb, hip, oracle, sc.cloths[0], sc.options[0], sc.static_triangles)
    g.set_stretch_mode(1)
    r.set_stretch_mode(1)
    worst = run_and_compare(g, r, sc.frames, sc.substeps, sc.iterations, diagnostics=True, every=10)
    print(f"config 1 coloured, 120 frames: max |hip - oracle| = {worst:.3e}")
    assert worst == 0.0
    assert g.diagnostics().abi41_hard_projection_fallbacks == r.diagnostics().abi41_hard_projection_fallbacks


def test_config3_stacks_colored_batched(nb, hip, oracle, scenes):
    from test_gpu_configs import _make
    sc = scenes.config3(seed=0)
    picked = [0, 1, 2, 3, 40, 41, 42, 43]          # stack 0 and stack 10 (over the box)
    bat, cross_b = _make(nb, hip, sc, picked)
    ref, cross_r = _make(nb, oracle, sc, picked)
    for s in bat + ref:
        s.set_stretch_mode(1)
    batch = nb.SolverBatch(bat)
    for f in range(24):
        cross_b.refresh(device=True)
        batch.step(sc.substeps, sc.iterations)
        cross_r.refresh(device=True)
        for r in ref:
            r.step(sc.substeps, sc.iterations, diagnostics=False, synchronize=True)
        if (f + 1) % 8 == 0:
            nb.synchronize(hip)
            for c, (b, r) in enumerate(zip(bat, ref)):
                for x, y in zip(b.download_state(), r.download_state()):
                    assert np.array_equal(x, y), ("coloured batch vs oracle", f, picked[c])
            print(f"config 3 coloured stacks 0 and 10, frame {f + 1}: bit-exact", flush=True)
    batch.close()
