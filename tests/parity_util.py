"""Helpers shared by the GPU parity tests: run the HIP library and the oracle
on the same inputs and compare state."""
import numpy as np


def make_pair(nb, hip, oracle, cloth, opts, static):
    g = nb.NativeXpbdSolver(cloth, opts, static, lib=hip)
    r = nb.NativeXpbdSolver(cloth, opts, static, lib=oracle)
    return g, r


def state_diff(g, r):
    a = g.download_state()
    b = r.download_state()
    return max(float(np.max(np.abs(x.astype(np.float64) - y.astype(np.float64)))) if x.size else 0.0 for x, y in zip(a, b))


def run_and_compare(g, r, frames, substeps, iterations, *, diagnostics=False, every=1):
    worst = 0.0
    for f in range(frames):
        g.step(substeps, iterations, diagnostics=diagnostics, synchronize=True)
        r.step(substeps, iterations, diagnostics=diagnostics, synchronize=True)
        if (f + 1) % every == 0 or f == frames - 1:
            worst = max(worst, state_diff(g, r))
    return worst


def rel_close(a, b, rel=1.0e-4, abs_tol=1.0e-5):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return bool(np.all(np.abs(a - b) <= np.maximum(abs_tol, rel * np.abs(b))))
