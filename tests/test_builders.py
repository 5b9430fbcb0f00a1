"""The vectorised constraint builders against a direct loop restatement of
xpbd_core.py:715-747 (set of sorted edges; first-seen opposite vertex per
edge, paired with every later different one), on regular, irregular and
non-manifold triangle lists."""
import numpy as np


def loop_edges(tri):
    edges = sorted({(min(a, b), max(a, b)) for t in tri.tolist() for a, b in ((t[0], t[1]), (t[1], t[2]), (t[2], t[0]))})
    return np.asarray(edges, np.int64).reshape(-1, 2)


def loop_bends(tri):
    first, pairs = {}, set()
    for a, b, c in tri.tolist():
        for x, y, opp in ((a, b, c), (b, c, a), (c, a, b)):
            key = (min(x, y), max(x, y))
            prev = first.get(key)
            if prev is None:
                first[key] = opp
            elif prev != opp:
                pairs.add((min(prev, opp), max(prev, opp)))
    return np.asarray(sorted(pairs), np.int64).reshape(-1, 2)


def meshes(scenes):
    rng = np.random.Generator(np.random.PCG64(5))
    pos, tri = scenes.grid_mesh(9, 1.0)
    yield pos, tri
    pos, tri = scenes.icosphere(2, 1.0)
    yield pos, tri
    # random soup with shared, repeated and non-manifold edges
    tri = rng.integers(0, 30, size=(120, 3))
    tri = tri[(tri[:, 0] != tri[:, 1]) & (tri[:, 1] != tri[:, 2]) & (tri[:, 0] != tri[:, 2])]
    yield rng.normal(size=(30, 3)), tri


def test_vectorised_builders_match_the_loops(scenes):
    for pos, tri in meshes(scenes):
        rest = np.asarray(pos, np.float32)
        e, el = scenes.edge_constraints(tri, rest)
        assert np.array_equal(e, loop_edges(tri))
        assert np.array_equal(el, np.maximum(np.linalg.norm(rest[e[:, 1]].astype(np.float64) - rest[e[:, 0]], axis=1),
                                             1.0e-7).astype(np.float32))
        b, _ = scenes.bend_constraints(tri, rest)
        assert np.array_equal(b, loop_bends(tri))
