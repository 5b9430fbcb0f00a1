"""Build ABI41 solvers from raw arrays, the way the reference's native smoke
test does (native/smoke/abi41_smoke.cpp), for any library exporting the ABI."""
import ctypes

import numpy as np

from ssbl_xpbd_amd import abi
from ssbl_xpbd_amd.native_backend import NativeSolverError, last_error


def _fp(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _ip(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


class Raw:
    def __init__(self, lib, positions, inv_mass, *, edges=None, edge_rest=None, edge_offsets=None, bends=None, bend_rest=None,
                 bend_offsets=None, triangles=None, static_triangles=None, **cfg_fields):
        self.lib = lib
        self.p = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
        self.n = len(self.p)
        keep = {"p": self.p, "w": np.ascontiguousarray(inv_mass, np.float32)}
        cfg = abi.Config()
        cfg.vertex_count = self.n
        cfg.dt = 1.0 / 60.0
        cfg.damping = 1.0
        mesh = abi.Mesh(positions=_fp(keep["p"]), inv_mass=_fp(keep["w"]))
        if edges is not None:
            keep["e"] = np.ascontiguousarray(edges, np.int32).reshape(-1, 2)
            keep["er"] = np.ascontiguousarray(edge_rest, np.float32)
            cfg.edge_count = len(keep["e"])
            mesh.edges, mesh.edge_rest_lengths = _ip(keep["e"]), _fp(keep["er"])
            if edge_offsets is not None:
                keep["eo"] = np.ascontiguousarray(edge_offsets, np.int32)
                cfg.edge_color_count = len(keep["eo"]) - 1
                mesh.edge_color_offsets = _ip(keep["eo"])
        if bends is not None:
            keep["b"] = np.ascontiguousarray(bends, np.int32).reshape(-1, 2)
            keep["br"] = np.ascontiguousarray(bend_rest, np.float32)
            keep["bo"] = np.ascontiguousarray(bend_offsets, np.int32)
            cfg.bend_count = len(keep["b"])
            cfg.bend_color_count = len(keep["bo"]) - 1
            mesh.bends, mesh.bend_rest_lengths, mesh.bend_color_offsets = _ip(keep["b"]), _fp(keep["br"]), _ip(keep["bo"])
        if triangles is not None:
            keep["t"] = np.ascontiguousarray(triangles, np.int32).reshape(-1, 3)
            cfg.triangle_count = len(keep["t"])
            mesh.triangles = _ip(keep["t"])
        if static_triangles is not None:
            keep["s"] = np.ascontiguousarray(static_triangles, np.float32).reshape(-1, 9)
            cfg.static_triangle_count = len(keep["s"])
            mesh.static_triangles = _fp(keep["s"])
        for k, v in cfg_fields.items():
            if isinstance(v, (tuple, list)):
                v = (ctypes.c_float * 3)(*v)
            setattr(cfg, k, v)
        self.h = lib.ssbl_create_solver(ctypes.byref(cfg), ctypes.byref(mesh))
        if not self.h:
            raise NativeSolverError(last_error(lib))

    def check(self, ok):
        if not ok:
            raise NativeSolverError(last_error(self.lib))

    def pins(self, idx, pos, weights=None):
        i = np.ascontiguousarray(idx, np.int32)
        p = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        w = None if weights is None else np.ascontiguousarray(weights, np.float32)
        self.check(self.lib.ssbl_update_pin_targets(self.h, _ip(i), _fp(p), _fp(w), len(i)))

    def set_positions(self, pos):
        p = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        self.check(self.lib.ssbl_update_positions(self.h, _fp(p), p.size))

    def step(self, substeps, iterations, fetch=1):
        self.check(self.lib.ssbl_step_solver_ex(self.h, substeps, iterations, fetch, 1))

    def positions(self):
        out = np.empty((self.n, 3), np.float32)
        self.check(self.lib.ssbl_download_positions(self.h, _fp(out), out.size))
        return out

    def state(self):
        p, q, v = (np.empty((self.n, 3), np.float32) for _ in range(3))
        self.check(self.lib.ssbl_ext_download_state(self.h, _fp(p), _fp(q), _fp(v), p.size))
        return p, q, v

    def diag(self):
        d = abi.Diagnostics()
        self.check(self.lib.ssbl_get_diagnostics(self.h, ctypes.byref(d)))
        return d

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.ssbl_destroy_solver(self.h)
            self.h = None
