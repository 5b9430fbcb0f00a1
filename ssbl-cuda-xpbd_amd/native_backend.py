"""ctypes bridge to the ABI41 library -- the host-side mirror of the reference's
native_backend.py (NativeXpbdSolver, native_backend.py:902-1463).

The default library is the gfx950 HIP build next to this file,
libssbl_xpbd_hip_abi41.so; SSBL_NATIVE_DLL_PATH overrides the path exactly as
in the reference (native_backend.py:511-514).  A missing or unloadable library
raises NativeBackendUnavailable -- there is no fallback implementation.

Any library exporting the ABI41 symbols can be passed explicitly (`lib=`); the
tests use that to drive the CPU oracle through the same class.
"""
from __future__ import annotations

import ctypes
import hashlib
import os
from pathlib import Path

import numpy as np

from . import abi
from .scenes import ClothBuildData, SolverOptions

LIBRARY_NAME = "libssbl_xpbd_hip_abi41.so"
_CACHE: dict[str, ctypes.CDLL] = {}


class NativeBackendUnavailable(RuntimeError):
    pass


class NativeSolverError(RuntimeError):
    pass


def dll_path() -> str:
    override = os.environ.get("SSBL_NATIVE_DLL_PATH")
    if override:
        return override
    return str(Path(__file__).resolve().parent / LIBRARY_NAME)


def load_library(path: str | None = None) -> ctypes.CDLL:
    path = str(path or dll_path())
    if path in _CACHE:
        return _CACHE[path]
    if not os.path.exists(path):
        raise NativeBackendUnavailable(f"missing ABI41 library {path}; build it with __graft_entry__.build()")
    try:
        lib = ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)
    except OSError as exc:
        raise NativeBackendUnavailable(f"unable to load ABI41 library {path}: {exc}") from exc
    abi.bind(lib, with_ext=True)
    _CACHE[path] = lib
    return lib


def content_signature(arr) -> tuple:
    """dtype, shape and a 16-byte BLAKE2b digest of an array's bytes: the key the
    frame-input update uses to skip re-uploading unchanged colliders
    (native_backend.py:644-677)."""
    arr = np.ascontiguousarray(arr)
    h = hashlib.blake2b(digest_size=16)
    h.update(arr.dtype.str.encode("ascii", errors="ignore"))
    h.update(np.asarray(arr.shape, dtype=np.int64).tobytes())
    if arr.size:
        h.update(arr.tobytes())
    return (arr.dtype.str, tuple(int(v) for v in arr.shape), h.digest())


def last_error(lib) -> str:
    raw = lib.ssbl_last_error()
    return raw.decode("utf-8", "replace") if raw else ""


def _fptr(arr):
    if arr is None:
        return None
    return arr.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _iptr(arr):
    if arr is None:
        return None
    return arr.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a if shape is None else a.reshape(shape)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def config_from_options(cloth: ClothBuildData, options: SolverOptions, static_triangles: np.ndarray) -> abi.Config:
    """SolverOptions -> SsblXpbdConfig (native_backend.py:712-766)."""
    cfg = abi.Config()
    cfg.vertex_count = len(cloth.positions_world)
    cfg.edge_count = len(cloth.edges)
    cfg.bend_count = len(cloth.bends)
    cfg.lra_count = len(cloth.lra_edges)
    cfg.triangle_count = len(cloth.triangles)
    cfg.static_triangle_count = len(static_triangles)
    cfg.edge_color_count = max(len(cloth.edge_color_offsets) - 1, 0)
    cfg.bend_color_count = max(len(cloth.bend_color_offsets) - 1, 0)
    cfg.lra_color_count = max(len(cloth.lra_color_offsets) - 1, 0)
    for name in ("dt", "damping", "stretch_compliance", "bend_compliance", "lra_compliance", "collision_margin",
                 "ground_height", "sphere_radius", "cloth_thickness", "self_collision_distance", "rest_volume",
                 "volume_compliance", "pressure_strength", "volume_target_scale", "contact_friction",
                 "contact_tangent_damping", "contact_compliance", "stretch_optimization_strength", "static_sdf_voxel_size"):
        setattr(cfg, name, float(getattr(options, name)))
    for name in ("use_ground", "use_wall", "use_sphere", "self_collision", "self_collision_mode", "self_collision_interval",
                 "max_self_collision_neighbors", "use_volume_pressure", "volume_solve_interval", "self_probe_interval",
                 "self_surface_pair_interval", "jitter_stabilizer_enabled", "fast_self_collision_passes",
                 "stretch_optimization_enabled", "static_sdf_band_voxels", "static_sdf_max_resolution"):
        setattr(cfg, name, int(getattr(options, name)))
    cfg.gravity = (ctypes.c_float * 3)(*map(float, options.gravity))
    cfg.wall_origin = (ctypes.c_float * 3)(*map(float, options.wall_origin))
    cfg.wall_normal = (ctypes.c_float * 3)(*map(float, options.wall_normal))
    cfg.sphere_center = (ctypes.c_float * 3)(*map(float, options.sphere_center))
    return cfg


def runtime_colliders_from_options(options: SolverOptions) -> abi.RuntimeColliders:
    rc = abi.RuntimeColliders()
    rc.use_ground = int(options.use_ground)
    rc.ground_height = float(options.ground_height)
    rc.use_wall = int(options.use_wall)
    rc.wall_origin = (ctypes.c_float * 3)(*map(float, options.wall_origin))
    rc.wall_normal = (ctypes.c_float * 3)(*map(float, options.wall_normal))
    rc.use_sphere = int(options.use_sphere)
    rc.sphere_center = (ctypes.c_float * 3)(*map(float, options.sphere_center))
    rc.sphere_radius = float(options.sphere_radius)
    return rc


class NativeXpbdSolver:
    """One ABI41 solver handle (native_backend.py:902-1463)."""

    def __init__(self, cloth: ClothBuildData, options: SolverOptions, static_triangles: np.ndarray | None = None,
                 *, lib: ctypes.CDLL | None = None):
        self._lib = lib if lib is not None else load_library()
        caps = self._lib.ssbl_capabilities()
        if options.stretch_optimization_enabled and not caps & abi.CAP_STRETCH_OPTIMIZATION:
            raise NativeSolverError("loaded ABI41 library does not support hard stretch optimization")
        if not caps & abi.CAP_PIN_WEIGHTS:
            raise NativeSolverError("loaded ABI41 library does not support vertex-group pin weights")
        static = _f32(np.zeros((0, 9)) if static_triangles is None else static_triangles, (-1, 9))
        self.vertex_count = len(cloth.positions_world)
        self.triangles = np.array(cloth.triangles, dtype=np.int32).reshape(-1, 3)
        self._keep = dict(
            positions=_f32(cloth.positions_world, (-1, 3)), inv_mass=_f32(cloth.inv_mass), edges=_i32(cloth.edges),
            edge_rest=_f32(cloth.edge_rest_lengths), edge_off=_i32(cloth.edge_color_offsets), bends=_i32(cloth.bends),
            bend_rest=_f32(cloth.bend_rest_lengths), bend_off=_i32(cloth.bend_color_offsets), lra=_i32(cloth.lra_edges),
            lra_rest=_f32(cloth.lra_rest_lengths), lra_off=_i32(cloth.lra_color_offsets), tris=_i32(cloth.triangles),
            static=static)
        k = self._keep
        cfg = config_from_options(cloth, options, static)
        mesh = abi.Mesh(
            positions=_fptr(k["positions"]), inv_mass=_fptr(k["inv_mass"]), edges=_iptr(k["edges"]),
            edge_rest_lengths=_fptr(k["edge_rest"]), edge_color_offsets=_iptr(k["edge_off"]), bends=_iptr(k["bends"]),
            bend_rest_lengths=_fptr(k["bend_rest"]), bend_color_offsets=_iptr(k["bend_off"]), lra_edges=_iptr(k["lra"]),
            lra_rest_lengths=_fptr(k["lra_rest"]), lra_color_offsets=_iptr(k["lra_off"]), triangles=_iptr(k["tris"]),
            static_triangles=_fptr(k["static"]))
        self._handle = self._lib.ssbl_create_solver(ctypes.byref(cfg), ctypes.byref(mesh))
        if not self._handle:
            raise NativeSolverError(last_error(self._lib))
        self._keep = None
        self._dynamic_triangle_signature = None
        self._dynamic_particle_signature = None
        if len(cloth.pin_indices) > 0:
            self.update_pin_targets(cloth.pin_indices, cloth.pin_targets_world, cloth.pin_weights)
        if cloth.start_positions is not None:
            self.update_positions(cloth.start_positions)

    @property
    def handle(self):
        return self._handle

    @property
    def lib(self):
        return self._lib

    def _check(self, ok) -> None:
        if not ok:
            raise NativeSolverError(last_error(self._lib))

    def close(self) -> None:
        if getattr(self, "_handle", None):
            self._lib.ssbl_destroy_solver(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reset(self) -> None:
        self._check(self._lib.ssbl_reset_solver(self._handle))

    def update_pin_targets(self, indices, positions, weights=None) -> None:
        idx = _i32(indices)
        pos = _f32(positions, (-1, 3))
        w = None if weights is None or len(weights) == 0 else _f32(weights)
        self._check(self._lib.ssbl_update_pin_targets(self._handle, _iptr(idx), _fptr(pos), _fptr(w), len(idx)))

    def update_runtime_colliders(self, options: SolverOptions) -> None:
        rc = runtime_colliders_from_options(options)
        self._check(self._lib.ssbl_update_runtime_colliders(self._handle, ctypes.byref(rc)))

    def update_positions(self, positions) -> None:
        p = _f32(positions, (-1, 3))
        self._check(self._lib.ssbl_update_positions(self._handle, _fptr(p), p.size))

    def update_static_triangles(self, static_triangles) -> None:
        t = _f32(static_triangles, (-1, 9))
        self._check(self._lib.ssbl_update_static_triangles(self._handle, _fptr(t), len(t)))

    def update_force_fields(self, fields: list[dict]) -> None:
        arr = (abi.ForceField * max(len(fields), 1))()
        for k, f in enumerate(fields):
            for key, value in f.items():
                cur = getattr(arr[k], key)
                setattr(arr[k], key, (ctypes.c_float * 3)(*value) if hasattr(cur, "__len__") else value)
        fi = abi.FrameInputs()
        fi.update_force_fields = 1
        fi.force_fields = ctypes.cast(arr, ctypes.POINTER(abi.ForceField))
        fi.force_field_count = len(fields)
        self._check(self._lib.ssbl_update_frame_inputs(self._handle, ctypes.byref(fi)))

    def set_static_sdf(self, values, origin, voxel: float) -> None:
        """A static SDF grid (values indexed [z, y, x], negative inside) in place of one built from triangles."""
        v = np.ascontiguousarray(values, np.float32)
        dims = np.asarray([v.shape[2], v.shape[1], v.shape[0]], np.int32)
        ov = np.asarray([origin[0], origin[1], origin[2], voxel], np.float32)
        flat = v.reshape(-1)
        self._check(self._lib.ssbl_ext_set_static_sdf(self._handle, _iptr(dims), _fptr(ov), _fptr(flat), flat.size))

    def update_dynamic_triangles(self, triangles) -> None:
        """Triangle soup (T x 3 x 3) this cloth collides with (native_backend.py ssbl_update_dynamic_triangles)."""
        t = _f32(triangles, (-1, 9))
        self._check(self._lib.ssbl_update_dynamic_triangles(self._handle, _fptr(t), len(t)))

    def update_dynamic_particles(self, positions, radii) -> None:
        p = _f32(positions, (-1, 3))
        r = _f32(radii)
        fi = abi.FrameInputs()
        fi.update_dynamic_particles = 1
        fi.dynamic_particle_positions = _fptr(p)
        fi.dynamic_particle_radii = _fptr(r)
        fi.dynamic_particle_count = len(p)
        self._check(self._lib.ssbl_update_frame_inputs(self._handle, ctypes.byref(fi)))

    def update_frame_inputs(self, *, pin_indices=None, pin_positions=None, pin_weights=None, update_pin: bool = False,
                            options: SolverOptions | None = None, update_runtime: bool = False, static_triangles=None,
                            update_static: bool = False, dynamic_triangles=None, update_dynamic: bool = False,
                            dynamic_triangle_vertices=None, dynamic_triangle_indices=None,
                            dynamic_triangle_topology_signature=None, dynamic_particles: dict | None = None,
                            update_dynamic_particles: bool = False, force_fields: list[dict] | None = None,
                            update_force_fields: bool = False) -> bool:
        """All of a frame's inputs in one ssbl_update_frame_inputs call
        (native_backend.py:1064-1279).

        Indexed dynamic triangles (vertices + indices) are sent as signed
        colliders.  Their topology counts as changed when the vertex count,
        triangle count or index content differ from the last upload, unless
        the caller passes its own signature.  A soup, or particles whose
        content equals the last upload, is not sent again.  Returns False
        when there was nothing to send."""
        pin_idx = _i32(pin_indices if pin_indices is not None else np.empty(0, np.int32))
        pin_pos = _f32(pin_positions if pin_positions is not None else np.empty((0, 3), np.float32), (-1, 3))
        pin_w = np.ones(len(pin_idx), np.float32) if pin_weights is None else _f32(pin_weights)
        if len(pin_idx) != len(pin_pos) or len(pin_idx) != len(pin_w):
            raise NativeSolverError("Pin index, target, and weight arrays must have matching lengths.")
        verts = _f32(dynamic_triangle_vertices if dynamic_triangle_vertices is not None else np.empty((0, 3), np.float32), (-1, 3))
        idx = _i32(dynamic_triangle_indices if dynamic_triangle_indices is not None else np.empty((0, 3), np.int32)).reshape(-1, 3)
        indexed = bool(update_dynamic and dynamic_triangle_vertices is not None and dynamic_triangle_indices is not None)
        if update_dynamic and not indexed and dynamic_triangles is None and len(verts) and len(idx):
            dynamic_triangles = verts[idx]
        soup = (_f32(dynamic_triangles, (-1, 9)) if (dynamic_triangles is not None and not indexed)
                else np.empty((0, 9), np.float32))
        tri_count = len(idx) if indexed else len(soup)
        signature, topology_changed = None, False
        if update_dynamic and indexed:
            signature = ((len(verts), tri_count, "provided", dynamic_triangle_topology_signature)
                         if dynamic_triangle_topology_signature is not None
                         else (len(verts), tri_count, content_signature(idx)))
            topology_changed = self._dynamic_triangle_signature != signature
        elif update_dynamic:
            signature = (tri_count, content_signature(soup))
            if signature == self._dynamic_triangle_signature:
                update_dynamic = False
        part = dynamic_particles or {}
        ppos = _f32(part.get("positions", np.empty((0, 3), np.float32)), (-1, 3))
        prad = _f32(part.get("radii", np.empty(0, np.float32))) if len(ppos) else np.empty(0, np.float32)
        if len(ppos) and len(prad) != len(ppos):
            raise NativeSolverError("Dynamic particle arrays must have matching lengths.")
        psig = None
        if update_dynamic_particles:
            psig = (("provided", part["signature"]) if part.get("signature") is not None
                    else (len(ppos), content_signature(ppos), content_signature(prad)))
            if psig == self._dynamic_particle_signature:
                update_dynamic_particles = False
        if not (update_pin or update_runtime or update_static or update_dynamic or update_dynamic_particles or update_force_fields):
            return False
        static = _f32(static_triangles if static_triangles is not None else np.empty((0, 9), np.float32), (-1, 9))
        fields = list(force_fields or [])
        field_arr = (abi.ForceField * max(len(fields), 1))()
        for k, f in enumerate(fields):
            for key, value in f.items():
                cur = getattr(field_arr[k], key)
                setattr(field_arr[k], key, (ctypes.c_float * 3)(*value) if hasattr(cur, "__len__") else value)
        fi = abi.FrameInputs()
        fi.update_pin_targets = int(update_pin)
        fi.pin_indices, fi.pin_positions, fi.pin_weights, fi.pin_count = _iptr(pin_idx), _fptr(pin_pos), _fptr(pin_w), len(pin_idx)
        fi.update_runtime_colliders = int(update_runtime and options is not None)
        if options is not None:
            fi.runtime_colliders = runtime_colliders_from_options(options)
        fi.update_static_triangles = int(update_static)
        fi.static_triangles, fi.static_triangle_count = _fptr(static), len(static)
        fi.update_dynamic_triangles = int(update_dynamic and not indexed)
        fi.dynamic_triangles, fi.dynamic_triangle_count = _fptr(soup), 0 if indexed else tri_count
        fi.update_dynamic_indexed_triangles = int(update_dynamic and indexed)
        fi.dynamic_triangle_vertices, fi.dynamic_triangle_vertex_count = _fptr(verts), len(verts)
        fi.dynamic_triangle_indices = _iptr(idx)
        fi.dynamic_indexed_triangle_count = tri_count if indexed else 0
        fi.dynamic_triangle_topology_changed = int(topology_changed)
        fi.update_dynamic_particles = int(update_dynamic_particles)
        fi.dynamic_particle_positions, fi.dynamic_particle_radii, fi.dynamic_particle_count = _fptr(ppos), _fptr(prad), len(ppos)
        fi.update_force_fields = int(update_force_fields)
        fi.force_fields = ctypes.cast(field_arr, ctypes.POINTER(abi.ForceField))
        fi.force_field_count = len(fields)
        self._check(self._lib.ssbl_update_frame_inputs(self._handle, ctypes.byref(fi)))
        if update_dynamic:
            self._dynamic_triangle_signature = signature
        if update_dynamic_particles:
            self._dynamic_particle_signature = psig
        return True

    def attach_global_dynamic_scene(self, scene, target_source_id: int = 0, target_collision_layer: int = 0,
                                    cross_mode: int = 1) -> None:
        self._check(self._lib.ssbl_attach_global_dynamic_scene(self._handle, scene.handle, int(target_source_id),
                                                               int(target_collision_layer), int(cross_mode)))

    def clear_global_dynamic_scene(self) -> None:
        self._check(self._lib.ssbl_clear_global_dynamic_scene(self._handle))

    def step(self, substeps: int, iterations: int, diagnostics: bool = True, synchronize: bool = True) -> None:
        self._check(self._lib.ssbl_step_solver_ex(self._handle, int(substeps), int(iterations), int(bool(diagnostics)),
                                                   int(bool(synchronize or diagnostics))))

    def download_positions(self) -> np.ndarray:
        out = np.empty((self.vertex_count, 3), np.float32)
        self._check(self._lib.ssbl_download_positions(self._handle, _fptr(out), out.size))
        return out

    def download_state(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        p, q, v = (np.empty((self.vertex_count, 3), np.float32) for _ in range(3))
        self._check(self._lib.ssbl_ext_download_state(self._handle, _fptr(p), _fptr(q), _fptr(v), p.size))
        return p, q, v

    def diagnostics(self) -> abi.Diagnostics:
        d = abi.Diagnostics()
        self._check(self._lib.ssbl_get_diagnostics(self._handle, ctypes.byref(d)))
        return d

    def set_stretch_mode(self, mode: int) -> None:
        self._check(self._lib.ssbl_ext_set_stretch_mode(self._handle, int(mode)))

    def self_hash_cells(self) -> np.ndarray:
        out = np.empty((self.vertex_count, 3), np.int32)
        self._check(self._lib.ssbl_ext_self_hash_cells(self._handle, _iptr(out), out.size))
        return out

    def self_neighbors(self) -> tuple[np.ndarray, np.ndarray]:
        counts = np.empty(self.vertex_count, np.uint32)
        idx = np.empty((self.vertex_count, 32), np.uint32)
        self._check(self._lib.ssbl_ext_self_neighbors(
            self._handle, counts.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)),
            idx.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)), self.vertex_count))
        return counts, idx

    def static_sdf(self) -> tuple[np.ndarray, np.ndarray]:
        dims = np.zeros(3, np.int32)
        ov = np.zeros(4, np.float32)
        self._check(self._lib.ssbl_ext_static_sdf(self._handle, _iptr(dims), _fptr(ov), None, 0))
        vals = np.empty(int(np.prod(dims)), np.float32)
        self._check(self._lib.ssbl_ext_static_sdf(self._handle, _iptr(dims), _fptr(ov), _fptr(vals), vals.size))
        return vals.reshape(dims[2], dims[1], dims[0]), ov


class GlobalDynamicScene:
    """Colliders shared by the cloths of a session (ssbl_*_global_dynamic_scene)."""

    CROSS_OFF, CROSS_ALL, CROSS_LOWER_LAYERS = 0, 1, 2

    def __init__(self, lib=None):
        self._lib = lib if lib is not None else load_library()
        self._handle = self._lib.ssbl_create_global_dynamic_scene()
        if not self._handle:
            raise NativeSolverError(last_error(self._lib))

    @property
    def handle(self):
        return self._handle

    def update(self, *, particle_positions, particle_radii, particle_source_ids, particle_collision_layers,
               triangle_vertices, triangle_source_ids, triangle_collision_layers, collision_margin: float,
               cloth_thickness: float, particle_prev_positions=None) -> None:
        keep = dict(pp=_f32(particle_positions, (-1, 3)), pr=_f32(particle_radii), ps=_i32(particle_source_ids),
                    pl=_i32(particle_collision_layers), tv=_f32(triangle_vertices, (-1, 9)), ts=_i32(triangle_source_ids),
                    tl=_i32(triangle_collision_layers),
                    pq=None if particle_prev_positions is None else _f32(particle_prev_positions, (-1, 3)))
        inp = abi.GlobalDynamicSceneInputs()
        inp.particle_positions = _fptr(keep["pp"])
        inp.particle_prev_positions = _fptr(keep["pq"])
        inp.particle_radii = _fptr(keep["pr"])
        inp.particle_source_ids = _iptr(keep["ps"])
        inp.particle_collision_layers = _iptr(keep["pl"])
        inp.particle_count = len(keep["pp"])
        inp.triangle_vertices = _fptr(keep["tv"])
        inp.triangle_source_ids = _iptr(keep["ts"])
        inp.triangle_collision_layers = _iptr(keep["tl"])
        inp.triangle_count = len(keep["tv"])
        inp.collision_margin = float(collision_margin)
        inp.cloth_thickness = float(cloth_thickness)
        if not self._lib.ssbl_update_global_dynamic_scene(self._handle, ctypes.byref(inp)):
            raise NativeSolverError(last_error(self._lib))

    def update_from_solvers(self, solvers, source_ids, collision_layers, particle_radii, collision_margin: float,
                            cloth_thickness: float) -> None:
        """Every solver's current vertices (particles) and mesh triangles, gathered by the library."""
        handles = (ctypes.c_void_p * len(solvers))(*[s.handle for s in solvers])
        ids, layers, radii = _i32(source_ids), _i32(collision_layers), _f32(particle_radii)
        if not self._lib.ssbl_ext_update_global_scene_from_handles(self._handle, handles, _iptr(ids), _iptr(layers),
                                                                   _fptr(radii), len(solvers), float(collision_margin),
                                                                   float(cloth_thickness)):
            raise NativeSolverError(last_error(self._lib))

    def close(self) -> None:
        if getattr(self, "_handle", None):
            self._lib.ssbl_destroy_global_dynamic_scene(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SolverBatch:
    """Many handles stepped together with one launch per solver phase."""

    def __init__(self, solvers: list[NativeXpbdSolver]):
        if not solvers:
            raise ValueError("empty batch")
        self._lib = solvers[0].lib
        self.solvers = list(solvers)
        handles = (ctypes.c_void_p * len(solvers))(*[s.handle for s in solvers])
        self._handle = self._lib.ssbl_ext_create_batch(handles, len(solvers))
        if not self._handle:
            raise NativeSolverError(last_error(self._lib))

    def set_members(self, solvers: list[NativeXpbdSolver]) -> None:
        """Points the batch at new members, reusing its device tables."""
        if not solvers:
            raise ValueError("empty batch")
        handles = (ctypes.c_void_p * len(solvers))(*[s.handle for s in solvers])
        if not self._lib.ssbl_ext_batch_set_members(self._handle, handles, len(solvers)):
            raise NativeSolverError(last_error(self._lib))
        self.solvers = list(solvers)

    def step(self, substeps: int, iterations: int, synchronize: bool = False, diagnostics: bool = False) -> None:
        """One batched step; with diagnostics, every member's counters are
        read back afterwards (each solver's diagnostics() then holds them)."""
        if diagnostics:
            ok = self._lib.ssbl_ext_step_batch_ex(self._handle, int(substeps), int(iterations), 1, 1)
        else:
            ok = self._lib.ssbl_ext_step_batch(self._handle, int(substeps), int(iterations), int(bool(synchronize)))
        if not ok:
            raise NativeSolverError(last_error(self._lib))

    def graph_stats(self) -> tuple[int, int]:
        """(captures, replays) of the batch's step graph."""
        return graph_stats(self._lib, batch=self._handle)

    def tuned_pool_limit(self) -> int:
        """The particle pass's pooling limit the batch's tuning chose (-1: not yet)."""
        return int(self._lib.ssbl_ext_tuned_pool_limit(None, self._handle))

    def time_pass(self, launches: int = 64) -> float:
        """Mean ms of `launches` back-to-back fused constraint passes; the
        members' state is restored afterwards.  The batch must have been
        stepped since it was created or given new members."""
        out = np.zeros(1, np.float32)
        if not self._lib.ssbl_ext_time_batch_pass(self._handle, int(launches), _fptr(out)):
            raise NativeSolverError(last_error(self._lib))
        return float(out[0])

    def time_iteration(self, iterations: int = 64) -> float:
        """Mean ms of one iteration over all members, a batch stepped as two
        parts running its parts' passes concurrently as its steps do."""
        out = np.zeros(1, np.float32)
        if not self._lib.ssbl_ext_time_batch_iteration(self._handle, int(iterations), _fptr(out)):
            raise NativeSolverError(last_error(self._lib))
        return float(out[0])

    def close(self) -> None:
        if getattr(self, "_handle", None):
            self._lib.ssbl_ext_destroy_batch(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def synchronize(lib=None) -> None:
    lib = lib or load_library()
    if not lib.ssbl_ext_synchronize():
        raise NativeSolverError(last_error(lib))


def set_graph_mode(lib, enabled: bool) -> bool:
    """Turns graph replay of the substep loop on or off; returns the previous mode."""
    return bool(lib.ssbl_ext_set_graph_mode(int(bool(enabled))))


def graph_stats(lib, handle=None, batch=None) -> tuple[int, int]:
    """(captures, replays) of one handle's step graph or one batch's."""
    captures, replays = ctypes.c_longlong(0), ctypes.c_longlong(0)
    if not lib.ssbl_ext_graph_stats(handle, batch, ctypes.byref(captures), ctypes.byref(replays)):
        raise NativeSolverError(last_error(lib))
    return int(captures.value), int(replays.value)


def set_batch_parts(lib, parts: int) -> int:
    """1 or 2 concurrent parts for batches created afterwards; returns the previous setting."""
    return int(lib.ssbl_ext_set_batch_parts(int(parts)))


def set_autotune(lib, enabled: bool) -> bool:
    """Turns launch tuning on or off; returns the previous mode."""
    return bool(lib.ssbl_ext_set_autotune(int(bool(enabled))))
