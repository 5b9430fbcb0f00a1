"""Serving loop for cloth sessions on one GPU.

A session is one cloth advanced frame by frame. It has a wind force field,
an optional moving sphere collider, and a number of frames to run. The server:

* admits waiting sessions in arrival order while the HBM budget and the
  session slots allow, building their handles (host set-up and uploads) on
  builder threads beside the tick loop;
* keeps every admitted session's state resident on the device (one ABI41
  handle each) until its last frame;
* steps resident sessions together each tick (continuous batching): a tick
  is one frame of each session it takes, least recently stepped first, up
  to a per-tick vertex capacity, run as one batched step
  (ssbl_ext_step_batch), so one launch per solver phase covers them all.
  Sessions that arrive while a tick runs join at the next tick. Finished
  sessions leave, and their memory frees room for the queue;
* records each session's per-frame step latency (from the start of the
  tick's input updates to the end of its step), its frame interval (from the
  end of its previous frame) and the aggregate vertex-substeps/s;
* writes a PC2 cache of every session that asks for one (``cache_path``):
  the start positions, then one sample per frame, committed when the session
  ends (the reference's realtime auto-cache, session_manager.py:4650-4728).

The per-frame order follows the reference's session manager
(session_manager.py:2520-2591): frame inputs of every slot, then the steps,
then downloads (every frame for cached sessions, otherwise at a session's
end).

Batched stepping gives each handle the same result as stepping it alone
(ssbl_ext_step_batch), so a session's result does not depend on the mix it
was served in.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field, replace
from functools import lru_cache

import numpy as np

from . import native_backend as nb
from .pc2 import Pc2Writer
from .scenes import ClothBuildData, SolverOptions, build_cloth, grid_mesh, y_up_to_z_up


# ------------------------------------------------------------- requests
@dataclass
class SpherePath:
    """Lissajous path of the sphere collider: center + amplitude * sin(2 pi f t + phase)."""
    center: tuple
    amplitude: tuple
    frequency: tuple
    phase: tuple
    radius: float

    def at(self, t: float) -> tuple:
        return tuple(float(c + a * math.sin(2.0 * math.pi * f * t + p))
                     for c, a, f, p in zip(self.center, self.amplitude, self.frequency, self.phase))


@dataclass
class SessionRequest:
    sid: int
    arrival: float                      # seconds after start (wall clock) or tick index (tick clock)
    cloth: ClothBuildData
    options: SolverOptions
    frames: int
    wind: dict | None = None            # SsblXpbdForceField fields of a wind field (type 1)
    sphere: SpherePath | None = None
    cache_path: str | None = None       # PC2 file of the session's frames
    start_frame: int = 1

    @property
    def vertex_count(self) -> int:
        return len(self.cloth.positions_world)


def wind_field(direction, strength: float) -> dict:
    """A wind force field (A41:1433-1447): force along `direction` scaled by the facing of the surface."""
    d = np.asarray(direction, np.float64)
    d = d / max(float(np.linalg.norm(d)), 1.0e-12)
    return {"type": 1, "strength": float(strength), "direction": tuple(float(x) for x in d), "axis": (0.0, 0.0, 1.0),
            "apply_to_location": 1, "noise": 0.0, "flow": 0.0}


def _next_pow2(v: int) -> int:
    r = 1
    while r < max(v, 1):
        r <<= 1
    return r


def estimate_device_bytes(cloth: ClothBuildData, options: SolverOptions) -> int:
    """Upper bound of the device memory one handle holds, for admission.

    Sums the library's allocations per element (float4 state x 5, the
    dynamic contacts' 8-byte key, edge incidences at both endpoints plus the
    8-slot table, triangles and their
    incidences, bends, pins, per-vertex pressure terms, the static SDF at its
    maximum resolution, the three self-collision hashes with their tables,
    and the PCG vectors), plus a per-allocation allowance. The admission
    test checks it against ssbl_ext_device_bytes.
    """
    v = len(cloth.positions_world)
    e = len(cloth.edges)
    t = len(cloth.triangles)
    b = len(cloth.bends)
    lra = len(cloth.lra_edges)
    p = len(cloth.pin_indices)
    core = (5 * 16 + 6 * 4 + 8 * 8 + 2 * 4 + 8) * v + 16 * e + 8 * e + (16 + 12) * t + (8 + 4 + 32 + 16) * b + 32 * lra + 24 * p
    res = max(int(options.static_sdf_max_resolution), 2)
    sdf = 4 * res ** 3
    self_bytes = 0
    if options.self_collision:
        def hash_bytes(n, cap):
            return 8 * _next_pow2(max(1024, 4 * n)) + 16 * cap
        self_bytes = (hash_bytes(v, v) + hash_bytes(t, 64 * t) + hash_bytes(e, 64 * e)
                      + (16 + 4 + 128 + 24 + 4) * v + 16 * t + 8 * e + 4096)
    pcg = (144 * v + 60 * e) if options.stretch_optimization_enabled else 0
    return int(1.1 * (core + sdf + self_bytes + pcg)) + 64 * 4096 + (1 << 20)


# ------------------------------------------------------------- sessions
@dataclass
class SessionResult:
    sid: int
    vertex_count: int
    frames: int
    positions: np.ndarray
    frame_ms: list          # step latency of each frame: the duration of the batched step that advanced it
    frame_gap_ms: list      # time from the previous frame's end (admission for the first) to this frame's end
    arrival: float
    admitted: float
    finished: float
    device_bytes: int
    cache_path: str | None = None


@dataclass
class _Active:
    req: SessionRequest
    solver: nb.NativeXpbdSolver
    device_bytes: int
    admitted: float
    order: int                       # admission order (ties of the round robin)
    last_done: float = 0.0           # wall time of the last frame's end (admission before the first)
    last_tick: int = -1              # last tick that stepped the session
    frame: int = 0
    frame_ms: list = field(default_factory=list)
    frame_gap_ms: list = field(default_factory=list)
    cache: Pc2Writer | None = None


# sessions counted as "10k-vertex" in the report: the size classes between 7k and 14k vertices
TEN_K = (7000, 14000)


@dataclass
class ServeReport:
    results: dict
    ticks: int
    wall_s: float
    vertex_substeps: int
    frame_ms_p50: float
    frame_ms_p99: float
    peak_device_bytes: int
    peak_active_vertices: int     # most vertices stepped in one tick
    peak_sessions: int            # most sessions resident at once
    mean_queue_wait: float
    unfinished: int
    budget_waits: int = 0         # times an arrived session waited for device memory
    step_ms_10k: tuple = (0.0, 0.0)   # percentiles over every frame served (finished sessions or not)
    gap_ms: tuple = (0.0, 0.0)
    gap_ms_10k: tuple = (0.0, 0.0)
    sessions_10k: int = 0             # finished 10k-vertex sessions

    @property
    def vertex_substeps_per_s(self) -> float:
        return self.vertex_substeps / self.wall_s if self.wall_s > 0 else 0.0

    def summary(self) -> dict:
        return {"sessions_completed": len(self.results), "sessions_unfinished": self.unfinished, "ticks": self.ticks,
                "wall_s": self.wall_s, "vertex_substeps_per_s": self.vertex_substeps_per_s,
                "frame_ms_p50": self.frame_ms_p50, "frame_ms_p99": self.frame_ms_p99,
                "frame_ms_p50_10k": self.step_ms_10k[0], "frame_ms_p99_10k": self.step_ms_10k[1],
                "sessions_10k": self.sessions_10k,
                "frame_interval_ms_p50": self.gap_ms[0], "frame_interval_ms_p99": self.gap_ms[1],
                "frame_interval_ms_p50_10k": self.gap_ms_10k[0], "frame_interval_ms_p99_10k": self.gap_ms_10k[1],
                "peak_device_bytes": self.peak_device_bytes, "peak_tick_vertices": self.peak_active_vertices,
                "peak_resident_sessions": self.peak_sessions, "mean_queue_wait": self.mean_queue_wait,
                "budget_waits": self.budget_waits}


def _pcts(values) -> tuple:
    return (float(np.percentile(values, 50)), float(np.percentile(values, 99))) if len(values) else (0.0, 0.0)


class ClothServer:
    """Admission by HBM budget and continuous batching over ABI41 handles of one library.

    Two limits, as in request-serving systems that keep state resident:
      residency  -- sessions are admitted in arrival order while the device
                    memory of the resident and building sessions (their
                    estimates until built, then what the handle holds) stays
                    within `hbm_budget_bytes` and at most `max_sessions` are
                    resident; a finished session frees its memory for the
                    queue;
      scheduling -- each tick steps resident sessions, least recently
                    stepped first, until `max_batch_vertices` vertices are in
                    the tick (at least one session), as one batched step.  The
                    cap bounds a tick's duration, which is every stepped
                    session's step latency; with more resident work than one
                    tick takes, the sessions take turns.
    With `builders` > 0 (wall clock only), handles are created by that many
    threads beside the tick loop: host set-up and uploads of a new session
    never delay a tick.  Graph replay is turned off while serving: the
    batch's members change from tick to tick, so a capture would rarely be
    replayed.
    """

    def __init__(self, *, hbm_budget_bytes: int, substeps: int, iterations: int, max_sessions: int = 4096,
                 max_batch_vertices: int = 1 << 62, static_triangles: np.ndarray | None = None, lib=None, builders: int = 0):
        self.lib = lib if lib is not None else nb.load_library()
        self.budget = int(hbm_budget_bytes)
        self.substeps = int(substeps)
        self.iterations = int(iterations)
        self.max_sessions = int(max_sessions)
        self.max_batch_vertices = int(max_batch_vertices)
        self.static_triangles = static_triangles
        self.builders = int(builders)
        self.pending: list[SessionRequest] = []
        self.building: list = []                  # (request, estimate, future) in admission order
        self.resident: list[_Active] = []
        self.results: dict[int, SessionResult] = {}
        self.used_bytes = 0                       # resident sessions' bytes plus building sessions' estimates
        self.ticks = 0
        self.vertex_substeps = 0
        self.peak_bytes = 0
        self.peak_vertices = 0
        self.peak_sessions = 0
        self.budget_waits = 0
        self._order = 0
        self.frame_log: list = []                 # (vertex count, step ms, interval ms) of every frame served
        self._pool = None
        self._batch: nb.SolverBatch | None = None
        self._batch_members: list = []

    # -- admission
    def submit(self, req: SessionRequest) -> None:
        if estimate_device_bytes(req.cloth, req.options) > self.budget or req.vertex_count > self.max_batch_vertices:
            raise ValueError(f"session {req.sid} can never fit the server's budget")
        self.pending.append(req)
        self.pending.sort(key=lambda r: (r.arrival, r.sid))

    def _build(self, req: SessionRequest):
        solver = nb.NativeXpbdSolver(req.cloth, req.options, self.static_triangles, lib=self.lib)
        if req.wind is not None:
            solver.update_force_fields([req.wind])
        held = int(self.lib.ssbl_ext_device_bytes(solver.handle))
        if held < 0:                                # library without device accounting
            held = estimate_device_bytes(req.cloth, req.options)
        return solver, held

    def _start_builds(self, now: float) -> None:
        """Starts building arrived sessions in arrival order while they fit."""
        while self.pending and self.pending[0].arrival <= now:
            req = self.pending[0]
            est = estimate_device_bytes(req.cloth, req.options)
            if len(self.resident) + len(self.building) >= self.max_sessions:
                break
            if self.used_bytes + est > self.budget:
                self.budget_waits += 1
                break
            self.pending.pop(0)
            self.used_bytes += est
            fut = self._pool.submit(self._build, req) if self._pool is not None else self._build(req)
            self.building.append((req, est, fut))

    def _collect_built(self, now: float, wait: bool = False) -> None:
        """Makes built sessions resident (in admission order)."""
        while self.building:
            req, est, fut = self.building[0]
            if self._pool is not None:
                if not wait and not fut.done():
                    break
                solver, held = fut.result()
            else:
                solver, held = fut
            self.building.pop(0)
            self.used_bytes += held - est
            act = _Active(req, solver, held, now, self._order, last_done=time.perf_counter())
            self._order += 1
            if req.cache_path is not None:
                start = req.cloth.start_positions if req.cloth.start_positions is not None else req.cloth.positions_world
                act.cache = Pc2Writer(req.cache_path, req.vertex_count, req.start_frame)
                act.cache.add_sample(start)
            self.resident.append(act)
        self.peak_bytes = max(self.peak_bytes, self.used_bytes)
        self.peak_sessions = max(self.peak_sessions, len(self.resident))

    def admit(self, now: float, wait: bool = False) -> None:
        self._start_builds(now)
        self._collect_built(now, wait)

    # -- scheduling
    def _choose(self) -> list:
        order = sorted(self.resident, key=lambda s: (s.last_tick, s.order))
        chosen, verts = [], 0
        for s in order:
            if chosen and verts + s.req.vertex_count > self.max_batch_vertices:
                continue
            chosen.append(s)
            verts += s.req.vertex_count
        return chosen

    def _frame_inputs(self, s: _Active) -> None:
        if s.req.sphere is None:
            return
        t = s.frame * float(s.req.options.dt)
        s.solver.update_runtime_colliders(replace(s.req.options, use_sphere=True, sphere_center=s.req.sphere.at(t),
                                                  sphere_radius=s.req.sphere.radius))

    def tick(self, now: float) -> bool:
        """One frame of the sessions chosen for this tick, as one batched step."""
        self.admit(now)
        if not self.resident:
            return False
        chosen = self._choose()
        t0 = time.perf_counter()
        for s in chosen:
            self._frame_inputs(s)
        if self._batch is None:
            self._batch = nb.SolverBatch([s.solver for s in chosen])
            self._batch_members = [id(s) for s in chosen]
        elif self._batch_members != [id(s) for s in chosen]:
            self._batch.set_members([s.solver for s in chosen])   # the batch's tables are reused
            self._batch_members = [id(s) for s in chosen]
        self._batch.step(self.substeps, self.iterations, synchronize=True)
        t1 = time.perf_counter()
        ms = (t1 - t0) * 1.0e3
        verts = sum(s.req.vertex_count for s in chosen)
        self.peak_vertices = max(self.peak_vertices, verts)
        self.vertex_substeps += verts * self.substeps
        done = []
        for s in chosen:
            if s.cache is not None:
                s.cache.add_sample(s.solver.download_positions())
            s.frame += 1
            s.frame_ms.append(ms)
            s.frame_gap_ms.append((t1 - s.last_done) * 1.0e3)
            self.frame_log.append((s.req.vertex_count, ms, s.frame_gap_ms[-1]))
            s.last_done = t1
            s.last_tick = self.ticks
            if s.frame >= s.req.frames:
                done.append(s)
        self.ticks += 1
        if done:
            self._batch_members = []     # the next tick points the batch at its new members
            for s in done:
                self._retire(s, now)
        return True

    def _retire(self, s: _Active, now: float) -> None:
        pos = s.solver.download_positions()
        path = s.cache.commit() if s.cache is not None else None
        s.solver.close()
        self.resident.remove(s)
        self.used_bytes -= s.device_bytes
        self.results[s.req.sid] = SessionResult(s.req.sid, s.req.vertex_count, s.frame, pos, s.frame_ms, s.frame_gap_ms,
                                                s.req.arrival, s.admitted, now, s.device_bytes, path)

    def close(self) -> None:
        if self._batch is not None:
            self._batch.close()
            self._batch = None
        if self._pool is not None:
            self._collect_built(0.0, wait=True)
        for s in list(self.resident):
            if s.cache is not None:       # unfinished: the reference drops an uncommitted cache
                s.cache.discard()
            s.solver.close()
        self.resident.clear()

    # -- driver
    def run(self, requests, *, clock: str = "wall", max_seconds: float | None = None, max_ticks: int | None = None) -> ServeReport:
        """Serves `requests` until all finish or a limit is reached.

        clock "wall": arrivals are seconds after the start; "tick": arrivals
        are tick indices and handles are built in the tick loop
        (deterministic, for tests)."""
        for r in requests:
            self.submit(r)
        threaded = clock == "wall" and self.builders > 0
        if threaded:
            from concurrent.futures import ThreadPoolExecutor
            self._pool = ThreadPoolExecutor(max_workers=self.builders)
        graph_before = nb.set_graph_mode(self.lib, False)
        start = time.perf_counter()
        tick_index = 0
        try:
            while self.pending or self.building or self.resident:
                now = (time.perf_counter() - start) if clock == "wall" else float(tick_index)
                if max_seconds is not None and clock == "wall" and now > max_seconds:
                    break
                if max_ticks is not None and tick_index >= max_ticks:
                    break
                ran = self.tick(now)
                tick_index += 1
                if not ran and clock == "wall":
                    time.sleep(0.001 if self.building else min(max((self.pending[0].arrival - now) if self.pending else 0.0, 0.0), 0.01))
            wall = time.perf_counter() - start
            unfinished = len(self.resident) + len(self.building) + len(self.pending)
        finally:
            self.close()
            if self._pool is not None:
                self._pool.shutdown(wait=True)
                self._pool = None
            nb.set_graph_mode(self.lib, graph_before)
        # every frame served, of finished and unfinished sessions alike
        log = np.asarray(self.frame_log, np.float64).reshape(-1, 3)
        tenk = (log[:, 0] >= TEN_K[0]) & (log[:, 0] <= TEN_K[1])
        waits = [r.admitted - r.arrival for r in self.results.values()]
        p50, p99 = _pcts(log[:, 1])
        n10k = len({r.sid for r in self.results.values() if TEN_K[0] <= r.vertex_count <= TEN_K[1]})
        return ServeReport(dict(self.results), self.ticks, wall, self.vertex_substeps, p50, p99, self.peak_bytes,
                           self.peak_vertices, self.peak_sessions, float(np.mean(waits)) if waits else 0.0, unfinished,
                           self.budget_waits, _pcts(log[tenk, 1]), _pcts(log[:, 2]), _pcts(log[tenk, 2]), n10k)


def run_alone(req: SessionRequest, substeps: int, iterations: int, static_triangles=None, lib=None) -> np.ndarray:
    """The same session stepped by itself with ssbl_step_solver_ex (no batch)."""
    solver = nb.NativeXpbdSolver(req.cloth, req.options, static_triangles, lib=lib)
    try:
        if req.wind is not None:
            solver.update_force_fields([req.wind])
        for f in range(req.frames):
            if req.sphere is not None:
                t = f * float(req.options.dt)
                solver.update_runtime_colliders(replace(req.options, use_sphere=True, sphere_center=req.sphere.at(t),
                                                        sphere_radius=req.sphere.radius))
            solver.step(substeps, iterations, diagnostics=False, synchronize=False)
        return solver.download_positions()
    finally:
        solver.close()


# ------------------------------------------------------ served workload
# The reference's default material hardness (xpbd_core.py:20) turns the
# hard-stretch PCG on with strength 0.4 (xpbd_core.py:238-239).
DEFAULT_HARDNESS = 0.4


def session_options(hardness: float = DEFAULT_HARDNESS) -> SolverOptions:
    """Solver options of a served session: ground with friction 0.3, 5 mm
    margin and thickness, and the PCG stretch solve at the given hardness."""
    return SolverOptions(use_ground=True, contact_friction=0.3, collision_margin=0.005, cloth_thickness=0.005,
                         stretch_optimization_enabled=hardness > 0.0, stretch_optimization_strength=float(hardness))


@lru_cache(maxsize=64)
def _grid_template(side: int, spacing: float) -> ClothBuildData:
    pos, tris = grid_mesh(side, (side - 1) * spacing)
    pos[:, 1] += 1.5
    pin = np.zeros(side * side, bool)
    pin[side * (side - 1)] = True
    pin[side * side - 1] = True
    return build_cloth(y_up_to_z_up(pos), tris, pin)


def session_cloth(side: int, seed: int, spacing: float = 0.01, jitter: float = 1.0e-4) -> ClothBuildData:
    """A side x side grid (spacing 1 cm, 1.5 m up, two far corners pinned), seed-jittered start."""
    tpl = _grid_template(int(side), float(spacing))
    rng = np.random.Generator(np.random.PCG64(seed))
    start = (tpl.positions_world.astype(np.float64) + rng.normal(0.0, jitter, size=tpl.positions_world.shape)).astype(np.float32)
    return replace(tpl, start_positions=start)


def served_requests(*, seed: int = 0, rate: float = 200.0, window_s: float = 1.0, vmin: int = 1000, vmax: int = 250000,
                    fmin: int = 60, fmax: int = 600, size_classes: int = 16, max_sessions: int | None = None,
                    arrival_unit: str = "s") -> list[SessionRequest]:
    """Poisson arrivals of grid-cloth sessions: sizes log-uniform in [vmin, vmax]
    vertices (snapped to `size_classes` geometric grid sides so meshes can be
    shared), frames uniform in [fmin, fmax], a wind field with a seeded
    direction and strength U[0, 20] N, and a sphere on a seeded Lissajous path."""
    rng = np.random.Generator(np.random.PCG64(seed))
    smin, smax = math.sqrt(vmin), math.sqrt(vmax)
    classes = np.unique(np.rint(np.geomspace(smin, smax, size_classes)).astype(int))
    reqs, t, sid = [], 0.0, 0
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t > window_s or (max_sessions is not None and sid >= max_sessions):
            break
        v = math.exp(rng.uniform(math.log(vmin), math.log(vmax)))
        side = int(classes[np.argmin(np.abs(classes - math.sqrt(v)))])
        frames = int(rng.integers(fmin, fmax + 1))
        d = rng.normal(size=3)
        d[2] = abs(d[2]) * 0.25
        wind = wind_field(d, float(rng.uniform(0.0, 20.0)))
        half = 0.5 * (side - 1) * 0.01
        sphere = SpherePath(center=(0.0, 0.0, 0.9), amplitude=(0.6 * half, 0.6 * half, 0.2),
                            frequency=tuple(float(x) for x in rng.uniform(0.1, 0.5, size=3)),
                            phase=tuple(float(x) for x in rng.uniform(0.0, 2.0 * math.pi, size=3)), radius=0.3)
        opts = session_options()
        reqs.append(SessionRequest(sid, t if arrival_unit == "s" else float(int(t)), session_cloth(side, seed * 100003 + sid),
                                   opts, frames, wind, sphere))
        sid += 1
    return reqs
