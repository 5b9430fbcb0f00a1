"""Benchmark: cloth vertex-substeps/s of the XPBD substep path on one MI355X.

A step is one frame: every cloth of the workload advances by one
ssbl_step_solver_ex(substeps, iterations) worth of work.  The headline
workload is the 64-cloth uneven batch (492,741 vertices, 10 substeps x 8
iterations, ground + 128^3 box SDF, cross-cloth collision over four layers).
Two forms run on the same GPU and the same inputs:
  plain   -- one ABI41 handle per cloth, stepped one after another (the
             reference's per-slot loop, session_manager.py:2528-2591), with
             the cross-cloth scene packed on the host every frame;
  batched -- the same handles stepped together through ssbl_ext_step_batch
             (one launch per solver phase covers every cloth), the scene
             gathered on the device.
`value` is the batched form's whole-job vertex-substeps/s over all `--steps`
frames (default: the workload's 120) from the workload's start state, with all
state resident in HBM: the warm-up frames run first, then every handle is put
back to its start state (ssbl_reset_solver + ssbl_update_positions) before the
timed frames.  Per-frame latency percentiles, the per-kernel roofline entries
and the CPU baseline come from separate passes over the same frames.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
import __graft_entry__  # noqa: E402

__graft_entry__.load_package()
from ssbl_xpbd_amd import cross_cloth, native_backend as nb, scenes  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"# {msg}", file=sys.stderr, flush=True)


# ---------------------------------------------------------------- bytes
def pass_bytes_per_vertex(scene) -> float:
    """Algorithmic bytes of one fused constraint/contact pass per vertex.

    The pass timed for the roofline is an iteration other than a substep's
    last (stretch gather, pins, contacts; no velocity update).  Its unique
    compulsory traffic: read position (12) + inverse mass (4) + previous
    position (12) + the incident edges' records (12 bytes per edge: two
    indices and a rest length, E/V edges per vertex); write position and
    previous position (24).  Velocity is read and written only by pinned or
    contacting vertices and is not counted.
    """
    v = scene.vertex_count
    e = sum(len(c.edges) for c in scene.cloths)
    return 12 + 4 + 12 + 12.0 * e / v + 24


def scene_hash_sizes(scene):
    """Triangle and particle counts of the cross-cloth scene and the bucket
    counts of its two hashes (next power of two of 4n, at least 1024)."""
    nt = sum(len(c.triangles) for c in scene.cloths)
    npart = scene.vertex_count
    buckets = lambda n: 1 << max(10, math.ceil(math.log2(max(4 * n, 1))))  # noqa: E731
    return nt, npart, buckets(nt), buckets(npart)


def kernel_bytes(name: str, scene, frames: int, entries: float, counters: dict) -> float | None:
    """Algorithmic bytes of all launches of one timed scope over `frames`
    frames of the cross-cloth workload (scene rebuilt once per frame; the
    dynamic contact pass runs once per substep):
      tri_hash_count   per triangle its 64-byte span record, 4-byte large
                       flag and 4-byte layer (its bucket range); per bucket a
                       4-byte count written and read back;
      tri_hash_scatter per triangle the record, flag and 8-byte source/layer;
                       per entry (triangle, cell) a 16-byte record written;
                       per bucket its offset and count (12 bytes);
      tri_bins         per bucket the offsets (4) and a 16-byte bin written;
                       every entry's 16-byte record read;
      part_hash        per particle its 16-byte position, 4-byte source, layer,
                       cell (16) and 48-byte packed entry; per bucket 12 bytes;
      dynamic_triangles per vertex and launch its position (16) and
                       16-byte triangle bin; per candidate triangle entry its
                       16-byte entry, 64-byte bounds record and 48-byte corners
      dynamic_contacts per vertex and launch position, previous position and
                       velocity read and written (72), its triangle contact
                       key (8) and coarse summary entry (8); per triangle contact
                       the winner's record and corners again (112), per probed
                       particle entry 48 bytes (counts from the diagnostics of
                       the same frames)."""
    nt, npart, tb, pb = scene_hash_sizes(scene)
    if name == "tri_hash_count":
        return frames * (nt * 72.0 + tb * 8.0)
    if name == "tri_hash_scatter":
        return frames * (nt * 76.0 + entries * 16.0 + tb * 12.0)
    if name == "tri_bins":
        return frames * (tb * 20.0 + entries * 16.0)
    if name == "part_hash":
        return frames * (npart * 88.0 + pb * 12.0)
    if name == "dynamic_triangles":
        launches = frames * scene.substeps
        return launches * scene.vertex_count * 32.0 + counters["dynamic_triangle_candidate_count"] * 128.0
    if name == "dynamic_contacts":
        launches = frames * scene.substeps
        return (launches * scene.vertex_count * 88.0 + counters["fast_triangle_pair_contacts"] * 112.0
                + counters["dynamic_particle_candidate_count"] * 48.0)
    return None


def measured_traffic(workload: str):
    """HBM bytes per launch of the timed kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE passes over this workload (scaled by the copy
    calibration), or None when none were taken."""
    path = ROOT / "ssbl-cuda-xpbd_amd" / "measured_traffic.json"
    try:
        rec = json.loads(path.read_text()).get(workload)
    except (OSError, ValueError):
        return None
    if not rec:
        return None
    return 1024.0 * (rec["fetch_kib_per_dispatch"] * rec["fetch_scale"] + rec["write_kib_per_dispatch"] * rec["write_scale"])


# ---------------------------------------------------------------- drivers
def create(scene, lib=None):
    return [nb.NativeXpbdSolver(c, o, scene.static_triangles, lib=lib) for c, o in zip(scene.cloths, scene.options)]


def restart(solvers, scene):
    """Every handle back to the workload's start state: rest positions, zero
    velocity, then the seed-jittered start (what creation left)."""
    for s, c in zip(solvers, scene.cloths):
        s.reset()
        if c.start_positions is not None:
            s.update_positions(c.start_positions)


def cross_scene(solvers, scene, lib=None):
    if not scene.layers or len(solvers) < 2:
        return None
    return cross_cloth.CrossClothScene(solvers, scene.layers[:len(solvers)], contact_distance=scene.contact_distance, lib=lib)


def run_plain(solvers, scene, frames, cross=None):
    """The reference's per-slot loop: scene packed on the host, then every slot stepped."""
    for _ in range(frames):
        if cross is not None:
            cross.refresh(device=False)
        for s in solvers:
            s.step(scene.substeps, scene.iterations, diagnostics=False, synchronize=False)


def run_batched(batch, scene, frames, cross=None, diagnostics=False, counters=None):
    """Batched frames; with `counters`, every member's diagnostics are summed
    into it after each frame, and "scene_entries" gets the scene's triangle
    hash entries of each frame."""
    for _ in range(frames):
        if cross is not None:
            cross.refresh(device=True)
            if counters is not None:
                counters["scene_entries"] += int(cross.lib.ssbl_ext_global_scene_entries(cross.scene.handle))
        batch.step(scene.substeps, scene.iterations, diagnostics=diagnostics)
        if counters is not None:
            for s in batch.solvers:
                d = s.diagnostics()
                for k in counters:
                    if k != "scene_entries":
                        counters[k] += int(getattr(d, k))


def timed(fn, lib):
    nb.synchronize(lib)
    t0 = time.perf_counter()
    fn()
    nb.synchronize(lib)
    return time.perf_counter() - t0


def latencies(step_frame, lib, frames):
    out = []
    for _ in range(frames):
        nb.synchronize(lib)
        t0 = time.perf_counter()
        step_frame()
        nb.synchronize(lib)
        out.append((time.perf_counter() - t0) * 1e3)
    return np.asarray(out)


def kernel_stats(lib, fn):
    lib.ssbl_ext_kernel_timing(1)
    fn()
    nb.synchronize(lib)
    lib.ssbl_ext_kernel_timing(0)
    buf = ctypes.create_string_buffer(1 << 16)
    if not lib.ssbl_ext_kernel_stats(buf, len(buf)):
        return {}
    return json.loads(buf.value.decode())


def pct(x, q):
    return float(np.percentile(x, q)) if len(x) else 0.0


# ------------------------------------------------------------ CPU baseline
def cpu_baseline(scene, budget_s: float):
    """The CPU oracle (FP32) on the first frames of the same workload: every
    cloth, the cross-cloth scene packed every frame, stepped once on one thread
    and once with the cloths spread over the host's cores (each cloth's step
    is independent of the others', so the threads change nothing in the
    results).  Frames: as many as the single-thread pass fits in `budget_s`
    (at least one); the all-core pass runs the same frames."""
    lib = nb.load_library(str(ROOT / "oracle" / "liboracle_abi41.so"))
    host_cores = len(os.sched_getaffinity(0))
    threads = max(1, min(host_cores, 16))
    solvers = create(scene, lib=lib)
    cross = cross_scene(solvers, scene, lib=lib)
    work_frame = scene.vertex_count * scene.substeps

    def frame(pool=None):
        if cross is not None:
            cross.refresh(device=True)
        if pool is None:
            for s in solvers:
                s.step(scene.substeps, scene.iterations, diagnostics=False, synchronize=False)
        else:
            list(pool.map(lambda s: s.step(scene.substeps, scene.iterations, diagnostics=False, synchronize=False),
                          sorted(solvers, key=lambda s: -s.vertex_count)))

    frames = 0
    t0 = time.perf_counter()
    while frames < 1 or (time.perf_counter() - t0 < budget_s and frames < scene.frames):
        frame()
        frames += 1
    t_single = time.perf_counter() - t0
    single_pos = [s.download_positions() for s in solvers]
    restart(solvers, scene)
    with ThreadPoolExecutor(max_workers=threads) as pool:
        t0 = time.perf_counter()
        for _ in range(frames):
            frame(pool)
        t_multi = time.perf_counter() - t0
    same = all(np.array_equal(a, s.download_positions()) for a, s in zip(single_pos, solvers))
    if cross is not None:
        cross.close()
    for s in solvers:
        s.close()
    return {"value": frames * work_frame / t_multi, "unit": "vertex-substeps/s", "cores": threads, "kind": "port",
            "sample": f"oracle FP32, all {len(scene.cloths)} cloths ({scene.vertex_count} vertices), frames 1-{frames} of the "
                      f"workload, cross-cloth scene packed every frame; cloths spread over {threads} threads",
            "single_thread_value": frames * work_frame / t_single, "host_cores": host_cores,
            "threads_equal_single_thread": bool(same)}


# ---------------------------------------------------------- secondary lines
def single_scene_line(scene, frames: int, warmup: int, lib, *, pins=None, batched=True) -> dict:
    """Throughput and per-frame latency of a small scene stepped from its start
    state: batched when it has several cloths, else the handle alone.  `pins`
    updates pin targets before frame f (a moving boundary)."""
    solvers = create(scene)
    batch = nb.SolverBatch(solvers) if batched and len(solvers) > 1 else None

    def frame(f):
        if pins is not None:
            pins(solvers, f)
        if batch is not None:
            batch.step(scene.substeps, scene.iterations)
        else:
            for s in solvers:
                s.step(scene.substeps, scene.iterations, diagnostics=False, synchronize=False)

    for f in range(warmup):
        frame(f)
    restart(solvers, scene)
    t = timed(lambda: [frame(f) for f in range(frames)], lib)
    restart(solvers, scene)
    lat = []
    for f in range(frames):
        nb.synchronize(lib)
        t0 = time.perf_counter()
        frame(f)
        nb.synchronize(lib)
        lat.append((time.perf_counter() - t0) * 1e3)
    finite = all(np.isfinite(s.download_positions()).all() for s in solvers)
    if batch is not None:
        batch.close()
    for s in solvers:
        s.close()
    return {"workload": scene.name, "cloths": len(scene.cloths), "vertices": scene.vertex_count, "substeps": scene.substeps,
            "iterations": scene.iterations, "frames": frames, "of_frames": scene.frames,
            "vertex_substeps_per_s": frames * scene.vertex_count * scene.substeps / t, "ms_per_frame": t / frames * 1e3,
            "frame_ms_p50": pct(lat, 50), "frame_ms_p99": pct(lat, 99), "form": "batched" if batch is not None else "alone",
            "finite": bool(finite)}


def stress_line(frames: int, warmup: int, lib) -> dict:
    """The closed-mesh stress case (655,362-vertex balloon with pressure and a
    1,000,000-vertex towel twisted by its turning end rows; strict self
    collision and PCG on both; 20 x 10): the first `frames` frames."""
    sc = scenes.config5(frames=max(frames, 1))
    towel = sc.cloths[1]
    dt = sc.options[1].dt

    def pins(solvers, f):
        solvers[1].update_pin_targets(towel.pin_indices, scenes.towel_pin_targets(towel, (f + 1) * dt), towel.pin_weights)

    return single_scene_line(sc, frames, warmup, lib, pins=pins)


def serving_line(reqs, window_s: float, max_s: float, budget: int, capacity: int, builders: int, lib) -> dict:
    """Served cloth sessions: Poisson(200/s) arrivals over `window_s` seconds,
    sizes log-uniform in [1k, 250k] vertices, 60-600 frames each, wind and a
    Lissajous sphere per session, PCG on; admitted against `budget` bytes of
    HBM by `builders` threads beside the tick loop, stepped by continuous
    batching (10 substeps x 8 iterations) with at most `capacity` vertices per
    tick; the run stops at `max_s` seconds.  Percentiles are over every frame
    served; "10k" is the 7k-14k-vertex sessions."""
    from ssbl_xpbd_amd import serving
    srv = serving.ClothServer(hbm_budget_bytes=budget, substeps=10, iterations=8, max_batch_vertices=capacity, lib=lib,
                              builders=builders)
    rep = srv.run(reqs, clock="wall", max_seconds=max_s)
    out = rep.summary()
    out.update({"workload": "served_sessions", "arrival_rate_per_s": 200.0, "arrival_window_s": window_s, "run_cap_s": max_s,
                "sessions_requested": len(reqs), "substeps": 10, "iterations": 8, "hbm_budget_bytes": budget,
                "tick_vertex_capacity": capacity, "builder_threads": builders, "frames_served": len(srv.frame_log),
                "pcg": "on (strength 0.4, the reference's default hardness)"})
    log(f"serving (budget {budget / 1e9:.0f} GB): {len(rep.results)}/{len(reqs)} sessions finished in {rep.wall_s:.1f} s, "
        f"{len(srv.frame_log)} frames, step p50 {rep.frame_ms_p50:.2f} ms p99 {rep.frame_ms_p99:.2f} ms, "
        f"10k p99 {rep.step_ms_10k[1]:.2f} ms, peak {rep.peak_device_bytes / 1e9:.1f} GB, budget waits {rep.budget_waits}")
    return out


def step_only(which: str, scene, args, lib) -> None:
    """Steps one workload without timing it, for a profiler or counter pass."""
    if which == "serving":
        from ssbl_xpbd_amd import serving
        reqs = serving.served_requests(seed=args.seed, rate=200.0, window_s=args.serving_window)
        out = serving_line(reqs, args.serving_window, args.serving_max_s, 200 * 10**9, args.serving_capacity,
                           args.serving_builders, lib)
        print(json.dumps({"only": which, "frames_served": out["frames_served"]}), flush=True)
        return
    if which == "batched":
        sc = scene
    elif which == "selfcol":
        sc = scenes.config2(seed=args.seed)
    else:
        sc = scenes.config5(frames=max(args.warmup + args.steps, 1))
    frames = args.warmup + (args.steps if args.steps > 0 else sc.frames)
    solvers = create(sc)
    batch = nb.SolverBatch(solvers) if len(solvers) > 1 else None
    cross = cross_scene(solvers, sc) if which == "batched" else None
    towel = sc.cloths[1] if which == "stress" else None
    for f in range(frames):
        if cross is not None:
            cross.refresh(device=True)
        if towel is not None:
            solvers[1].update_pin_targets(towel.pin_indices, scenes.towel_pin_targets(towel, (f + 1) * sc.options[1].dt),
                                          towel.pin_weights)
        if batch is not None:
            batch.step(sc.substeps, sc.iterations)
        else:
            for s in solvers:
                s.step(sc.substeps, sc.iterations, diagnostics=False, synchronize=False)
        if f % 10 == 9:
            nb.synchronize(lib)
            log(f"{which}: {f + 1} of {frames} frames")
    nb.synchronize(lib)
    print(json.dumps({"only": which, "workload": sc.name, "frames": frames}), flush=True)


# ------------------------------------------------------------------- main
def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0, help="timed frames (0: the workload's 120)")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--plain-latency-frames", type=int, default=20)
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-secondary", action="store_true", help="skip the other workloads' lines")
    ap.add_argument("--stress-frames", type=int, default=6, help="frames of the 1.66M-vertex stress case (0 = skip)")
    ap.add_argument("--serving-window", type=float, default=6.0,
                    help="seconds of Poisson(200/s) session arrivals in the served-sessions runs (0 = skip them)")
    ap.add_argument("--serving-max-s", type=float, default=15.0, help="wall-clock cap of each served-sessions run")
    ap.add_argument("--serving-capacity", type=int, default=1_000_000, help="most vertices stepped in one served tick")
    ap.add_argument("--serving-builders", type=int, default=4, help="threads creating served sessions' handles")
    ap.add_argument("--serving-small-budget-gb", type=float, default=16.0,
                    help="HBM budget of the second served run, small enough to bind")
    ap.add_argument("--graph", type=int, default=1, help="graph replay of the substep loop (0 = direct launches)")
    ap.add_argument("--batch-parts", type=int, default=2, help="concurrent parts of a large batch (1 = one launch table)")
    ap.add_argument("--only", choices=("batched", "selfcol", "stress", "serving"), default=None,
                    help="only step one workload, untimed, and exit (for profiler and counter passes): the batched "
                         "64-cloth form, the 10k-vertex self-collision cloth or the stress case for warmup + steps "
                         "frames, or one served-sessions run")
    ap.add_argument("--batched-only", action="store_true", help="the same as --only batched")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device copy of this many MiB (counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, device="cuda", dtype=torch.float32)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": n * 4, "copy_calibration_bytes_written": n * 4}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported"}))
        raise SystemExit(2)

    scene = scenes.config3(seed=args.seed)
    steps = args.steps if args.steps > 0 else scene.frames
    lib = nb.load_library()
    nb.set_graph_mode(lib, bool(args.graph))
    nb.set_batch_parts(lib, args.batch_parts)
    log(f"workload {scene.name}: {len(scene.cloths)} cloths, {scene.vertex_count} vertices, "
        f"{scene.substeps}x{scene.iterations}, {steps} frames; device {lib.ssbl_ext_backend_name().decode()}")
    if args.batched_only:
        args.only = "batched"
    if args.only:
        step_only(args.only, scene, args, lib)
        return
    plain = create(scene)
    members = create(scene)
    cross_p, cross_b = cross_scene(plain, scene), cross_scene(members, scene)
    batch = nb.SolverBatch(members)
    work_per_frame = scene.vertex_count * scene.substeps

    # plain form: warm up, back to the start, all timed frames
    run_plain(plain, scene, args.warmup, cross_p)
    restart(plain, scene)
    t_plain = timed(lambda: run_plain(plain, scene, steps, cross_p), lib)
    log(f"plain: {steps} frames in {t_plain:.3f} s")
    lat_p = latencies(lambda: run_plain(plain, scene, 1, cross_p), lib, args.plain_latency_frames)

    # batched form: the value
    run_batched(batch, scene, args.warmup, cross_b)
    restart(members, scene)
    t_batch = timed(lambda: run_batched(batch, scene, steps, cross_b), lib)
    log(f"batched: {steps} frames in {t_batch:.3f} s")
    # per-frame latency over the same frames (a synchronisation after each)
    restart(members, scene)
    lat_b = latencies(lambda: run_batched(batch, scene, 1, cross_b), lib, steps)
    # the same frames once more with every launch timed by its own dispatch,
    # stepped as one launch table (one launch per phase over all cloths, the
    # launch the kernel figures below describe; the two-part form's launches
    # overlap, so their durations do not add up to a frame)
    restart(members, scene)
    previous = nb.set_batch_parts(lib, 1)
    one_table = nb.SolverBatch(members)
    nb.set_batch_parts(lib, previous)
    stats = kernel_stats(lib, lambda: run_batched(one_table, scene, steps, cross_b))
    one_table.close()
    # and once more with every member's counters read back after each frame;
    # at five points of the run the fused pass is timed back to back (the
    # measurement restores the state, so the frames are those of the timed run)
    restart(members, scene)
    counters = {"dynamic_triangle_candidate_count": 0, "dynamic_particle_candidate_count": 0, "fast_triangle_pair_contacts": 0,
                "dynamic_particle_contacts": 0, "static_sdf_contact_count": 0, "scene_entries": 0}
    probes, probes_parts, done = [], [], 0
    for at in sorted({0, steps // 4, steps // 2, (3 * steps) // 4, steps}):
        run_batched(batch, scene, at - done, cross_b, diagnostics=True, counters=counters)
        done = at
        probes.append(batch.time_pass(32))
        probes_parts.append(batch.time_iteration(32))
    avg_ms = float(np.mean(probes))
    kernels = {}
    for name, rec in stats.items():
        per = rec["total_ms"] / max(rec["launches"], 1)
        kernels[name] = {"launches": rec["launches"], "total_ms": rec["total_ms"], "avg_us": per * 1e3}
    roof_more = []
    for name in ("tri_hash_count", "tri_hash_scatter", "tri_bins", "part_hash", "dynamic_triangles", "dynamic_contacts"):
        if name not in kernels:
            continue
        b = kernel_bytes(name, scene, steps, counters["scene_entries"] / steps, counters)
        k = kernels[name]
        achieved = b / (k["total_ms"] * 1e-3) / 1e9
        roof_more.append({"kernel": name, "bound": "hbm", "algorithmic_bytes_per_launch": b / k["launches"],
                          "avg_launch_us": k["avg_us"], "achieved": achieved, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS})
    vp = kernels.get("vertex_pass_stretch")
    bytes_per_launch = pass_bytes_per_vertex(scene) * scene.vertex_count
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    contacts = (counters["fast_triangle_pair_contacts"] + counters["dynamic_particle_contacts"]) / steps
    value = steps * work_per_frame / t_batch
    plain_value = steps * work_per_frame / t_plain
    line = {
        "metric": "cloth vertex-substeps/s, batched vs plain on same GPU; p50/p99 step latency",
        "value": value,
        "unit": "vertex-substeps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": t_batch / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded grid cloths)",
        "config": {
            "workload": scene.name,
            "cloths": len(scene.cloths),
            "vertices": scene.vertex_count,
            "substeps": scene.substeps,
            "iterations": scene.iterations,
            "frames": f"1-{steps} of {scene.frames}, from the start state",
            "graph_replay": bool(args.graph),
            "cross_cloth_collision": True,
            "cross_cloth_layers": 4,
            "cross_cloth_contacts_per_frame": contacts,
            "plain_vertex_substeps_per_s": plain_value,
            "batched_over_plain": value / plain_value,
            "frame_ms_p50_batched": pct(lat_b, 50),
            "frame_ms_p99_batched": pct(lat_b, 99),
            "frame_ms_p50_plain": pct(lat_p, 50),
            "frame_ms_p99_plain": pct(lat_p, 99),
            "kernels": kernels,
            "kernels_stepped_as": "one launch table, every launch stamped by its own dispatch (same frames)",
            "batch_parts": args.batch_parts,
            "scene_hash_entries_per_frame": counters["scene_entries"] / steps,
            "step_graph_captures_replays": list(batch.graph_stats()),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_vertex_pass<stretch> (fused constraint/contact pass, all cloths)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(scene.name),
            "algorithmic_bytes_per_launch": bytes_per_launch,
            "avg_launch_ms": avg_ms,
            "avg_launch_ms_at": {f"frame_{a}": p for a, p in zip(sorted({0, steps // 4, steps // 2, (3 * steps) // 4, steps}), probes)},
            "avg_launch_ms_dispatch_stamped": vp["total_ms"] / vp["launches"] if vp else None,
            # the batch steps as two concurrent parts: one iteration over all
            # cloths is two launches on two streams; its time, same frames
            "two_part_iteration": {
                "avg_iteration_ms": float(np.mean(probes_parts)),
                "achieved": bytes_per_launch / (float(np.mean(probes_parts)) * 1e-3) / 1e9,
                "frac": bytes_per_launch / (float(np.mean(probes_parts)) * 1e-3) / 1e9 / HBM_PEAK_GBS,
            },
            "other_kernels": roof_more,
        },
    }
    batch.close()
    for c in (cross_p, cross_b):
        c.close()
    for s in plain + members:
        s.close()
    if not args.no_secondary:
        sec = {}
        sec["plain_32x32"] = single_scene_line(scenes.config1(seed=args.seed), 120, 10, lib)
        log(f"plain_32x32: {sec['plain_32x32']['ms_per_frame']:.2f} ms/frame")
        sec["selfcol_100x100"] = single_scene_line(scenes.config2(seed=args.seed), 240, 5, lib)
        log(f"selfcol_100x100: {sec['selfcol_100x100']['ms_per_frame']:.2f} ms/frame")
        if args.stress_frames > 0:
            sec["stress_balloon_towel"] = stress_line(args.stress_frames, 1, lib)
            log(f"stress: {sec['stress_balloon_towel']['ms_per_frame']:.1f} ms/frame")
        if args.serving_window > 0:
            from ssbl_xpbd_amd import serving
            t0 = time.perf_counter()
            reqs = serving.served_requests(seed=args.seed, rate=200.0, window_s=args.serving_window)
            log(f"served requests: {len(reqs)} sessions, {sum(r.vertex_count for r in reqs)} vertices, "
                f"built in {time.perf_counter() - t0:.1f} s")
            for key, budget in (("serving", 200 * 10**9), ("serving_budget_bound", args.serving_small_budget_gb * 10**9)):
                sec[key] = serving_line(reqs, args.serving_window, args.serving_max_s, budget, args.serving_capacity,
                                        args.serving_builders, lib)
        line["config"]["other_workloads"] = sec
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(scene, args.cpu_budget)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
