"""Benchmark: cloth vertex-substeps/s of the XPBD substep path on one MI355X.

A step is one frame: every cloth of the workload advances by one
ssbl_step_solver_ex(substeps, iterations) worth of work.  Two forms run on the
same GPU and the same inputs:
  plain   -- one ABI41 handle per cloth, stepped one after another (the
             reference's per-slot loop, session_manager.py:2528-2591);
  batched -- the same handles stepped together through ssbl_ext_step_batch
             (one launch per solver phase covers every cloth).
`value` is the batched form's whole-job vertex-substeps/s with all state
resident in HBM.  Per-frame latency percentiles come from a separate pass that
synchronises after every frame.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
import __graft_entry__  # noqa: E402

__graft_entry__.load_package()
from ssbl_xpbd_amd import cross_cloth, native_backend as nb, scenes  # noqa: E402

HBM_PEAK_GBS = 8000.0


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


def make_workload(name: str, seed: int):
    if name == "plain_32x32":
        return scenes.config1(seed=seed)
    if name == "batched_64_uneven":
        return scenes.config3(seed=seed)
    if name == "selfcol_100x100":
        return scenes.config2(seed=seed)
    raise SystemExit(f"unknown workload {name}")


def single_cloth_line(name: str, seed: int, frames: int, warmup: int, lib) -> dict:
    """Throughput and per-frame latency of one cloth stepped alone."""
    sc = make_workload(name, seed)
    (s,) = create(sc)
    run_plain([s], sc, warmup)
    t = timed(lambda: run_plain([s], sc, frames), lib)
    lat = latencies(lambda: run_plain([s], sc, 1), lib, frames)
    d = s.diagnostics()
    return {"workload": sc.name, "vertices": sc.vertex_count, "substeps": sc.substeps, "iterations": sc.iterations,
            "vertex_substeps_per_s": frames * sc.vertex_count * sc.substeps / t,
            "frame_ms_p50": float(np.percentile(lat, 50)), "frame_ms_p99": float(np.percentile(lat, 99)),
            "self_collision": bool(sc.options[0].self_collision), "finite": int(d.finite_flag)}


def serving_line(seed: int, window_s: float, max_s: float, lib) -> dict:
    """Served cloth sessions: Poisson(200/s) arrivals over `window_s` seconds,
    sizes log-uniform in [1k, 250k] vertices, 60-600 frames each, wind and a
    Lissajous sphere per session, admitted against a 200 GB HBM budget and
    stepped by continuous batching (10 substeps x 8 iterations)."""
    from ssbl_xpbd_amd import serving
    t0 = time.perf_counter()
    reqs = serving.served_requests(seed=seed, rate=200.0, window_s=window_s)
    build_s = time.perf_counter() - t0
    srv = serving.ClothServer(hbm_budget_bytes=200 * 10**9, substeps=10, iterations=8, max_sessions=512,
                              max_batch_vertices=16 * 10**6, lib=lib)
    rep = srv.run(reqs, clock="wall", max_seconds=max_s)
    out = rep.summary()
    out.update({"workload": "served_sessions", "arrival_rate_per_s": 200.0, "arrival_window_s": window_s,
                "sessions_requested": len(reqs), "request_build_s": build_s, "substeps": 10, "iterations": 8,
                "hbm_budget_bytes": 200 * 10**9})
    print(f"# serving: {len(rep.results)}/{len(reqs)} sessions in {rep.wall_s:.1f} s, "
          f"p50 {rep.frame_ms_p50:.2f} ms p99 {rep.frame_ms_p99:.2f} ms", file=sys.stderr, flush=True)
    return out


def create(scene, lib=None):
    return [nb.NativeXpbdSolver(c, o, scene.static_triangles, lib=lib) for c, o in zip(scene.cloths, scene.options)]


def cross_scene(solvers, scene, lib=None):
    if not scene.layers or len(solvers) < 2:
        return None
    return cross_cloth.CrossClothScene(solvers, scene.layers[:len(solvers)], contact_distance=scene.contact_distance, lib=lib)


def cross_contacts(solvers, scene):
    """Dynamic (cross-cloth) contacts of one frame, summed over the cloths."""
    total = 0
    for s in solvers:
        s.step(scene.substeps, scene.iterations, diagnostics=True, synchronize=True)
        d = s.diagnostics()
        total += int(d.fast_triangle_pair_contacts + d.dynamic_particle_contacts)
    return total


def run_plain(solvers, scene, frames, cross=None):
    """The reference's per-slot loop: scene packed on the host, then every slot stepped."""
    for _ in range(frames):
        if cross is not None:
            cross.refresh(device=False)
        for s in solvers:
            s.step(scene.substeps, scene.iterations, diagnostics=False, synchronize=False)


def run_batched(batch, scene, frames, cross=None):
    for _ in range(frames):
        if cross is not None:
            cross.refresh(device=True)
        batch.step(scene.substeps, scene.iterations)


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
    buf = __import__("ctypes").create_string_buffer(1 << 16)
    if not lib.ssbl_ext_kernel_stats(buf, len(buf)):
        return {}
    return json.loads(buf.value.decode())


def cpu_baseline(scene, budget_s: float):
    """The CPU oracle (FP32, single thread) on a bounded sample of the workload."""
    lib = nb.load_library(str(ROOT / "oracle" / "liboracle_abi41.so"))
    if scene.layers and len(scene.cloths) >= 8:
        picked = list(range(8))            # the first two stacks of four (cross-cloth collision within them)
    else:
        order = np.argsort([len(c.positions_world) for c in scene.cloths])
        picked = [int(k) for k in order[::max(1, len(order) // 8)]]
    solvers = [nb.NativeXpbdSolver(scene.cloths[k], scene.options[k], scene.static_triangles, lib=lib) for k in picked]
    cross = None
    if scene.layers and len(solvers) > 1:
        cross = cross_cloth.CrossClothScene(solvers, [scene.layers[k] for k in picked], contact_distance=scene.contact_distance, lib=lib)
    verts = sum(s.vertex_count for s in solvers)
    frames = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s and frames < scene.frames:
        if cross is not None:
            cross.refresh(device=True)
        for s in solvers:
            s.step(scene.substeps, scene.iterations, diagnostics=False, synchronize=False)
        frames += 1
    dt = time.perf_counter() - t0
    work = verts * scene.substeps * frames
    return {"value": work / dt, "unit": "vertex-substeps/s", "cores": 1, "kind": "port",
            "sample": f"oracle FP32 single thread, {len(solvers)} of {len(scene.cloths)} cloths ({verts} vertices) x {frames} frames"
                      + (", cross-cloth scene packed every frame" if cross is not None else "")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30, help="timed frames")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default=os.environ.get("SSBL_BENCH_WORKLOAD", "batched_64_uneven"))
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--latency-frames", type=int, default=20)
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serving-window", type=float, default=0.5,
                    help="seconds of Poisson(200/s) session arrivals in the served-sessions run (0 = skip it)")
    ap.add_argument("--serving-max-s", type=float, default=25.0, help="wall-clock cap of the served-sessions run")
    ap.add_argument("--batched-only", action="store_true",
                    help="only step the batched form (warmup + steps frames) and exit: for counter passes")
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

    scene = make_workload(args.workload, args.seed)
    lib = nb.load_library()
    print(f"# workload {scene.name}: {len(scene.cloths)} cloths, {scene.vertex_count} vertices, "
          f"{scene.substeps}x{scene.iterations}; device {lib.ssbl_ext_backend_name().decode()}", file=sys.stderr, flush=True)
    if args.batched_only:
        members = create(scene)
        batch = nb.SolverBatch(members)
        run_batched(batch, scene, args.warmup + args.steps, cross_scene(members, scene))
        nb.synchronize(lib)
        print(json.dumps({"batched_only_frames": args.warmup + args.steps, "workload": scene.name}), flush=True)
        return
    plain = create(scene)
    members = create(scene)
    cross_p, cross_b = cross_scene(plain, scene), cross_scene(members, scene)
    batch = nb.SolverBatch(members) if len(members) > 1 else None
    work_per_frame = scene.vertex_count * scene.substeps

    run_plain(plain, scene, args.warmup, cross_p)
    t_plain = timed(lambda: run_plain(plain, scene, args.steps, cross_p), lib)
    print(f"# plain: {args.steps} frames in {t_plain:.3f} s", file=sys.stderr, flush=True)
    if batch is not None:
        run_batched(batch, scene, args.warmup, cross_b)
        t_batch = timed(lambda: run_batched(batch, scene, args.steps, cross_b), lib)
        step_frame = lambda: run_batched(batch, scene, 1, cross_b)  # noqa: E731
    else:
        t_batch = t_plain
        step_frame = lambda: run_plain(plain, scene, 1, cross_p)  # noqa: E731
    print(f"# batched: {args.steps} frames in {t_batch:.3f} s", file=sys.stderr, flush=True)
    lat_b = latencies(step_frame, lib, args.latency_frames)
    lat_p = latencies(lambda: run_plain(plain, scene, 1, cross_p), lib, args.latency_frames)
    contacts = cross_contacts(members, scene) if cross_b is not None else None

    secondary = single_cloth_line("selfcol_100x100", args.seed, 10, 2, lib) if args.workload != "selfcol_100x100" else None
    serving = serving_line(args.seed, args.serving_window, args.serving_max_s, lib) if args.serving_window > 0 else None
    stats = kernel_stats(lib, step_frame)
    # the dominant kernel's mean duration: back-to-back launches between two
    # events on the library stream (per-launch instrumentation slows each
    # launch, so the per-kernel stats above are only a breakdown)
    timing_batch = batch if batch is not None else nb.SolverBatch(plain)
    avg_ms = timing_batch.time_pass(64)
    bytes_per_launch = pass_bytes_per_vertex(scene) * scene.vertex_count
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    traffic = measured_traffic(scene.name)
    value = args.steps * work_per_frame / t_batch
    plain_value = args.steps * work_per_frame / t_plain
    line = {
        "metric": "cloth vertex-substeps/s, batched vs plain on same GPU; p50/p99 step latency",
        "value": value,
        "unit": "vertex-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_batch / args.steps * 1e3,
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
            "cross_cloth_collision": cross_b is not None,
            "cross_cloth_layers": 4 if cross_b is not None else 0,
            "cross_cloth_contacts_per_frame": contacts,
            "plain_vertex_substeps_per_s": plain_value,
            "batched_over_plain": value / plain_value,
            "frame_ms_p50_batched": float(np.percentile(lat_b, 50)),
            "frame_ms_p99_batched": float(np.percentile(lat_b, 99)),
            "frame_ms_p50_plain": float(np.percentile(lat_p, 50)),
            "frame_ms_p99_plain": float(np.percentile(lat_p, 99)),
            "kernels": stats,
            "tuned_pool_limit": batch.tuned_pool_limit() if batch is not None else None,
            "step_graph_captures_replays": list(batch.graph_stats()) if batch is not None else None,
            "selfcol_10k": secondary,
            "serving": serving,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_vertex_pass<stretch> (fused constraint/contact pass, all cloths)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "algorithmic_bytes_per_launch": bytes_per_launch,
            "avg_launch_ms": avg_ms,
        },
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(scene, args.cpu_budget)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
