/*
 * ssbl_xpbd_hip_ext.h -- entry points this library adds beside the ABI41
 * contract.  None of them has a counterpart in the reference; they serve
 * callers the reference has no interface for:
 *   - parity tests, which need more state than ssbl_download_positions gives
 *     (prev/vel, self-collision hash cells and candidate pairs, the static SDF);
 *   - batched and served stepping, which advances many handles with one
 *     launch per solver phase instead of one ssbl_step_solver_ex per handle;
 *   - measurement, which needs device-side kernel durations.
 * The same symbols are exported by the CPU oracle library so one host harness
 * drives both.  Conventions follow the ABI41 functions: 1 = ok, 0 = error with
 * ssbl_last_error() set.
 */
#ifndef SSBL_XPBD_HIP_EXT_H
#define SSBL_XPBD_HIP_EXT_H

#include "ssbl_xpbd_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Direct-stretch semantics.  0 (default): the reference's single uncoloured
 * pass (ssbl_xpbd_cuda_abi41.cu:1927-1980) made deterministic as a Jacobi
 * sweep -- every edge reads the positions from before the pass and each vertex
 * sums its corrections in ascending edge order.  1: graph-coloured
 * Gauss-Seidel, one pass per edge colour in colour order (the colouring of
 * the mesh's edge_color_offsets, as the reference's per-colour bend and cap
 * launches use theirs, ssbl_xpbd_cuda_abi41.cu:9207-9214), each edge reading
 * the positions its colour's predecessors left.  Mode 1 fails unless the
 * handle has edge colour offsets with no vertex twice in one colour.  The
 * members of a batch must share the mode when the batch is created. */
enum { SSBL_EXT_STRETCH_JACOBI = 0, SSBL_EXT_STRETCH_COLORED = 1 };
SSBL_API int ssbl_ext_set_stretch_mode(void* handle, int mode);

/* Copies pos, prev and vel (each float[3V], any pointer may be NULL). */
SSBL_API int ssbl_ext_download_state(void* handle, float* pos, float* prev, float* vel, int max_floats);

/* Self-collision introspection, valid after a step that ran self collision.
 * cells: int[3V] vertex cell coordinates of the last vertex-hash build.
 * counts: uint[V], indices: uint[V*32] -- the last vertex-vertex candidate table. */
SSBL_API int ssbl_ext_self_hash_cells(void* handle, int* cells, int max_ints);
SSBL_API int ssbl_ext_self_neighbors(void* handle, unsigned int* counts, unsigned int* indices, int max_vertices);

/* Static SDF grid: dims int[3], origin_and_voxel float[4]; values float[dx*dy*dz]
 * (values may be NULL to query the shape).  Returns 0 when no SDF is built. */
SSBL_API int ssbl_ext_static_sdf(void* handle, int* dims, float* origin_and_voxel, float* values, int max_values);

/* Static SDF given as a grid instead of built from static triangles (the
 * reference builds it from triangles only, ssbl_xpbd_cuda_abi41.cu:6123-6218):
 * dims int[3] (each >= 2), origin_and_voxel float[4], values float[dx*dy*dz]
 * with x fastest, negative inside.  Sampling and contacts are those of a built
 * grid; static triangles uploaded before are kept for the normal fallback. */
SSBL_API int ssbl_ext_set_static_sdf(void* handle, const int* dims, const float* origin_and_voxel, const float* values, int count);

/* Batched stepping: every handle advances by the same substeps/iterations; the
 * result of each handle equals what ssbl_step_solver_ex on that handle alone
 * produces.  The batch refers to the handles; it does not own them. */
SSBL_API void* ssbl_ext_create_batch(void* const* handles, int count);
SSBL_API int ssbl_ext_destroy_batch(void* batch);
/* Points an existing batch at new members (same rules as create): its device
 * tables and pinned staging are reused, grown only when needed, and its
 * captured graphs dropped.  Waits for the library's queued work.  For loops
 * whose set of stepped handles changes from step to step (serving). */
SSBL_API int ssbl_ext_batch_set_members(void* batch, void* const* handles, int count);
SSBL_API int ssbl_ext_step_batch(void* batch, int substeps, int iterations, int force_sync);
/* The same step with every member's diagnostics counted and read back when
 * fetch_diagnostics is nonzero (ssbl_get_diagnostics on each handle
 * afterwards), as ssbl_step_solver_ex does for one handle; waits for the
 * step then. */
SSBL_API int ssbl_ext_step_batch_ex(void* batch, int substeps, int iterations, int fetch_diagnostics, int force_sync);

/* Runs `launches` (rounded up to even) back-to-back fused constraint passes of
 * the batch (stretch gather, pins, analytic and SDF contacts) between two
 * events on the library stream and writes their mean duration in ms.  The
 * members' state is copied aside first and restored afterwards, so the call
 * leaves it unchanged: a measurement call (it allocates the copy).  Fails
 * for a batch not stepped since it was created or given new members (the
 * pass needs a prepared step's member views). */
SSBL_API int ssbl_ext_time_batch_pass(void* batch, int launches, float* avg_ms);
/* The same for `iterations` iterations of a batch stepped as two parts
 * (ssbl_ext_set_batch_parts): each iteration is its two parts' passes on
 * their two streams, as in its steps, and the mean is the time of one
 * iteration over all members (one table's pass otherwise). */
SSBL_API int ssbl_ext_time_batch_iteration(void* batch, int iterations, float* avg_ms);

/* Global dynamic scene from cloth handles, gathered on the device: particle k
 * of handle h is that handle's vertex k at its current position with radius
 * particle_radii[h]; triangles are each handle's mesh triangles at current
 * positions; source ids and layers per handle.  The same scene as
 * ssbl_update_global_dynamic_scene built from host copies of those arrays
 * (session_manager.py:3242-3300 packs them on the host), with no host round
 * trip: the gather and the hashes run on the library stream, ordered after
 * the handles' queued steps.  Returns without waiting. */
SSBL_API int ssbl_ext_update_global_scene_from_handles(void* scene, void* const* handles, const int* source_ids,
                                                      const int* collision_layers, const float* particle_radii, int count,
                                                      float collision_margin, float cloth_thickness);

/* (triangle, cell) entries of a global scene's triangle hash after its last
 * build, for measurement (the algorithmic bytes of the hash build); waits for
 * the library's queued work.  -1 for an invalid scene. */
SSBL_API long long ssbl_ext_global_scene_entries(void* scene);

/* Device memory (bytes) the handle holds: state, constraint tables, SDF,
 * self-collision hashes and PCG arrays.  -1 for an invalid handle.  A
 * serving loop sums these against its HBM budget. */
SSBL_API long long ssbl_ext_device_bytes(void* handle);

/* Waits for all work the library queued. */
SSBL_API int ssbl_ext_synchronize(void);

/* Graph replay of the substep loop (on by default).  ssbl_step_solver_ex and
 * ssbl_ext_step_batch capture the launches of a step into a HIP graph the
 * first time and replay it on later steps with the same substeps,
 * iterations and set of active passes; steps with self collision or the
 * hard-stretch solve, and steps while kernel timing is on, launch directly.
 * Results are the same either way.  mode 0 turns replay off, nonzero on;
 * returns the previous mode.  No counterpart in the reference, which launches
 * every kernel of a substep from the host (ssbl_xpbd_cuda_abi41.cu:9087-9398). */
SSBL_API int ssbl_ext_set_graph_mode(int mode);
/* Captures and replays so far of one handle's graph (batch NULL) or one
 * batch's (handle NULL). */
SSBL_API int ssbl_ext_graph_stats(void* handle, void* batch, long long* captures, long long* replays);

/* Concurrent batch parts (on by default: parts 2).  A batch of two or more
 * cloths and at least 64 tiles of 256 vertices is split, when created or
 * given new members, into two parts of about equal vertex counts (largest
 * cloths dealt first to the part with fewer vertices), and each step runs the
 * two parts' launch sequences on two streams, forked from and joined back to
 * the library stream: the parts' cloths share nothing within a step, so one
 * part's launches use the GPU while the other's wait for memory or finish on
 * a few slow waves.  Results are the same as one table's.  parts 1 keeps
 * batches created afterwards as one table; returns the previous setting.
 * Steps with launch tuning on run as one table.  No counterpart in the
 * reference, which steps its cloths one by one. */
SSBL_API int ssbl_ext_set_batch_parts(int parts);

/* Launch tuning (off by default).  A handle or batch whose steps run the
 * dynamic collider contacts times the first of those steps and every 16th
 * after it, with the particle pass's pooling limit (the most probing
 * vertices of a wave whose searches are pooled into one list) varied over
 * {0, 8, 16, 32, 64} from substep to substep; a later step reads the timings
 * once they are complete (no wait) and keeps the fastest limit until the
 * next timed step.  Every limit gives the same contacts.  mode 0 keeps the
 * default limit 16 (or the last choice); returns the previous mode.  No
 * counterpart in the reference, whose launch shapes are fixed. */
SSBL_API int ssbl_ext_set_autotune(int mode);
/* The limit chosen for one handle (batch NULL) or one batch (handle NULL):
 * -1 while not yet tuned, -2 on invalid arguments. */
SSBL_API int ssbl_ext_tuned_pool_limit(void* handle, void* batch);

/* Device-side kernel timing.  While enabled, every launch carries a start and
 * a stop event stamped by its own dispatch, on the stream it runs on.  ssbl_ext_kernel_stats
 * writes a JSON object {"kernel": {"launches": n, "total_ms": t}, ...} and
 * resets the counters. */
SSBL_API int ssbl_ext_kernel_timing(int enable);
SSBL_API int ssbl_ext_kernel_stats(char* json_out, int capacity);

/* Allocation-free variant of ssbl_step_solver_ex (same arguments, same
 * results).  No counterpart in the reference: it is for callers that cannot
 * allocate device memory mid-run (serving loops, harnesses that forbid it,
 * code around graph capture).  It calls no hipMalloc, hipFree or other
 * allocating routine: the hard-stretch PCG solve takes its scratch (iteration
 * vectors, 3x3 blocks, reduction partials, scalars) from a static device array
 * of 96 MiB per GPU, held until the process ends, instead of the handle's own
 * arrays, and the step launches directly (no graph capture or instantiation,
 * no launch tuning).  A handle whose solve needs more scratch than that
 * (grid cloths above about 330k vertices) is rejected with an error.  The
 * array is shared by every call on one GPU and keeps the last call's
 * contents; calls are ordered on the library's stream, and calls that
 * overlap on one GPU from several host threads race.  Everything a step may
 * allocate besides is sized when the handle and its pins are uploaded. */
SSBL_API int ssbl_ext_step_solver_ex_noalloc(void* handle, int substeps, int iterations, int fetch_diagnostics, int force_sync);

/* Implementation name and device description, e.g. "hip gfx950 ...". */
SSBL_API const char* ssbl_ext_backend_name(void);

#ifdef __cplusplus
}
#endif

#endif  /* SSBL_XPBD_HIP_EXT_H */
