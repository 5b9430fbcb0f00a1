"""Cross-cloth collision for the cloths of a session: one global dynamic
scene holding every cloth's vertices (as particles of radius
`contact_distance`) and triangles, refreshed once per frame before the steps,
each cloth attached with its own source id and collision layer
(session_manager.py:3242-3325).

refresh(device=False) packs the scene on the host from downloaded positions,
as the reference's session manager does; refresh(device=True) has the
library gather the same arrays on the device from the handles' current
positions (ssbl_ext_update_global_scene_from_handles), with no host round
trip.  Both produce the same scene.
"""
from __future__ import annotations

import numpy as np

from . import native_backend as nb


class CrossClothScene:
    def __init__(self, solvers, layers, *, contact_distance: float, mode: int = nb.GlobalDynamicScene.CROSS_LOWER_LAYERS,
                 lib=None):
        self.solvers = list(solvers)
        self.lib = lib if lib is not None else (self.solvers[0].lib if self.solvers else nb.load_library())
        self.scene = nb.GlobalDynamicScene(lib=self.lib)
        self.ids = list(range(1, len(self.solvers) + 1))
        self.layers = [int(x) for x in layers]
        self.radii = [float(contact_distance)] * len(self.solvers)
        self.margin = float(contact_distance)
        for s, i, layer in zip(self.solvers, self.ids, self.layers):
            s.attach_global_dynamic_scene(self.scene, i, layer, mode)

    def refresh(self, device: bool = True) -> None:
        if device:
            self.scene.update_from_solvers(self.solvers, self.ids, self.layers, self.radii, self.margin, self.margin)
            return
        pos = [s.download_positions() for s in self.solvers]
        tris = [p[s.triangles].reshape(-1, 9) for s, p in zip(self.solvers, pos)]
        self.scene.update(
            particle_positions=np.concatenate(pos),
            particle_radii=np.concatenate([np.full(len(p), r, np.float32) for p, r in zip(pos, self.radii)]),
            particle_source_ids=np.concatenate([np.full(len(p), i, np.int32) for p, i in zip(pos, self.ids)]),
            particle_collision_layers=np.concatenate([np.full(len(p), g, np.int32) for p, g in zip(pos, self.layers)]),
            triangle_vertices=np.concatenate(tris),
            triangle_source_ids=np.concatenate([np.full(len(t), i, np.int32) for t, i in zip(tris, self.ids)]),
            triangle_collision_layers=np.concatenate([np.full(len(t), g, np.int32) for t, g in zip(tris, self.layers)]),
            collision_margin=self.margin, cloth_thickness=self.margin)

    def close(self) -> None:
        for s in self.solvers:
            if s.handle:
                s.clear_global_dynamic_scene()
        self.scene.close()
