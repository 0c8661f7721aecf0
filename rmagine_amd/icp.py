"""Batched RCC / ICP pose correction on the device.

The reference runs one hypothesis with a host round trip per inner iteration
(tests/optix/correction_rcc.cpp:151-207; CPU twin tests/embree/correction_rcc.cpp:139-180):
    for outer: model = simulate(Tbm_est)
               Tpre = I
               for inner: stats = statistics_p2l(Tpre, dataset, model, params)
                          Tpre = Tpre * umeyama_transform(stats)
               Tbm_est = Tbm_est * Tpre
Here every step runs on the device for all hypotheses at once: one simulate over all
hypothesis poses, segmented FP64 p2l statistics (the observed dataset scan shared by every
hypothesis, the model scan per hypothesis), a per-hypothesis FP64 Umeyama solve and the
pose updates, with no host synchronisation.  After the inner iterations every outer
iteration also evaluates, per hypothesis, the point-to-plane Gauss-Newton system at the
corrected pose (build_linear_system_p2l, src/rmagine_core/include/rmagine/math/
optimization.h:170-200): `grad` holds J^T w r (the gradient of (1/2) sum w r^2 over the pose
update, rotation first), J^T w J and sum w r^2 of the last outer iteration.
`run(..., graph=True)` records the outer iterations into a HIP graph (one per iteration
count) and replays it.
"""
from __future__ import annotations

import torch

from . import math as M
from .simulation import Bundle, HipGraph, HipMap, resize_memory_bundle, simulator_for


class BatchedRCC:
    def __init__(self, map: HipMap, model, dataset_points: torch.Tensor, dataset_mask: torch.Tensor,
                 Tbm_init: torch.Tensor, n_inner: int = 5, max_dist: float = 100.0, kernel_scale: float = 5.0,
                 device="cuda"):
        self.device = torch.device(device)
        self.sim = simulator_for(model, map, device=device)
        self.W, self.H = model.getWidth(), model.getHeight()
        self.n = Tbm_init.shape[0]
        self.n_inner = n_inner
        self.params = M.UmeyamaReductionConstraints(max_dist=max_dist)
        self.Tbm = Tbm_init.clone().contiguous()
        self.dataset = M.PointCloudView(dataset_points.contiguous(), mask=dataset_mask.contiguous(), segment_stride=0)
        self.model_scan = resize_memory_bundle(Bundle(), self.W, self.H, self.n, ("hits", "points", "normals"),
                                               device=self.device)
        self.model_view = M.PointCloudView(self.model_scan.points, mask=self.model_scan.hits,
                                           normals=self.model_scan.normals)
        self.Tpre = torch.empty((self.n, 8), dtype=torch.float32, device=self.device)
        self.Tnext = torch.empty_like(self.Tpre)
        self.stats = torch.empty((self.n, 16), dtype=torch.float64, device=self.device)
        self.kernel_scale = float(kernel_scale)
        self.grad = torch.zeros((self.n, M.P2L_SYSTEM_COLS), dtype=torch.float64, device=self.device)
        ws = M._workspace(self.W * self.H, self.n, self.device)
        self.workspace = ws
        self.graphs: dict[int, HipGraph] = {}
        self._stream: torch.cuda.Stream | None = None

    def _statistics(self, stream):
        from . import _lib, types as T
        import ctypes
        seg = self.W * self.H
        dc, mc = self.dataset.to_c(seg), self.model_view.to_c(seg)
        prm = T.rmh_umeyama_constraints(self.params.max_dist, self.params.dataset_id, self.params.model_id)
        ws = self.workspace
        _lib.call("rmh_statistics", M.P2L, M._ptr(self.Tpre), self.n, ctypes.byref(dc), ctypes.byref(mc), seg,
                  self.n, ctypes.byref(prm), M._ptr(self.stats), M._ptr(ws), 0 if ws is None else ws.numel(),
                  M.stream_handle(stream))

    def iteration(self, stream=None) -> None:
        """One outer iteration for every hypothesis (no allocation, no synchronisation)."""
        from . import _lib
        self.sim.simulate(self.Tbm, self.model_scan, stream=stream)
        _lib.call("rmh_transform_set_identity", M._ptr(self.Tpre), self.n, M.stream_handle(stream))
        for _ in range(self.n_inner):
            self._statistics(stream)
            M.umeyama_transform(self.stats, out=self.Tnext, stream=stream)
            M.transform_mul(self.Tpre, self.Tnext, out=self.Tpre, stream=stream)
        # the gradient at the corrected pose, before the pose update consumes Tpre
        M.linear_system_p2l(self.Tpre, self.dataset, self.model_view, self.params, kernel_scale=self.kernel_scale,
                            segment_size=self.W * self.H, n_segments=self.n, out=self.grad,
                            workspace=self.workspace, stream=stream)
        M.transform_mul(self.Tbm, self.Tpre, out=self.Tbm, stream=stream)

    def capture(self, n_outer: int, stream: torch.cuda.Stream) -> HipGraph:
        g = HipGraph.capture(lambda s: [self.iteration(s) for _ in range(n_outer)], stream)
        self.graphs[n_outer] = g
        return g

    def gradient(self) -> torch.Tensor:
        """(n, 6) J^T w r of the last outer iteration (rotation xyz, then translation xyz)."""
        return self.grad[:, 36:42]

    def run(self, n_outer: int = 10, graph: bool = False, stream: torch.cuda.Stream | None = None) -> torch.Tensor:
        if graph:
            if stream is None:
                if self._stream is None:
                    self._stream = torch.cuda.Stream()
                stream = self._stream
            g = self.graphs.get(n_outer)
            if g is None:
                g = self.capture(n_outer, stream)
            # the replay reads Tbm and the dataset: wait for work queued on the caller's stream
            stream.wait_stream(torch.cuda.current_stream())
            g.launch(stream)
            stream.synchronize()
        else:
            for _ in range(n_outer):
                self.iteration(stream)
            torch.cuda.synchronize()
        return self.Tbm
