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
pose updates, with no host synchronisation.  `run(..., graph=True)` records the outer
iterations once into a HIP graph and replays it.
"""
from __future__ import annotations

import torch

from . import math as M
from .simulation import Bundle, HipGraph, HipMap, resize_memory_bundle, simulator_for


class BatchedRCC:
    def __init__(self, map: HipMap, model, dataset_points: torch.Tensor, dataset_mask: torch.Tensor,
                 Tbm_init: torch.Tensor, n_inner: int = 5, max_dist: float = 100.0, device="cuda"):
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
        ws = M._workspace(self.W * self.H, self.n, self.device)
        self.workspace = ws
        self.graph = None

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
        M.transform_mul(self.Tbm, self.Tpre, out=self.Tbm, stream=stream)

    def capture(self, n_outer: int, stream: torch.cuda.Stream) -> HipGraph:
        self.graph = HipGraph.capture(lambda s: [self.iteration(s) for _ in range(n_outer)], stream)
        return self.graph

    def run(self, n_outer: int = 10, graph: bool = False, stream: torch.cuda.Stream | None = None) -> torch.Tensor:
        if graph:
            stream = stream or torch.cuda.Stream()
            if self.graph is None:
                self.capture(n_outer, stream)
            self.graph.launch(stream)
            stream.synchronize()
        else:
            for _ in range(n_outer):
                self.iteration(stream)
            torch.cuda.synchronize()
        return self.Tbm
