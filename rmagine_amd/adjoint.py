"""Differentiable range-residual loss w.r.t. 6-DoF sensor poses (librmh's rmh_pose_grad).

The reference has no differentiable path (closest pieces: jacobian_and_residual_p2l,
src/rmagine_core/include/rmagine/math/optimization.h:113-137; se3_exp / so3 Jacobians,
src/rmagine_core/include/rmagine/math/lie.tcc:252-277).

Pose parametrisation: a base pose T0 = (R0, t0) is perturbed by xi = (tau, omega) as
    t = t0 + tau,   R = Exp(omega) R0          (left perturbation in the map frame).
The kernel returns dL/d(tau, omega_local) at the perturbed pose, where omega_local is a
further left perturbation Exp(omega_local) R.  Since Exp(omega + d) = Exp(J_l(omega) d) Exp(omega)
to first order, dL/domega = J_l(omega)^T dL/domega_local, with J_l the SO(3) left Jacobian
I + (1 - cos a)/a^2 W + (a - sin a)/a^3 W^2, W = [omega]x, a = |omega|.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from . import types as T
from .simulation import HipMap, SimulatorHip, stream_handle


def _ptr(t):
    return ctypes.c_void_p(0 if t is None else int(t.data_ptr()))


def workspace_bytes(sim: SimulatorHip, n_poses: int) -> int:
    n = ctypes.c_size_t()
    _lib.call("rmh_pose_grad_workspace_size", sim.map().handle, sim._sensor.handle, n_poses, ctypes.byref(n))
    return n.value


def pose_grad(sim: SimulatorHip, Tbm: torch.Tensor, target: torch.Tensor, target_mask: torch.Tensor | None = None,
              per_pose_targets: bool = False, huber_delta: float = 0.05, out: torch.Tensor | None = None,
              workspace: torch.Tensor | None = None, stream=None) -> torch.Tensor:
    """Loss and gradient per pose: (N, 8) float64 {L, dL/dtau xyz, dL/domega xyz, rays used}.

    Tbm: (N, 8) float32 device poses; target: (W*H,) or (N*W*H,) float32 device ranges.
    """
    if sim.map() is None:
        raise _lib.NoMapError(3, "pose_grad: no map set")
    if sim._sensor is None:
        raise RuntimeError("pose_grad: no sensor model set")
    for name, t, dt in (("Tbm", Tbm, torch.float32), ("target", target, torch.float32)):
        if not (isinstance(t, torch.Tensor) and t.device.type == "cuda" and t.is_contiguous() and t.dtype == dt):
            raise ValueError(f"{name} must be a contiguous {dt} device tensor")
    N = Tbm.shape[0]
    rays = sim._sensor.W * sim._sensor.H
    need = rays * (N if per_pose_targets else 1)
    if target.numel() != need or (target_mask is not None and target_mask.numel() != need):
        raise ValueError(f"target (and mask) must hold {need} entries")
    if out is None:
        out = torch.empty((N, 8), dtype=torch.float64, device=Tbm.device)
    nbytes = workspace_bytes(sim, N)
    if workspace is None or workspace.numel() < nbytes:
        workspace = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=Tbm.device)
    Tsb = T.transform_to_c(sim._Tsb)
    _lib.call("rmh_pose_grad", sim.map().handle, sim._sensor.handle, ctypes.byref(Tsb), _ptr(Tbm), N, _ptr(target),
              _ptr(target_mask), int(bool(per_pose_targets)), float(huber_delta), _ptr(out), _ptr(workspace),
              workspace.numel(), stream_handle(stream))
    return out


def _skew(w: torch.Tensor) -> torch.Tensor:
    z = torch.zeros_like(w[:, 0])
    return torch.stack([torch.stack([z, -w[:, 2], w[:, 1]], 1),
                        torch.stack([w[:, 2], z, -w[:, 0]], 1),
                        torch.stack([-w[:, 1], w[:, 0], z], 1)], 1)


def so3_left_jacobian(w: torch.Tensor) -> torch.Tensor:
    """J_l(w) for (N, 3) float64 rotation vectors."""
    a = torch.linalg.norm(w, dim=1)
    W = _skew(w)
    W2 = W @ W
    small = a < 1e-6
    a_ = torch.where(small, torch.ones_like(a), a)
    c1 = torch.where(small, 0.5 - a * a / 24.0, (1 - torch.cos(a_)) / a_ ** 2)
    c2 = torch.where(small, 1.0 / 6.0 - a * a / 120.0, (a_ - torch.sin(a_)) / a_ ** 3)
    I = torch.eye(3, dtype=w.dtype, device=w.device).expand_as(W)
    return I + c1[:, None, None] * W + c2[:, None, None] * W2


def perturb(T0: torch.Tensor, xi: torch.Tensor) -> torch.Tensor:
    """(N,8) float32 poses (t0 + tau, Exp(omega) R0) for (N,6) float64 xi = (tau, omega)."""
    xi = xi.to(torch.float64)
    tau, w = xi[:, :3], xi[:, 3:]
    a = torch.linalg.norm(w, dim=1, keepdim=True)
    half = 0.5 * a
    s = torch.where(a > 0, torch.sin(half) / torch.where(a > 0, a, torch.ones_like(a)), 0.5 * torch.ones_like(a))
    qw = torch.cat([w * s, torch.cos(half)], 1)  # (x, y, z, w)
    q0 = T0[:, 0:4].to(torch.float64)
    x1, y1, z1, w1 = qw.unbind(1)
    x2, y2, z2, w2 = q0.unbind(1)
    q = torch.stack([w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                     w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
                     w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2,
                     w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2], 1)
    out = torch.zeros_like(T0)
    out[:, 0:4] = q.to(torch.float32)
    out[:, 4:7] = (T0[:, 4:7].to(torch.float64) + tau).to(torch.float32)
    return out.contiguous()


class PoseRangeLoss(torch.autograd.Function):
    """Per-pose Huber range loss L(xi) of poses perturb(T0, xi) against a target scan.

    forward(xi (N,6), T0 (N,8), sim, target, target_mask, huber_delta) -> (N,) float64.
    The backward pass uses the adjoint computed in the same kernel launch as the forward.
    """

    @staticmethod
    def forward(ctx, xi, T0, sim, target, target_mask=None, huber_delta=0.05):
        Tp = perturb(T0, xi.detach())
        res = pose_grad(sim, Tp, target, target_mask, per_pose_targets=target.numel() != sim._sensor.W * sim._sensor.H,
                        huber_delta=huber_delta)
        J = so3_left_jacobian(xi.detach().to(torch.float64)[:, 3:])
        g = torch.empty((xi.shape[0], 6), dtype=torch.float64, device=xi.device)
        g[:, :3] = res[:, 1:4]
        g[:, 3:] = torch.einsum("nji,nj->ni", J, res[:, 4:7])
        ctx.save_for_backward(g)
        ctx.xi_dtype = xi.dtype
        return res[:, 0]

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return (grad_out.to(torch.float64)[:, None] * g).to(ctx.xi_dtype), None, None, None, None, None


def pose_range_loss(xi, T0, sim, target, target_mask=None, huber_delta=0.05):
    return PoseRangeLoss.apply(xi, T0, sim, target, target_mask, huber_delta)
