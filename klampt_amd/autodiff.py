"""torch.autograd bridges: the SDF hinge penalty and the batched minimum
distance.

Plays the role of Klampt's ``ADModule`` torch wrapper
(``Python/klampt/math/autodiff/pytorch.py:72-115``) around the distance
gradient chain ``GeomPointDistance`` x ``WorldPosition``
(``Python/klampt/math/autodiff/geometry_ad.py:250-293``,
``kinematics_ad.py:37-67``), but evaluates the whole batch in two HIP kernels:
the forward computes only the loss (nothing per link is stored) and the
backward recomputes forward kinematics and the SDF gathers to form dL/dq by
an FK-Jacobian adjoint.
"""
from __future__ import annotations

import numpy as np
import torch

from .engine import Geom, Robot


class _SDFPenalty(torch.autograd.Function):
    @staticmethod
    def forward(ctx, Q, state):
        robot, sdf, pts, ppl, dsafe = state
        loss = robot.sdf_penalty_dev(sdf, Q, pts, ppl, dsafe)
        ctx.save_for_backward(Q)
        ctx.state = state
        return loss

    @staticmethod
    def backward(ctx, dloss):
        (Q,) = ctx.saved_tensors
        robot, sdf, pts, ppl, dsafe = ctx.state
        dQ = robot.sdf_penalty_grad_dev(sdf, Q, pts, ppl, dsafe, dloss=dloss.contiguous().float())
        return dQ, None


class SDFPenalty:
    """L(q) = sum over links k and their surface points p of
    max(0, dsafe - sdf(T_k(q) p))^2 for a batch of configurations.

    ``local_points``: list (one entry per link) of (m_k, 3) arrays in the link
    frame.  Calling the object with a (N, n) float32 CUDA tensor returns the
    (N,) loss tensor; ``loss.sum().backward()`` fills ``Q.grad``.
    """

    def __init__(self, robot: Robot, sdf: Geom, local_points, dsafe=0.05, device="cuda"):
        self.robot, self.sdf, self.dsafe = robot, sdf, float(dsafe)
        pts = [np.asarray(p, np.float32).reshape(-1, 3) for p in local_points]
        assert len(pts) == robot.n
        self.ppl = np.array([p.shape[0] for p in pts], np.int32)
        self.pts = torch.from_numpy(np.concatenate(pts) if sum(self.ppl) else np.zeros((0, 3), np.float32)).to(device)

    def __call__(self, Q):
        if not (Q.is_cuda and Q.dtype == torch.float32):
            raise TypeError("SDFPenalty expects a float32 CUDA tensor")
        return _SDFPenalty.apply(Q.contiguous(), (self.robot, self.sdf, self.pts, self.ppl, self.dsafe))


class _MinDistance(torch.autograd.Function):
    @staticmethod
    def forward(ctx, Q, state):
        world, robot, upper = state
        d, cp, g = world.distance_jacobian_batch_dev(Q, upper, robot=robot)
        dmin, arg = torch.min(d, dim=1)
        hit = torch.isfinite(dmin)
        # the closest pair's joint-space gradient; 0 where every pair lies
        # beyond the bound (the value is then the bound itself)
        grad = torch.gather(g, 1, arg.view(-1, 1, 1).expand(-1, 1, g.shape[2])).squeeze(1)
        grad = torch.where(hit[:, None], grad, torch.zeros_like(grad))
        ctx.save_for_backward(grad)
        return torch.where(hit, dmin, torch.full_like(dmin, float(upper)))

    @staticmethod
    def backward(ctx, dout):
        (grad,) = ctx.saved_tensors
        return dout[:, None] * grad, None


class MinDistance:
    """Smallest distance between the robot's link meshes and everything they
    are checked against (the kg_world_pair_list pairs: environment items and
    self pairs), per configuration, with its gradient in q: the batched form
    of ``MinDistance`` over the link and obstacle geometries
    (``Python/klampt/math/autodiff/geometry_ad.py:613-800``) chained with the
    links' kinematics.  As there, a distance above ``upperBound`` is not
    searched for: the value is then ``upperBound`` and the gradient 0; the
    gradient is that of the closest pair (a subgradient where pairs tie),
    dd/dq = nhat . (J_a(cp_a) - J_b(cp_b)) from kg_distance_jacobian_batch,
    0 where the geometries touch.  Q: (N, n) float32 CUDA tensor."""

    def __init__(self, world, robot_index=0, upperBound=1.0):
        if not (upperBound > 0 and np.isfinite(upperBound)):
            raise ValueError("upperBound must be positive and finite")
        self.world, self.robot, self.upper = world, int(robot_index), float(upperBound)

    def __call__(self, Q):
        if not (Q.is_cuda and Q.dtype == torch.float32):
            raise TypeError("MinDistance expects a float32 CUDA tensor")
        return _MinDistance.apply(Q.contiguous(), (self.world, self.robot, self.upper))
