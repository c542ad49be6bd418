"""torch.autograd bridge for the SDF hinge penalty.

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
