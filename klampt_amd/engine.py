"""Object wrappers over the C ABI: contexts, geometries, robots, worlds.

Host (numpy) batch calls stage through the library's own device buffers;
``*_dev`` methods take torch tensors already resident on the GPU and run on
the current torch stream.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import call, lib, DistSettings, DistResult, Params

_pd = ctypes.POINTER(ctypes.c_double)
_pf = ctypes.POINTER(ctypes.c_float)
_pi = ctypes.POINTER(ctypes.c_int32)
_pu8 = ctypes.POINTER(ctypes.c_uint8)


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


def _torch_stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _tp(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


class Context:
    """One GPU device and one HIP stream (``kg_ctx``)."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        call("kg_ctx_create", device, ctypes.byref(h))
        self.h = h
        self.device = device

    def synchronize(self):
        call("kg_ctx_synchronize", self.h)

    def kernel_times(self, reset=True):
        """(total seconds, count) of the FP32 traversal kernel over the
        feasibility batches since the last reset (HIP events)."""
        ns, cnt = ctypes.c_int64(), ctypes.c_int32()
        call("kg_ctx_kernel_times", self.h, 1 if reset else 0, ctypes.byref(ns), ctypes.byref(cnt))
        return ns.value * 1e-9, cnt.value

    def params(self):
        """The context's engine parameters (kg_params) as a dict."""
        p = Params()
        call("kg_ctx_get_params", self.h, ctypes.byref(p))
        return {name: int(getattr(p, name)) for name, _ in Params._fields_}

    def set_params(self, **changes):
        """Change some engine parameters; returns the previous values (a dict
        that set_params(**old) restores)."""
        old = self.params()
        new = dict(old)
        for k, v in changes.items():
            if k not in new:
                raise KeyError(f"unknown engine parameter {k!r}")
            new[k] = int(v)
        p = Params(**new)
        call("kg_ctx_set_params", self.h, ctypes.byref(p))
        return old

    def status_dev(self, out=None):
        """Status of the last batched call as a 4-int device tensor [spill
        flag, uncertain results, FP64 re-checks, 0], written in stream order
        on the current torch stream (kg_ctx_status_dev)."""
        import torch
        if out is None:
            out = torch.empty(4, dtype=torch.int32, device=f"cuda:{self.device}")
        call("kg_ctx_status_dev", self.h, _tp(out), _torch_stream())
        return out

    def set_timing(self, level):
        """HIP events recorded per batch: 0 none, 1 (default) the traversal
        kernel (kernel_times), 2 also the broad phase and the FP64 pass
        (last_stats times)."""
        call("kg_ctx_set_timing", self.h, int(level))

    def close(self):
        if self.h:
            lib().kg_ctx_destroy(self.h)
            self.h = None


class Geom:
    """A collision geometry: triangle mesh (with its OBB tree) or SDF grid."""

    def __init__(self, ctx, h, kind):
        self.ctx, self.h, self.kind = ctx, h, kind

    @classmethod
    def trimesh(cls, ctx, verts, tris):
        V, T = _d(verts).reshape(-1, 3), _i(tris).reshape(-1, 3)
        h = ctypes.c_void_p()
        call("kg_geom_create_trimesh", ctx.h, _p(V, _pd), V.shape[0], _p(T, _pi), T.shape[0], ctypes.byref(h))
        return cls(ctx, h, "TriangleMesh")

    @classmethod
    def pointcloud(cls, ctx, points):
        P = _d(points).reshape(-1, 3)
        h = ctypes.c_void_p()
        call("kg_geom_create_pointcloud", ctx.h, _p(P, _pd), P.shape[0], ctypes.byref(h))
        return cls(ctx, h, "PointCloud")

    @classmethod
    def sdf(cls, ctx, bmin, bmax, values):
        vals = _f(values)
        dims = _i(vals.shape)
        h = ctypes.c_void_p()
        call("kg_geom_create_sdf", ctx.h, _p(_d(bmin), _pd), _p(_d(bmax), _pd), _p(dims, _pi), _p(vals, _pf),
             ctypes.byref(h))
        g = cls(ctx, h, "ImplicitSurface")
        g.dims = tuple(int(x) for x in vals.shape)
        return g

    def bvh(self):
        """The mesh's OBB tree as an (nnodes, 16) float32 array (centre,
        half extents, column-major axes, child as int32 bits; see
        kg_geom_get_bvh)."""
        n = ctypes.c_int32()
        call("kg_geom_get_bvh", self.h, None, 0, ctypes.byref(n))
        out = np.zeros((max(1, n.value), 16), np.float32)
        call("kg_geom_get_bvh", self.h, _p(out, _pf), n.value, ctypes.byref(n))
        return out[:n.value].copy()

    def to_sdf(self, resolution=0.0, pad=2):
        """TriangleMesh -> ImplicitSurface (kg_geom_mesh_to_sdf)."""
        h = ctypes.c_void_p()
        call("kg_geom_mesh_to_sdf", self.h, float(resolution), int(pad), ctypes.byref(h))
        g = Geom(self.ctx, h, "ImplicitSurface")
        g.dims = g.sdf_grid()[0]
        return g

    def sdf_grid(self, values=False):
        """(dims, bmin, bmax[, values]) of an ImplicitSurface."""
        dims, lo, hi = np.zeros(3, np.int32), np.zeros(3), np.zeros(3)
        call("kg_geom_get_sdf", self.h, _p(dims, _pi), _p(lo, _pd), _p(hi, _pd), None)
        dims = tuple(int(x) for x in dims)
        if not values:
            return dims, lo, hi
        v = np.zeros(dims, np.float32)
        call("kg_geom_get_sdf", self.h, _p(np.zeros(3, np.int32), _pi), _p(np.zeros(3), _pd), _p(np.zeros(3), _pd),
             _p(v, _pf))
        return dims, lo, hi, v

    def clone(self):
        """Another handle on the same data (kg_geom_clone): meshes share the
        BVH; the transform and margin start as this handle's."""
        h = ctypes.c_void_p()
        call("kg_geom_clone", self.h, ctypes.byref(h))
        g = Geom(self.ctx, h, self.kind)
        if hasattr(self, "dims"):
            g.dims = self.dims
        return g

    def set_transform(self, R, t):
        call("kg_geom_set_transform", self.h, _p(_d(R), _pd), _p(_d(t), _pd))

    def get_transform(self):
        R, t = np.zeros(9), np.zeros(3)
        call("kg_geom_get_transform", self.h, _p(R, _pd), _p(t, _pd))
        return R, t

    def get_bb(self):
        lo, hi = np.zeros(3), np.zeros(3)
        call("kg_geom_get_bb", self.h, _p(lo, _pd), _p(hi, _pd))
        return lo, hi

    def set_margin(self, margin):
        call("kg_geom_set_collision_margin", self.h, float(margin))

    def get_margin(self):
        m = ctypes.c_double()
        call("kg_geom_get_collision_margin", self.h, ctypes.byref(m))
        return m.value

    def collides(self, other):
        out = ctypes.c_int32()
        call("kg_geom_collide", self.h, other.h, ctypes.byref(out))
        return bool(out.value)

    def within_distance(self, other, tol):
        out = ctypes.c_int32()
        call("kg_geom_within_distance", self.h, other.h, float(tol), ctypes.byref(out))
        return bool(out.value)

    def distance(self, other, relErr=0.0, absErr=0.0, upperBound=float("inf")):
        s = DistSettings(relErr, absErr, upperBound)
        r = DistResult()
        call("kg_geom_distance", self.h, other.h, ctypes.byref(s), ctypes.byref(r))
        return r

    def distance_point(self, pt, relErr=0.0, absErr=0.0, upperBound=float("inf")):
        s = DistSettings(relErr, absErr, upperBound)
        r = DistResult()
        call("kg_geom_distance_point", self.h, _p(_d(pt), _pd), ctypes.byref(s), ctypes.byref(r))
        return r

    def sdf_eval(self, pts):
        P = _f(pts).reshape(-1, 3)
        d = np.zeros(P.shape[0], np.float32)
        g = np.zeros((P.shape[0], 3), np.float32)
        call("kg_sdf_eval", self.h, _p(P, _pf), P.shape[0], _p(d, _pf), _p(g, _pf))
        return d, g

    def close(self):
        if self.h:
            lib().kg_geom_destroy(self.h)
            self.h = None


class Robot:
    """Kinematic tree with per-link meshes (``kg_robot``)."""

    def __init__(self, ctx, parents, T0, axes, jtype, qmin, qmax):
        self.ctx = ctx
        self.n = len(parents)
        h = ctypes.c_void_p()
        call("kg_robot_create", ctx.h, self.n, _p(_i(parents), _pi), _p(_d(T0), _pd), _p(_d(axes), _pd),
             _p(_i(jtype), _pi), _p(_d(qmin), _pd), _p(_d(qmax), _pd), ctypes.byref(h))
        self.h = h
        self.link_geoms = [None] * self.n

    @classmethod
    def from_spec(cls, ctx, spec):
        r = cls(ctx, spec.parents, spec.T0, spec.axes, spec.jtype, spec.qmin, spec.qmax)
        for l, m in enumerate(spec.link_meshes):
            if m is not None:
                r.set_link_geom(l, Geom.trimesh(ctx, m.verts, m.tris))
        if spec.self_pairs is not None:
            for i in range(r.n):
                for j in range(i + 1, r.n):
                    r.enable_self_collision(i, j, False)
            for i, j in spec.self_pairs:
                r.enable_self_collision(int(i), int(j), True)
        return r

    def set_link_geom(self, link, geom):
        call("kg_robot_set_link_geom", self.h, link, geom.h if geom is not None else None)
        self.link_geoms[link] = geom

    def enable_self_collision(self, i, j, enabled=True):
        call("kg_robot_enable_self_collision", self.h, i, j, 1 if enabled else 0)

    def set_self_pairs(self, pairs):
        P = _i(pairs).reshape(-1, 2)
        call("kg_robot_set_self_pairs", self.h, _p(P, _pi), P.shape[0])

    def self_collision_enabled(self, i, j):
        out = ctypes.c_int32()
        call("kg_robot_self_collision_enabled", self.h, i, j, ctypes.byref(out))
        return bool(out.value)

    def set_config(self, q):
        call("kg_robot_set_config", self.h, _p(_d(q), _pd))

    def get_config(self):
        q = np.zeros(self.n)
        call("kg_robot_get_config", self.h, _p(q, _pd))
        return q

    def link_transform(self, link):
        R, t = np.zeros(9), np.zeros(3)
        call("kg_robot_get_link_transform", self.h, link, _p(R, _pd), _p(t, _pd))
        return R, t

    def self_collides(self):
        out = ctypes.c_int32()
        call("kg_robot_self_collides", self.h, ctypes.byref(out))
        return bool(out.value)

    def fk_batch64(self, Q):
        Q = _d(Q).reshape(-1, self.n)
        T = np.zeros((Q.shape[0], self.n, 12))
        call("kg_fk_batch64", self.h, _p(Q, _pd), Q.shape[0], _p(T, _pd))
        return T

    def fk_batch(self, Q):
        Q = _f(Q).reshape(-1, self.n)
        T = np.zeros((Q.shape[0], self.n, 12), np.float32)
        call("kg_fk_batch", self.h, _p(Q, _pf), Q.shape[0], _p(T, _pf))
        return T

    def fk_batch_dev(self, Q, out=None):
        import torch
        assert Q.is_cuda and Q.dtype == torch.float32 and Q.is_contiguous()
        N = Q.shape[0]
        if out is None:
            out = torch.empty((N, self.n, 12), dtype=torch.float32, device=Q.device)
        call("kg_fk_batch_dev", self.h, _tp(Q), N, _tp(out), _torch_stream())
        return out


    # ---- SDF hinge penalty (sum over links/points of max(0, dsafe - d)^2)
    def sdf_penalty(self, sdf, Q, pts, pts_per_link, dsafe=0.05):
        Q = _f(Q).reshape(-1, self.n)
        P, ppl = _f(pts).reshape(-1, 3), _i(pts_per_link)
        loss = np.zeros(Q.shape[0], np.float32)
        call("kg_sdf_penalty_fwd", self.h, sdf.h, _p(Q, _pf), Q.shape[0], _p(P, _pf), _p(ppl, _pi), float(dsafe),
             _p(loss, _pf))
        return loss

    def sdf_penalty_grad(self, sdf, Q, pts, pts_per_link, dsafe=0.05, dloss=None):
        Q = _f(Q).reshape(-1, self.n)
        P, ppl = _f(pts).reshape(-1, 3), _i(pts_per_link)
        dl = None if dloss is None else _f(dloss)
        dQ = np.zeros_like(Q)
        call("kg_sdf_penalty_bwd", self.h, sdf.h, _p(Q, _pf), Q.shape[0], _p(P, _pf), _p(ppl, _pi), float(dsafe),
             _p(dl, _pf), _p(dQ, _pf))
        return dQ

    def sdf_penalty_dev(self, sdf, Q, pts, pts_per_link, dsafe=0.05, out=None):
        import torch
        ppl = _i(pts_per_link)
        if out is None:
            out = torch.empty(Q.shape[0], dtype=torch.float32, device=Q.device)
        call("kg_sdf_penalty_fwd_dev", self.h, sdf.h, _tp(Q), Q.shape[0], _tp(pts), _p(ppl, _pi), float(dsafe),
             _tp(out), _torch_stream())
        return out

    def sdf_penalty_grad_dev(self, sdf, Q, pts, pts_per_link, dsafe=0.05, dloss=None, out=None):
        import torch
        ppl = _i(pts_per_link)
        if out is None:
            out = torch.empty_like(Q)
        call("kg_sdf_penalty_bwd_dev", self.h, sdf.h, _tp(Q), Q.shape[0], _tp(pts), _p(ppl, _pi), float(dsafe),
             _tp(dloss), _tp(out), _torch_stream())
        return out


class World:
    """World model with Klampt's ID layout and default collision mask."""

    def __init__(self, ctx):
        self.ctx = ctx
        h = ctypes.c_void_p()
        call("kg_world_create", ctx.h, ctypes.byref(h))
        self.h = h
        self.robots = []
        self._keep = []

    def add_terrain(self, geom):
        out = ctypes.c_int32()
        call("kg_world_add_terrain", self.h, geom.h, ctypes.byref(out))
        self._keep.append(geom)
        return out.value

    def add_rigid_object(self, geom, xf12):
        x = _d(xf12)
        out = ctypes.c_int32()
        call("kg_world_add_rigid_object", self.h, geom.h, _p(x[:9].copy(), _pd), _p(x[9:].copy(), _pd),
             ctypes.byref(out))
        self._keep.append(geom)
        return out.value

    def set_terrain_geom(self, index, geom):
        call("kg_world_set_terrain_geom", self.h, index, geom.h)
        self._keep.append(geom)

    def set_rigid_object_geom(self, index, geom):
        call("kg_world_set_rigid_object_geom", self.h, index, geom.h)
        self._keep.append(geom)

    def set_rigid_object_transform(self, index, xf12):
        x = _d(xf12)
        call("kg_world_set_rigid_object_transform", self.h, index, _p(x[:9].copy(), _pd), _p(x[9:].copy(), _pd))

    def add_robot(self, robot):
        out = ctypes.c_int32()
        call("kg_world_add_robot", self.h, robot.h, ctypes.byref(out))
        self.robots.append(robot)
        return len(self.robots) - 1

    def num_ids(self):
        return lib().kg_world_num_ids(self.h)

    def set_collision_enabled(self, a, b, enabled):
        call("kg_world_set_collision_enabled", self.h, a, b, 1 if enabled else 0)

    def collision_enabled(self, a, b):
        out = ctypes.c_int32()
        call("kg_world_get_collision_enabled", self.h, a, b, ctypes.byref(out))
        return bool(out.value)

    def pair_list(self, robot=0):
        n = ctypes.c_int32()
        call("kg_world_pair_list", self.h, robot, None, 0, ctypes.byref(n))
        out = np.zeros((max(1, n.value), 2), np.int32)
        call("kg_world_pair_list", self.h, robot, _p(out, _pi), n.value, ctypes.byref(n))
        return out[:n.value].copy()

    def feasible_batch(self, Q, robot=0, flags=False, first_pair=False):
        """IsFeasible over a batch.  float64 configurations (Klampt's Config)
        are checked as given (kg_feasible_batch_f64); anything else as
        float32 (kg_feasible_batch)."""
        n = self.robots[robot].n
        dbl = np.asarray(Q).dtype == np.float64
        Q = (_d(Q) if dbl else _f(Q)).reshape(-1, n)
        N = Q.shape[0]
        feas = np.zeros(N, np.uint8)
        env = np.zeros(N, np.uint8) if flags else None
        slf = np.zeros(N, np.uint8) if flags else None
        fp = np.zeros((N, 2), np.int32) if first_pair else None
        if dbl:
            call("kg_feasible_batch_f64", self.h, robot, _p(Q, _pd), N, _p(feas, _pu8), _p(env, _pu8),
                 _p(slf, _pu8), _p(fp, _pi))
        else:
            call("kg_feasible_batch", self.h, robot, _p(Q, _pf), N, _p(feas, _pu8), _p(env, _pu8), _p(slf, _pu8),
                 _p(fp, _pi))
        return feas, env, slf, fp

    def feasible_batch_dev(self, Q, robot=0, feasible=None, env=None, self_hit=None, first_pair=None):
        """Device-tensor IsFeasible on the current torch stream; Q float32 or
        float64 (checked as given)."""
        import torch
        assert Q.is_cuda and Q.dtype in (torch.float32, torch.float64) and Q.is_contiguous()
        N = Q.shape[0]
        if feasible is None:
            feasible = torch.empty(N, dtype=torch.uint8, device=Q.device)
        name = "kg_feasible_batch_f64_dev" if Q.dtype == torch.float64 else "kg_feasible_batch_dev"
        call(name, self.h, robot, _tp(Q), N, _tp(feasible), _tp(env), _tp(self_hit), _tp(first_pair),
             _torch_stream())
        return feasible

    def feasible_batch_dev_noalloc(self, Q, robot=0, feasible=None, env=None, self_hit=None, first_pair=None):
        """kg_feasible_batch_dev_noalloc: the same check with every workspace
        in the library's fixed scratch array (nothing allocated on the
        device; the outputs are the caller's).  Q float32."""
        import torch
        assert Q.is_cuda and Q.dtype == torch.float32 and Q.is_contiguous()
        N = Q.shape[0]
        assert feasible is not None
        call("kg_feasible_batch_dev_noalloc", self.h, robot, _tp(Q), N, _tp(feasible), _tp(env), _tp(self_hit),
             _tp(first_pair), _torch_stream())
        return feasible

    def num_obstacles(self):
        out = ctypes.c_int32()
        call("kg_world_num_obstacles", self.h, ctypes.byref(out))
        return out.value

    def sweep_batch(self, Q0, Q1, dcheck, robot=0):
        """KineTrajOpt sweep terms (kg_sweep_batch) for segments Q0[k] -> Q1[k]
        (float64): d (N, n, O), info (N, n, O, 16: alpha, normal, pswept,
        pobs, p0, p1) and coef (N, n, O, 2, n: normal . J at p0 / at p1)."""
        n = self.robots[robot].n
        Q0, Q1 = _d(Q0).reshape(-1, n), _d(Q1).reshape(-1, n)
        N, O = Q0.shape[0], self.num_obstacles()
        d = np.full((N, n, O), np.inf)
        info = np.zeros((N, n, O, 16))
        coef = np.zeros((N, n, O, 2, n))
        call("kg_sweep_batch", self.h, robot, _p(Q0, _pd), _p(Q1, _pd), N, float(dcheck), _p(d, _pd), _p(info, _pd),
             _p(coef, _pd))
        return d, info, coef

    def sweep_batch_dev(self, Q0, Q1, dcheck, robot=0):
        """kg_sweep_batch_dev on float64 CUDA tensors; returns device tensors."""
        import torch
        n = self.robots[robot].n
        N, O = Q0.shape[0], self.num_obstacles()
        d = torch.empty((N, n, O), dtype=torch.float64, device=Q0.device)
        info = torch.empty((N, n, O, 16), dtype=torch.float64, device=Q0.device)
        coef = torch.empty((N, n, O, 2, n), dtype=torch.float64, device=Q0.device)
        call("kg_sweep_batch_dev", self.h, robot, _tp(Q0.contiguous()), _tp(Q1.contiguous()), N, float(dcheck),
             _tp(d), _tp(info), _tp(coef), _torch_stream())
        return d, info, coef

    def distance_batch(self, Q, robot=0):
        n = self.robots[robot].n
        Q = _f(Q).reshape(-1, n)
        npairs = len(self.pair_list(robot))
        out = np.zeros((Q.shape[0], npairs), np.float32)
        call("kg_distance_batch", self.h, robot, _p(Q, _pf), Q.shape[0], _p(out, _pf))
        return out

    def distance_batch_dev(self, Q, robot=0, out=None):
        import torch
        N = Q.shape[0]
        npairs = len(self.pair_list(robot))
        if out is None:
            out = torch.empty((N, npairs), dtype=torch.float32, device=Q.device)
        call("kg_distance_batch_dev", self.h, robot, _tp(Q), N, _tp(out), _torch_stream())
        return out

    def distance_jacobian_batch_dev(self, Q, dcheck, robot=0, out=None):
        """Device-tensor form of distance_jacobian_batch: returns (d, cp, grad)
        torch tensors on Q's device (or fills out = (d, cp, grad))."""
        import torch
        N, n = Q.shape
        npairs = len(self.pair_list(robot))
        if out is None:
            out = (torch.empty((N, npairs), dtype=torch.float32, device=Q.device),
                   torch.empty((N, npairs, 6), dtype=torch.float32, device=Q.device),
                   torch.empty((N, npairs, n), dtype=torch.float32, device=Q.device))
        d, cp, g = out
        call("kg_distance_jacobian_batch_dev", self.h, robot, _tp(Q), N, float(dcheck), _tp(d), _tp(cp), _tp(g),
             _torch_stream())
        return d, cp, g

    def distance_jacobian_batch(self, Q, dcheck, robot=0):
        """Pair distances within dcheck, closest points (N, npairs, 6) and
        joint-space gradients (N, npairs, n) (kg_distance_jacobian_batch)."""
        Qf = _f(Q)
        N, n = Qf.shape
        npairs = len(self.pair_list(robot))
        d = np.zeros((N, npairs), np.float32)
        cp = np.zeros((N, npairs, 6), np.float32)
        g = np.zeros((N, npairs, n), np.float32)
        call("kg_distance_jacobian_batch", self.h, robot, _p(Qf, _pf), N, float(dcheck), _p(d, _pf), _p(cp, _pf),
             _p(g, _pf))
        return d, cp, g

    def edge_batch(self, A, B, eps=0.01, norm_inf=False, robot=0):
        """First infeasible interior sample u per segment (-1 if none); the
        segment length is the L2 norm unless norm_inf (the L-infinity norm).
        float64 endpoints (Klampt's Config) are interpolated as given
        (kg_edge_batch_f64); anything else as float32."""
        n = self.robots[robot].n
        dbl = np.asarray(A).dtype == np.float64 and np.asarray(B).dtype == np.float64
        cv, pt, name = (_d, _pd, "kg_edge_batch_f64") if dbl else (_f, _pf, "kg_edge_batch")
        A, B = cv(A).reshape(-1, n), cv(B).reshape(-1, n)
        assert A.shape == B.shape
        out = np.zeros(A.shape[0], np.float32)
        call(name, self.h, robot, _p(A, pt), _p(B, pt), A.shape[0], float(eps), 1 if norm_inf else 0, _p(out, _pf))
        return out

    def edge_batch_dev(self, A, B, eps=0.01, norm_inf=False, robot=0, max_steps=0, out=None, nsamples=None):
        """Device-tensor edge checks on the current torch stream (float32 or
        float64 endpoints); returns first_u (-2 for a segment longer than
        max_steps steps, see kg_edge_batch_dev)."""
        import torch
        assert A.is_cuda and B.is_cuda and A.dtype == B.dtype and A.shape == B.shape
        assert A.is_contiguous() and B.is_contiguous()
        nseg = A.shape[0]
        if out is None:
            out = torch.empty(nseg, dtype=torch.float32, device=A.device)
        name = "kg_edge_batch_f64_dev" if A.dtype == torch.float64 else "kg_edge_batch_dev"
        call(name, self.h, robot, _tp(A), _tp(B), nseg, float(eps), 1 if norm_inf else 0, int(max_steps), _tp(out),
             _tp(nsamples), _torch_stream())
        return out

    def last_stats(self):
        s = np.zeros(10, np.int64)
        call("kg_world_last_stats", self.h, s.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), 10)
        keys = ["fp64_rechecks", "pair_items", "spills", "t_prep_ns", "t_trav32_ns", "t_trav64_ns",
                "box_tests", "tri_tests", "uncertain", "exact_pairs"]
        return {k: int(v) for k, v in zip(keys, s)}


def world_from_scene(ctx, scene):
    """World for a ``klampt_amd.synthetic.Scene``: terrains, rigid objects,
    then the robot (Klampt ID order)."""
    w = World(ctx)
    for t in scene.terrains:
        w.add_terrain(Geom.trimesh(ctx, t.verts, t.tris))
    for mesh, xf in scene.objects:
        w.add_rigid_object(Geom.trimesh(ctx, mesh.verts, mesh.tris), xf)
    r = Robot.from_spec(ctx, scene.robot)
    w.add_robot(r)
    return w, r


def _multi_args(worlds, robots, counts):
    nb = len(worlds)
    W = (ctypes.c_void_p * nb)(*[w.h.value if hasattr(w.h, "value") else w.h for w in worlds])
    R = _i(robots)
    C = _i(counts)
    return nb, W, R, C


def feasible_multi(worlds, robots, counts, Qs):
    """IsFeasible over a heterogeneous batch in one launch: Qs[b] is a
    (counts[b], n_b) array for robot robots[b] of worlds[b]."""
    nb, W, R, C = _multi_args(worlds, robots, counts)
    Q = np.concatenate([_f(q).ravel() for q in Qs]) if nb else np.zeros(0, np.float32)
    out = np.zeros(int(np.sum(counts)), np.uint8)
    call("kg_feasible_multi", nb, W, _p(R, _pi), _p(C, _pi), _p(_f(Q), _pf), _p(out, _pu8))
    return np.split(out, np.cumsum(counts)[:-1])


def feasible_multi_dev(worlds, robots, counts, dQ, out=None):
    import torch
    nb, W, R, C = _multi_args(worlds, robots, counts)
    if out is None:
        out = torch.empty(int(np.sum(counts)), dtype=torch.uint8, device=dQ.device)
    call("kg_feasible_multi_dev", nb, W, _p(R, _pi), _p(C, _pi), _tp(dQ), _tp(out), _torch_stream())
    return out


def autotune(worlds, robots, counts, dQ):
    """Pick the fastest traversal settings for this batch; returns
    (cpw, leaf_min, scap, waves_per_cu, ms)."""
    nb, W, R, C = _multi_args(worlds, robots, counts)
    ch = np.zeros(4, np.int32)
    ms = ctypes.c_float()
    call("kg_autotune", nb, W, _p(R, _pi), _p(C, _pi), _tp(dQ), _p(ch, _pi), ctypes.byref(ms))
    return tuple(int(x) for x in ch) + (ms.value,)


class Serve:
    """kg_serve_*: a feasibility batch over several robots and the edge
    batches of several robots, captured by the library into one HIP graph and
    replayed by run().  The device buffers are fixed at creation: write new
    configurations / segment endpoints into ``Q``, ``A``, ``B`` (float32 CUDA
    tensors, concatenated per batch / group) and call run(); results land in
    ``feasible`` (uint8) and ``first_u`` (float32)."""

    def __init__(self, ctx, feas=(), edges=(), eps=0.01, norm_inf=False, max_steps=0, device="cuda"):
        import torch
        self.ctx = ctx
        fw = [w for w, _, _ in feas]
        fr, fc = [int(r) for _, r, _ in feas], [int(c) for _, _, c in feas]
        ew = [w for w, _, _ in edges]
        er, ec = [int(r) for _, r, _ in edges], [int(c) for _, _, c in edges]
        nq = sum(c * w.robots[r].n for w, r, c in feas)
        ne = sum(c * w.robots[r].n for w, r, c in edges)
        self.Q = torch.zeros(max(1, nq), dtype=torch.float32, device=device)
        self.feasible = torch.zeros(max(1, sum(fc)), dtype=torch.uint8, device=device)
        self.A = torch.zeros(max(1, ne), dtype=torch.float32, device=device)
        self.B = torch.zeros(max(1, ne), dtype=torch.float32, device=device)
        self.first_u = torch.zeros(max(1, sum(ec)), dtype=torch.float32, device=device)
        self.feas_counts, self.edge_counts = fc, ec
        nb, W, R, C = _multi_args(fw, fr, fc)
        ne_, EW, ER, EC = _multi_args(ew, er, ec)
        torch.cuda.synchronize()
        h = ctypes.c_void_p()
        call("kg_serve_create", nb, W, _p(R, _pi), _p(C, _pi), _tp(self.Q), _tp(self.feasible), ne_, EW,
             _p(ER, _pi), _p(EC, _pi), _tp(self.A), _tp(self.B), float(eps), 1 if norm_inf else 0, int(max_steps),
             _tp(self.first_u), ctypes.byref(h))
        self.h = h

    def run(self):
        call("kg_serve_run", self.h, _torch_stream())

    def captures(self):
        out = ctypes.c_int32()
        call("kg_serve_captures", self.h, ctypes.byref(out))
        return out.value

    def close(self):
        if self.h:
            call("kg_serve_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def set_tuning(ctx, cpw=0, leaf_min=0, scap=0, waves_per_cu=0):
    call("kg_ctx_set_tuning", ctx.h, cpw, leaf_min, scap, waves_per_cu)
