"""ctypes wrapper of the FP64 CPU oracle (``oracle/kg_oracle.cpp``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` as the checker.  The product
library (``klampt_amd``) never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "build", "liboracle.so")
_lib = None

_d = ctypes.POINTER(ctypes.c_double)
_i = ctypes.POINTER(ctypes.c_int32)
_u8 = ctypes.POINTER(ctypes.c_uint8)
_vp = ctypes.c_void_p


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = ctypes.CDLL(_SO)
        sig = {
            "or_world_new": (_vp, []),
            "or_world_free": (None, [_vp]),
            "or_set_mutation": (None, [ctypes.c_int]),
            "or_add_mesh": (ctypes.c_int, [_vp, _d, ctypes.c_int, _i, ctypes.c_int]),
            "or_add_terrain": (ctypes.c_int, [_vp, ctypes.c_int]),
            "or_add_object": (ctypes.c_int, [_vp, ctypes.c_int, _d]),
            "or_add_robot": (ctypes.c_int, [_vp, ctypes.c_int, _i, _d, _d, _i, _d, _d, _i]),
            "or_set_self_collision": (None, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
            "or_init_default": (None, [_vp]),
            "or_num_ids": (ctypes.c_int, [_vp]),
            "or_set_collision_enabled": (None, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
            "or_get_collision_enabled": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
            "or_set_robot_config": (None, [_vp, ctypes.c_int, _d]),
            "or_fk": (None, [_vp, ctypes.c_int, _d, _d]),
            "or_feasible_batch": (None, [_vp, ctypes.c_int, _d, ctypes.c_int, _u8, _u8, _u8, _i,
                                         ctypes.c_int, ctypes.c_int]),
            "or_pair_list": (ctypes.c_int, [_vp, ctypes.c_int, _i, ctypes.c_int]),
            "or_isfeasible_batch": (None, [_vp, ctypes.c_int, _d, ctypes.c_int, _u8, ctypes.c_int]),
            "or_distance_batch": (None, [_vp, ctypes.c_int, _d, ctypes.c_int, _d, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int]),
            "or_edge_batch": (None, [_vp, ctypes.c_int, _d, _d, ctypes.c_int, ctypes.c_double,
                                     ctypes.c_int, _d, _i, _i, ctypes.c_int]),
            "or_bisection_order": (None, [ctypes.c_int, _i]),
            "or_mesh_collide": (ctypes.c_int, [_vp, ctypes.c_int, _d, ctypes.c_int, _d, ctypes.c_int]),
            "or_mesh_distance": (ctypes.c_double, [_vp, ctypes.c_int, _d, ctypes.c_int, _d, ctypes.c_int,
                                                   _d, _d, _i]),
            "or_mesh_separation": (ctypes.c_double, [_vp, ctypes.c_int, _d, ctypes.c_int, _d]),
            "or_tritri": (ctypes.c_int, [_d, _d]),
            "or_tritri_dist": (ctypes.c_double, [_d, _d]),
            "or_tritri_depth": (ctypes.c_double, [_d, _d]),
            "or_world_aabb": (None, [_vp, ctypes.c_int, _d, _d, _d]),
            "or_sdf_eval": (None, [_d, _i, _d, _d, _d, ctypes.c_int, _d, _d]),
            "or_mesh_to_sdf": (None, [_d, ctypes.c_int, _i, ctypes.c_int, _d, ctypes.c_double, _i, _d, _i,
                                      ctypes.c_int, ctypes.c_int]),
            "or_sdf_penalty": (None, [_vp, ctypes.c_int, _d, _i, _d, _d, _d, ctypes.c_int, _d, _i,
                                      ctypes.c_double, _d, _d, ctypes.c_double, ctypes.c_int]),
            "or_mesh_set_obb": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_float), ctypes.c_int]),
            "or_hull_distance": (ctypes.c_double, [_d, ctypes.c_int, _d, ctypes.c_int, _d, ctypes.c_int, _d, _d, _d,
                                                   _i]),
            "or_hull_support": (None, [_d, ctypes.c_int, _d, _d, _d]),
            "or_feasible_counted": (ctypes.c_int, [_vp, ctypes.c_int, _d, ctypes.c_int, _u8,
                                                   ctypes.POINTER(ctypes.c_int64), ctypes.c_int]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def _dd(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _ii(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def nthreads_default():
    return max(1, min(16, os.cpu_count() or 1))


class OracleWorld:
    """FP64 restatement of a Klampt ``WorldModel`` + ``WorldPlannerSettings``."""

    def __init__(self):
        self.L = lib()
        self.h = self.L.or_world_new()
        self._robots = []
        self.meshes = []  # oracle mesh index per add_mesh call, in order

    def __del__(self):
        try:
            self.L.or_world_free(self.h)
        except Exception:
            pass

    def add_mesh(self, mesh):
        V = _dd(mesh.verts)
        T = _ii(mesh.tris)
        g = self.L.or_add_mesh(self.h, _p(V, _d), V.shape[0], _p(T, _i), T.shape[0])
        self.meshes.append(g)
        return g

    def set_obb_tree(self, g, nodes):
        """Give mesh g an OBB tree: (nnodes, 16) float32 records (centre,
        half extents, column-major axes, child as int32 bits), e.g. the
        engine's own tree (engine.Geom.bvh), so that feasible_counted
        traverses the same tree."""
        nodes = np.ascontiguousarray(nodes, np.float32)
        assert nodes.ndim == 2 and nodes.shape[1] == 16
        rc = self.L.or_mesh_set_obb(self.h, g, nodes.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), nodes.shape[0])
        if rc != 0:
            raise ValueError("invalid OBB tree for mesh %d" % g)

    def feasible_counted(self, r, Q, nthreads=None):
        """IsFeasible with Klampt's early returns on the OBB trees given by
        set_obb_tree, with the work of each check: (feasible[N], counts[N, 5])
        with counts = environment box tests, environment-node visits,
        environment triangle tests, self box tests, self triangle tests."""
        Q = _dd(Q)
        N = Q.shape[0]
        feas = np.zeros(N, np.uint8)
        cnt = np.zeros((N, 5), np.int64)
        rc = self.L.or_feasible_counted(self.h, r, _p(Q, _d), N, _p(feas, _u8),
                                        cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                        nthreads or nthreads_default())
        if rc != 0:
            raise ValueError("a mesh of the check has no OBB tree (set_obb_tree)")
        return feas, cnt

    def add_terrain(self, mesh):
        return self.L.or_add_terrain(self.h, self.add_mesh(mesh))

    def add_object(self, mesh, xf12):
        x = _dd(xf12)
        return self.L.or_add_object(self.h, self.add_mesh(mesh), _p(x, _d))

    def add_robot(self, spec):
        geoms = [self.add_mesh(m) if m is not None else -1 for m in spec.link_meshes]
        g = _ii(geoms)
        par, T0, ax = _ii(spec.parents), _dd(spec.T0), _dd(spec.axes)
        jt, lo, hi = _ii(spec.jtype), _dd(spec.qmin), _dd(spec.qmax)
        r = self.L.or_add_robot(self.h, spec.n, _p(par, _i), _p(T0, _d), _p(ax, _d), _p(jt, _i),
                                _p(lo, _d), _p(hi, _d), _p(g, _i))
        if spec.self_pairs is not None:
            for i in range(spec.n):
                for j in range(i + 1, spec.n):
                    self.L.or_set_self_collision(self.h, r, i, j, 0)
            for i, j in spec.self_pairs:
                self.L.or_set_self_collision(self.h, r, int(min(i, j)), int(max(i, j)), 1)
        self._robots.append(spec)
        return r

    def init_default(self):
        self.L.or_init_default(self.h)

    def num_ids(self):
        return self.L.or_num_ids(self.h)

    def set_collision_enabled(self, a, b, en):
        self.L.or_set_collision_enabled(self.h, a, b, 1 if en else 0)

    def set_robot_config(self, r, q):
        q = _dd(q)
        self.L.or_set_robot_config(self.h, r, _p(q, _d))

    def fk(self, r, q):
        q = _dd(q)
        n = self._robots[r].n
        out = np.zeros((n, 12))
        self.L.or_fk(self.h, r, _p(q, _d), _p(out, _d))
        return out

    def feasible_batch(self, r, Q, brute=False, nthreads=None):
        Q = _dd(Q)
        N = Q.shape[0]
        feas = np.zeros(N, np.uint8)
        env = np.zeros(N, np.uint8)
        slf = np.zeros(N, np.uint8)
        fp = np.zeros((N, 2), np.int32)
        self.L.or_feasible_batch(self.h, r, _p(Q, _d), N, _p(feas, _u8), _p(env, _u8), _p(slf, _u8),
                                 _p(fp, _i), 1 if brute else 0, nthreads or nthreads_default())
        return feas, env, slf, fp

    def isfeasible_batch(self, r, Q, nthreads=None):
        """IsFeasible with Klampt's early returns (the timed CPU path)."""
        Q = _dd(Q)
        feas = np.zeros(Q.shape[0], np.uint8)
        self.L.or_isfeasible_batch(self.h, r, _p(Q, _d), Q.shape[0], _p(feas, _u8), nthreads or nthreads_default())
        return feas

    def pair_list(self, r):
        out = np.zeros((4096, 2), np.int32)
        n = self.L.or_pair_list(self.h, r, _p(out, _i), 4096)
        return out[:n].copy()

    def distance_batch(self, r, Q, brute=False, signed=False, nthreads=None):
        Q = _dd(Q)
        np_ = len(self.pair_list(r))
        out = np.zeros((Q.shape[0], np_))
        self.L.or_distance_batch(self.h, r, _p(Q, _d), Q.shape[0], _p(out, _d), 1 if brute else 0,
                                 1 if signed else 0, nthreads or nthreads_default())
        return out

    def edge_batch(self, r, A, B, eps=0.01, norm_inf=True, nthreads=None):
        A, B = _dd(A), _dd(B)
        ns = A.shape[0]
        fu = np.zeros(ns)
        nsamp = np.zeros(ns, np.int32)
        nchk = np.zeros(ns, np.int32)
        self.L.or_edge_batch(self.h, r, _p(A, _d), _p(B, _d), ns, eps, 1 if norm_inf else 0, _p(fu, _d),
                             _p(nsamp, _i), _p(nchk, _i), nthreads or nthreads_default())
        return fu, nsamp, nchk

    def mesh_collide(self, ga, xa, gb, xb, brute=False):
        xa, xb = _dd(xa), _dd(xb)
        return bool(self.L.or_mesh_collide(self.h, ga, _p(xa, _d), gb, _p(xb, _d), 1 if brute else 0))

    def mesh_distance(self, ga, xa, gb, xb, brute=False):
        xa, xb = _dd(xa), _dd(xb)
        cp1, cp2 = np.zeros(3), np.zeros(3)
        el = np.zeros(2, np.int32)
        d = self.L.or_mesh_distance(self.h, ga, _p(xa, _d), gb, _p(xb, _d), 1 if brute else 0,
                                    _p(cp1, _d), _p(cp2, _d), _p(el, _i))
        return d, cp1, cp2, el

    def mesh_separation(self, ga, xa, gb, xb):
        xa, xb = _dd(xa), _dd(xb)
        return self.L.or_mesh_separation(self.h, ga, _p(xa, _d), gb, _p(xb, _d))

    def world_aabb(self, g, x):
        x = _dd(x)
        lo, hi = np.zeros(3), np.zeros(3)
        self.L.or_world_aabb(self.h, g, _p(x, _d), _p(lo, _d), _p(hi, _d))
        return lo, hi

    def sdf_penalty(self, r, grid, Q, pts, pts_per_link, dsafe=0.05, grad=True, fd_h=0.0, nthreads=None):
        vals = _dd(grid.values)
        dims = _ii(grid.values.shape)
        bmin, bmax = _dd(grid.bmin), _dd(grid.bmax)
        Q, P, ppl = _dd(Q), _dd(pts), _ii(pts_per_link)
        N, n = Q.shape
        loss = np.zeros(N)
        g = np.zeros((N, n)) if grad else None
        self.L.or_sdf_penalty(self.h, r, _p(vals, _d), _p(dims, _i), _p(bmin, _d), _p(bmax, _d), _p(Q, _d), N,
                              _p(P, _d), _p(ppl, _i), dsafe, _p(loss, _d), _p(g, _d), fd_h,
                              nthreads or nthreads_default())
        return loss, g


def tritri(a, b):
    a, b = _dd(a), _dd(b)
    return bool(lib().or_tritri(_p(a, _d), _p(b, _d)))


def tritri_dist(a, b):
    a, b = _dd(a), _dd(b)
    return lib().or_tritri_dist(_p(a, _d), _p(b, _d))


def tritri_depth(a, b):
    a, b = _dd(a), _dd(b)
    return lib().or_tritri_depth(_p(a, _d), _p(b, _d))


# Planted policy mistakes (test hook, see or_set_mutation in kg_oracle.cpp).
MUT_NONE, MUT_STABLE_ERASE, MUT_LINKS_FIRST, MUT_NO_REJECT, MUT_ENV_MAJOR = range(5)


class mutation:
    """Context manager: run the oracle with one planted policy mistake, so a
    test can show that its known answers detect it.  Single-threaded use."""

    def __init__(self, m):
        self.m = int(m)

    def __enter__(self):
        lib().or_set_mutation(self.m)
        return self

    def __exit__(self, *exc):
        lib().or_set_mutation(MUT_NONE)
        return False


def bisection_order(n):
    out = np.zeros(max(0, n - 1), np.int32)
    if n >= 2:
        lib().or_bisection_order(n, _p(out, _i))
    return out


def sdf_eval(grid, pts):
    vals = _dd(grid.values)
    dims = _ii(grid.values.shape)
    bmin, bmax = _dd(grid.bmin), _dd(grid.bmax)
    P = _dd(pts)
    d = np.zeros(P.shape[0])
    g = np.zeros((P.shape[0], 3))
    lib().or_sdf_eval(_p(vals, _d), _p(dims, _i), _p(bmin, _d), _p(bmax, _d), _p(P, _d), P.shape[0],
                      _p(d, _d), _p(g, _d))
    return d, g


def mesh_to_sdf(mesh, bmin, h, dims, brute=False, with_winding=False, nthreads=None):
    """Signed distance grid of a triangle mesh on the given cell-centred grid
    (exact FP64 distances, sign from the winding number along +z; the
    definition is at or_mesh_to_sdf in kg_oracle.cpp)."""
    V = _dd(mesh.verts)
    T = _ii(mesh.tris)
    dims = _ii(dims)
    out = np.zeros(tuple(int(x) for x in dims))
    wind = np.zeros(out.shape, np.int32) if with_winding else None
    lib().or_mesh_to_sdf(_p(V, _d), V.shape[0], _p(T, _i), T.shape[0], _p(_dd(bmin), _d), float(h), _p(dims, _i),
                         _p(out, _d), _p(wind, _i) if wind is not None else None, 1 if brute else 0,
                         nthreads or nthreads_default())
    return (out, wind) if with_winding else out


def world_from_scene(scene):
    """OracleWorld for a ``klampt_amd.synthetic.Scene``: terrains, then rigid
    objects, then the robot (Klampt ID order), with the default mask."""
    w = OracleWorld()
    for t in scene.terrains:
        w.add_terrain(t)
    for mesh, xf in scene.objects:
        w.add_object(mesh, xf)
    r = w.add_robot(scene.robot)
    w.init_default()
    return w, r


def hull_distance(A, xa, B, xb, xa2=None):
    """GJK distance between the convex hull of points A placed by xa (and,
    when xa2 is given, also by xa2: the hull of both copies) and the hull of
    points B placed by xb.  Returns (d, cp_a, cp_b, iterations)."""
    A, B = _dd(A).reshape(-1, 3), _dd(B).reshape(-1, 3)
    X = _dd(np.r_[xa, xa2] if xa2 is not None else xa)
    xb = _dd(xb)
    ca, cb = np.zeros(3), np.zeros(3)
    it = np.zeros(1, np.int32)
    d = lib().or_hull_distance(_p(A, _d), A.shape[0], _p(X, _d), 2 if xa2 is not None else 1, _p(B, _d),
                               B.shape[0], _p(xb, _d), _p(ca, _d), _p(cb, _d), _p(it, _i))
    return d, ca, cb, int(it[0])


def hull_support(A, x, d):
    """Support point of the hull of points A placed by x in direction d."""
    A = _dd(A).reshape(-1, 3)
    out = np.zeros(3)
    lib().or_hull_support(_p(A, _d), A.shape[0], _p(_dd(x), _d), _p(_dd(d), _d), _p(out, _d))
    return out


def sweep_terms(ow, r, link_points, l, q0, q1, B, xb, dcheck, fd_h=1e-5):
    """KineTrajOpt's sweep terms for one (segment, link, obstacle), restating
    find_collision_pair / add_sweep (Python/klampt/plan/kinetrajopt/
    kinetrajopt.py:556-583,603-611) in FP64: the GJK distance between the hull
    of the link's points at both configurations and the obstacle's hull B
    (placed by xb), kept below dcheck; the normal from the obstacle to the
    swept hull; the support points p0, p1 of the link's hull at q0 and at q1
    in -normal; alpha = |p1 - pswept| / (|p0 - pswept| + |p1 - pswept|); and
    normal . J at p0 (q0) and at p1 (q1), with J the position Jacobian of the
    link point by central differences of the oracle's FK (step fd_h).
    Returns None when d >= dcheck, (d,) when the hulls overlap, else
    (d, alpha, normal, pswept, pobs, p0, p1, coef0, coef1)."""
    q0, q1 = _dd(q0), _dd(q1)
    X0, X1 = ow.fk(r, q0)[l], ow.fk(r, q1)[l]
    d, ps, po, _ = hull_distance(link_points, X0, B, xb, xa2=X1)
    if not d < dcheck:
        return None
    if not d > 0:
        return (d,)
    nrm = (ps - po) / d
    p0 = hull_support(link_points, X0, -nrm)
    p1 = hull_support(link_points, X1, -nrm)
    d0, d1 = np.linalg.norm(p0 - ps), np.linalg.norm(p1 - ps)
    alpha = d1 / (d0 + d1) if d0 + d1 > 0 else 0.5

    return (d, alpha, nrm, ps, po, p0, p1, normal_row(ow, r, l, q0, p0, nrm, fd_h),
            normal_row(ow, r, l, q1, p1, nrm, fd_h))


def normal_row(ow, r, l, q, p, nrm, fd_h=1e-5):
    """nrm . J, J the position Jacobian at configuration q of the point of
    link l that is at world position p (robotsim positionJacobian), by
    central differences of the oracle's FK with step fd_h."""
    q = _dd(q)
    X = ow.fk(r, q)[l]
    R, t = X[:9].reshape(3, 3).T, X[9:]
    pl = R.T @ (np.asarray(p, np.float64) - t)
    out = np.zeros(len(q))
    for j in range(len(q)):
        qp, qm = q.copy(), q.copy()
        qp[j] += fd_h
        qm[j] -= fd_h
        Xp, Xm = ow.fk(r, qp)[l], ow.fk(r, qm)[l]
        a = Xp[:9].reshape(3, 3).T @ pl + Xp[9:]
        b = Xm[:9].reshape(3, 3).T @ pl + Xm[9:]
        out[j] = np.asarray(nrm) @ ((a - b) / (2 * fd_h))
    return out
