"""Klampt-shaped Python API over the C ABI.

Mirrors the names, argument meaning and error behaviour of the Klampt SWIG
module for the collision / distance / feasibility path
(``Python/klampt/src/geometry.h``, ``robotmodel.h``; behaviour in
``Python/klampt/src/robotsim.cpp``):

* ``Geometry3D.collides`` returns False for empty geometry (robotsim.cpp:1660);
* ``Geometry3D.distance`` raises on empty geometry (robotsim.cpp:1786-1787)
  and on unsupported pairs (the ``Inf`` -> raise rule, robotsim.cpp:1793-1795);
* ``RobotModel.setConfig`` raises on a wrong length (robotsim.cpp:5314-5316).

Rotations are Klampt's column-major 9-lists (``Python/klampt/math/so3.py``).
"""
from __future__ import annotations

import math
from typing import List, Optional, Sequence

import numpy as np

from . import engine
from ._lib import KlamptGPUError

_default_ctx = None


def default_context():
    """The process-wide GPU context (device 0) used by this module."""
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = engine.Context(0)
    return _default_ctx


class DistanceQuerySettings:
    """``DistanceQuerySettings`` (geometry.h:631-638)."""

    def __init__(self):
        self.relErr = 0.0
        self.absErr = 0.0
        self.upperBound = math.inf


class DistanceQueryResult:
    """``DistanceQueryResult`` (geometry.h:663-673)."""

    def __init__(self, r):
        self.d = r.d
        self.hasClosestPoints = bool(r.hasClosestPoints)
        self.hasGradients = bool(r.hasGradients)
        self.cp1 = list(r.cp1) if self.hasClosestPoints else []
        self.cp2 = list(r.cp2) if self.hasClosestPoints else []
        self.grad1 = list(r.grad1) if self.hasGradients else []
        self.grad2 = list(r.grad2) if self.hasGradients else []
        self.elem1 = r.elem1
        self.elem2 = r.elem2


class TriangleMesh:
    """Vertices (V,3) and triangle indices (T,3)."""

    def __init__(self, vertices=None, indices=None):
        self.vertices = np.zeros((0, 3)) if vertices is None else np.asarray(vertices, np.float64).reshape(-1, 3)
        self.indices = np.zeros((0, 3), np.int32) if indices is None else np.asarray(indices, np.int32).reshape(-1, 3)


class ImplicitSurface:
    """Cell-centred signed distance grid (geometry.h:322-391): values has shape
    dims, cell (i,j,k) covers [bmin + (i,j,k) h, bmin + (i+1,j+1,k+1) h)."""

    def __init__(self, bmin=(0, 0, 0), bmax=(1, 1, 1), values=None):
        self.bmin = np.asarray(bmin, np.float64)
        self.bmax = np.asarray(bmax, np.float64)
        self.values = np.zeros((1, 1, 1), np.float32) if values is None else np.asarray(values, np.float32)


class Geometry3D:
    """A TriangleMesh or ImplicitSurface with a current transform."""

    def __init__(self, data=None, ctx=None):
        self._ctx = ctx or default_context()
        self._g = None
        self._type = ""
        if data is not None:
            self.set(data)

    def set(self, data):
        if isinstance(data, TriangleMesh):
            self._g = engine.Geom.trimesh(self._ctx, data.vertices, data.indices)
            self._type = "TriangleMesh"
            self._mesh = data
        elif isinstance(data, ImplicitSurface):
            self._g = engine.Geom.sdf(self._ctx, data.bmin, data.bmax, data.values)
            self._type = "ImplicitSurface"
        else:
            raise ValueError("Geometry3D supports TriangleMesh and ImplicitSurface data")

    def type(self) -> str:
        return self._type

    def empty(self) -> bool:
        return self._g is None or (self._type == "TriangleMesh" and self._mesh.indices.shape[0] == 0)

    def setCurrentTransform(self, R: Sequence[float], t: Sequence[float]):
        self._g.set_transform(np.asarray(R, np.float64), np.asarray(t, np.float64))

    def getCurrentTransform(self):
        R, t = self._g.get_transform()
        return list(R), list(t)

    def setCollisionMargin(self, margin: float):
        """Fattens the geometry for collision testing (robotsim.cpp:1561-1566)."""
        self._g.set_margin(margin)

    def getCollisionMargin(self) -> float:
        return self._g.get_margin()

    def getBB(self):
        lo, hi = self._g.get_bb()
        return list(lo), list(hi)

    def collides(self, other: "Geometry3D") -> bool:
        if self.empty() or other.empty():
            return False
        return self._g.collides(other._g)

    def withinDistance(self, other: "Geometry3D", tol: float) -> bool:
        if self.empty() or other.empty():
            return False
        return self._g.within_distance(other._g, tol)

    def distance(self, other: "Geometry3D") -> DistanceQueryResult:
        return self.distance_ext(other, DistanceQuerySettings())

    def distance_ext(self, other: "Geometry3D", settings: DistanceQuerySettings) -> DistanceQueryResult:
        if self.empty():
            raise RuntimeError("Geometry3D.distance: Geometry is empty")
        if other.empty():
            raise RuntimeError("Geometry3D.distance: Other geometry is empty")
        try:
            r = self._g.distance(other._g, settings.relErr, settings.absErr, settings.upperBound)
        except KlamptGPUError as e:
            raise RuntimeError(str(e)) from None
        return DistanceQueryResult(r)

    def distance_simple(self, other: "Geometry3D", relErr=0.0, absErr=0.0) -> float:
        return self.distance(other).d

    def distance_point(self, pt) -> DistanceQueryResult:
        return self.distance_point_ext(pt, DistanceQuerySettings())

    def distance_point_ext(self, pt, settings: DistanceQuerySettings) -> DistanceQueryResult:
        if self.empty():
            raise RuntimeError("Geometry3D.distance_point: Geometry is empty")
        try:
            r = self._g.distance_point(pt, settings.relErr, settings.absErr, settings.upperBound)
        except KlamptGPUError as e:
            raise RuntimeError(str(e)) from None
        return DistanceQueryResult(r)


class RobotModelLink:
    def __init__(self, robot: "RobotModel", index: int):
        self._robot, self.index = robot, index

    def getParent(self) -> int:
        return int(self._robot._spec_parents[self.index])

    def getIndex(self) -> int:
        return self.index

    def geometry(self) -> Geometry3D:
        return self._robot._geoms[self.index]

    def setGeometry(self, geom: Optional[Geometry3D]):
        self._robot._geoms[self.index] = geom
        self._robot._r.set_link_geom(self.index, None if geom is None or geom.empty() else geom._g)

    def getTransform(self):
        R, t = self._robot._r.link_transform(self.index)
        return list(R), list(t)


class RobotModel:
    """Kinematic tree + link geometries (``RobotModel``, robotmodel.h:537)."""

    def __init__(self, parents, T0, axes, jtype, qmin, qmax, ctx=None):
        self._ctx = ctx or default_context()
        self._spec_parents = np.asarray(parents, np.int32)
        self._r = engine.Robot(self._ctx, parents, T0, axes, jtype, qmin, qmax)
        self._qmin, self._qmax = list(map(float, qmin)), list(map(float, qmax))
        self._geoms: List[Optional[Geometry3D]] = [Geometry3D(ctx=self._ctx) for _ in range(len(parents))]

    @classmethod
    def from_spec(cls, spec, ctx=None):
        rob = cls(spec.parents, spec.T0, spec.axes, spec.jtype, spec.qmin, spec.qmax, ctx)
        for l, m in enumerate(spec.link_meshes):
            if m is not None:
                rob.link(l).setGeometry(Geometry3D(TriangleMesh(m.verts, m.tris), ctx=rob._ctx))
        return rob

    def numLinks(self) -> int:
        return self._r.n

    def link(self, i: int) -> RobotModelLink:
        if not 0 <= i < self._r.n:
            raise IndexError("invalid link index")
        return RobotModelLink(self, i)

    def getJointLimits(self):
        return list(self._qmin), list(self._qmax)

    def setConfig(self, q: Sequence[float]):
        if len(q) != self._r.n:
            raise RuntimeError("Invalid size of configuration")
        self._r.set_config(np.asarray(q, np.float64))

    def getConfig(self) -> List[float]:
        return list(self._r.get_config())

    def selfCollides(self) -> bool:
        return self._r.self_collides()

    def enableSelfCollision(self, link1: int, link2: int, value: bool):
        self._r.enable_self_collision(link1, link2, value)

    def selfCollisionEnabled(self, link1: int, link2: int) -> bool:
        return self._r.self_collision_enabled(link1, link2)


class WorldModel:
    """Terrains, rigid objects and robots with Klampt's ID layout
    (``Cpp/Modeling/World.cpp:47-53,156-180``)."""

    def __init__(self, ctx=None):
        self._ctx = ctx or default_context()
        self._w = engine.World(self._ctx)
        self._terrains, self._objects, self._robots = [], [], []

    def addTerrain(self, geom: Geometry3D) -> int:
        self._terrains.append(geom)
        return self._w.add_terrain(geom._g)

    def addRigidObject(self, geom: Geometry3D, R=None, t=None) -> int:
        R = [1, 0, 0, 0, 1, 0, 0, 0, 1] if R is None else R
        t = [0, 0, 0] if t is None else t
        self._objects.append(geom)
        return self._w.add_rigid_object(geom._g, np.r_[np.asarray(R, float), np.asarray(t, float)])

    def addRobot(self, robot: RobotModel) -> int:
        self._robots.append(robot)
        self._w.robots.append(robot._r)
        from ._lib import call
        import ctypes
        out = ctypes.c_int32()
        call("kg_world_add_robot", self._w.h, robot._r.h, ctypes.byref(out))
        return out.value

    def numRobots(self) -> int:
        return len(self._robots)

    def robot(self, i: int) -> RobotModel:
        return self._robots[i]

    def numIDs(self) -> int:
        return self._w.num_ids()


class RobotCSpace:
    """Robot C-space feasibility and visibility
    (``Python/klampt/plan/robotcspace.py:31-118``; C++
    ``SingleRobotCSpace::IsFeasible`` / ``PathChecker``,
    ``Cpp/Planning/RobotCSpace.cpp:786-837``).  ``eps`` is the edge
    resolution (``collisionEpsilon`` default 0.01,
    ``Cpp/Planning/PlannerSettings.cpp:86-87``)."""

    def __init__(self, robot: RobotModel, world: WorldModel, eps: float = 0.01, norm_inf: bool = False):
        if robot not in world._robots:
            raise ValueError("robot is not in the world")
        self.robot, self.world = robot, world
        self.index = world._robots.index(robot)
        self.eps, self.norm_inf = eps, norm_inf

    def isFeasible(self, q) -> bool:
        return bool(self.world._w.feasible_batch(np.asarray([q], np.float32), robot=self.index)[0][0])

    def isVisible(self, a, b) -> bool:
        u = self.world._w.edge_batch(np.asarray([a], np.float32), np.asarray([b], np.float32), self.eps,
                                     self.norm_inf, robot=self.index)
        return bool(u[0] < 0)

    # batched forms
    def feasible_batch(self, Q, flags=False, first_pair=False):
        return self.world._w.feasible_batch(Q, robot=self.index, flags=flags, first_pair=first_pair)

    def edge_batch(self, A, B):
        return self.world._w.edge_batch(A, B, self.eps, self.norm_inf, robot=self.index)

    def distance_batch(self, Q):
        return self.world._w.distance_batch(Q, robot=self.index)
