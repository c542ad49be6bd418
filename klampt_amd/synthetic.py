"""Seeded synthetic workloads for the collision / distance / feasibility path.

Every coordinate produced here is rounded to the nearest float32 value and kept
as float64, so the FP32 device copies and the FP64 oracle see the *same*
geometry.  Conventions follow Klampt:

* rotations are column-major 3x3 (``Python/klampt/math/so3.py:1-16``);
* a link transform is ``T_World[i] = T_World[parent] * T0_Parent[i] * J(w_i, q_i)``
  (``Cpp/docs/Manual-Modeling.md:84-117``, ``Cpp/Modeling/Robot.cpp:957-981``);
* joint type 0 is revolute, 1 is prismatic (KrisLibrary ``RobotLink3D::Type``).

Workloads (sizes as named by their callers):

* ``arm7``: 7-DOF serial revolute arm, link offsets 0.3 m along alternating z/x,
  joint axes alternating z/x, limits [-pi, pi], 512-triangle capsule per link
  (radius 0.05 m).
* ``boxes_scene``: 16 rotated boxes (12 triangles each), sides U[0.1, 0.4] m,
  centres uniform in a 2 m cube -> rigid objects.
* ``clutter_soup``: ~200k-triangle soup of rotated boxes and ellipsoids with
  triangle edges of 1-3 cm -> one terrain.
* ``hetero_robots``: robots with 6..30 DOF (tree-branched above 12 DOF) and
  link meshes of 200..5000 triangles.
* ``sdf_scene``: 1M-point noisy planes/spheres cloud -> 256^3 cell-centred SDF.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np


def f32r(x):
    """Round to float32-representable values, returned as float64."""
    return np.asarray(x, dtype=np.float64).astype(np.float32).astype(np.float64)


# ---------------------------------------------------------------- rotations

def rot_colmajor(R):
    """3x3 numpy matrix -> Klampt 9-list (column-major)."""
    return np.asarray(R, dtype=np.float64).T.reshape(9)


def random_rotation(rng):
    """Uniform random rotation (quaternion method)."""
    u1, u2, u3 = rng.random(3)
    q = np.array([
        math.sqrt(1 - u1) * math.sin(2 * math.pi * u2),
        math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
        math.sqrt(u1) * math.sin(2 * math.pi * u3),
        math.sqrt(u1) * math.cos(2 * math.pi * u3),
    ])
    x, y, z, w = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


# ---------------------------------------------------------------- meshes

@dataclass
class Mesh:
    verts: np.ndarray   # (V,3) float64, float32-representable
    tris: np.ndarray    # (T,3) int32
    # identity of the file data this mesh came from (None: built in memory);
    # meshes with equal keys may share one device copy
    key: Optional[tuple] = field(default=None, compare=False, repr=False)

    @property
    def ntris(self):
        return int(self.tris.shape[0])


def capsule_mesh(radius=0.05, length=0.3, axis=2, m=16, k=8) -> Mesh:
    """Capsule around the segment [0, length] on ``axis`` with 4*m*k triangles
    (m azimuthal segments, k rings per hemisphere; 16 x 8 -> 512)."""
    verts = []
    # local frame: capsule along +z, then permuted to the requested axis
    top_pole = (0.0, 0.0, length + radius)
    bot_pole = (0.0, 0.0, -radius)
    verts.append(top_pole)
    verts.append(bot_pole)
    ring_index = {}
    for hemi in (0, 1):  # 0 top, 1 bottom
        for j in range(1, k + 1):
            phi = j * (math.pi / 2) / k
            rr = radius * math.sin(phi)
            dz = radius * math.cos(phi)
            zc = (length + dz) if hemi == 0 else (-dz)
            for s in range(m):
                th = 2 * math.pi * s / m
                ring_index[(hemi, j, s)] = len(verts)
                verts.append((rr * math.cos(th), rr * math.sin(th), zc))
    tris = []
    for hemi in (0, 1):
        pole = 0 if hemi == 0 else 1
        for s in range(m):
            a = ring_index[(hemi, 1, s)]
            b = ring_index[(hemi, 1, (s + 1) % m)]
            tris.append((pole, a, b) if hemi == 0 else (pole, b, a))
        for j in range(1, k):
            for s in range(m):
                a = ring_index[(hemi, j, s)]
                b = ring_index[(hemi, j, (s + 1) % m)]
                c = ring_index[(hemi, j + 1, s)]
                d = ring_index[(hemi, j + 1, (s + 1) % m)]
                if hemi == 0:
                    tris.append((a, c, d)); tris.append((a, d, b))
                else:
                    tris.append((a, d, c)); tris.append((a, b, d))
    for s in range(m):  # cylinder between the two equators
        a = ring_index[(0, k, s)]
        b = ring_index[(0, k, (s + 1) % m)]
        c = ring_index[(1, k, s)]
        d = ring_index[(1, k, (s + 1) % m)]
        tris.append((a, c, d)); tris.append((a, d, b))
    V = np.array(verts, dtype=np.float64)
    perm = {2: [0, 1, 2], 0: [2, 0, 1], 1: [1, 2, 0]}[axis]
    V2 = np.zeros_like(V)
    for src, dst in enumerate(perm):
        V2[:, dst] = V[:, src]
    return Mesh(f32r(V2), np.array(tris, dtype=np.int32))


def box_mesh(half, R=None, c=None, edge=None) -> Mesh:
    """Box with half extents ``half``; faces subdivided so edges are <= ``edge``
    (12 triangles when ``edge`` is None)."""
    half = np.asarray(half, dtype=np.float64)
    R = np.eye(3) if R is None else np.asarray(R)
    c = np.zeros(3) if c is None else np.asarray(c, dtype=np.float64)
    verts, tris = [], []
    for axis in range(3):
        for sign in (-1.0, 1.0):
            u_ax, v_ax = [a for a in range(3) if a != axis]
            nu = 1 if edge is None else max(1, int(math.ceil(2 * half[u_ax] / edge)))
            nv = 1 if edge is None else max(1, int(math.ceil(2 * half[v_ax] / edge)))
            base = len(verts)
            for iu in range(nu + 1):
                for iv in range(nv + 1):
                    p = np.zeros(3)
                    p[axis] = sign * half[axis]
                    p[u_ax] = -half[u_ax] + 2 * half[u_ax] * iu / nu
                    p[v_ax] = -half[v_ax] + 2 * half[v_ax] * iv / nv
                    verts.append(p)
            # outward orientation: (u x v) has sign +/- along axis
            flip = (sign > 0) != (((axis + 1) % 3) == u_ax)
            for iu in range(nu):
                for iv in range(nv):
                    a = base + iu * (nv + 1) + iv
                    b = a + 1
                    cc = a + (nv + 1)
                    d = cc + 1
                    if flip:
                        tris.append((a, d, cc)); tris.append((a, b, d))
                    else:
                        tris.append((a, cc, d)); tris.append((a, d, b))
    V = np.array(verts) @ R.T + c
    return Mesh(f32r(V), np.array(tris, dtype=np.int32))


def ellipsoid_mesh(semi, R=None, c=None, edge=0.02) -> Mesh:
    """UV ellipsoid with triangle edges of roughly ``edge``."""
    semi = np.asarray(semi, dtype=np.float64)
    R = np.eye(3) if R is None else np.asarray(R)
    c = np.zeros(3) if c is None else np.asarray(c, dtype=np.float64)
    amax = float(semi.max())
    m = max(8, int(round(2 * math.pi * amax / edge)))
    k = max(4, int(round(math.pi * amax / edge)))
    verts = [(0.0, 0.0, 1.0), (0.0, 0.0, -1.0)]
    for j in range(1, k):
        phi = math.pi * j / k
        for s in range(m):
            th = 2 * math.pi * s / m
            verts.append((math.sin(phi) * math.cos(th), math.sin(phi) * math.sin(th), math.cos(phi)))
    tris = []
    ring = lambda j, s: 2 + (j - 1) * m + (s % m)
    for s in range(m):
        tris.append((0, ring(1, s), ring(1, s + 1)))
        tris.append((1, ring(k - 1, s + 1), ring(k - 1, s)))
    for j in range(1, k - 1):
        for s in range(m):
            a, b = ring(j, s), ring(j, s + 1)
            cc, d = ring(j + 1, s), ring(j + 1, s + 1)
            tris.append((a, cc, d)); tris.append((a, d, b))
    V = np.array(verts) * semi
    V = V @ R.T + c
    return Mesh(f32r(V), np.array(tris, dtype=np.int32))


def merge_meshes(meshes: List[Mesh]) -> Mesh:
    Vs, Ts, off = [], [], 0
    for m in meshes:
        Vs.append(m.verts)
        Ts.append(m.tris + off)
        off += m.verts.shape[0]
    return Mesh(np.concatenate(Vs), np.concatenate(Ts).astype(np.int32))


# ---------------------------------------------------------------- robots

REVOLUTE, PRISMATIC = 0, 1


@dataclass
class RobotSpec:
    name: str
    parents: np.ndarray          # (n,) int32, -1 = attached to world
    T0: np.ndarray               # (n,12) float64: column-major R (9) then t (3)
    axes: np.ndarray             # (n,3) float64, unit
    jtype: np.ndarray            # (n,) int32
    qmin: np.ndarray             # (n,) float64
    qmax: np.ndarray             # (n,) float64
    link_meshes: List[Optional[Mesh]] = field(default_factory=list)
    self_pairs: Optional[np.ndarray] = None   # (k,2) int32 or None = Klampt default

    @property
    def n(self):
        return int(self.parents.shape[0])

    def default_self_pairs(self):
        """All pairs except link vs parent (``Cpp/docs/Manual-FileTypes.md:212``)."""
        out = []
        for i in range(self.n):
            for j in range(i + 1, self.n):
                if self.parents[j] == i or self.parents[i] == j:
                    continue
                out.append((i, j))
        return np.array(out, dtype=np.int32).reshape(-1, 2)


def _xf(R=None, t=(0, 0, 0)):
    R = np.eye(3) if R is None else np.asarray(R)
    return np.concatenate([rot_colmajor(R), np.asarray(t, dtype=np.float64)])


def arm7(link_offset=0.3, radius=0.05, tris_per_link=512) -> RobotSpec:
    """7-DOF serial revolute arm: offsets 0.3 m along z (even links) / x (odd
    links), axes alternating z / x, capsule meshes along the next link offset."""
    n = 7
    parents = np.array([-1, 0, 1, 2, 3, 4, 5], dtype=np.int32)
    T0, axes, meshes = [], [], []
    dirv = lambda k: 2 if k % 2 == 0 else 0
    m = 16
    k = tris_per_link // (4 * m)
    for i in range(n):
        t = np.zeros(3)
        t[dirv(i)] = link_offset
        T0.append(_xf(None, t))
        ax = np.zeros(3)
        ax[2 if i % 2 == 0 else 0] = 1.0
        axes.append(ax)
        meshes.append(capsule_mesh(radius, link_offset, axis=dirv(i + 1), m=m, k=k))
    return RobotSpec(
        name="arm7",
        parents=parents,
        T0=f32r(np.array(T0)),
        axes=np.array(axes),
        jtype=np.zeros(n, dtype=np.int32),
        qmin=np.full(n, -math.pi),
        qmax=np.full(n, math.pi),
        link_meshes=meshes,
    )


def uniform_configs(robot: RobotSpec, n, seed=0):
    """Configurations uniform in the joint limits, as float32."""
    rng = np.random.default_rng(seed)
    lo = robot.qmin.astype(np.float32)
    hi = robot.qmax.astype(np.float32)
    Q = rng.random((n, robot.n), dtype=np.float32) * (hi - lo) + lo
    return np.clip(Q, lo, hi).astype(np.float32)


# ---------------------------------------------------------------- scenes

@dataclass
class Scene:
    robot: RobotSpec
    terrains: List[Mesh] = field(default_factory=list)
    objects: List[tuple] = field(default_factory=list)   # (Mesh, xf12)


def boxes_scene(seed=0, nboxes=16) -> Scene:
    """7-DOF arm + 16 rotated boxes (rigid objects) in the 2 m cube [-1,1]^3."""
    rng = np.random.default_rng(seed)
    robot = arm7()
    sc = Scene(robot)
    for _ in range(nboxes):
        half = rng.uniform(0.1, 0.4, 3) / 2
        R = random_rotation(rng)
        c = rng.uniform(-1, 1, 3)
        mesh = box_mesh(half)                      # local frame
        sc.objects.append((mesh, f32r(_xf(R, c))))
    return sc


def clutter_soup(seed=0, target_tris=200_000, center=(1.3, 0.0, 0.3), side=2.0,
                 edge_range=(0.01, 0.03)) -> Mesh:
    """Soup of randomly rotated boxes and ellipsoids with triangle edges in
    ``edge_range``, centres uniform in a cube of ``side`` metres around
    ``center``, until at least ``target_tris`` triangles."""
    rng = np.random.default_rng(seed)
    meshes, total = [], 0
    center = np.asarray(center, dtype=np.float64)
    while total < target_tris:
        e = rng.uniform(*edge_range)
        R = random_rotation(rng)
        c = center + rng.uniform(-side / 2, side / 2, 3)
        if rng.random() < 0.5:
            half = rng.uniform(0.02, 0.10, 3)
            mesh = box_mesh(half, R, c, edge=e)
        else:
            semi = rng.uniform(0.02, 0.10, 3)
            mesh = ellipsoid_mesh(semi, R, c, edge=e)
        meshes.append(mesh)
        total += mesh.ntris
    return merge_meshes(meshes)


def soup_scene(seed=0, target_tris=200_000) -> Scene:
    """7-DOF arm + ~200k-triangle clutter soup loaded as one terrain."""
    sc = Scene(arm7())
    sc.terrains.append(clutter_soup(seed, target_tris))
    return sc


# ---------------------------------------------------------------- edges

def edge_segments(robot: RobotSpec, nseg, eps=0.01, smin=10, smax=500, seed=0):
    """Segments with L-infinity length giving between smin and smax samples at
    step ``eps`` (endpoints re-drawn until the count falls in range)."""
    rng = np.random.default_rng(seed)
    lo = robot.qmin.astype(np.float32)
    hi = robot.qmax.astype(np.float32)
    A = np.empty((nseg, robot.n), np.float32)
    B = np.empty((nseg, robot.n), np.float32)
    i = 0
    while i < nseg:
        a = (rng.random(robot.n, dtype=np.float32) * (hi - lo) + lo).astype(np.float32)
        # choose a target sample count log-uniformly, then a direction
        target = math.exp(rng.uniform(math.log(smin), math.log(smax)))
        d = rng.normal(size=robot.n)
        d /= np.abs(d).max()
        b = np.clip(a + d * target * eps, lo, hi).astype(np.float32)
        n = math.ceil(float(np.abs(b.astype(np.float64) - a.astype(np.float64)).max()) / eps)
        if smin <= n - 1 <= smax:
            A[i], B[i] = a, b
            i += 1
    return A, B


# ---------------------------------------------------------------- heterogeneous robots

def hetero_robots(seed=0, nrobots=64, dof_range=(6, 30), tri_range=(200, 5000),
                  cfg_range=(100, 10000)):
    """Robots with uneven DOF, branching above 12 DOF, link meshes of uneven
    triangle counts, and log-uniform configuration counts."""
    rng = np.random.default_rng(seed)
    robots, counts = [], []
    for r in range(nrobots):
        n = int(rng.integers(dof_range[0], dof_range[1] + 1))
        parents = np.empty(n, np.int32)
        T0, axes, meshes = [], [], []
        offset = rng.uniform(0.08, 0.2)
        for i in range(n):
            if i == 0:
                parents[i] = -1
            elif n > 12 and i >= 3:
                parents[i] = int(rng.integers(max(0, i - 3), i))
            else:
                parents[i] = i - 1
            t = np.zeros(3)
            t[2 if i % 2 == 0 else 0] = offset if i > 0 else 0.1
            T0.append(_xf(None, t))
            ax = np.zeros(3)
            ax[2 if i % 2 == 0 else 0] = 1.0
            axes.append(ax)
        for i in range(n):
            ntri = int(rng.integers(tri_range[0], tri_range[1] + 1))
            m = max(8, int(round(math.sqrt(ntri / 4))))
            k = max(1, int(round(ntri / (4 * m))))
            # capsule along the direction of the first child offset
            nxt = 2 if (i + 1) % 2 == 0 else 0
            meshes.append(capsule_mesh(0.03, offset, axis=nxt, m=m, k=k))
        base = np.array([rng.uniform(-1.0, 0.2), rng.uniform(-1.0, 1.0), 0.0])
        T0[0] = _xf(None, base + np.array([0, 0, 0.1]))
        rob = RobotSpec(
            name=f"hetero{r}", parents=parents, T0=f32r(np.array(T0)), axes=np.array(axes),
            jtype=np.zeros(n, np.int32), qmin=np.full(n, -math.pi), qmax=np.full(n, math.pi),
            link_meshes=meshes)
        robots.append(rob)
        counts.append(int(round(math.exp(rng.uniform(math.log(cfg_range[0]), math.log(cfg_range[1]))))))
    return robots, counts


# ---------------------------------------------------------------- SDF

@dataclass
class SDFGrid:
    bmin: np.ndarray     # (3,)
    bmax: np.ndarray     # (3,)
    values: np.ndarray   # (m,n,p) float32, C-order, cell-centred


def noisy_cloud(seed=0, npoints=1_000_000, sigma=0.002):
    """Points on a floor plane, two walls and spheres with Gaussian noise,
    plus an analytic inside test for signing the distance field."""
    rng = np.random.default_rng(seed)
    spheres = [(np.array([0.9, 0.3, 0.6]), 0.25), (np.array([0.5, -0.7, 0.9]), 0.18),
               (np.array([-0.6, 0.6, 0.4]), 0.3), (np.array([0.2, 0.9, 1.2]), 0.12)]
    planes = [(np.array([0, 0, 1.0]), 0.0), (np.array([1.0, 0, 0]), -1.2),
              (np.array([0, -1.0, 0]), -1.2)]
    per = npoints // (len(spheres) + len(planes))
    pts = []
    for c, r in spheres:
        d = rng.normal(size=(per, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        pts.append(c + r * d)
    for n, off in planes:
        uv = rng.uniform(-1.25, 1.25, (per, 2))
        # build a basis of the plane n.x = off
        a = np.array([1.0, 0, 0]) if abs(n[0]) < 0.9 else np.array([0, 1.0, 0])
        u = np.cross(n, a); u /= np.linalg.norm(u)
        v = np.cross(n, u)
        base = n * off
        p = base + uv[:, :1] * u + uv[:, 1:] * v
        p[:, 2] = np.clip(p[:, 2], -0.2, 2.2) if n[2] == 0 else p[:, 2]
        pts.append(p)
    P = np.concatenate(pts)
    P = P + rng.normal(scale=sigma, size=P.shape)
    rest = npoints - P.shape[0]
    if rest > 0:
        P = np.concatenate([P, P[:rest] + rng.normal(scale=sigma, size=(rest, 3))])

    def inside(X):
        ins = np.zeros(X.shape[0], bool)
        for c, r in spheres:
            ins |= np.linalg.norm(X - c, axis=1) < r
        for n, off in planes:
            ins |= (X @ n) < off
        return ins
    return P.astype(np.float32), inside


def sdf_from_cloud(points, inside, dims=256, voxel=0.01, bmin=(-1.28, -1.28, -0.3)):
    """Cell-centred signed distance grid: Euclidean distance transform to the
    occupied cells, negative where ``inside`` holds at the cell centre."""
    from scipy import ndimage
    bmin = np.asarray(bmin, dtype=np.float64)
    bmax = bmin + dims * voxel
    idx = np.floor((points.astype(np.float64) - bmin) / voxel).astype(np.int64)
    ok = np.all((idx >= 0) & (idx < dims), axis=1)
    occ = np.zeros((dims, dims, dims), bool)
    occ[idx[ok, 0], idx[ok, 1], idx[ok, 2]] = True
    dist = ndimage.distance_transform_edt(~occ).astype(np.float64) * voxel
    g = (np.arange(dims) + 0.5) * voxel
    vals = np.empty((dims, dims, dims), np.float32)
    for i in range(dims):  # slab-wise signing keeps memory bounded
        X = np.stack(np.meshgrid(bmin[0] + g[i], bmin[1] + g, bmin[2] + g, indexing="ij"), -1).reshape(-1, 3)
        s = np.where(inside(X), -1.0, 1.0).reshape(dims, dims)
        vals[i] = (dist[i] * s).astype(np.float32)
    return SDFGrid(bmin, bmax, vals)


def surface_samples(mesh: Mesh, n=64, seed=0):
    """Area-weighted surface samples of a mesh in its local frame (the
    sampling rule of ``Python/klampt/model/geometry.py:1064-1176``)."""
    rng = np.random.default_rng(seed)
    V, T = mesh.verts, mesh.tris
    e1 = V[T[:, 1]] - V[T[:, 0]]
    e2 = V[T[:, 2]] - V[T[:, 0]]
    area = np.linalg.norm(np.cross(e1, e2), axis=1)
    pick = rng.choice(len(T), n, p=area / area.sum())
    uv = rng.random((n, 2))
    flip = uv.sum(1) > 1
    uv[flip] = 1 - uv[flip]
    P = V[T[pick, 0]] + uv[:, :1] * e1[pick] + uv[:, 1:] * e2[pick]
    return f32r(P)


def trajectories(robot: RobotSpec, ntraj, nway, seed=0):
    """Straight joint-space trajectories (``ntraj`` x ``nway`` waypoints)."""
    rng = np.random.default_rng(seed)
    lo, hi = robot.qmin, robot.qmax
    a = rng.uniform(lo, hi, (ntraj, 1, robot.n))
    b = rng.uniform(lo, hi, (ntraj, 1, robot.n))
    s = np.linspace(0, 1, nway).reshape(1, nway, 1)
    Q = a * (1 - s) + b * s
    return Q.reshape(ntraj * nway, robot.n).astype(np.float32)


# ---------------------------------------------------------------- writers

def write_off(path, mesh: Mesh):
    with open(path, "w") as f:
        f.write("OFF\n%d %d 0\n" % (mesh.verts.shape[0], mesh.tris.shape[0]))
        for v in mesh.verts:
            f.write("%.9g %.9g %.9g\n" % tuple(v))
        for t in mesh.tris:
            f.write("3 %d %d %d\n" % tuple(t))


def write_rob(path, robot: RobotSpec, geom_files):
    """Klampt .rob file (``Cpp/docs/Manual-FileTypes.md:163-236``)."""
    def tp(x):
        R = x[:9].reshape(3, 3).T  # back to row-major matrix
        cols = [R[:, 0], R[:, 1], R[:, 2], x[9:12]]
        return "  ".join(" ".join("%.9g" % v for v in c) for c in cols)
    with open(path, "w") as f:
        f.write("links " + " ".join('"link%d"' % i for i in range(robot.n)) + "\n")
        f.write("parents " + " ".join(str(int(p)) for p in robot.parents) + "\n")
        f.write("TParent " + " \\\n  ".join(tp(x) for x in robot.T0) + "\n")
        f.write("axis " + "  ".join(" ".join("%g" % v for v in a) for a in robot.axes) + "\n")
        f.write("jointtype " + " ".join("r" if t == 0 else "p" for t in robot.jtype) + "\n")
        f.write("qmin " + " ".join("%.17g" % v for v in robot.qmin) + "\n")
        f.write("qmax " + " ".join("%.17g" % v for v in robot.qmax) + "\n")
        f.write("geometry " + " ".join('"%s"' % os.path.basename(g) for g in geom_files) + "\n")
        f.write("joint normal " + "\njoint normal ".join(str(i) for i in range(robot.n)) + "\n")
