"""On-disk formats that feed the collision engine: triangle meshes (.off,
Wavefront .obj, .stl), Klampt robot files (.rob), URDF, Klampt rigid-object
(.obj) and terrain (.env) files, and world XML.

Every loader returns plain host data (``synthetic.Mesh`` / ``RobotSpec`` /
``Scene``-compatible records, float64 throughout) with each file-level
geometry transform (``geomscale``, ``geomtransform``, URDF collision origins
and mesh scales, world-XML ``scale`` / rotations) baked into the vertices, as
Klampt's ``ManagedGeometry::TransformGeometry`` does.  ``klampt_amd.robotsim``
turns these records into device objects; meshes that come from the same file
with the same baked transform share one device BVH there (``kg_geom_clone``),
the caching ``ManagedGeometry::Load`` does (``Cpp/Modeling/ManagedGeometry.cpp:111-164``).

Behaviour followed, by file:

* ``.rob``: ``RobotModel::LoadRob`` (``Cpp/Modeling/Robot.cpp:216-1383``):
  line syntax (``#`` comments, backslash-newline continuation, lower-cased
  keys, quoted strings), the link-count vote and size checks, defaults
  (parents i-1, axis z, limits +-inf, q 0), ``*deg`` keys (prismatic limits
  left unconverted), the base ``translation`` / ``rotation`` applied to
  root links, ``geomscale`` / ``geommargin`` (one value or one per link),
  ``geomtransform`` and the ``selfcollision`` / ``noselfcollision`` rules
  (default: every pair except link-parent).  The file syntax is
  ``Cpp/docs/Manual-FileTypes.md:163-236``; transforms are row-major
  rotation then translation, as ``RobotModel::Save`` writes them
  (``Robot.cpp:1503-1512``).
* URDF: ``RobotModel::LoadURDF`` and ``URDFConverter``
  (``Robot.cpp:2566-3400``, ``Cpp/IO/URDFConverter.cpp:44-320``,
  ``Cpp/IO/urdf_model.h:144-222``, ``Cpp/IO/urdf_joint.cpp:101-470``):
  depth-first link order with children in joint-name order, fixed base when
  the root link is named ``world`` (else a 6-link floating base), joint
  origins as parent transforms, first collision geometry per link.
* world XML: ``XmlWorld`` (``Cpp/IO/XmlWorld.cpp:96-370,615-700``) and the
  rigid-object / terrain files (``Cpp/Modeling/RigidObject.cpp:35-238``,
  ``Cpp/Modeling/Terrain.cpp:27-72``).

Not supported, each raising ``LoadError`` with the reason: D-H parameters
(their convention lives in KrisLibrary, outside the reference), ``mount``,
whole-robot ``scale`` (the reference aborts on it too), URDF floating or
planar joints below the root, non-tree link orders that put a parent after
its child (the device kinematics need parents first), robots over 64 links.
URDF box / cylinder / sphere primitives become closed triangle meshes (box
exact, cylinder and sphere tessellated): their surfaces collide as the
primitive's do, but a body wholly inside a primitive is not reported.
"""
from __future__ import annotations

import hashlib
import math
import os
import re
import struct
import xml.etree.ElementTree as ET
from collections import OrderedDict
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import synthetic as S


class LoadError(IOError):
    """A file that cannot be read, or uses a feature this loader rejects."""


def _ext(path):
    return os.path.splitext(path)[1][1:].lower()


def _resolve(base_dir, fn):
    """``ResolveFileReference``: relative names are relative to the file
    that names them."""
    if os.path.isabs(fn) or not base_dir:
        return fn
    return os.path.join(base_dir, fn)


def _file_key(path):
    st = os.stat(path)
    return (os.path.realpath(path), st.st_mtime_ns, st.st_size)


# ---------------------------------------------------------------- transforms

def rot_x(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[1, 0, 0], [0, c, -s], [0, s, c]], np.float64)


def rot_y(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]], np.float64)


def rot_z(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]], np.float64)


def rot_rpy(roll, pitch, yaw):
    """Fixed-axis roll-pitch-yaw: Rz(yaw) Ry(pitch) Rx(roll) (URDF ``rpy``;
    world-XML ``rotateRPY``, XmlWorld.cpp:109-118)."""
    return rot_z(yaw) @ rot_y(pitch) @ rot_x(roll)


def rot_moment(w):
    """Rotation of the exponential-map vector w (Rodrigues' formula)."""
    w = np.asarray(w, np.float64)
    th = float(np.linalg.norm(w))
    if th == 0.0:
        return np.eye(3)
    k = w / th
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) + math.sin(th) * K + (1 - math.cos(th)) * (K @ K)


def xf12(R, t):
    """(3x3 matrix, translation) -> the 12-value transform the engine takes
    (column-major rotation, then translation)."""
    return np.concatenate([S.rot_colmajor(R), np.asarray(t, np.float64).reshape(3)])


def xf_mat(x):
    """12-value transform -> (3x3 matrix, translation)."""
    x = np.asarray(x, np.float64)
    return x[:9].reshape(3, 3).T.copy(), x[9:12].copy()


def transform_mesh(mesh: S.Mesh, M) -> S.Mesh:
    """Apply a 3x4 / 4x4 affine map to the vertices (``TransformGeometry``)."""
    M = np.asarray(M, np.float64)
    V = mesh.verts @ M[:3, :3].T + M[:3, 3]
    return S.Mesh(V, mesh.tris.copy())


# ---------------------------------------------------------------- meshes

def _fan(poly, nv, where):
    if len(poly) < 3:
        raise LoadError("%s: face with fewer than 3 vertices" % where)
    for i in poly:
        if not 0 <= i < nv:
            raise LoadError("%s: vertex index %d out of range (%d vertices)" % (where, i, nv))
    return [(poly[0], poly[k], poly[k + 1]) for k in range(1, len(poly) - 1)]


def _read_off(path):
    """Object File Format: optional ``OFF`` header (``COFF`` / ``NOFF`` /
    ``CNOFF`` too: only the first 3 values of a vertex line are used), counts,
    vertex lines, then faces ``n i0 .. i(n-1)`` triangulated as fans."""
    with open(path, "r", errors="replace") as f:
        text = f.read()
    if "#" not in text:
        # plain "OFF" with triangles only (the common case): one token pass
        m = re.match(r"\s*OFF\s+(\d+)\s+(\d+)\s+(\d+)\s", text)
        if m:
            nv, nf = int(m.group(1)), int(m.group(2))
            body = np.fromstring(text[m.end():], dtype=np.float64, sep=" ")
            if body.size == 3 * nv + 4 * nf:
                V = body[:3 * nv].reshape(-1, 3)
                F = body[3 * nv:].reshape(-1, 4)
                if (F[:, 0] == 3).all() and np.array_equal(F, np.floor(F)):
                    T = F[:, 1:].astype(np.int64)
                    if T.size and (T.min() < 0 or T.max() >= nv):
                        raise LoadError("%s: vertex index out of range (%d vertices)" % (path, nv))
                    return S.Mesh(V.copy(), T.astype(np.int32))
    lines = []
    for ln in text.splitlines():
        ln = ln.split("#", 1)[0].strip()
        if ln:
            lines.append(ln)
    if not lines:
        raise LoadError("%s: empty OFF file" % path)
    it = iter(lines)
    first = next(it).split()
    if first[0].upper().endswith("OFF"):
        hdr = first[1:] or next(it, "").split()
    else:
        hdr = first
    try:
        nv, nf = int(hdr[0]), int(hdr[1])
        vlines = [next(it) for _ in range(nv)]
        flines = [next(it) for _ in range(nf)]
        vt = np.fromstring(" ".join(vlines), dtype=np.float64, sep=" ") if nv else np.zeros(0)
        if vt.size == 3 * nv:   # plain "x y z" lines (the common case): one pass
            V = vt.reshape(-1, 3)
        else:
            V = np.array([[float(x) for x in ln.split()[:3]] for ln in vlines], np.float64).reshape(-1, 3)
    except (StopIteration, ValueError, IndexError) as e:
        raise LoadError("%s: malformed OFF file (%s)" % (path, e or "truncated")) from None
    # all-triangle face blocks (the common case) parse in one pass
    tok = " ".join(flines).split()
    F = None
    if len(tok) == 4 * nf:
        try:
            F = np.array(tok, np.int64).reshape(-1, 4) if nf else np.zeros((0, 4), np.int64)
        except ValueError:
            F = None   # not all integers: the per-face parse below reports it
    if F is not None:
        if (F[:, 0] == 3).all():
            T = F[:, 1:]
            if T.size and (T.min() < 0 or T.max() >= nv):
                raise LoadError("%s: vertex index out of range (%d vertices)" % (path, nv))
            return S.Mesh(V, T.astype(np.int32))
    tris = []
    try:
        for fi, ln in enumerate(flines):
            tok = ln.split()
            k = int(tok[0])
            tris += _fan([int(x) for x in tok[1:1 + k]], nv, "%s face %d" % (path, fi))
    except (ValueError, IndexError) as e:
        raise LoadError("%s: malformed OFF face (%s)" % (path, e)) from None
    return S.Mesh(V, np.array(tris, np.int32).reshape(-1, 3))


def _read_wavefront(path):
    """Wavefront OBJ: ``v x y z [w]`` and ``f a b c ..`` (``a``, ``a/t``,
    ``a//n``, ``a/t/n``; 1-based, negative = relative to the end), polygons
    triangulated as fans; every other statement is ignored."""
    V, tris = [], []
    with open(path, "r", errors="replace") as f:
        for lineno, ln in enumerate(f, 1):
            tok = ln.split("#", 1)[0].split()
            if not tok:
                continue
            try:
                if tok[0] == "v":
                    V.append([float(x) for x in tok[1:4]])
                elif tok[0] == "f":
                    poly = []
                    for t in tok[1:]:
                        i = int(t.split("/")[0])
                        poly.append(i - 1 if i > 0 else len(V) + i)
                    tris += _fan(poly, len(V), "%s:%d" % (path, lineno))
            except ValueError:
                raise LoadError("%s:%d: malformed statement" % (path, lineno)) from None
    Va = np.array(V, np.float64).reshape(-1, 3)
    if Va.shape[0] != len(V):
        raise LoadError("%s: vertex with fewer than 3 coordinates" % path)
    return S.Mesh(Va, np.array(tris, np.int32).reshape(-1, 3))


def _weld(P):
    """Triangle soup (3T x 3) -> shared vertices: identical coordinates become
    one vertex (the triangles themselves are unchanged)."""
    V, inv = np.unique(P, axis=0, return_inverse=True)
    return S.Mesh(V.astype(np.float64), inv.reshape(-1, 3).astype(np.int32))


def _read_stl(path):
    """STL, binary (80-byte header, uint32 count, 50-byte facets) or ASCII
    (``facet`` / ``vertex`` lines)."""
    size = os.path.getsize(path)
    with open(path, "rb") as f:
        data = f.read()
    if size >= 84:
        n = struct.unpack_from("<I", data, 80)[0]
        if size == 84 + 50 * n:
            rec = np.frombuffer(data, dtype=np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")]),
                                count=n, offset=84)
            return _weld(rec["v"].reshape(-1, 3).astype(np.float64))
    text = data.decode("ascii", errors="replace")
    if not text.lstrip().lower().startswith("solid"):
        raise LoadError("%s: neither binary nor ASCII STL" % path)
    P = [[float(x) for x in m.groups()] for m in
         re.finditer(r"vertex\s+(\S+)\s+(\S+)\s+(\S+)", text)]
    if len(P) % 3:
        raise LoadError("%s: facet without 3 vertices" % path)
    return _weld(np.array(P, np.float64).reshape(-1, 3))


_MESH_CACHE: "OrderedDict[tuple, S.Mesh]" = OrderedDict()
_MESH_CACHE_MAX = 64   # parsed files kept (least recently used dropped)


def load_mesh(path) -> S.Mesh:
    """Triangle mesh from .off, Wavefront .obj or .stl.  The returned mesh
    carries ``key`` (file identity), so loads of one file share device data;
    parsed files are cached by path, modification time and size."""
    if not os.path.isfile(path):
        raise LoadError("%s: no such file" % path)
    key = _file_key(path)
    hit = _MESH_CACHE.get(key)
    if hit is not None:
        _MESH_CACHE.move_to_end(key)
    else:
        ext = _ext(path)
        if ext == "off":
            hit = _read_off(path)
        elif ext == "obj":
            hit = _read_wavefront(path)
        elif ext == "stl":
            hit = _read_stl(path)
        else:
            raise LoadError("%s: unsupported mesh format .%s (supported: .off, .obj, .stl)" % (path, ext))
        _MESH_CACHE[key] = hit
        while len(_MESH_CACHE) > _MESH_CACHE_MAX:
            _MESH_CACHE.popitem(last=False)
    return S.Mesh(hit.verts.copy(), hit.tris.copy(), key=key)


def keyed_transform(mesh: S.Mesh, M) -> S.Mesh:
    """``transform_mesh`` that keeps a cache key: the source key plus the
    bytes of the map, so equal files under equal maps still share."""
    M = np.asarray(M, np.float64)
    out = transform_mesh(mesh, M)
    if mesh.key is not None:
        out.key = mesh.key + (hashlib.sha1(np.ascontiguousarray(M[:3, :4]).tobytes()).hexdigest(),)
    return out


def save_mesh(path, mesh: S.Mesh):
    """Write .off or Wavefront .obj (17 significant digits: exact round trip)
    or binary .stl (float32 coordinates)."""
    ext = _ext(path)
    V, T = np.asarray(mesh.verts, np.float64), np.asarray(mesh.tris, np.int64)
    if ext == "off":
        with open(path, "w") as f:
            f.write("OFF\n%d %d 0\n" % (V.shape[0], T.shape[0]))
            f.writelines("%.17g %.17g %.17g\n" % tuple(v) for v in V)
            f.writelines("3 %d %d %d\n" % tuple(t) for t in T)
    elif ext == "obj":
        with open(path, "w") as f:
            f.writelines("v %.17g %.17g %.17g\n" % tuple(v) for v in V)
            f.writelines("f %d %d %d\n" % tuple(t + 1) for t in T)
    elif ext == "stl":
        rec = np.zeros(T.shape[0], dtype=np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")]))
        P = V[T]
        nrm = np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0])
        ln = np.linalg.norm(nrm, axis=1, keepdims=True)
        rec["n"] = np.where(ln > 0, nrm / np.maximum(ln, 1e-300), 0)
        rec["v"] = P
        with open(path, "wb") as f:
            f.write(b"klampt_amd binary STL".ljust(80, b" "))
            f.write(struct.pack("<I", T.shape[0]))
            f.write(rec.tobytes())
    else:
        raise LoadError("%s: unsupported mesh format .%s" % (path, ext))


def mesh_to_points(mesh: S.Mesh, dispersion: float = 0.0) -> np.ndarray:
    """TriangleMesh -> PointCloud points (``Geometry3D.convert("PointCloud",
    dispersion)``, Python/klampt/src/geometry.h:959-961): every vertex, plus,
    on each triangle whose longest edge exceeds the dispersion, the points
    a + (i/k)(b-a) + (j/k)(c-a), i + j <= k, with k = ceil(longest edge /
    dispersion), so every surface point lies within the dispersion of a
    returned point.  dispersion 0: the average longest triangle edge."""
    V = np.asarray(mesh.verts, np.float64)
    T = np.asarray(mesh.tris, np.int64)
    if T.shape[0] == 0:
        return V.copy()
    P = V[T]
    L = np.max(np.linalg.norm(P - np.roll(P, -1, axis=1), axis=-1), axis=1)
    d = float(dispersion) if dispersion > 0 else float(L.mean())
    if not d > 0:
        return V.copy()
    out = [V]
    K = np.ceil(L / d).astype(np.int64)
    for k in np.unique(K[K > 1]):
        sel = np.nonzero(K == k)[0]
        ij = np.array([(i, j) for i in range(k + 1) for j in range(k + 1 - i)
                       if (i, j) not in ((0, 0), (k, 0), (0, k))], np.float64) / k
        A, B, C = P[sel, 0], P[sel, 1], P[sel, 2]
        pts = A[:, None] + ij[None, :, :1] * (B - A)[:, None] + ij[None, :, 1:] * (C - A)[:, None]
        out.append(pts.reshape(-1, 3))
    return np.concatenate(out)


# ---------------------------------------------------------------- primitives

def box_primitive(size) -> S.Mesh:
    """Axis-aligned box of the given side lengths centred at the origin."""
    return S.box_mesh(np.asarray(size, np.float64) / 2)


def cylinder_primitive(radius, length, m=32) -> S.Mesh:
    """Closed cylinder along z centred at the origin, m-gon cross-section
    inscribed in the circle."""
    a = 2 * np.pi * np.arange(m) / m
    ring = np.stack([radius * np.cos(a), radius * np.sin(a)], 1)
    V = np.concatenate([np.c_[ring, np.full(m, -length / 2)], np.c_[ring, np.full(m, length / 2)],
                        [[0, 0, -length / 2], [0, 0, length / 2]]])
    T = []
    for i in range(m):
        j = (i + 1) % m
        T += [(i, j, m + j), (i, m + j, m + i), (2 * m, j, i), (2 * m + 1, m + i, m + j)]
    return S.Mesh(V, np.array(T, np.int32))


def sphere_primitive(radius, m=24, k=12) -> S.Mesh:
    """UV sphere (m longitudes, k latitude bands) with vertices on the sphere."""
    V = [[0, 0, -radius]]
    for b in range(1, k):
        th = np.pi * b / k - np.pi / 2
        for i in range(m):
            ph = 2 * np.pi * i / m
            V.append([radius * math.cos(th) * math.cos(ph), radius * math.cos(th) * math.sin(ph),
                      radius * math.sin(th)])
    V.append([0, 0, radius])
    top = len(V) - 1
    T = []
    for i in range(m):
        T.append((0, 1 + (i + 1) % m, 1 + i))
    for b in range(k - 2):
        r0, r1 = 1 + b * m, 1 + (b + 1) * m
        for i in range(m):
            j = (i + 1) % m
            T += [(r0 + i, r0 + j, r1 + j), (r0 + i, r1 + j, r1 + i)]
    r = 1 + (k - 2) * m
    for i in range(m):
        T.append((r + i, r + (i + 1) % m, top))
    return S.Mesh(np.array(V, np.float64), np.array(T, np.int32))


# ---------------------------------------------------------------- robots

@dataclass
class RobotFile:
    """A loaded robot: the kinematic spec the engine takes (link meshes with
    their file transforms applied, explicit self-collision pairs) plus what
    the file says around it."""
    spec: S.RobotSpec
    name: str = "robot"
    link_names: List[str] = field(default_factory=list)
    q: Optional[np.ndarray] = None
    geom_files: List[str] = field(default_factory=list)   # resolved paths, "" = none
    margins: Optional[np.ndarray] = None                   # per-link collision margin
    joints: List[tuple] = field(default_factory=list)      # (type, link index[, base index])

    def link_index(self, name):
        """``RobotModel::LinkIndex`` (Robot.cpp:184-195): an integer string is
        an index, anything else a link name (-1 when unknown)."""
        s = str(name).strip()
        if re.fullmatch(r"[+-]?\d+", s):
            return int(s)
        return self.link_names.index(s) if s in self.link_names else -1


def _rob_lines(text):
    """Logical lines of a .rob file, following the reader loop of
    Robot.cpp:269-307: ``#`` starts a comment that runs to the end of the
    physical line; a backslash before a newline (or CR LF) joins the next
    line; before any other character both characters are kept."""
    out, cur, comment, i, n = [], [], False, 0, len(text)
    while i < n:
        c = text[i]
        i += 1
        if c == "\n":
            out.append("".join(cur))
            cur, comment = [], False
            continue
        if comment or c == "#":
            comment = True
            continue
        if c == "\\":
            c2 = text[i] if i < n else ""
            i += 1
            if c2 == "\r":
                c2 = text[i] if i < n else ""
                i += 1
            elif c2 != "\n":
                cur.append("\\")
            cur.append(c2)
            continue
        cur.append(c)
    if cur:
        out.append("".join(cur))
    return out


_TOKEN = re.compile(r'"((?:[^"\\]|\\.)*)"|(\S+)')


def _tokens(s):
    """Whitespace-separated items, double-quoted strings kept whole
    (``SafeInputString``)."""
    return [m.group(1) if m.group(1) is not None else m.group(2) for m in _TOKEN.finditer(s)]


def _floats(tok, where):
    out = []
    for t in tok:
        try:
            out.append(float(t))
        except ValueError:
            break  # the reference's stream read stops at the first non-number
    return out


def _groups(vals, k):
    m = len(vals) // k
    return [vals[k * i:k * i + k] for i in range(m)]


def _rigid_rowmajor(v):
    """12 values: row-major rotation, then translation -> (R, t)."""
    return np.array(v[:9], np.float64).reshape(3, 3), np.array(v[9:12], np.float64)


def _vote(sizes):
    """The link count: the most frequent nonzero size (ties: the smaller)."""
    cnt = {}
    for s in sizes:
        if s:
            cnt[s] = cnt.get(s, 0) + 1
    if not cnt:
        return 0
    best = max(cnt.values())
    return min(s for s, c in cnt.items() if c == best)


def _self_pairs(n, parents, link_index, on, off, where):
    """Self-collision pairs after the ``selfcollision`` / ``noselfcollision``
    rules of Robot.cpp:1275-1306 (all pairs except link-parent unless pairs
    are listed)."""
    def idx(a, b):
        i, j = link_index(a), link_index(b)
        if not (0 <= i < n and 0 <= j < n):
            raise LoadError("%s: invalid self-collision link pair %s, %s (range 0..%d)" % (where, a, b, n - 1))
        if i > j:
            i, j = j, i
        if i == j:
            raise LoadError("%s: invalid self-collision pair %s, %s" % (where, a, b))
        return i, j
    if on:
        enabled = set(idx(a, b) for a, b in on)
    else:
        enabled = set((i, j) for i in range(n) for j in range(i + 1, n)
                      if parents[j] != i and parents[i] != j)
    for a, b in off:
        enabled.discard(idx(a, b))
    return np.array(sorted(enabled), np.int32).reshape(-1, 2)


def _check_tree(parents, where):
    n = len(parents)
    if n < 1 or n > 64:
        raise LoadError("%s: %d links; the engine takes 1..64" % (where, n))
    for l, p in enumerate(parents):
        if not -1 <= p < l:
            raise LoadError("%s: link %d has parent %d; parents must precede their children" % (where, l, p))


_ROB_VECTOR_KEYS = {
    "qmin", "qmindeg", "qmax", "qmaxdeg", "q", "qdeg", "velmax", "velmaxdeg", "velmin", "velmindeg",
    "torquemax", "accmax", "accmaxdeg", "powermax", "servop", "servoi", "servod", "dryfriction",
    "viscousfriction", "alpha", "alphadeg", "a", "d", "theta", "thetadeg", "mass", "geomscale", "geommargin",
}


def load_rob(path) -> RobotFile:
    """Klampt .rob file (``RobotModel::LoadRob``, Robot.cpp:216-1383)."""
    if not os.path.isfile(path):
        raise LoadError("%s: no such file" % path)
    base = os.path.dirname(path)
    with open(path, "r", errors="replace") as f:
        text = f.read()
    v: Dict[str, list] = {k: [] for k in _ROB_VECTOR_KEYS}
    parents, names, jtypes, tparent, axes, com, inertia = [], [], [], [], [], [], []
    geom, selfc, noselfc, gxf, joints = [], [], [], [], []
    base_R, base_t = np.eye(3), np.zeros(3)
    for lineno, line in enumerate(_rob_lines(text), 1):
        tok = _tokens(line)
        if not tok:
            continue
        key, rest = tok[0].lower(), tok[1:]
        where = "%s: line %d" % (path, lineno)
        if key == "parents":
            parents += [int(float(x)) for x in rest]
        elif key == "links":
            names += rest
        elif key == "jointtype":
            for t in rest:
                if t not in ("r", "p"):
                    raise LoadError("%s: invalid joint type %s" % (where, t))
                jtypes.append(S.REVOLUTE if t == "r" else S.PRISMATIC)
        elif key in ("qmindeg", "qmaxdeg"):
            for i, x in enumerate(_floats(rest, where)):
                prism = i < len(jtypes) and jtypes[i] == S.PRISMATIC
                v[key].append(x if prism else math.radians(x))
        elif key in ("qdeg", "velmaxdeg", "velmindeg", "accmaxdeg", "alphadeg", "thetadeg"):
            v[key] += [math.radians(x) for x in _floats(rest, where)]
        elif key in _ROB_VECTOR_KEYS:
            v[key] += _floats(rest, where)
        elif key == "tparent":
            tparent += [_rigid_rowmajor(g) for g in _groups(_floats(rest, where), 12)]
        elif key == "axis":
            axes += _groups(_floats(rest, where), 3)
        elif key == "translation":
            base_t = np.array(_floats(rest, where)[:3], np.float64)
        elif key == "rotation":
            base_R = np.array(_floats(rest, where)[:9], np.float64).reshape(3, 3)
        elif key == "com":
            com += _groups(_floats(rest, where), 3)
        elif key == "inertiadiag":
            inertia += _groups(_floats(rest, where), 3)
        elif key == "inertia":
            inertia += _groups(_floats(rest, where), 9)
        elif key in ("automass", "autotorque", "drivers", "driver", "property", "nocollision"):
            if key == "nocollision" and rest:
                raise LoadError("%s: no mechanism to turn off environment collisions" % where)
        elif key == "collision":
            if rest:
                raise LoadError("%s: no mechanism to select environment collisions" % where)
        elif key == "geometry":
            geom += rest
        elif key == "scale":
            s = _floats(rest, where)
            if s and s[0] != 1.0:
                raise LoadError("%s: whole-robot scale is not supported" % where)
        elif key in ("selfcollision", "noselfcollision"):
            pairs = [(rest[2 * i], rest[2 * i + 1]) for i in range(len(rest) // 2)]
            (selfc if key == "selfcollision" else noselfc).extend(pairs)
        elif key == "geomtransform":
            vals = _floats(rest, where)
            if len(vals) < 17:
                raise LoadError("%s: invalid geomTransform" % where)
            M = np.array(vals[1:17], np.float64).reshape(4, 4)
            if M[3, 3] != 1.0:
                raise LoadError("%s: invalid geomTransform, not a homogeneous transform" % where)
            gxf.append((int(vals[0]), M))
        elif key == "joint":
            if not rest:
                continue
            jt = rest[0].lower()
            if jt not in ("weld", "normal", "spin", "floating", "ballandsocket", "floatingplanar"):
                raise LoadError("%s: invalid joint type %s" % (where, rest[0]))
            ids = [int(float(x)) for x in rest[1:3]]
            if jt in ("floating", "ballandsocket", "floatingplanar") and (not ids or ids[0] <= 0):
                raise LoadError("%s: %s joint needs a link index > 0" % (where, jt))
            joints.append((jt,) + tuple(ids))
        elif key == "mount":
            raise LoadError("%s: mount is not supported" % where)
        else:
            raise LoadError("%s: invalid robot property %s" % (where, tok[0]))

    sizes = {"parents": len(parents), "joint types": len(jtypes), "link names": len(names),
             "masses": len(v["mass"]), "COMs": len(com), "inertias": len(inertia),
             "D-H a": len(v["a"]), "D-H d": len(v["d"]), "D-H alpha": len(v["alpha"]) + len(v["alphadeg"]),
             "D-H theta": len(v["theta"]) + len(v["thetadeg"]), "link transforms": len(tparent),
             "axes": len(axes), "configuration variables": len(v["q"]) + len(v["qdeg"]),
             "joint limits (min)": len(v["qmin"]) + len(v["qmindeg"]),
             "joint limits (max)": len(v["qmax"]) + len(v["qmaxdeg"]),
             "velocity limits (min)": len(v["velmin"]) + len(v["velmindeg"]),
             "velocity limits (max)": len(v["velmax"]) + len(v["velmaxdeg"]),
             "torque limits": len(v["torquemax"]), "power limits": len(v["powermax"])}
    n = _vote(sizes.values())
    bad = [k for k, s in sizes.items() if s and s != n]
    if geom and len(geom) != n:
        bad.append("geometry files")
    for k in ("geomscale", "geommargin"):
        if v[k] and len(v[k]) not in (1, n):
            bad.append(k)
    if bad:
        raise LoadError("%s: wrong number of %s (the links vote for %d)" % (path, ", ".join(bad), n))
    if n == 0:
        raise LoadError("%s: no links" % path)
    if not tparent:
        if v["a"] and v["d"] and (v["alpha"] or v["alphadeg"]) and (v["theta"] or v["thetadeg"]):
            raise LoadError("%s: D-H parameters are not supported; give tparent" % path)
        raise LoadError("%s: no D-H parameters or link transforms specified" % path)
    if not parents:
        parents = [i - 1 for i in range(n)]
    _check_tree(parents, path)
    if not names:
        names = ["Link_%d" % i for i in range(n)]
    q = np.array(v["q"] + v["qdeg"], np.float64) if (v["q"] or v["qdeg"]) else np.zeros(n)
    qmin = np.array(v["qmin"] + v["qmindeg"], np.float64) if (v["qmin"] or v["qmindeg"]) else np.full(n, -np.inf)
    qmax = np.array(v["qmax"] + v["qmaxdeg"], np.float64) if (v["qmax"] or v["qmaxdeg"]) else np.full(n, np.inf)
    ax = np.array(axes, np.float64).reshape(-1, 3) if axes else np.tile([0.0, 0.0, 1.0], (n, 1))
    T0 = np.zeros((n, 12))
    for i, (R, t) in enumerate(tparent):
        if parents[i] < 0:
            R, t = base_R @ R, base_R @ t + base_t
        T0[i] = xf12(R, t)
    jt = np.array(jtypes, np.int32) if jtypes else np.zeros(n, np.int32)

    gscale = v["geomscale"] * n if len(v["geomscale"]) == 1 else v["geomscale"]
    gmargin = v["geommargin"] * n if len(v["geommargin"]) == 1 else v["geommargin"]
    meshes: List[Optional[S.Mesh]] = [None] * n
    files = [""] * n
    for i, g in enumerate(geom):
        if not g:
            continue
        files[i] = _resolve(base, g)
        m = load_mesh(files[i])
        if gscale and gscale[i] != 1.0:
            m = keyed_transform(m, np.diag([gscale[i]] * 3 + [1.0]))
        meshes[i] = m
    for i, M in gxf:
        if not 0 <= i < n:
            raise LoadError("%s: geomtransform link %d out of range" % (path, i))
        if meshes[i] is not None:
            meshes[i] = keyed_transform(meshes[i], M)
    rf = RobotFile(S.RobotSpec(os.path.splitext(os.path.basename(path))[0], np.array(parents, np.int32), T0,
                               ax, jt, qmin, qmax, meshes),
                   name=os.path.splitext(os.path.basename(path))[0], link_names=list(names), q=q,
                   geom_files=files, margins=np.array(gmargin if gmargin else [0.0] * n, np.float64),
                   joints=joints)
    rf.spec.self_pairs = _self_pairs(n, parents, rf.link_index, selfc, noselfc, path)
    return rf


def save_rob(path, rf: RobotFile):
    """Write a .rob file of the items this module reads, in the layout
    ``RobotModel::Save`` writes (Robot.cpp:1468-1540): row-major TParent,
    quoted link names and geometry files (relative to the .rob file where
    possible), then the self-collision pairs when they differ from the
    default set."""
    sp = rf.spec
    n = sp.n
    base = os.path.dirname(os.path.abspath(path))

    def tp(x):
        R, t = xf_mat(x)
        return "  ".join(" ".join("%.17g" % e for e in row) for row in R) + "  " + " ".join("%.17g" % e for e in t)

    def rel(g):
        if not g:
            return ""
        try:
            return os.path.relpath(os.path.abspath(g), base)
        except ValueError:
            return os.path.abspath(g)
    names = rf.link_names or ["Link_%d" % i for i in range(n)]
    q = rf.q if rf.q is not None else np.zeros(n)
    with open(path, "w") as f:
        f.write("links\t" + " ".join('"%s"' % s for s in names) + "\n\n")
        f.write("parents\t" + " ".join(str(int(p)) for p in sp.parents) + "\n\n")
        f.write("axis\t" + "\t".join(" ".join("%.17g" % e for e in a) for a in sp.axes) + "\n\n")
        f.write("jointtype\t" + " ".join("p" if t == S.PRISMATIC else "r" for t in sp.jtype) + "\n\n")
        f.write("Tparent\t" + " \\\n\t".join(tp(x) for x in sp.T0) + "\n\n")
        f.write("q\t" + " ".join("%.17g" % e for e in q) + "\n\n")
        f.write("qmin\t" + " ".join("%.17g" % e for e in sp.qmin) + "\n\n")
        f.write("qmax\t" + " ".join("%.17g" % e for e in sp.qmax) + "\n\n")
        files = rf.geom_files or [""] * n
        f.write("geometry\t" + " ".join('"%s"' % rel(g) for g in files) + "\n\n")
        if rf.margins is not None and np.any(rf.margins != 0):
            f.write("geommargin\t" + " ".join("%.17g" % e for e in rf.margins) + "\n\n")
        if sp.self_pairs is not None:
            default = set(map(tuple, sp.default_self_pairs().tolist()))
            pairs = set(map(tuple, np.asarray(sp.self_pairs).reshape(-1, 2).tolist()))
            if pairs != default:
                f.write("selfcollision\t" + "  ".join("%d %d" % p for p in sorted(pairs)) + "\n\n")
        for j in rf.joints or [("normal", i) for i in range(n)]:
            f.write("joint " + " ".join(str(x) for x in j) + "\n")


# ---------------------------------------------------------------- URDF

def _vec(s, k=3, default=None):
    if s is None:
        return default
    vals = [float(x) for x in s.split()]
    if len(vals) != k:
        raise LoadError("expected %d numbers, got %r" % (k, s))
    return np.array(vals, np.float64)


def _pose(el):
    """``<origin xyz rpy>`` -> (R, t); identity when absent."""
    if el is None:
        return np.eye(3), np.zeros(3)
    t = _vec(el.get("xyz"), 3, np.zeros(3))
    r = _vec(el.get("rpy"), 3, np.zeros(3))
    return rot_rpy(*r), t


def _urdf_geom(el, base, package_root, flip_yz, where):
    """First ``<geometry>`` child -> mesh in the link frame, or None."""
    g = el.find("geometry") if el is not None else None
    if g is None or len(g) == 0:
        return None
    shape = g[0]
    R, t = _pose(el.find("origin"))
    M = np.eye(4)
    M[:3, :3], M[:3, 3] = R, t
    if shape.tag == "mesh":
        fn = shape.get("filename", "")
        if fn.startswith("package://"):
            fn = (package_root or "") + "/" + fn[10:]
        path = _resolve(base, fn)
        if not os.path.isfile(path) and os.path.isfile(fn):
            path = fn
        mesh = load_mesh(path)
        sc = _vec(shape.get("scale"), 3, np.ones(3))
        Sm = np.diag(list(sc) + [1.0])
        if flip_yz:
            Ryz = np.zeros((4, 4))
            Ryz[0, 0] = 1
            Ryz[1, 2] = -1
            Ryz[2, 1] = 1
            Ryz[3, 3] = 1
            Sm = Sm @ Ryz
        return keyed_transform(mesh, M @ Sm)
    if shape.tag == "box":
        mesh = box_primitive(_vec(shape.get("size"), 3))
    elif shape.tag == "cylinder":
        mesh = cylinder_primitive(float(shape.get("radius")), float(shape.get("length")))
    elif shape.tag == "sphere":
        mesh = sphere_primitive(float(shape.get("radius")))
    else:
        raise LoadError("%s: unknown URDF geometry <%s>" % (where, shape.tag))
    return transform_mesh(mesh, M)


def _pairs_attr(e):
    if e.get("pairs") is not None:
        tok = _tokens(e.get("pairs"))
        return [(tok[2 * i], tok[2 * i + 1]) for i in range(len(tok) // 2)]
    if e.get("group1") is not None and e.get("group2") is not None:
        return [(a, b) for a in _tokens(e.get("group1")) for b in _tokens(e.get("group2"))]
    raise LoadError("<klampt><%s> has neither pairs nor group1/group2" % e.tag)


def load_urdf(path) -> RobotFile:
    """URDF robot (``RobotModel::LoadURDF``, Robot.cpp:2566-3400)."""
    if not os.path.isfile(path):
        raise LoadError("%s: no such file" % path)
    base = os.path.dirname(path)
    try:
        root = ET.parse(path).getroot()
    except ET.ParseError as e:
        raise LoadError("%s: XML error: %s" % (path, e)) from None
    if root.tag != "robot":
        raise LoadError("%s: no <robot> element" % path)
    kl = root.find("klampt")
    world_frame, package_root, use_vis, flip_yz = "world", "", False, False
    selfc, noselfc = [], []
    if kl is not None:
        use_vis = bool(int(float(kl.get("use_vis_geom", "0"))))
        flip_yz = bool(int(float(kl.get("flip_yz", "0"))))
        world_frame = kl.get("world_frame", "world")
        if kl.get("package_root") is not None:
            pr = kl.get("package_root")
            package_root = pr if (pr.startswith("/") or not base) else base + "/" + pr
        for e in kl.findall("selfcollision"):
            selfc += _pairs_attr(e)
        for e in kl.findall("noselfcollision"):
            noselfc += _pairs_attr(e)
        if kl.find("mount") is not None:
            raise LoadError("%s: <klampt><mount> is not supported" % path)
    links = {}
    for e in root.findall("link"):
        if e.get("name") is None:
            raise LoadError("%s: link without a name" % path)
        links[e.get("name")] = e
    joints = {}
    for e in root.findall("joint"):
        name = e.get("name")
        jt = e.get("type")
        if jt not in ("revolute", "continuous", "prismatic", "fixed", "floating", "planar"):
            raise LoadError("%s: joint %s has no known type %s" % (path, name, jt))
        pe, ce = e.find("parent"), e.find("child")
        if pe is None or ce is None or not pe.get("link") or not ce.get("link"):
            raise LoadError("%s: joint %s is missing a parent and/or child link" % (path, name))
        if pe.get("link") not in links or ce.get("link") not in links:
            raise LoadError("%s: joint %s names an undefined link" % (path, name))
        lim = e.find("limit")
        if lim is None and jt in ("revolute", "prismatic"):
            raise LoadError("%s: %s joint %s does not specify limits" % (path, jt, name))
        ax = None
        if jt not in ("fixed", "floating"):
            ae = e.find("axis")
            ax = np.array([1.0, 0.0, 0.0]) if ae is None else _vec(ae.get("xyz"), 3, np.zeros(3))
        joints[name] = dict(type=jt, parent=pe.get("link"), child=ce.get("link"), origin=_pose(e.find("origin")),
                            axis=ax, limits=None if lim is None else (float(lim.get("lower", 0)),
                                                                      float(lim.get("upper", 0))))
    # tree: children of each link in joint-name order (std::map iteration)
    children = {k: [] for k in links}
    parent_joint = {}
    for jname in sorted(joints):
        j = joints[jname]
        children[j["parent"]].append(j["child"])
        parent_joint[j["child"]] = jname
    roots = sorted(k for k in links if k not in parent_joint)
    if len(roots) != 1:
        raise LoadError("%s: %s" % (path, "no root link" if not roots else "two root links: %s" % roots[:2]))
    order, stack = [], [(roots[0], -1)]
    while stack:  # depth-first, children in list order
        name, par = stack.pop()
        idx = len(order)
        order.append((name, par))
        for c in reversed(children[name]):
            stack.append((c, idx))
    floating = roots[0] != world_frame
    off = 5 if floating else -1
    n = len(order) + off
    if not 1 <= n <= 64:
        raise LoadError("%s: %d links; the engine takes 1..64" % (path, n))
    names, parents = [""] * n, [0] * n
    T0, axes = np.zeros((n, 12)), np.zeros((n, 3))
    jtype, qmin, qmax = np.zeros(n, np.int32), np.full(n, -np.inf), np.full(n, np.inf)
    meshes: List[Optional[S.Mesh]] = [None] * n
    files = [""] * n
    joint_list = []
    if floating:
        for i in range(6):
            names[i], parents[i], T0[i] = "base%d" % i, i - 1, xf12(np.eye(3), np.zeros(3))
        jtype[:3] = S.PRISMATIC
        axes[:6] = [[1, 0, 0], [0, 1, 0], [0, 0, 1], [0, 0, 1], [0, 1, 0], [1, 0, 0]]
        joint_list.append(("floating", 5, -1))
    for k, (lname, par) in enumerate(order):
        if not floating and k == 0:
            continue
        # the floating base's last link (base5) is the root link itself
        li, pi = k + off, (par + off if par >= 0 else (4 if floating else -1))
        names[li], parents[li] = lname, pi
        if k > 0 or not floating:
            T0[li] = xf12(np.eye(3), np.zeros(3))
            axes[li] = [0, 0, 1]
        le = links[lname]
        jn = parent_joint.get(lname)
        if jn is not None:
            j = joints[jn]
            if j["type"] in ("floating", "planar"):
                raise LoadError("%s: %s joint %s below the root is not supported" % (path, j["type"], jn))
            R, t = j["origin"]
            T0[li] = xf12(R, t)
            if j["axis"] is not None and np.linalg.norm(j["axis"]) > 0:
                axes[li] = j["axis"]
            if j["type"] == "prismatic":
                jtype[li] = S.PRISMATIC
            if j["limits"] is not None:
                qmin[li], qmax[li] = j["limits"]
            if j["type"] == "continuous":
                qmin[li], qmax[li] = -np.inf, np.inf
            if j["type"] == "fixed":
                qmin[li] = qmax[li] = 0.0
            joint_list.append(({"fixed": "weld", "continuous": "spin"}.get(j["type"], "normal"), li))
        src = le.find("visual" if use_vis else "collision")
        mesh = _urdf_geom(src, base, package_root, flip_yz, "%s link %s" % (path, lname))
        if mesh is not None:
            meshes[li] = mesh
            shape = src.find("geometry")[0]
            files[li] = shape.get("filename", "") if shape.tag == "mesh" else "<%s>" % shape.tag
    _check_tree(parents, path)
    rf = RobotFile(S.RobotSpec(root.get("name", "robot"), np.array(parents, np.int32), T0, axes, jtype, qmin, qmax,
                               meshes),
                   name=os.path.splitext(os.path.basename(path))[0], link_names=names,
                   q=np.zeros(n), geom_files=files, margins=np.zeros(n), joints=joint_list)
    rf.spec.self_pairs = _self_pairs(n, parents, rf.link_index, selfc, noselfc, path)
    return rf


def load_robot(path) -> RobotFile:
    """.rob or .urdf by extension (``RobotModel::Load``, Robot.cpp:197-214)."""
    ext = _ext(path)
    if ext == "rob":
        return load_rob(path)
    if ext == "urdf":
        return load_urdf(path)
    raise LoadError("%s: robot files must have a .rob or .urdf extension" % path)


# ---------------------------------------------------------------- objects, terrains

@dataclass
class RigidObjectFile:
    mesh: S.Mesh
    xf: np.ndarray                # 12-value transform of the object
    name: str = "Object"
    margin: float = 0.0


@dataclass
class TerrainFile:
    mesh: S.Mesh
    name: str = "Terrain"
    margin: float = 0.0


def _simple_file(path):
    """KrisLibrary ``SimpleFile``: one item per line, ``name value ..``,
    values numbers or (quoted) strings, ``#`` comments."""
    items: Dict[str, list] = {}
    with open(path, "r", errors="replace") as f:
        for lineno, ln in enumerate(f, 1):
            tok = _tokens(ln.split("#", 1)[0])
            if not tok:
                continue
            vals = []
            for t in tok[1:]:
                try:
                    vals.append(float(t))
                except ValueError:
                    vals.append(t)
            items[tok[0]] = vals
    return items


def _nums(items, key, path, sizes):
    vals = items[key]
    if not all(isinstance(x, float) for x in vals) or (sizes and len(vals) not in sizes):
        raise LoadError("%s: invalid %s" % (path, key))
    return vals


def load_rigid_object(path) -> RigidObjectFile:
    """Rigid object: a Klampt object file for .obj (``mesh``, ``geomscale``,
    ``geomtranslate``, ``T``, ``position``, ...; RigidObject.cpp:35-222),
    else a mesh file with the identity transform."""
    name = os.path.splitext(os.path.basename(path))[0]
    if _ext(path) != "obj":
        return RigidObjectFile(load_mesh(path), xf12(np.eye(3), np.zeros(3)), name)
    if not os.path.isfile(path):
        raise LoadError("%s: no such file" % path)
    items = _simple_file(path)
    allowed = {"mesh", "geomscale", "geomtranslate", "T", "position", "velocity", "angularVelocity", "mass",
               "inertia", "com", "kFriction", "kRestitution", "kStiffness", "kDamping", "autoMass"}
    unknown = set(items) - allowed
    if unknown:
        raise LoadError("%s: not a Klampt object file (unknown items %s)" % (path, sorted(unknown)[:3]))
    if len(items.get("mesh", [])) != 1 or not isinstance(items["mesh"][0], str):
        raise LoadError("%s: object file needs one mesh file name" % path)
    mesh = load_mesh(_resolve(os.path.dirname(path), items["mesh"][0]))
    G = np.eye(4)
    if "geomscale" in items:
        s = _nums(items, "geomscale", path, (1, 3))
        G[0, 0], G[1, 1], G[2, 2] = (s * 3)[:3] if len(s) == 1 else s
    if "geomtranslate" in items:
        G[:3, 3] = _nums(items, "geomtranslate", path, (3,))
    if not np.array_equal(G, np.eye(4)):
        mesh = keyed_transform(mesh, G)
    R, t = np.eye(3), np.zeros(3)
    if "T" in items:
        T = _nums(items, "T", path, (12, 16))
        if len(T) == 12:
            R, t = _rigid_rowmajor(T)
        else:
            M = np.array(T).reshape(4, 4)
            R, t = M[:3, :3].copy(), M[:3, 3].copy()
    if "position" in items:
        t = np.array(_nums(items, "position", path, (3,)))
    return RigidObjectFile(mesh, xf12(R, t), name)


def load_terrain(path) -> TerrainFile:
    """Terrain: a .env file (``mesh`` item; Terrain.cpp:27-72) or a mesh file."""
    name = os.path.splitext(os.path.basename(path))[0]
    if _ext(path) != "env":
        return TerrainFile(load_mesh(path), name)
    if not os.path.isfile(path):
        raise LoadError("%s: no such file" % path)
    items = _simple_file(path)
    if len(items.get("mesh", [])) != 1:
        raise LoadError("%s: terrain file doesn't contain a mesh file" % path)
    return TerrainFile(load_mesh(_resolve(os.path.dirname(path), str(items["mesh"][0]))), name)


# ---------------------------------------------------------------- world XML

def _attr_vec(e, name, k):
    s = e.get(name)
    if s is None:
        return None
    vals = s.split()
    try:
        if len(vals) < k:
            return None
        return np.array([float(x) for x in vals[:k]], np.float64)
    except ValueError:
        return None


def read_transform(e, allow_scale=False):
    """``ReadTransform`` (XmlWorld.cpp:96-190): ``translation`` or
    ``position``; ``rotateRPY`` (roll pitch yaw), ``rotateMoment`` (replaces
    it), then ``rotateX``, ``rotateY``, ``rotateZ`` each premultiplied in that
    order; with ``allow_scale`` a scalar or 3-vector ``scale`` multiplies the
    matrix columns.  Returns a 4x4 matrix, or None when no attribute is set."""
    R, t, any_ = np.eye(3), np.zeros(3), False
    for key in ("translation", "position"):
        p = _attr_vec(e, key, 3)
        if p is not None:
            t, any_ = p, True
            break
    rpy = _attr_vec(e, "rotateRPY", 3)
    if rpy is not None:
        R, any_ = rot_rpy(*rpy), True
    m = _attr_vec(e, "rotateMoment", 3)
    if m is not None:
        R, any_ = rot_moment(m), True
    for key, f in (("rotateX", rot_x), ("rotateY", rot_y), ("rotateZ", rot_z)):
        a = _attr_vec(e, key, 1)
        if a is not None:
            R, any_ = f(a[0]) @ R, True
    M = np.eye(4)
    M[:3, :3], M[:3, 3] = R, t
    if allow_scale and e.get("scale") is not None:
        s = _attr_vec(e, "scale", 3)
        if s is None:
            s = _attr_vec(e, "scale", 1)
            s = None if s is None else np.repeat(s, 3)
        if s is not None:
            M[:3, :3] = M[:3, :3] * s[None, :]
            any_ = True
    return M if any_ else None


def _klampt_vector(s, where):
    """KrisLibrary vector text: the count, then the values."""
    tok = s.split()
    try:
        n = int(float(tok[0]))
        vals = [float(x) for x in tok[1:1 + n]]
    except (ValueError, IndexError):
        raise LoadError("%s: malformed configuration %r" % (where, s)) from None
    if len(vals) != n:
        raise LoadError("%s: configuration has %d of %d values" % (where, len(vals), n))
    return np.array(vals, np.float64)


@dataclass
class WorldFile:
    robots: List[RobotFile] = field(default_factory=list)
    objects: List[RigidObjectFile] = field(default_factory=list)
    terrains: List[TerrainFile] = field(default_factory=list)

    def scene(self, robot=0) -> S.Scene:
        """The world as a one-robot ``synthetic.Scene`` (terrains, rigid
        objects, robot) for the engine's scene builders and the oracle."""
        if len(self.robots) != 1 or robot != 0:
            raise ValueError("scene() takes a world with exactly one robot")
        sc = S.Scene(self.robots[0].spec)
        sc.terrains = [t.mesh for t in self.terrains]
        sc.objects = [(o.mesh, o.xf) for o in self.objects]
        return sc


def load_world(path) -> WorldFile:
    """World XML (``XmlWorld::GetWorld``, XmlWorld.cpp:615-700): ``<robot>``,
    ``<rigidObject>`` and ``<terrain>`` elements, in that order."""
    if not os.path.isfile(path):
        raise LoadError("%s: no such file" % path)
    try:
        root = ET.parse(path).getroot()
    except ET.ParseError as e:
        raise LoadError("%s: XML error: %s" % (path, e)) from None
    if root.tag != "world":
        raise LoadError("%s: root element is <%s>, not <world>" % (path, root.tag))
    base = os.path.dirname(path)
    w = WorldFile()
    for e in root.findall("robot"):
        fn = e.get("file")
        if not fn:
            raise LoadError("%s: <robot> without a file attribute" % path)
        rf = load_robot(_resolve(base, fn))
        rf.name = e.get("name", "Robot")
        where = "%s: robot %s" % (path, rf.name)
        if e.get("config") is not None:
            q = _klampt_vector(e.get("config"), where)
            if q.shape[0] != rf.spec.n:
                raise LoadError("%s: configuration size %d != %d links" % (where, q.shape[0], rf.spec.n))
            rf.q = q
        if e.get("configfile") is not None:
            cf = _resolve(base, e.get("configfile"))
            if not os.path.isfile(cf):
                raise LoadError("%s: could not open config file %s" % (where, cf))
            q = _klampt_vector(open(cf).read(), where)
            if q.shape[0] != rf.spec.n:
                raise LoadError("%s: configuration size %d != %d links" % (where, q.shape[0], rf.spec.n))
            rf.q = q
        M = read_transform(e)
        if M is not None:
            for i in range(rf.spec.n):
                if rf.spec.parents[i] == -1:
                    R, t = xf_mat(rf.spec.T0[i])
                    rf.spec.T0[i] = xf12(M[:3, :3] @ R, M[:3, :3] @ t + M[:3, 3])
        w.robots.append(rf)
    for e in root.findall("rigidObject"):
        name = e.get("name", "Object")
        obj = None
        if e.get("file"):
            obj = load_rigid_object(_resolve(base, e.get("file")))
        g = e.find("geometry")
        if g is not None:
            gfn = g.get("file") or g.get("mesh")
            if gfn:
                mesh = load_mesh(_resolve(base, gfn))
                obj = RigidObjectFile(mesh, obj.xf if obj else xf12(np.eye(3), np.zeros(3)))
            if obj is not None:
                M = read_transform(g, allow_scale=True)
                if M is not None:
                    obj.mesh = keyed_transform(obj.mesh, M)
                if g.get("margin") is not None:
                    obj.margin = float(g.get("margin"))
        if obj is None or obj.mesh.tris.shape[0] == 0:
            raise LoadError("%s: rigid object %s has no geometry" % (path, name))
        M = read_transform(e)
        if M is not None:
            obj.xf = xf12(M[:3, :3], M[:3, 3])
        obj.name = name
        w.objects.append(obj)
    for e in root.findall("terrain"):
        fn = e.get("file")
        if not fn:
            raise LoadError("%s: <terrain> without a file attribute" % path)
        ter = load_terrain(_resolve(base, fn))
        M = read_transform(e, allow_scale=True)
        if M is not None:
            ter.mesh = keyed_transform(ter.mesh, M)
        if e.get("margin") is not None:
            ter.margin = float(e.get("margin"))
        ter.name = e.get("name", ter.name)
        w.terrains.append(ter)
    return w
