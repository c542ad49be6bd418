"""TriangleMesh -> ImplicitSurface on the GPU (kg_geom_mesh_to_sdf) against
the FP64 oracle on the same grid: distances within 1e-5 m, the inside/outside
classification (winding number along +z) identical in every cell; the grid
layout rule and the analytic SDF of the reference's cube.off fixture."""
import os

import numpy as np
import pytest

import oracle
from klampt_amd import engine, io, robotsim, synthetic as S

pytestmark = pytest.mark.gpu
CUBE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "objects", "cube.off")
TOL = 1e-5


def _convert(mesh, h, pad=2):
    g = engine.Geom.trimesh(robotsim.default_context(), mesh.verts, mesh.tris)
    s = g.to_sdf(h, pad)
    return s.sdf_grid(values=True)


def _box_sdf(P, lo, hi):
    c, e = (np.asarray(lo) + hi) / 2, (np.asarray(hi) - lo) / 2
    q = np.abs(P - c) - e
    return np.linalg.norm(np.maximum(q, 0), axis=-1) + np.minimum(q.max(-1), 0)


def test_cube_grid_layout_and_analytic_values():
    cube = io.load_mesh(CUBE)
    dims, lo, hi, v = _convert(cube, 0.1)
    assert dims == (14, 14, 14)
    assert np.allclose(lo, -0.2, atol=1e-15) and np.allclose(hi, 1.2, atol=1e-12)
    P = lo + (np.indices(dims).transpose(1, 2, 3, 0) + 0.5) * 0.1
    assert np.abs(v - _box_sdf(P, [0, 0, 0], [1, 1, 1])).max() < 1e-6
    dims1, lo1, _, _ = _convert(cube, 0.1, pad=0)
    assert dims1 == (10, 10, 10) and np.allclose(lo1, 0)


def test_default_resolution_is_average_triangle_diameter():
    m = S.ellipsoid_mesh([0.2, 0.3, 0.25], edge=0.05)
    P = m.verts[m.tris]
    h = np.mean(np.max(np.linalg.norm(P - np.roll(P, -1, axis=1), axis=-1), axis=1))
    dims, lo, hi, _ = _convert(m, 0.0)
    want = np.ceil((m.verts.max(0) - m.verts.min(0)) / h).astype(int) + 4
    assert dims == tuple(want) and np.allclose(lo, m.verts.min(0) - 2 * h)


def _meshes():
    rng = np.random.default_rng(7)
    inv = S.box_mesh([0.2, 0.2, 0.2], c=[0.9, 0, 0])
    yield "ellipsoid", S.ellipsoid_mesh([0.2, 0.3, 0.25], S.random_rotation(rng), c=[0.1, 0, 0], edge=0.03), 0.02
    yield "boxes", S.merge_meshes([S.box_mesh([0.3, 0.3, 0.3]), S.box_mesh([0.3, 0.3, 0.3], c=[0.2, 0.1, 0.05]),
                                   S.Mesh(inv.verts, inv.tris[:, ::-1].copy())]), 0.025
    yield "capsule", S.capsule_mesh(0.05, 0.3, axis=0), 0.01
    yield "soup", S.clutter_soup(seed=2, target_tris=20_000, center=(0, 0, 0), side=1.0), 0.02


@pytest.mark.parametrize("name,mesh,h", list(_meshes()), ids=[n for n, _, _ in _meshes()])
def test_matches_oracle(name, mesh, h):
    dims, lo, hi, v = _convert(mesh, h)
    o, w = oracle.mesh_to_sdf(mesh, lo, h, dims, with_winding=True)
    err = np.abs(v.astype(np.float64) - o)
    assert err.max() <= TOL, (name, float(err.max()))
    # the classification is exact: every cell negated exactly where the
    # oracle's winding number is nonzero (signbit also covers -0.0)
    assert np.array_equal(np.signbit(v), w != 0), name
    assert (w != 0).any() and (w == 0).any()


def test_robotsim_convert_and_query():
    a = robotsim.Geometry3D()
    assert a.loadFile(CUBE)
    a.setCurrentTransform([1, 0, 0, 0, 1, 0, 0, 0, 1], [2.0, 0, 0])
    s = a.convert("ImplicitSurface", 0.05)
    assert s.type() == "ImplicitSurface"
    assert s.getCurrentTransform()[1] == [2.0, 0, 0]
    grid = s.getImplicitSurface()
    assert grid.values.shape == (24, 24, 24) and np.allclose(grid.bmin, -0.1)
    r = s.distance_point([2.5, 0.5, 1.05])         # 0.05 above the top face, inside the grid
    assert abs(r.d - 0.05) < 1e-5
    with pytest.raises(RuntimeError, match="Cannot perform the geometry conversion"):
        s.convert("TriangleMesh")
    with pytest.raises(RuntimeError, match="nonnegative"):
        a.convert("ImplicitSurface", -1)


def test_edge_cases_empty_flat_and_open_meshes():
    from klampt_amd._lib import KlamptGPUError
    ctx = robotsim.default_context()
    empty = engine.Geom.trimesh(ctx, np.zeros((0, 3)), np.zeros((0, 3), np.int32))
    with pytest.raises(KlamptGPUError, match="empty"):
        empty.to_sdf(0.1)
    # one triangle in the z = 0 plane: an open, flat mesh (dims along z = 2 pad)
    tri = S.Mesh(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float64), np.array([[0, 1, 2]], np.int32))
    dims, lo, hi, v = _convert(tri, 0.1)
    assert dims == (14, 14, 4)
    o, w = oracle.mesh_to_sdf(tri, lo, 0.1, dims, with_winding=True)
    assert np.abs(v - o).max() <= TOL and np.array_equal(np.signbit(v), w != 0)
    with pytest.raises(KlamptGPUError):
        engine.Geom.trimesh(ctx, tri.verts, tri.tris).to_sdf(-1.0)
