"""The oracle's TriangleMesh -> ImplicitSurface grid, pinned by closed forms:
the reference's cube.off fixture against the analytic box distance (with
cell centres exactly on the cube's face diagonals, where the winding count
must take each shared edge once), winding numbers of overlapping and
inverted boxes against analytic containment, and the AABB-tree search
against brute force."""
import os

import numpy as np
import pytest

import oracle
from klampt_amd import io, synthetic as S

CUBE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "objects", "cube.off")


def _centres(bmin, h, dims):
    return np.asarray(bmin) + (np.indices(dims).transpose(1, 2, 3, 0) + 0.5) * h


def _box_sdf(P, lo, hi):
    c, e = (np.asarray(lo) + hi) / 2, (np.asarray(hi) - lo) / 2
    q = np.abs(P - c) - e
    return np.linalg.norm(np.maximum(q, 0), axis=-1) + np.minimum(q.max(-1), 0)


@pytest.mark.parametrize("bmin,h,dims", [((-0.2, -0.2, -0.2), 0.1, (14, 14, 14)),
                                         ((-0.25, -0.25, -0.3), 0.125, (12, 12, 13))])
def test_reference_cube_matches_analytic_sdf(bmin, h, dims):
    cube = io.load_mesh(CUBE)
    out, wind = oracle.mesh_to_sdf(cube, bmin, h, dims, with_winding=True)
    P = _centres(bmin, h, dims)
    assert np.abs(out - _box_sdf(P, [0, 0, 0], [1, 1, 1])).max() < 1e-14
    inside = (P > 0).all(-1) & (P < 1).all(-1)
    assert np.array_equal(wind != 0, inside)
    assert set(np.unique(wind)) <= {0, 1}


def test_overlapping_and_inverted_boxes_winding():
    a = S.box_mesh([0.3, 0.3, 0.3], c=[0, 0, 0])
    b = S.box_mesh([0.3, 0.3, 0.3], c=[0.2, 0.1, 0.05])
    inv = S.box_mesh([0.2, 0.2, 0.2], c=[0.9, 0, 0])
    inv = S.Mesh(inv.verts, inv.tris[:, ::-1].copy())
    m = S.merge_meshes([a, b, inv])
    bmin, h, dims = (-0.45, -0.4, -0.4), 0.05, (34, 16, 16)
    out, wind = oracle.mesh_to_sdf(m, bmin, h, dims, with_winding=True)
    P = _centres(bmin, h, dims)
    ina = (np.abs(P) < 0.3).all(-1)
    inb = (np.abs(P - [0.2, 0.1, 0.05]) < 0.3).all(-1)
    inn = (np.abs(P - [0.9, 0, 0]) < 0.2).all(-1)
    assert np.array_equal(wind, ina.astype(int) + inb.astype(int) - inn.astype(int))
    assert np.array_equal(out < 0, ina | inb | inn)
    # distance: to the nearest triangle of any of the three boxes (their
    # corners are float32-rounded, so the bounds come from the vertices)
    d = np.min([np.abs(_box_sdf(P, x.verts.min(0), x.verts.max(0))) for x in (a, b, inv)], axis=0)
    assert np.abs(np.abs(out) - d).max() < 1e-14


def test_tree_search_equals_brute_force():
    m = S.merge_meshes([S.ellipsoid_mesh([0.2, 0.3, 0.25], S.random_rotation(np.random.default_rng(1)),
                                         c=[0.1, 0, 0], edge=0.06),
                        S.box_mesh([0.1, 0.2, 0.05], c=[-0.3, 0.2, 0.1])])
    bmin, h, dims = (-0.5, -0.45, -0.4), 0.037, (24, 25, 22)
    a, wa = oracle.mesh_to_sdf(m, bmin, h, dims, with_winding=True)
    b, wb = oracle.mesh_to_sdf(m, bmin, h, dims, brute=True, with_winding=True)
    assert np.array_equal(a, b) and np.array_equal(wa, wb)
    assert (a < 0).any() and (a > 0).any()
