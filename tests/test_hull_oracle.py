"""The oracle's convex-hull distance (GJK, used for KineTrajOpt's swept
link hulls) pinned without sharing its code: against the minimum over the
triangle pairs of scipy's hull triangulations (the oracle's triangle
distance, itself pinned in test_oracle.py), closest points on the two hulls
that realise the distance, overlap giving 0, and the hull of two placed
copies equal to the hull of their union."""
import numpy as np
import pytest
from scipy.spatial import ConvexHull

import oracle

I12 = np.r_[np.eye(3).ravel(), 0.0, 0.0, 0.0]


def _xf(R, t):
    return np.r_[R.T.ravel(), t]  # column-major R, then t


def _rot(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def _brute(A, B):
    ha, hb = ConvexHull(A), ConvexHull(B)
    return min(oracle.tritri_dist(A[sa].ravel(), B[sb].ravel()) for sa in ha.simplices for sb in hb.simplices)


def _inside(P, pts, tol=1e-9):
    h = ConvexHull(P)
    return np.all(h.equations[:, :3] @ pts + h.equations[:, 3] <= tol)


def test_gjk_equals_brute_force_over_hull_triangles():
    rng = np.random.default_rng(1)
    for k in range(40):
        A = rng.normal(size=(24, 3)) * rng.uniform(0.2, 1, 3)
        B = rng.normal(size=(24, 3)) * rng.uniform(0.2, 1, 3)
        Ra, Rb = _rot(rng), _rot(rng)
        ta, tb = rng.uniform(-1, 1, 3), rng.uniform(-3, 3, 3)
        d, ca, cb, _ = oracle.hull_distance(A, _xf(Ra, ta), B, _xf(Rb, tb))
        Aw, Bw = A @ Ra.T + ta, B @ Rb.T + tb
        want = _brute(Aw, Bw)
        assert abs(d - want) < 1e-9, k
        if want > 1e-9:
            assert abs(np.linalg.norm(ca - cb) - d) < 1e-9, k
            assert _inside(Aw, ca) and _inside(Bw, cb), k
        else:
            assert d == 0


def test_swept_hull_is_the_hull_of_both_copies():
    rng = np.random.default_rng(2)
    for k in range(20):
        A = rng.normal(size=(16, 3)) * 0.3
        B = rng.normal(size=(16, 3)) * 0.3 + rng.uniform(-2, 2, 3)
        x1 = _xf(_rot(rng), rng.uniform(-1, 1, 3))
        x2 = _xf(_rot(rng), rng.uniform(-1, 1, 3))
        d, ca, cb, _ = oracle.hull_distance(A, x1, B, I12, xa2=x2)
        R1, t1 = x1[:9].reshape(3, 3).T, x1[9:]
        R2, t2 = x2[:9].reshape(3, 3).T, x2[9:]
        U = np.vstack([A @ R1.T + t1, A @ R2.T + t2])
        d2, _, _, _ = oracle.hull_distance(U, I12, B, I12)
        assert abs(d - d2) < 1e-9 and abs(d - _brute(U, B)) < 1e-9, k
        if d > 1e-9:
            assert _inside(U, ca), k


def test_overlap_and_support():
    A = np.array([[x, y, z] for x in (0, 1) for y in (0, 1) for z in (0, 1)], float)
    d, _, _, _ = oracle.hull_distance(A, I12, A + 0.5, I12)
    assert d == 0
    d, ca, cb, _ = oracle.hull_distance(A, I12, A + np.array([3.0, 0, 0]), I12)
    assert abs(d - 2.0) < 1e-12 and abs(ca[0] - 1) < 1e-12 and abs(cb[0] - 3) < 1e-12
    assert np.allclose(oracle.hull_support(A, I12, [1, 1, 1]), [1, 1, 1])
    assert np.allclose(oracle.hull_support(A, _xf(np.eye(3), [2, 0, 0]), [-1, -2, -3]), [2, 0, 0])
