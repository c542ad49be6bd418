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


def test_sweep_terms_known_answer():
    """One link point (1, 0, 0) on a revolute joint about z at the origin,
    swept from q = 0 to q = 0.1 rad past the box [2.5, 3.5] x [-0.5, 0.5]^2:
    the closest point of the swept segment is its start, so d = 1.5, the
    normal is -x, alpha = |p1 - p0| / (0 + |p1 - p0|) = 1, and normal . (z x p)
    is 0 at p0 = (1, 0, 0) and sin(0.1) at p1 = (cos 0.1, sin 0.1, 0)."""
    from klampt_amd import synthetic as S
    spec = S.RobotSpec(name="pin", parents=np.array([-1], np.int32), T0=I12[None, :].copy(),
                       axes=np.array([[0.0, 0.0, 1.0]]), jtype=np.array([0], np.int32),
                       qmin=np.array([-np.pi]), qmax=np.array([np.pi]), link_meshes=[None])
    ow = oracle.OracleWorld()
    r = ow.add_robot(spec)
    ow.init_default()
    box = np.array([[x, y, z] for x in (2.5, 3.5) for y in (-0.5, 0.5) for z in (-0.5, 0.5)])
    got = oracle.sweep_terms(ow, r, np.array([[1.0, 0.0, 0.0]]), 0, [0.0], [0.1], box, I12, dcheck=3.0)
    d, alpha, nrm, ps, po, p0, p1, c0, c1 = got
    assert abs(d - 1.5) < 1e-12 and np.allclose(nrm, [-1, 0, 0], atol=1e-12)
    assert np.allclose(ps, [1, 0, 0], atol=1e-12) and abs(po[0] - 2.5) < 1e-12
    assert np.allclose(p1, [np.cos(0.1), np.sin(0.1), 0], atol=1e-12) and abs(alpha - 1) < 1e-12
    assert abs(c0[0]) < 1e-9 and abs(c1[0] - np.sin(0.1)) < 1e-9
    # beyond the bound: no terms; the same point at both ends: alpha 1/2
    assert oracle.sweep_terms(ow, r, np.array([[1.0, 0.0, 0.0]]), 0, [0.0], [0.1], box, I12, dcheck=1.0) is None
    same = oracle.sweep_terms(ow, r, np.array([[1.0, 0.0, 0.0]]), 0, [0.3], [0.3], box, I12, dcheck=3.0)
    assert abs(same[1] - 0.5) < 1e-12 and np.allclose(same[7], same[8])


def test_coplanar_points_swept_in_their_plane():
    """Flat hulls: a square of side 1 in z = 0 at two placements in that
    plane (turned about z by at most 0.3 rad and shifted by at most 0.3, so
    both cover the point (0.5, 0.5)) against the cube [0, 1]^2 x [0.5, 1.5]
    is 0.5 away; against the cube lowered to straddle z = 0 it overlaps."""
    rng = np.random.default_rng(7)
    sq = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0], [0.5, 0.5, 0], [0.3, 0.8, 0]], float)
    cube = np.array([[x, y, z] for x in (0, 1) for y in (0, 1) for z in (0.5, 1.5)], float)

    def Rz(t):
        return np.array([[np.cos(t), -np.sin(t), 0], [np.sin(t), np.cos(t), 0], [0, 0, 1.0]])

    for k in range(20):
        a, b = rng.uniform(-0.3, 0.3, 2)
        x1 = _xf(Rz(a), np.r_[rng.uniform(-0.3, 0.3, 2), 0.0])
        x2 = _xf(Rz(b), np.r_[rng.uniform(-0.3, 0.3, 2), 0.0])
        d, ca, cb, _ = oracle.hull_distance(sq, x1, cube, I12, xa2=x2)
        assert abs(d - 0.5) < 1e-12, (k, d)
        assert abs(ca[2]) < 1e-12 and abs(cb[2] - 0.5) < 1e-12
        d2, _, _, _ = oracle.hull_distance(sq, x1, cube - [0, 0, 1.0], I12, xa2=x2)
        assert d2 == 0


def test_flat_simplex_is_not_an_overlap():
    """A capsule link swept past a box (the C1 arm's link 1, a trajectory step
    of the sweep-term workload, box 4), where GJK builds a flat tetrahedron:
    the distance is the brute-force minimum over the hull triangles (0.0146 m),
    not 0."""
    from klampt_amd import synthetic as S
    sc = S.boxes_scene(0)
    ow, orb = oracle.world_from_scene(sc)
    rng = np.random.default_rng(3)
    Q0 = S.uniform_configs(sc.robot, 48, seed=3).astype(np.float64)
    step = rng.normal(size=Q0.shape)
    step *= rng.uniform(0.02, 0.3, (48, 1)) / np.abs(step).max(axis=1, keepdims=True)
    Q1 = Q0 + step
    P = sc.robot.link_meshes[1].verts
    X0, X1 = ow.fk(orb, Q0[3])[1], ow.fk(orb, Q1[3])[1]
    B, xb = sc.objects[4][0].verts, np.asarray(sc.objects[4][1], np.float64)
    d, _, _, _ = oracle.hull_distance(P, X0, B, xb, xa2=X1)

    def place(X, V):
        return V @ X[:9].reshape(3, 3) + X[9:]

    want = _brute(np.vstack([place(X0, P), place(X1, P)]), place(xb, B))
    assert want > 0.01 and abs(d - want) < 1e-9, (d, want)
