"""Pins the FP64 oracle against independent references (CPU only).

The reference's own tests hold no vectors for this path, so each oracle
function is checked against something that does not share its code:
scipy's rotation library (FK), LP separability (triangle intersection),
dense barycentric sampling (triangle distance), brute force over all triangle
pairs (BVH traversals), closed forms (boxes, planar SDFs), and an exhaustive
scan of samples (edge checking).
"""
import math

import numpy as np
import pytest
from scipy.optimize import linprog
from scipy.spatial.transform import Rotation

import oracle
from klampt_amd import synthetic as S


# ------------------------------------------------------------------ FK

def _xf_np(x12):
    R = x12[:9].reshape(3, 3).T
    return R, x12[9:]


def test_fk_matches_scipy_rotations():
    """T_World[i] = T_World[p] * T0[i] * Rot(w_i, q_i) composed with scipy."""
    spec = S.arm7()
    w = oracle.OracleWorld()
    r = w.add_robot(spec)
    rng = np.random.default_rng(3)
    for _ in range(20):
        q = rng.uniform(-math.pi, math.pi, spec.n)
        T = w.fk(r, q)
        Rw, tw = np.eye(3), np.zeros(3)
        for i in range(spec.n):
            R0, t0 = _xf_np(spec.T0[i])
            J = Rotation.from_rotvec(spec.axes[i] * q[i]).as_matrix()
            tw = Rw @ t0 + tw
            Rw = Rw @ R0 @ J
            Ro, to = _xf_np(T[i])
            assert np.allclose(Ro, Rw, atol=1e-12)
            assert np.allclose(to, tw, atol=1e-12)
            assert np.allclose(Ro.T @ Ro, np.eye(3), atol=1e-12)
            assert abs(np.linalg.det(Ro) - 1) < 1e-12


def test_fk_special_cases():
    spec = S.arm7()
    w = oracle.OracleWorld()
    r = w.add_robot(spec)
    T = w.fk(r, np.zeros(7))
    # q = 0: translations accumulate 0.3 along z, x, z, x, ...
    expect = np.cumsum([[0, 0, 0.3] if i % 2 == 0 else [0.3, 0, 0] for i in range(7)], axis=0)
    assert np.allclose(T[:, 9:], S.f32r(expect), atol=1e-7)
    # single joint 0 at pi/2 about z: x -> y exactly (up to cos(pi/2) rounding)
    q = np.zeros(7)
    q[0] = math.pi / 2
    R, _ = _xf_np(w.fk(r, q)[0])
    assert np.allclose(R, [[0, -1, 0], [1, 0, 0], [0, 0, 1]], atol=1e-15)


# ------------------------------------------------------------------ triangles

def _separable(t1, t2):
    """Strict separability of two triangles by LP: find (w,b) with
    w.p >= b+1 on t1 and w.q <= b-1 on t2."""
    A, ub = [], []
    for p in t1:
        A.append([-p[0], -p[1], -p[2], 1.0]); ub.append(-1.0)
    for q in t2:
        A.append([q[0], q[1], q[2], -1.0]); ub.append(-1.0)
    res = linprog(np.zeros(4), A_ub=np.array(A), b_ub=np.array(ub), bounds=[(None, None)] * 4, method="highs")
    return res.status == 0


def _sampled_dist(t1, t2, k=40):
    g = np.array([(i / k, j / k) for i in range(k + 1) for j in range(k + 1 - i)])
    P = t1[0] + g[:, :1] * (t1[1] - t1[0]) + g[:, 1:] * (t1[2] - t1[0])
    Q = t2[0] + g[:, :1] * (t2[1] - t2[0]) + g[:, 1:] * (t2[2] - t2[0])
    d = np.sqrt(((P[:, None, :] - Q[None, :, :]) ** 2).sum(-1))
    return d.min()


def test_tritri_against_lp_separability():
    rng = np.random.default_rng(11)
    n_hit = n_miss = 0
    for _ in range(600):
        t1 = rng.normal(size=(3, 3))
        t2 = rng.normal(size=(3, 3)) * rng.uniform(0.3, 1.0) + rng.normal(size=3) * 0.8
        hit = oracle.tritri(t1.ravel(), t2.ravel())
        assert hit == (not _separable(t1, t2))
        n_hit += hit
        n_miss += not hit
    assert n_hit > 50 and n_miss > 50


def test_tritri_known_cases():
    base = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], float)
    # piercing
    assert oracle.tritri(base.ravel(), np.array([[0.2, 0.2, -1], [0.2, 0.2, 1], [0.3, 0.25, 1]]).ravel())
    # parallel, separated by 0.1
    assert not oracle.tritri(base.ravel(), (base + [0, 0, 0.1]).ravel())
    # coplanar overlapping / disjoint
    assert oracle.tritri(base.ravel(), (base + [0.2, 0.2, 0]).ravel())
    assert not oracle.tritri(base.ravel(), (base + [1.2, 0, 0]).ravel())
    # coplanar containment
    assert oracle.tritri(base.ravel(), (base * 0.1 + [0.1, 0.1, 0]).ravel())
    # touching: vertex on the face (closed sets)
    assert oracle.tritri(base.ravel(), np.array([[0.2, 0.2, 0], [0.2, 0.3, 1], [0.4, 0.2, 1]]).ravel())


def test_tritri_distance_against_sampling():
    rng = np.random.default_rng(5)
    for _ in range(150):
        t1 = rng.normal(size=(3, 3))
        t2 = rng.normal(size=(3, 3)) + rng.normal(size=3) * 2.5
        d = oracle.tritri_dist(t1.ravel(), t2.ravel())
        ds = _sampled_dist(t1, t2)
        # sampling over-estimates by at most the grid's covering radius
        h = max(np.linalg.norm(t1[i] - t1[j]) for i in range(3) for j in range(3)) / 40
        h2 = max(np.linalg.norm(t2[i] - t2[j]) for i in range(3) for j in range(3)) / 40
        assert d <= ds + 1e-12
        assert d >= ds - (h + h2) - 1e-12
        assert (d == 0) == oracle.tritri(t1.ravel(), t2.ravel())


def test_tritri_distance_closed_form():
    base = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], float)
    assert abs(oracle.tritri_dist(base.ravel(), (base + [0, 0, 0.25]).ravel()) - 0.25) < 1e-15
    # edge-edge: skew segments 0.5 apart
    t2 = np.array([[0.5, -1, 0.5], [0.5, 1, 0.5], [0.5, 0, 2.0]])
    assert abs(oracle.tritri_dist(base.ravel(), t2.ravel()) - 0.5) < 1e-12


# ------------------------------------------------------------------ meshes

def _two_mesh_world(m1, m2):
    w = oracle.OracleWorld()
    return w, w.add_mesh(m1), w.add_mesh(m2)


def _xf(R, t):
    return np.concatenate([R.T.reshape(9), t])


def test_mesh_collide_and_distance_bvh_equal_brute_force():
    rng = np.random.default_rng(2)
    cap = S.capsule_mesh(0.05, 0.3, axis=2, m=8, k=3)
    box = S.box_mesh([0.1, 0.05, 0.08], edge=0.04)
    w, ga, gb = _two_mesh_world(cap, box)
    hits = 0
    for _ in range(300):
        xa = _xf(S.random_rotation(rng), rng.uniform(-0.2, 0.2, 3))
        xb = _xf(S.random_rotation(rng), rng.uniform(-0.2, 0.2, 3))
        c1 = w.mesh_collide(ga, xa, gb, xb, brute=False)
        c2 = w.mesh_collide(ga, xa, gb, xb, brute=True)
        assert c1 == c2
        hits += c1
        if _ < 60:
            d1 = w.mesh_distance(ga, xa, gb, xb, brute=False)[0]
            d2 = w.mesh_distance(ga, xa, gb, xb, brute=True)[0]
            assert abs(d1 - d2) < 1e-12
            assert (d1 == 0) == c1
    assert 30 < hits < 270


def test_box_box_closed_forms():
    cube = S.box_mesh([0.5, 0.5, 0.5])
    w, ga, gb = _two_mesh_world(cube, cube)
    I = np.eye(3)
    for gap in [0.3, 0.01, 1e-5]:
        xb = _xf(I, np.array([1.0 + gap, 0.2, -0.1]))
        d = w.mesh_distance(ga, _xf(I, np.zeros(3)), gb, xb)[0]
        assert abs(d - gap) < 1e-12
        assert not w.mesh_collide(ga, _xf(I, np.zeros(3)), gb, xb)
    xb = _xf(I, np.array([0.99, 0.2, -0.1]))
    assert w.mesh_collide(ga, _xf(I, np.zeros(3)), gb, xb)
    # diagonal separation: corner to corner along (1,1,1)
    xb = _xf(I, np.array([1.1, 1.1, 1.1]))
    d = w.mesh_distance(ga, _xf(I, np.zeros(3)), gb, xb)[0]
    assert abs(d - math.sqrt(3) * 0.1) < 1e-12


# ------------------------------------------------------------------ Klampt policy

def _tiny_world():
    """2-link arm with unit-box links + a terrain and a rigid object."""
    spec = S.RobotSpec(
        name="tiny", parents=np.array([-1, 0], np.int32),
        T0=np.array([np.r_[np.eye(3).T.ravel(), [0, 0, 0]], np.r_[np.eye(3).T.ravel(), [1.0, 0, 0]]]),
        axes=np.array([[0, 0, 1.0], [0, 0, 1.0]]), jtype=np.zeros(2, np.int32),
        qmin=np.full(2, -math.pi), qmax=np.full(2, math.pi),
        link_meshes=[S.box_mesh([0.3, 0.1, 0.1]), S.box_mesh([0.3, 0.1, 0.1])])
    return spec


def test_mask_defaults_follow_initialize_default():
    spec = _tiny_world()
    w = oracle.OracleWorld()
    w.add_terrain(S.box_mesh([5, 5, 0.05], c=[0, 0, -0.2]))
    w.add_object(S.box_mesh([0.1, 0.1, 0.1]), _xf(np.eye(3), np.array([3.0, 0, 0])))
    r = w.add_robot(spec)
    w.init_default()
    # IDs: terrain 0, object 1, robot 2, links 3,4
    en = lambda a, b: bool(w.L.or_get_collision_enabled(w.h, a, b))
    assert w.num_ids() == 5
    assert not en(0, 0) and not en(3, 3)
    assert en(2, 2)                        # robot may self-collide
    assert not en(3, 2) and not en(2, 4)   # link <-> own robot off
    assert not en(3, 0) and not en(0, 3)   # link 0 (parent -1) <-> terrain off
    assert en(4, 0) and en(0, 4)           # link 1 <-> terrain on
    assert en(3, 1) and en(1, 3)           # fixed link <-> rigid object stays on
    assert not en(3, 4)                    # link vs parent excluded from self pairs
    assert en(0, 1)                        # terrain <-> object


def test_bisection_order_visits_each_sample_once():
    for n in [2, 3, 10, 11, 57, 501]:
        o = oracle.bisection_order(n)
        assert sorted(o.tolist()) == list(range(1, n))
        assert o[0] == n // 2


def test_edge_first_u_equals_exhaustive_scan():
    sc = S.boxes_scene(0)
    w, r = oracle.world_from_scene(sc)
    A, B = S.edge_segments(sc.robot, 40, eps=0.01, smin=10, smax=120, seed=4)
    fu, ns, nchk = w.edge_batch(r, A, B, eps=0.01, norm_inf=True)
    for s in range(len(A)):
        n = ns[s]
        order = oracle.bisection_order(n)
        a, b = A[s].astype(np.float64), B[s].astype(np.float64)
        Q = np.array([a * (1 - i / n) + b * (i / n) for i in range(1, n)])
        feas = w.feasible_batch(r, Q, nthreads=1)[0]
        bad = [i for i in order if not feas[i - 1]]
        assert fu[s] == (bad[0] / n if bad else -1.0)
        assert n == math.ceil(np.abs(b - a).max() / 0.01)


# ------------------------------------------------------------------ SDF

def _plane_grid(n=(1.0, -2.0, 0.5), c=0.3, dims=(12, 10, 9)):
    n = np.asarray(n) / np.linalg.norm(n)
    bmin, bmax = np.array([-0.6, -0.5, -0.45]), np.array([0.6, 0.5, 0.45])
    h = (bmax - bmin) / dims
    ii = [(np.arange(dims[a]) + 0.5) * h[a] + bmin[a] for a in range(3)]
    X = np.stack(np.meshgrid(*ii, indexing="ij"), -1)
    vals = X @ n - c
    return S.SDFGrid(bmin, bmax, vals), n, c


def test_sdf_trilinear_reproduces_planes_exactly():
    g, n, c = _plane_grid()
    rng = np.random.default_rng(0)
    h = (g.bmax - g.bmin) / g.values.shape
    lo, hi = g.bmin + h / 2, g.bmax - h / 2          # box of cell centres
    P = rng.uniform(lo, hi, (500, 3))
    d, gr = oracle.sdf_eval(g, P)
    assert np.allclose(d, P @ n - c, atol=1e-13)
    assert np.allclose(gr, np.broadcast_to(n, gr.shape), atol=1e-12)
    # cell centres return the stored values
    C = g.bmin + h * (np.array([3, 4, 5]) + 0.5)
    assert abs(oracle.sdf_eval(g, C[None])[0][0] - g.values[3, 4, 5]) < 1e-15
    # outside: clamped to the nearest cell-centre face, gradient 0 across it
    Pout = np.array([[hi[0] + 0.3, 0.0, 0.0]])
    d, gr = oracle.sdf_eval(g, Pout)
    assert abs(d[0] - (np.array([hi[0], 0, 0]) @ n - c)) < 1e-13
    assert gr[0, 0] == 0 and abs(gr[0, 1] - n[1]) < 1e-12


def test_sdf_penalty_gradient_matches_central_differences_on_plane():
    """Klampt's distance-gradient rule (-2 h grad_d^T J_pos) restated with
    explicit Jacobians equals FP64 central differences of the loss where the
    SDF is smooth (a planar field, reproduced exactly by trilinear)."""
    spec = S.arm7()
    w = oracle.OracleWorld()
    r = w.add_robot(spec)
    # plane through the workspace, large grid so points stay inside
    dims = (40, 40, 40)
    bmin, bmax = np.array([-2.5, -2.5, -2.0]), np.array([2.5, 2.5, 3.0])
    h = (bmax - bmin) / dims
    ii = [(np.arange(dims[a]) + 0.5) * h[a] + bmin[a] for a in range(3)]
    X = np.stack(np.meshgrid(*ii, indexing="ij"), -1)
    n = np.array([0.3, -0.2, 1.0]); n /= np.linalg.norm(n)
    grid = S.SDFGrid(bmin, bmax, X @ n - 0.6)
    pts = np.concatenate([S.surface_samples(m, 16, seed=i) for i, m in enumerate(spec.link_meshes)])
    ppl = np.full(7, 16, np.int32)
    Q = S.trajectories(spec, 4, 8, seed=1).astype(np.float64) * 0.3
    loss, g = w.sdf_penalty(r, grid, Q, pts, ppl, dsafe=0.5, grad=True)
    _, gfd = w.sdf_penalty(r, grid, Q, pts, ppl, dsafe=0.5, grad=True, fd_h=1e-4)
    active = loss > 0
    assert active.sum() >= 8
    scale = np.abs(g).max(axis=1, keepdims=True) + 1e-300
    assert np.all(np.abs(g - gfd)[active] / scale[active] < 1e-5)
