"""Batched pair distances with closest points and joint-space gradients
(kg_distance_jacobian_batch, KineTrajOpt's point-collision step) against the
FP64 oracle: distances within 1e-5 m inside the bound and +inf outside it,
closest points that realise the distance, and gradients equal to central
finite differences of the oracle's distance."""
import numpy as np
import pytest

import oracle
from klampt_amd import engine, robotsim, synthetic as S

pytestmark = pytest.mark.gpu
DCHECK = 0.3


@pytest.fixture(scope="module")
def scene():
    sc = S.boxes_scene(0)
    w, r = engine.world_from_scene(robotsim.default_context(), sc)
    ow, orb = oracle.world_from_scene(sc)
    return sc, w, ow, orb


def test_distances_and_closest_points(scene):
    sc, w, ow, orb = scene
    Q = S.uniform_configs(sc.robot, 512, seed=81)
    d, cp, g = w.distance_jacobian_batch(Q, DCHECK)
    want = ow.distance_batch(orb, Q)
    assert d.shape == want.shape
    inside = want < DCHECK - 1e-5
    outside = want > DCHECK + 1e-5
    assert inside.sum() > 100 and outside.sum() > 100
    assert np.abs(d[inside] - want[inside]).max() <= 1e-5
    assert np.isinf(d[outside]).all()
    sep = np.linalg.norm(cp[..., :3] - cp[..., 3:], axis=-1)
    fin = np.isfinite(d)
    assert np.abs(sep[fin] - d[fin]).max() <= 1e-5
    assert (d == 0).sum() > 0 and (sep[d == 0] == 0).all()
    assert (g[~fin] == 0).all() and (cp[~fin] == 0).all()


def test_gradients_match_finite_differences(scene):
    sc, w, ow, orb = scene
    Q = S.uniform_configs(sc.robot, 48, seed=82).astype(np.float64)
    d, cp, g = w.distance_jacobian_batch(Q.astype(np.float32), DCHECK)
    Q32 = Q.astype(np.float32).astype(np.float64)
    n = sc.robot.n
    h = 1e-5
    errs = []
    for c in range(len(Q32)):
        P = np.repeat(Q32[c:c + 1], 2 * n, axis=0)
        for j in range(n):
            P[2 * j, j] += h
            P[2 * j + 1, j] -= h
        D = ow.distance_batch(orb, P)
        fd = (D[0::2] - D[1::2]) / (2 * h)          # (n, npairs)
        for p in range(d.shape[1]):
            if 0.01 < d[c, p] < DCHECK - 0.01:
                errs.append(np.abs(g[c, p] - fd[:, p]).max())
    errs = np.array(errs)
    assert len(errs) > 50
    assert np.median(errs) < 1e-4
    assert np.mean(errs < 2e-3) >= 0.99, np.sort(errs)[-5:]


def test_cspace_collision_pairs():
    from klampt_amd import plan
    sc = S.boxes_scene(0)
    W = robotsim.WorldModel()
    for mesh, xf in sc.objects:
        W.addRigidObject(robotsim.Geometry3D(robotsim.TriangleMesh(mesh.verts, mesh.tris)), xf[:9], xf[9:])
    R = robotsim.RobotModel.from_spec(sc.robot)
    W.addRobot(R)
    space = plan.RobotCSpace(R, W)
    Q = S.uniform_configs(sc.robot, 32, seed=83)
    terms = space.collision_pairs(Q, DCHECK)
    d, cp, g = W._w.distance_jacobian_batch(Q, DCHECK)
    assert len(terms) == int(np.isfinite(d).sum()) > 0
    ids = W._w.pair_list(0)
    for c, pair, dist, pa, pb, grad in terms[:50]:
        p = [tuple(x) for x in ids.tolist()].index(pair)
        assert dist == d[c, p] and np.allclose(grad, g[c, p]) and abs(np.linalg.norm(pa - pb) - dist) < 1e-5


def test_link_position_jacobian_matches_finite_differences():
    spec = S.hetero_robots(seed=4, nrobots=1, dof_range=(14, 14), tri_range=(100, 200))[0][0]
    spec.jtype[3] = 1      # one prismatic joint in the chain
    R = robotsim.RobotModel.from_spec(spec)
    q = S.uniform_configs(spec, 1, seed=5)[0].astype(np.float64)
    R.setConfig(list(q))
    link = R.link(spec.n - 1)
    pl = [0.05, -0.02, 0.1]
    J = np.asarray(link.getPositionJacobian(pl))
    assert np.allclose(link.getLocalPosition(link.getWorldPosition(pl)), pl, atol=1e-12)
    h = 1e-6
    for j in range(spec.n):
        qp, qm = q.copy(), q.copy()
        qp[j] += h
        qm[j] -= h
        R.setConfig(list(qp))
        a = np.asarray(link.getWorldPosition(pl))
        R.setConfig(list(qm))
        b = np.asarray(link.getWorldPosition(pl))
        assert np.allclose(J[:, j], (a - b) / (2 * h), atol=1e-6), j


def test_min_distance_autograd_value_and_gradient(scene):
    """autodiff.MinDistance: per configuration the smallest pair distance
    (the oracle's, within 1e-5 m, or the bound when every pair is beyond it)
    and dL/dq equal to FP64 central differences of the oracle's minimum
    where the closest pair is unique and away from contact."""
    import torch
    from klampt_amd import autodiff
    sc, w, ow, orb = scene
    Q = S.uniform_configs(sc.robot, 96, seed=84)
    D = ow.distance_batch(orb, Q.astype(np.float64))
    # a bound at the median of the positive minima: about half of the
    # separated configurations lie beyond it
    dmin = D.min(axis=1)
    bound = float(np.median(dmin[dmin > 0]))
    md = autodiff.MinDistance(w, 0, upperBound=bound)
    Qt = torch.from_numpy(Q).cuda().requires_grad_(True)
    val = md(Qt)
    val.sum().backward()
    v = val.detach().cpu().numpy()
    G = Qt.grad.cpu().numpy()
    want = np.minimum(dmin, bound)
    assert np.abs(v - want).max() <= 1e-5
    beyond = dmin > bound + 1e-5
    print(f"bound {bound:.4f} m: {int(beyond.sum())} of {len(Q)} configurations beyond it")
    assert beyond.any() and (G[beyond] == 0).all()
    n, h = sc.robot.n, 1e-5
    Q32 = Q.astype(np.float64)
    checked = 0
    for c in range(len(Q)):
        srt = np.sort(D[c])
        if not (0.01 < srt[0] < bound - 0.01 and srt[1] - srt[0] > 1e-3):
            continue
        P = np.repeat(Q32[c:c + 1], 2 * n, axis=0)
        for j in range(n):
            P[2 * j, j] += h
            P[2 * j + 1, j] -= h
        M = ow.distance_batch(orb, P).min(axis=1)
        fd = (M[0::2] - M[1::2]) / (2 * h)
        assert np.abs(G[c] - fd).max() < 2e-3, c
        checked += 1
    assert checked >= 10
