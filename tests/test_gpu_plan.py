"""Robot C-space and planners on the GPU engine: RobotCSpace feasibility and
visibility equal the FP64 oracle (and the golden edge fixture), and the
planned paths pass the oracle's own edge check."""
import numpy as np
import pytest

import fixtures_io as F
import oracle
from klampt_amd import plan, robotsim, synthetic as S

pytestmark = pytest.mark.gpu
BAND = 1e-6


def _robotsim_world(sc):
    W = robotsim.WorldModel()
    for t in sc.terrains:
        W.addTerrain(robotsim.Geometry3D(robotsim.TriangleMesh(t.verts, t.tris)))
    for mesh, xf in sc.objects:
        W.addRigidObject(robotsim.Geometry3D(robotsim.TriangleMesh(mesh.verts, mesh.tris)), xf[:9], xf[9:])
    R = robotsim.RobotModel.from_spec(sc.robot)
    W.addRobot(R)
    return W, R


@pytest.fixture(scope="module")
def boxes():
    sc = S.boxes_scene(0)
    W, R = _robotsim_world(sc)
    ow, orb = oracle.world_from_scene(sc)
    return sc, W, R, ow, orb


def test_robot_cspace_feasibility_matches_oracle(boxes):
    sc, W, R, ow, orb = boxes
    space = plan.RobotCSpace(R, W)
    Q = S.uniform_configs(sc.robot, 2048, seed=61)
    got = space.feasible_batch(Q)
    want = ow.feasible_batch(orb, Q)[0].astype(bool)
    bad = np.nonzero(got != want)[0]
    if len(bad):
        sep = ow.distance_batch(orb, Q[bad], signed=True)
        assert (np.abs(sep).min(axis=1) < BAND).all()
    assert space.feasible(Q[int(np.argmax(want))]) and not space.feasible(Q[int(np.argmin(want))])


def test_robot_cspace_visibility_matches_golden_edges():
    z = F.load("c1_edges.npz")
    sc = F.unpack_scene(z)
    W, R = _robotsim_world(sc)
    space = plan.RobotCSpace(R, W, eps=float(z["l2_eps"]))
    vis = space.visible_batch(z["A"], z["B"])
    assert np.array_equal(vis, z["l2_first_u"] < 0)


def test_added_constraint_is_applied(boxes):
    sc, W, R, ow, orb = boxes
    space = plan.RobotCSpace(R, W)
    Q = S.uniform_configs(sc.robot, 256, seed=62)
    base = space.feasible_batch(Q)
    space.addConstraint(lambda q: q[0] > 0.0, "q0 positive")
    assert np.array_equal(space.feasible_batch(Q), base & (Q[:, 0] > 0))


@pytest.mark.parametrize("ptype", ["prm", "lazyprm", "prm*", "lazyprm*", "rrt"])
def test_planned_path_passes_the_oracle(boxes, ptype):
    sc, W, R, ow, orb = boxes
    space = plan.RobotCSpace(R, W, eps=0.02)
    space.rng = np.random.default_rng(7)
    Q = S.uniform_configs(sc.robot, 4096, seed=63)
    feas = ow.feasible_batch(orb, Q)[0].astype(bool)
    Qf = Q[feas]
    # two feasible configurations 1.5-3 rad apart whose straight segment is
    # blocked (the first such pair in sample order)
    d = np.linalg.norm(Qf[:256, None] - Qf[None, :256], axis=-1)
    cand = np.argwhere((d > 1.5) & (d < 3.0))
    cand = cand[cand[:, 0] < cand[:, 1]]
    blocked = ~space.visible_batch(Qf[cand[:, 0]], Qf[cand[:, 1]])
    i, j = cand[int(np.argmax(blocked))]
    assert blocked.any()
    start, goal = Qf[i].astype(float), Qf[j].astype(float)
    opts = {"perturbationRadius": 0.5} if ptype == "rrt" else ({} if ptype.endswith("*") else {"knn": 15})
    mp = plan.MotionPlan(space, type=ptype, **opts)
    mp.setEndpoints(start, goal)
    for _ in range(8 if ptype != "rrt" else 64):
        mp.planMore(4096 if ptype != "rrt" else 512)
        if mp.getPath():
            break
    path = mp.getPath()
    st = mp.getStats()
    print(f"{ptype}: milestones {st['milestones']}, edges {st['edges']}, checked {st['edges_checked']} "
          f"in {st['check_calls']} calls, {st['time_s']:.2f} s, path {len(path) if path else None} milestones")
    assert path is not None
    P = np.asarray(path)
    assert np.allclose(P[0], start) and np.allclose(P[-1], goal)
    assert ow.feasible_batch(orb, P)[0].all()
    fu, _, _ = ow.edge_batch(orb, P[:-1], P[1:], eps=0.02, norm_inf=False)
    assert (fu < 0).all()


def test_gpu_nearest_neighbours_equal_kd_tree():
    """The roadmap's k-nearest-neighbour query on the GPU (exact FP64
    distances) returns the k-d tree's neighbours and distances."""
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(3)
    V = rng.uniform(-3, 3, (6400, 7))
    Q = V[rng.choice(len(V), 3000, replace=False)]
    d, j = plan._knn(Q, V, 11)
    d2, j2 = cKDTree(V).query(Q, k=11)
    assert np.allclose(d, d2, rtol=0, atol=1e-12)
    assert np.array_equal(j, j2)
