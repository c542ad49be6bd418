"""Planner layer on a CPU-only C-space (a disk obstacle in the unit square):
the C-space semantics of ``Python/klampt/plan/cspace.py`` (named feasibility
tests, bisection visibility at eps) and the batched PRM / lazy PRM
planners.  The robot C-space on the GPU is covered by tests/test_gpu_plan.py."""
import math

import numpy as np
import pytest

import oracle
from klampt_amd import plan


class Disk(plan.CSpace):
    def __init__(self, seed=0):
        super().__init__()
        self.setBounds([(0.0, 1.0), (0.0, 1.0)])
        self.eps = 1e-2
        self.rng = np.random.default_rng(seed)
        self.addFeasibilityTest(lambda x: (x[0] - 0.5) ** 2 + (x[1] - 0.5) ** 2 > 0.3 ** 2, "disk")


def test_bisection_order_matches_the_edge_checker():
    for n in range(0, 70):
        got = plan._bisection_order(n)
        assert sorted(got) == list(range(1, n)), n
        assert list(got) == list(oracle.bisection_order(n)), n


def test_visible_equals_a_scan_of_all_samples():
    sp = Disk(1)
    rng = np.random.default_rng(2)
    A, B = rng.uniform(0, 1, (200, 2)), rng.uniform(0, 1, (200, 2))
    got = sp.visible_batch(A, B)
    for s in range(200):
        n = int(math.ceil(np.linalg.norm(B[s] - A[s]) / sp.eps))
        want = all(sp.feasible(list(A[s] + (j / n) * (B[s] - A[s]))) for j in range(1, n))
        assert got[s] == want, s
    assert got.any() and not got.all()


def test_feasibility_tests_and_bounds():
    sp = Disk()
    assert sp.feasibilityTestNames == ["disk"]
    assert not sp.feasible([0.5, 0.5]) and sp.feasible([0.05, 0.05])
    assert sp.inBounds([0.0, 1.0]) and not sp.inBounds([1.0001, 0.5])
    assert sp.properties["volume"] == 1.0 and sp.properties["minimum"] == [0.0, 0.0]
    with pytest.raises(TypeError):
        sp.addFeasibilityTest(3)


@pytest.mark.parametrize("ptype", ["prm", "lazyprm", "prm*", "lazyprm*", "rrt"])
def test_planner_path_is_feasible(ptype):
    sp = Disk(3)
    start, goal = [0.1, 0.5], [0.9, 0.5]
    assert not sp.visible(start, goal)
    opts = {"perturbationRadius": 0.15} if ptype == "rrt" else {"knn": 8}
    mp = plan.MotionPlan(sp, type=ptype, **opts)
    mp.setEndpoints(start, goal)
    for _ in range(10):
        mp.planMore(300)
        if mp.getPath():
            break
    path = mp.getPath()
    assert path is not None and path[0] == start and path[-1] == goal
    assert all(sp.feasible(q) for q in path)
    assert sp.visible_batch(np.asarray(path[:-1]), np.asarray(path[1:])).all()
    # around the disk: longer than the straight line, shorter than the square's perimeter
    assert 0.8 < mp.pathCost(path) < 2.0
    V, E = mp.getRoadmap()
    assert len(V) == mp.getStats()["milestones"] and all(0 <= i < len(V) and 0 <= j < len(V) for i, j in E)


def test_lazy_checks_fewer_edges():
    stats = {}
    for ptype in ("prm", "lazyprm"):
        sp = Disk(4)
        mp = plan.MotionPlan(sp, type=ptype, knn=8)
        mp.setEndpoints([0.1, 0.5], [0.9, 0.5])
        mp.planMore(600)
        assert mp.getPath() is not None
        stats[ptype] = mp.getStats()
    assert stats["lazyprm"]["edges_checked"] < stats["prm"]["edges_checked"] / 3


def test_errors():
    sp = Disk()
    with pytest.raises(RuntimeError):
        plan.MotionPlan(sp).setEndpoints([0.5, 0.5], [0.9, 0.5])
    with pytest.raises(ValueError):
        plan.MotionPlan(sp, type="rrt*")
    with pytest.raises(ValueError):
        plan.MotionPlan(sp, bogus=1)
    with pytest.raises(RuntimeError):
        plan.MotionPlan(sp).planMore(10)


def test_prm_star_neighbour_count_grows_with_the_roadmap():
    sp = Disk(5)
    mp = plan.MotionPlan(sp, type="prm*")
    mp.setEndpoints([0.1, 0.5], [0.9, 0.5])
    k0 = mp._k()
    mp.planMore(400)
    k1 = mp._k()
    n, d = len(mp.V), 2
    assert k1 == math.ceil(math.e * 1.5 * math.log(n)) and k1 > k0


@pytest.mark.parametrize("bidirectional", [1, 0])
def test_rrt_tree_edges_are_short_and_valid(bidirectional):
    sp = Disk(6)
    mp = plan.MotionPlan(sp, type="rrt", perturbationRadius=0.1, bidirectional=bidirectional)
    mp.setEndpoints([0.1, 0.5], [0.9, 0.5])
    for _ in range(40):
        mp.planMore(64)
        if mp.getPath():
            break
    path = mp.getPath()
    assert path is not None
    P = np.asarray(path)
    assert np.all(np.linalg.norm(P[1:] - P[:-1], axis=1) <= 0.1 + 1e-9)
    assert sp.visible_batch(P[:-1], P[1:]).all() and all(sp.feasible(q) for q in path)
    # every tree node's parent is within the radius and in the same tree
    par = mp.parent
    for i in range(2, len(mp.V)):
        assert mp.tree[i] == mp.tree[par[i]] and np.linalg.norm(mp.V[i] - mp.V[par[i]]) <= 0.1 + 1e-9
    if not bidirectional:
        assert np.all(mp.tree[2:] == 0)
    with pytest.raises(ValueError):
        mp.addMilestone([0.2, 0.2])


def test_edge_store_dedup_find_and_states():
    """The roadmap's edge arrays: a batch with repeats and edges already
    present adds each new edge once, unchecked; find maps pairs to indices
    (-1 when absent); set adds or overwrites states; the dict view follows."""
    E = plan._EdgeStore()
    added = E.put(np.array([0, 1, 0, 2]), np.array([5, 3, 5, 4]), -1)
    assert len(E) == 3 and len(added) == 3
    again = E.put(np.array([1, 6]), np.array([3, 7]), -1)
    assert len(again) == 1 and len(E) == 4
    assert E.find(np.array([0, 2, 1, 3]), np.array([5, 4, 3, 9])).tolist()[3] == -1
    idx = E.find(np.array([0, 2]), np.array([5, 4]))
    assert (idx >= 0).all() and E.a[idx].tolist() == [0, 2] and E.b[idx].tolist() == [5, 4]
    E.set(np.array([0, 8]), np.array([5, 9]), np.array([True, False]))
    d = E.as_dict()
    assert d[(0, 5)] is True and d[(8, 9)] is False and d[(1, 3)] is None and len(d) == 5
    big = 1 << 20
    E.put(np.array([big]), np.array([big + 1]), 1)
    assert E.find(np.array([big]), np.array([big + 1]))[0] == 5
