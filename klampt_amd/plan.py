"""Planner integration: Klampt's CSpace, RobotCSpace and MotionPlan shapes on
the batched feasibility and edge-check engine.

Mirrors
* ``klampt.plan.cspace.CSpace`` (``Python/klampt/plan/cspace.py:12-215``):
  bounds, samplers, named feasibility tests, ``inBounds``, ``feasible``,
  ``isFeasible``, ``isVisible`` (subdivision at ``eps``), ``distance``,
  ``interpolate``, ``getStats``;
* ``klampt.plan.robotcspace.RobotCSpace``
  (``Python/klampt/plan/robotcspace.py:11-130``): joint limits, self and
  environment collision (``SingleRobotCSpace::IsFeasible``,
  ``Cpp/Planning/RobotCSpace.cpp:786-823``), ``addConstraint``;
* ``klampt.plan.cspace.MotionPlan`` (``cspace.py:216-440``): ``setEndpoints``,
  ``planMore``, ``getPath``, ``getRoadmap``, ``getStats``, ``pathCost``.

What is new here is the batched form of the two queries planners make,
``feasible_batch(Q)`` and ``visible_batch(A, B)``, and planners that use them:
``'prm'`` validates every candidate edge of a batch of milestones in one call;
``'lazyprm'`` (Bohlin and Kavraki's lazy PRM) searches the roadmap of
unvalidated edges and validates only the edges of the best path, all of them
in one call per round, until a path survives; ``'prm*'`` and ``'lazyprm*'``
connect each milestone to its k(n) = ceil(e (1 + 1/d) ln n) nearest
neighbours (Karaman and Frazzoli's PRM* rule, d the dimension) as the roadmap
grows; ``'rrt'`` grows one tree from the start and one from the goal
(``bidirectional``, Klampt's default) by a batch of extensions per
``planMore`` call -- every new node's feasibility in one call, every new edge
in one call, the connection attempts between the trees in one call.
"""
import math
import time

import numpy as np
# the roadmap's shortest-path search (loaded with the module, not on the
# first path query)
from scipy.sparse import csr_matrix
from scipy.sparse.csgraph import dijkstra

from . import robotsim


class CSpace:
    """A configuration space (``Python/klampt/plan/cspace.py:12-215``).
    ``feasible`` runs the added tests in order (bounds only when there are
    none); ``visible`` checks the straight segment at resolution ``eps`` in
    Klampt's bisection order (``Python/klampt/plan/robotcspace.py:216-228``
    discretisation, ``n = ceil(d / eps)`` samples at ``u = i / n``)."""

    def __init__(self):
        self.eps = 1e-3
        self.bound = [(0.0, 1.0)]
        self.properties = {}
        self.feasibilityTests = None
        self.feasibilityTestNames = None
        self.feasibilityTestDependencies = None
        self.rng = np.random.default_rng()
        self._stats = {"feasible_calls": 0, "feasible_configs": 0, "visible_calls": 0, "visible_edges": 0}

    def setBounds(self, bound):
        self.bound = [(float(a), float(b)) for a, b in bound]
        self.properties["minimum"] = [b[0] for b in self.bound]
        self.properties["maximum"] = [b[1] for b in self.bound]
        volume = 1.0
        for a, b in self.bound:
            if a != b:
                volume *= b - a
        self.properties["volume"] = volume

    def close(self):
        pass

    def sample(self):
        lo = np.array([b[0] for b in self.bound])
        hi = np.array([b[1] for b in self.bound])
        return list(self.rng.uniform(lo, hi))

    def sampleneighborhood(self, c, r):
        return [float(self.rng.uniform(max(b[0], ci - r), min(b[1], ci + r))) for ci, b in zip(c, self.bound)]

    def sample_batch(self, n):
        """n samples of ``sample()`` (uniform in the bounds unless overridden)."""
        lo = np.array([b[0] for b in self.bound])
        hi = np.array([b[1] for b in self.bound])
        return self.rng.uniform(lo, hi, size=(n, len(self.bound)))

    def addFeasibilityTest(self, func, name=None, dependencies=None):
        if self.feasibilityTests is None:
            self.feasibilityTests, self.feasibilityTestNames, self.feasibilityTestDependencies = [], [], []
        if name is not None and not isinstance(name, str):
            raise TypeError("feasibility test name must be a string")
        if not callable(func):
            raise TypeError("feasibility test must be callable")
        self.feasibilityTests.append(func)
        self.feasibilityTestNames.append(name if name is not None else f"test_{len(self.feasibilityTests) - 1}")
        if dependencies is not None:
            deps = dependencies if isinstance(dependencies, (list, tuple)) else [dependencies]
            for d in deps:
                self.feasibilityTestDependencies.append((self.feasibilityTestNames[-1], d))

    def inBounds(self, x):
        return all(a <= xi <= b for xi, (a, b) in zip(x, self.bound))

    def feasible(self, x):
        if self.feasibilityTests is None:
            return self.inBounds(x)
        return all(test(x) for test in self.feasibilityTests)

    def feasible_batch(self, Q):
        """Feasibility of each row of Q (bool array)."""
        return np.array([bool(self.feasible(list(q))) for q in np.asarray(Q, float)], bool)

    def isFeasible(self, x):
        self._stats["feasible_calls"] += 1
        self._stats["feasible_configs"] += 1
        return bool(self.feasible(x))

    def distance(self, a, b):
        return float(np.linalg.norm(np.asarray(b, float) - np.asarray(a, float)))

    def interpolate(self, a, b, u):
        a = np.asarray(a, float)
        return list(a + u * (np.asarray(b, float) - a))

    def visible(self, a, b):
        return bool(self.visible_batch(np.asarray([a], float), np.asarray([b], float))[0])

    def visible_batch(self, A, B):
        """Whether each segment A[i] -> B[i] is feasible at its interior
        samples u = j / n, j = 1..n-1, n = ceil(distance / eps), tested in
        Klampt's bisection order with an early exit (the endpoints are
        milestones and not re-tested)."""
        out = np.ones(len(A), bool)
        for s, (a, b) in enumerate(zip(np.asarray(A, float), np.asarray(B, float))):
            n = int(math.ceil(self.distance(a, b) / self.eps))
            for j in _bisection_order(n):
                if not self.feasible(self.interpolate(a, b, j / n)):
                    out[s] = False
                    break
        return out

    def isVisible(self, a, b):
        self._stats["visible_calls"] += 1
        self._stats["visible_edges"] += 1
        return self.visible(a, b)

    def getStats(self):
        return dict(self._stats)


def _bisection_order(n):
    """Interior sample indices 1..n-1 in level-order bisection (midpoints of
    ever finer halvings), the order the batched edge checker uses."""
    if n < 2:
        return []
    out, seen = [], set()
    level = [(0, n)]
    while level:
        nxt = []
        for lo, hi in level:
            if hi - lo < 2:
                continue
            mid = (lo + hi) // 2
            if mid not in seen:
                seen.add(mid)
                out.append(mid)
            nxt.append((lo, mid))
            nxt.append((mid, hi))
        level = nxt
    return out


class RobotCSpace(CSpace):
    """Collision-free motion of one robot of a world
    (``Python/klampt/plan/robotcspace.py:11-130``): joint limits (closed
    intervals), self collision and collision with the world's terrains,
    rigid objects and other robots, per Klampt's default collision mask.
    Both queries run batched on the GPU; constraints added with
    ``addConstraint`` are applied on top, on the host."""

    def __init__(self, robot: robotsim.RobotModel, world: robotsim.WorldModel, eps: float = 1e-3):
        super().__init__()
        self.robot, self.world = robot, world
        self._c = robotsim.RobotCSpace(robot, world, eps=eps, norm_inf=False)
        self.eps = eps
        qmin, qmax = robot.getJointLimits()
        self.setBounds(list(zip(qmin, qmax)))
        self.properties["geodesic"] = 1
        self._extra = []

    def addConstraint(self, checker, name=None):
        """An extra feasibility test f(q) -> bool, applied after the
        collision checks (``robotcspace.py:73-74``)."""
        if not callable(checker):
            raise TypeError("constraint must be callable")
        self._extra.append((checker, name or f"constraint_{len(self._extra)}"))

    def inJointLimits(self, x):
        return self.inBounds(x)

    def feasible(self, x):
        return bool(self.feasible_batch(np.asarray([x], float))[0])

    def feasible_batch(self, Q):
        # double configurations are checked as given (joint limits included)
        Q = np.asarray(Q, np.float64).reshape(-1, len(self.bound))
        self._stats["feasible_configs"] += len(Q)
        ok = self._c.feasible_batch(Q)[0].astype(bool)
        for f, _ in self._extra:
            for i in np.nonzero(ok)[0]:
                ok[i] = bool(f(list(map(float, Q[i]))))
        return ok

    def selfCollision(self, x=None):
        x = self.robot.getConfig() if x is None else x
        _, env, slf, _ = self._c.feasible_batch(np.asarray([x], np.float64), flags=True)
        return bool(slf[0])

    def envCollision(self, x=None):
        x = self.robot.getConfig() if x is None else x
        _, env, slf, _ = self._c.feasible_batch(np.asarray([x], np.float64), flags=True)
        return bool(env[0])

    def visible_batch(self, A, B):
        A = np.asarray(A, np.float32).reshape(-1, len(self.bound))
        B = np.asarray(B, np.float32).reshape(-1, len(self.bound))
        # endpoints on a joint limit stay inside it after the float32 cast
        A, B = self._f32_inside(A), self._f32_inside(B)
        self._stats["visible_edges"] += len(A)
        if len(A) == 0:
            return np.zeros(0, bool)
        ok = self._c.edge_batch(A, B) < 0
        if self._extra:
            ok &= CSpace.visible_batch(self, A, B)
        return ok

    def _f32_inside(self, X):
        """float32 values clamped to the float32 numbers inside the double
        joint limits (a double on a limit can round outside it)."""
        lo = np.array([b[0] for b in self.bound], np.float64)
        hi = np.array([b[1] for b in self.bound], np.float64)
        lo32, hi32 = lo.astype(np.float32), hi.astype(np.float32)
        lo32 = np.where(lo32.astype(np.float64) < lo, np.nextafter(lo32, np.float32(np.inf)), lo32)
        hi32 = np.where(hi32.astype(np.float64) > hi, np.nextafter(hi32, np.float32(-np.inf)), hi32)
        return np.clip(X, lo32, hi32).astype(np.float32)

    def collision_pairs(self, Q, dcheck):
        """The point-collision terms of a trajectory optimiser over
        waypoints Q (``KineTrajOpt.find_collision_pair``,
        ``Python/klampt/plan/kinetrajopt/kinetrajopt.py:530-556``), in one
        batched call: for every waypoint and every link / obstacle (or link
        / link) pair within dcheck, the distance, the closest points and the
        gradient of the distance in joint space (the normal times the
        position Jacobian).  Returns a list of (waypoint, (id_a, id_b), d,
        cp_a, cp_b, grad) in waypoint and pair order; world IDs as
        ``pair_list``."""
        Q = np.asarray(Q, np.float32).reshape(-1, len(self.bound))
        w = self.world._w
        d, cp, g = w.distance_jacobian_batch(Q, dcheck, robot=self._c.index)
        ids = w.pair_list(self._c.index)
        out = []
        for c, p in zip(*np.nonzero(np.isfinite(d))):
            out.append((int(c), tuple(int(x) for x in ids[p]), float(d[c, p]), cp[c, p, :3].astype(float),
                        cp[c, p, 3:].astype(float), g[c, p].astype(float)))
        return out


def _knn(Q, V, k):
    """Distances and indices of the k nearest rows of V (Euclidean) for each
    row of Q, nearest first: exact FP64 distances on the GPU through torch
    when one is present (in row blocks of bounded size), a k-d tree on the
    host otherwise."""
    try:
        import torch
        gpu = torch.cuda.is_available()
    except ImportError:
        gpu = False
    if gpu and len(V) * len(Q) > 1 << 16:
        # squared differences summed per coordinate, as the k-d tree does
        # (torch.cdist's direct mode gave rows of other query points on
        # 3,000 x 6,393 FP64 inputs here)
        dV = torch.from_numpy(np.ascontiguousarray(V, np.float64)).cuda()
        rows = max(1, (1 << 24) // max(1, len(V) * V.shape[1]))
        ds, js = [], []
        for r0 in range(0, len(Q), rows):
            dQ = torch.from_numpy(np.ascontiguousarray(Q[r0:r0 + rows], np.float64)).cuda()
            D2 = ((dQ[:, None, :] - dV[None, :, :]) ** 2).sum(dim=2)
            d2, j = torch.topk(D2, k, dim=1, largest=False, sorted=True)
            ds.append(torch.sqrt(d2).cpu().numpy())
            js.append(j.cpu().numpy())
        return np.concatenate(ds), np.concatenate(js)
    from scipy.spatial import cKDTree
    d, j = cKDTree(V).query(Q, k=k)
    return np.asarray(d).reshape(len(Q), -1), np.asarray(j).reshape(len(Q), -1)


class _EdgeStore:
    """Roadmap edges (i < j) as arrays: the endpoints and a state per edge
    (-1 not checked, 0 blocked, 1 valid), with their keys i * 2^32 + j kept
    sorted for membership tests, so that a batch of candidate links is
    added, deduplicated and validated without a Python loop per edge."""

    def __init__(self):
        self.a = np.zeros(0, np.int64)
        self.b = np.zeros(0, np.int64)
        self.state = np.zeros(0, np.int8)
        self._keys = np.zeros(0, np.int64)   # sorted
        self._pos = np.zeros(0, np.int64)    # edge index of each sorted key

    def __len__(self):
        return len(self.a)

    def find(self, a, b):
        """Edge index of each (a, b), a < b, or -1."""
        k = (np.asarray(a, np.int64) << 32) | np.asarray(b, np.int64)
        i = np.searchsorted(self._keys, k)
        i = np.minimum(i, max(0, len(self._keys) - 1))
        hit = (len(self._keys) > 0) & (self._keys[i] == k) if len(self._keys) else np.zeros(len(k), bool)
        return np.where(hit, self._pos[i] if len(self._pos) else -1, -1)

    def put(self, a, b, state):
        """Adds the edges (a, b), a < b, not present yet, with `state`
        (edges already present keep theirs); returns the indices of the
        edges added."""
        a = np.asarray(a, np.int64)
        b = np.asarray(b, np.int64)
        k = np.unique((a << 32) | b)
        new = k[self.find(k >> 32, k & 0xffffffff) < 0] if len(self._keys) else k
        n0 = len(self.a)
        idx = np.arange(n0, n0 + len(new))
        self.a = np.r_[self.a, new >> 32]
        self.b = np.r_[self.b, new & 0xffffffff]
        self.state = np.r_[self.state, np.full(len(new), state, np.int8)]
        keys = np.r_[self._keys, new]
        pos = np.r_[self._pos, idx]
        order = np.argsort(keys, kind="stable")
        self._keys, self._pos = keys[order], pos[order]
        return idx

    def set(self, a, b, state):
        """Adds or overwrites the edges (a, b), a < b, with `state` (one
        state per edge)."""
        state = np.broadcast_to(np.asarray(state, np.int8), np.shape(a))
        self.put(a, b, -1)
        self.state[self.find(a, b)] = state

    def as_dict(self):
        val = {-1: None, 0: False, 1: True}
        return {(int(i), int(j)): val[int(v)] for i, j, v in zip(self.a, self.b, self.state)}


class MotionPlan:
    """A planner over a CSpace (``Python/klampt/plan/cspace.py:216-440``).

    type 'prm': each ``planMore(n)`` draws n configurations in one batch,
    keeps the feasible ones, links each new milestone to its ``knn`` nearest
    milestones (within ``connectionThreshold``) and validates all those edges
    in one batch.  type 'lazyprm': the same links are kept unvalidated; after
    each batch the shortest start-goal path over the links not known to be
    blocked is found and its unvalidated edges are checked together, until
    a path survives or none remains.  'prm*' / 'lazyprm*': the same with
    k(n) = ceil(e (1 + 1/d) ln n) neighbours.  'rrt': each ``planMore(n)``
    draws n targets, extends the nearest tree node towards each by at most
    ``perturbationRadius`` (the start and goal trees alternate when
    ``bidirectional``), checks the new nodes and then their edges in one call
    each, and tries to link every new node to the other tree's nearest node
    within ``perturbationRadius`` in one more call; the first link found
    gives the path.
    """

    TYPES = ("prm", "lazyprm", "prm*", "lazyprm*", "rrt")

    def __init__(self, space: CSpace, type: str = None, **options):
        self.space = space
        self.type = type or "lazyprm"
        if self.type not in self.TYPES:
            raise ValueError(f"unsupported planner type {self.type!r} ({', '.join(self.TYPES)})")
        self.knn = int(options.pop("knn", 10))
        self.connectionThreshold = float(options.pop("connectionThreshold", math.inf))
        # RRT: the longest extension (Klampt's perturbationRadius) and whether
        # a second tree grows from the goal
        self.perturbationRadius = float(options.pop("perturbationRadius", 0.1))
        self.bidirectional = bool(int(options.pop("bidirectional", 1)))
        if options:
            raise ValueError(f"unknown planner options {sorted(options)}")
        self.lazy = self.type in ("lazyprm", "lazyprm*")
        self.star = self.type.endswith("*")
        self.rrt = self.type == "rrt"
        self.parent = None       # RRT: parent milestone index (-1 at a root)
        self.tree = None         # RRT: 0 start tree, 1 goal tree
        self.V = np.zeros((0, len(space.bound)))
        self._E = _EdgeStore()
        self.start = self.goal = None
        self.path_idx = None
        self._path_stale = False  # PRM: roadmap grown since the last path search
        self._stats = {"milestones": 0, "samples": 0, "edges_checked": 0, "check_calls": 0, "time_s": 0.0}

    def close(self):
        pass

    @property
    def edges(self):
        """The roadmap's edges: (i, j), i < j -> True (valid), False
        (blocked), None (not checked)."""
        return self._E.as_dict()

    def setEndpoints(self, start, goal):
        if not callable(goal) and not isinstance(goal, tuple):
            ok = self.space.feasible_batch(np.asarray([start, goal], float))
            if not ok[0]:
                raise RuntimeError("Start configuration is infeasible")
            if not ok[1]:
                raise RuntimeError("Goal configuration is infeasible")
        else:
            raise ValueError("goal sets are not supported: pass a configuration")
        self.V = np.asarray([start, goal], float)
        self.start, self.goal = 0, 1
        self._E = _EdgeStore()
        self.path_idx = None
        self._path_stale = False
        if self.rrt:
            self.parent = np.array([-1, -1])
            self.tree = np.array([0, 1])
            self._rrt_try_connect(np.array([0]))
            return
        self._connect(np.array([0, 1]))

    def addMilestone(self, x):
        if self.rrt:
            raise ValueError("the 'rrt' planner does not take milestones")
        self.V = np.vstack([self.V, np.asarray(x, float)[None]])
        k = len(self.V) - 1
        self._connect(np.array([k]))
        return k

    def _k(self):
        """Neighbours per milestone: knn, or the PRM* rule for the '*' types."""
        if not self.star:
            return self.knn
        d = self.V.shape[1]
        return max(1, int(math.ceil(math.e * (1.0 + 1.0 / d) * math.log(max(2, len(self.V))))))

    def _connect(self, new):
        """Candidate links of the milestones in `new` to their nearest
        milestones (all milestones, Euclidean)."""
        if len(self.V) < 2:
            return
        k = min(self._k() + 1, len(self.V))
        d, nb = _knn(self.V[new], self.V, k)
        i = np.repeat(np.asarray(new, np.int64), nb.shape[1])
        j = nb.ravel().astype(np.int64)
        keep = (j != i) & (d.ravel() <= self.connectionThreshold) & (j < len(self.V))
        i, j = i[keep], j[keep]
        cand = self._E.put(np.minimum(i, j), np.maximum(i, j), -1)
        if not self.lazy:
            self._validate(cand)

    def _validate(self, idx):
        """Checks the not yet checked edges among the edge indices `idx`
        in one batched call."""
        idx = np.asarray(idx, np.int64)
        idx = idx[self._E.state[idx] < 0]
        if len(idx) == 0:
            return
        ok = self.space.visible_batch(self.V[self._E.a[idx]], self.V[self._E.b[idx]])
        self._E.state[idx] = np.asarray(ok, bool).astype(np.int8)
        self._stats["edges_checked"] += len(idx)
        self._stats["check_calls"] += 1

    def _steer(self, near, target):
        """Configurations from `near` towards `target`, at most
        perturbationRadius away (row-wise)."""
        d = target - near
        L = np.linalg.norm(d, axis=1)
        scale = np.minimum(1.0, self.perturbationRadius / np.maximum(L, 1e-300))
        return near + d * scale[:, None]

    def _rrt_try_connect(self, new):
        """Links each node in `new` to the nearest node of the other tree
        when that one is within perturbationRadius (goal node for a
        unidirectional plan); all candidate links checked in one call."""
        if self.path_idx is not None or len(new) == 0:
            return
        pairs = []
        for t in (0, 1):
            mine = new[self.tree[new] == t]
            other = np.nonzero(self.tree == 1 - t)[0]
            if len(mine) == 0 or len(other) == 0:
                continue
            d, j = _knn(self.V[mine], self.V[other], 1)
            for a, dist, jj in zip(mine, d.ravel(), j.ravel()):
                if dist <= self.perturbationRadius:
                    pairs.append((int(a), int(other[jj])))
        if not pairs:
            return
        P = np.asarray(pairs)
        ok = self.space.visible_batch(self.V[P[:, 0]], self.V[P[:, 1]])
        self._stats["edges_checked"] += len(P)
        self._stats["check_calls"] += 1
        self._E.set(np.minimum(P[:, 0], P[:, 1]), np.maximum(P[:, 0], P[:, 1]), np.asarray(ok, bool))
        hits = [p for p, v in zip(pairs, ok) if v]
        if hits:
            def root_path(i):
                out = [i]
                while self.parent[out[-1]] >= 0:
                    out.append(int(self.parent[out[-1]]))
                return out
            best = None
            for a, b in hits:
                pa, pb = root_path(a), root_path(b)
                ps, pg = (pa, pb) if self.tree[a] == 0 else (pb, pa)
                cand = ps[::-1] + pg
                cost = sum(np.linalg.norm(self.V[x] - self.V[y]) for x, y in zip(cand[:-1], cand[1:]))
                if best is None or cost < best[0]:
                    best = (cost, cand)
            self.path_idx = best[1]

    def _rrt_more(self, iterations):
        """One batch of RRT extensions: `iterations` targets drawn at once,
        each extends the nearest node of the tree it is given to (the trees
        take turns when bidirectional), feasibility and edge checks batched."""
        n = int(iterations)
        Xt = self.space.sample_batch(n)
        trees = np.arange(n) % 2 if self.bidirectional else np.zeros(n, int)
        near = np.zeros(n, int)
        for t in (0, 1):
            sel = np.nonzero(trees == t)[0]
            nodes = np.nonzero(self.tree == t)[0]
            if len(sel) == 0:
                continue
            _, j = _knn(Xt[sel], self.V[nodes], 1)
            near[sel] = nodes[j.ravel()]
        Xn = self._steer(self.V[near], Xt)
        ok = self.space.feasible_batch(Xn)
        self._stats["samples"] += n
        idx = np.nonzero(ok)[0]
        if len(idx) == 0:
            return
        vis = self.space.visible_batch(self.V[near[idx]], Xn[idx])
        self._stats["edges_checked"] += len(idx)
        self._stats["check_calls"] += 1
        keep = idx[vis]
        if len(keep) == 0:
            return
        first = len(self.V)
        self.V = np.vstack([self.V, Xn[keep]])
        self.parent = np.r_[self.parent, near[keep]]
        self.tree = np.r_[self.tree, trees[keep] if self.bidirectional else np.zeros(len(keep), int)]
        kk = np.arange(first, len(self.V))
        self._E.set(np.minimum(kk, near[keep]), np.maximum(kk, near[keep]), True)
        self._rrt_try_connect(np.arange(first, len(self.V)))

    def planMore(self, iterations):
        """Draws `iterations` configurations in one batch and extends the
        roadmap (or the trees); then looks for a path (validating lazily for
        the lazy types)."""
        if self.start is None:
            raise RuntimeError("setEndpoints must be called first")
        t = time.perf_counter()
        if self.rrt:
            if self.path_idx is None:
                self._rrt_more(iterations)
            self._stats["milestones"] = len(self.V)
            self._stats["time_s"] += time.perf_counter() - t
            return
        Q = self.space.sample_batch(int(iterations))
        ok = self.space.feasible_batch(Q)
        self._stats["samples"] += len(Q)
        Q = Q[ok]
        if len(Q):
            first = len(self.V)
            self.V = np.vstack([self.V, Q])
            self._connect(np.arange(first, len(self.V)))
        self._search()
        self._stats["milestones"] = len(self.V)
        self._stats["time_s"] += time.perf_counter() - t

    def _shortest(self, unchecked_ok):
        """Shortest start-goal path over the valid edges (and the unchecked
        ones when unchecked_ok), as milestone indices, or None."""
        sel = self._E.state != 0 if unchecked_ok else self._E.state == 1
        if not sel.any():
            return None
        E = np.stack([self._E.a[sel], self._E.b[sel]], axis=1)
        w = np.linalg.norm(self.V[E[:, 0]] - self.V[E[:, 1]], axis=1) + 1e-12
        n = len(self.V)
        G = csr_matrix((np.r_[w, w], (np.r_[E[:, 0], E[:, 1]], np.r_[E[:, 1], E[:, 0]])), shape=(n, n))
        dist, pred = dijkstra(G, indices=self.start, return_predecessors=True)
        if not np.isfinite(dist[self.goal]):
            return None
        path = [self.goal]
        while path[-1] != self.start:
            path.append(int(pred[path[-1]]))
        return path[::-1]

    def _search(self):
        if not self.lazy:
            # the shortest path over the valid links is searched when it is
            # asked for (getPath), once, not after every batch
            self._path_stale = True
            return
        while True:
            p = self._shortest(True)
            if p is None:
                self.path_idx = None
                return
            pa, pb = np.asarray(p[:-1]), np.asarray(p[1:])
            idx = self._E.find(np.minimum(pa, pb), np.maximum(pa, pb))
            unknown = idx[self._E.state[idx] < 0]
            if len(unknown) == 0:
                self.path_idx = p
                return
            self._validate(unknown)

    def getPath(self, milestone1=None, milestone2=None):
        if milestone1 is not None or milestone2 is not None:
            raise ValueError("paths between arbitrary milestones are not supported")
        if self._path_stale:
            self.path_idx = self._shortest(False)
            self._path_stale = False
        if self.path_idx is None:
            return None
        return [list(map(float, self.V[i])) for i in self.path_idx]

    def getRoadmap(self):
        V = [list(map(float, v)) for v in self.V]
        ok = self._E.state == 1
        E = [(int(i), int(j)) for i, j in zip(self._E.a[ok], self._E.b[ok])]
        return V, E

    def getStats(self):
        s = dict(self._stats)
        s["edges"] = len(self._E)
        s["edges_valid"] = int((self._E.state == 1).sum())
        return s

    def pathCost(self, path):
        return sum(self.space.distance(a, b) for a, b in zip(path[:-1], path[1:]))
