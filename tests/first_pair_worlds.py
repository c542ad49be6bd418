"""Hand-built worlds whose first colliding pair depends on Klampt's broad-phase
order, with the expected pair worked out by hand.

``WorldPlannerSettings::CheckCollision(world, ids1, ids2)``
(Cpp/Planning/PlannerSettings.cpp:269-331) reduces set 2 (the environment)
against the AABB union of set 1 (the robot's links) and then set 1 against
the union of what is left of set 2; a rejected entry is replaced by the
current last entry (swap-erase, lines 292-317), and the entry at the same
index is examined again.  The pair loop is links outermost, environment
innermost, and the first colliding pair is returned (lines 320-328).  The
self check ``CheckCollision(world, ids)`` (lines 241-267) walks i < j over
the unreduced link list with the mask test (i,j) || (i,i).
``SingleRobotCSpace::CheckCollisionFree`` runs the environment check first
(Cpp/Planning/RobotCSpace.cpp:794-823).

Every link is a root link (parent -1) placed by its T0 translation, so at
q = 0 its world AABB is its box, and every environment entry is a rigid
object (link 0's terrain rule does not apply).  Boxes that overlap cross
transversally (no coplanar faces), so each collision is far from the
contact band.  Each world's docstring gives the derivation; the world IDs are
objects first, then the robot, then its links (Cpp/Modeling/World.cpp:47-53).
"""
from __future__ import annotations

import numpy as np

from klampt_amd import synthetic as S


def _robot(centres, halves, self_pairs=None):
    n = len(centres)
    T0 = np.array([np.r_[np.eye(3).ravel(), np.asarray(c, float)] for c in centres])
    return S.RobotSpec(
        name="kat", parents=np.full(n, -1, np.int32), T0=T0,
        axes=np.tile([0.0, 0.0, 1.0], (n, 1)), jtype=np.zeros(n, np.int32),
        qmin=np.full(n, -np.pi), qmax=np.full(n, np.pi),
        link_meshes=[S.box_mesh(h) for h in halves], self_pairs=self_pairs)


def _obj(centre, half):
    return (S.box_mesh(half), np.r_[np.eye(3).ravel(), np.asarray(centre, float)])


def world_env_swap():
    """Environment entries swap-erased.

    Objects O0 (5,0,0), O1 (0.15,0,0), O2 (0,0,5), O3 (-0.15,0,0), half
    extents (0.1,0.15,0.15); one link L0 at the origin, half 0.1.
    Env reduction against L0's box [-0.1,0.1]^3: i=0 O0 misses -> O3 moves to
    slot 0 ([O3,O1,O2]), O3 hits; i=1 O1 hits; i=2 O2 misses and is last ->
    [O3,O1].  L0 hits the union of O3 and O1.  Loop: L0 x O3 collides first.
    IDs: objects 0-3, robot 4, link 5.  Expected (5, 3); a stable erase or
    no reduction gives (5, 1)."""
    objs = [_obj((5, 0, 0), (.1, .15, .15)), _obj((.15, 0, 0), (.1, .15, .15)),
            _obj((0, 0, 5), (.1, .15, .15)), _obj((-.15, 0, 0), (.1, .15, .15))]
    return S.Scene(_robot([(0, 0, 0)], [(.1, .1, .1)]), [], objs), [], (5, 3)


def world_link_swap():
    """Link entries swap-erased.

    Object O0 at the origin, half 0.1.  Links L0 (5,0,0), L1 (0.12,0,0),
    L2 (0,0,0.6), L3 (-0.12,0,0), half 0.05.  The environment survives (it
    meets the links' union).  Link reduction against O0's box: i=0 L0 misses
    -> L3 to slot 0 ([L3,L1,L2]), L3 hits; i=1 L1 hits; i=2 L2 misses and is
    last -> [L3,L1].  Loop: L3 x O0 collides first.  No two links overlap.
    IDs: object 0, robot 1, links 2-5.  Expected (5, 0); a stable erase or
    no reduction gives (3, 0)."""
    rob = _robot([(5, 0, 0), (.12, 0, 0), (0, 0, .6), (-.12, 0, 0)], [(.05, .05, .05)] * 4)
    return S.Scene(rob, [], [_obj((0, 0, 0), (.1, .1, .1))]), [], (5, 0)


def world_reduction_order():
    """Environment reduced before the links.

    Objects O0 (5,0,0), O1 (0.1,0,0), O2 (-0.1,0,0), half (0.1,0.15,0.15);
    links L0 at the origin and L1 at (10,0,0), half 0.05.
    Klampt: env against the links' union x in [-0.05,10.05] (y,z within
    0.05): all three objects meet it -> [O0,O1,O2].  Links against the env
    union x in [-0.2,5.1]: L1 misses and is last -> [L0].  Loop: L0 x O0
    fails the pair AABB test, L0 x O1 collides.  IDs: objects 0-2, robot 3,
    links 4-5.  Expected (4, 1).  Reducing the links first leaves [L0], then
    O0 misses L0's box -> O2 to slot 0 ([O2,O1]) and the answer is (4, 2)."""
    objs = [_obj((5, 0, 0), (.1, .15, .15)), _obj((.1, 0, 0), (.1, .15, .15)), _obj((-.1, 0, 0), (.1, .15, .15))]
    rob = _robot([(0, 0, 0), (10, 0, 0)], [(.05, .05, .05)] * 2)
    return S.Scene(rob, [], objs), [], (4, 1)


def world_loop_nesting():
    """Links outermost in the pair loop.

    Objects O0 (0.1,0,1), O1 (0.1,0,0), half (0.1,0.15,0.15); links L0 at
    the origin and L1 at (0,0,1), half 0.05.  Nothing is reduced.  L0 meets
    only O1 and L1 only O0.  Links outermost: L0 x O0 fails the pair AABB
    test, L0 x O1 collides.  IDs: objects 0-1, robot 2, links 3-4.  Expected
    (3, 1); the environment outermost gives (4, 0)."""
    objs = [_obj((.1, 0, 1), (.1, .15, .15)), _obj((.1, 0, 0), (.1, .15, .15))]
    rob = _robot([(0, 0, 0), (0, 0, 1)], [(.05, .05, .05)] * 2)
    return S.Scene(rob, [], objs), [], (3, 1)


def _self_robot():
    # L0 (0,0,0) h .05 overlaps L3 (0.08,0,0) h (.05,.07,.07); L1 (1,0,0) h .05
    # overlaps L2 (1.08,0,0) h (.05,.07,.07); the two groups are 1 m apart
    return _robot([(0, 0, 0), (1, 0, 0), (1.08, 0, 0), (.08, 0, 0)],
                  [(.05, .05, .05), (.05, .05, .05), (.05, .07, .07), (.05, .07, .07)])


def world_self_order():
    """Self pairs in row-major (i<j) order.

    No environment; colliding pairs (L0,L3) and (L1,L2).  Row-major order
    (0,1),(0,2),(0,3): (0,3) comes first.  IDs: robot 0, links 1-4.
    Expected (1, 4); ordering by the second index gives (2, 3)."""
    return S.Scene(_self_robot(), [], []), [], (1, 4)


def world_self_mask():
    """The (i,j) pair disabled in the mask.

    As world_self_order with collisionEnabled(L0,L3) cleared: the next
    colliding pair in row-major order is (L1,L2).  Expected (2, 3)."""
    return S.Scene(_self_robot(), [], []), [(1, 4, False)], (2, 3)


def world_self_quirk():
    """Klampt's (i,j) || (i,i) self mask quirk (PlannerSettings.cpp:258).

    As world_self_mask, and collisionEnabled(L0,L0) set: the quirk enables
    every pair (L0, j) again, so (L0,L3) is checked and reported first.
    Expected (1, 4)."""
    return S.Scene(_self_robot(), [], []), [(1, 4, False), (1, 1, True)], (1, 4)


def world_env_before_self():
    """The environment check runs before the self check, and a link swapped
    into a rejected slot is examined (and rejected) in its turn.

    world_self_order plus an object at (1.08,0,0.12), half (0.1,0.15,0.15),
    crossing L1 and L2.  Env reduction: the object meets the links' union.
    Link reduction against the object's box x in [0.98,1.18]: i=0 L0 misses
    -> L3 to slot 0 ([L3,L1,L2]); L3 misses -> L2 to slot 0 ([L2,L1]); L2
    hits; L1 hits.  Loop: L2 x object collides first, and the environment
    pair is reported although the self pair (L0,L3) also collides.  IDs:
    object 0, robot 1, links 2-5.  Expected (4, 0); a stable erase gives
    (3, 0)."""
    return S.Scene(_self_robot(), [], [_obj((1.08, 0, .12), (.1, .15, .15))]), [], (4, 0)


WORLDS = [world_env_swap, world_link_swap, world_reduction_order, world_loop_nesting, world_self_order,
          world_self_mask, world_self_quirk, world_env_before_self]
