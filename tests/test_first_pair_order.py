"""The first colliding pair in Klampt's broad-phase order, pinned by
hand-worked known answers (tests/first_pair_worlds.py), and shown to detect
planted policy mistakes in the oracle."""
import numpy as np
import pytest

import oracle
from tests import first_pair_worlds as K


def _oracle_first_pair(make):
    sc, overrides, _ = make()
    w, r = oracle.world_from_scene(sc)
    for a, b, en in overrides:
        w.set_collision_enabled(a, b, en)
    feas, env, slf, fp = w.feasible_batch(r, np.zeros((1, sc.robot.n)), nthreads=1)
    return tuple(int(x) for x in fp[0]), int(env[0]), int(slf[0]), int(feas[0])


@pytest.mark.parametrize("make", K.WORLDS, ids=[f.__name__ for f in K.WORLDS])
def test_oracle_first_pair_matches_hand_worked_answer(make):
    expected = make()[2]
    fp, env, slf, feas = _oracle_first_pair(make)
    assert fp == expected
    assert feas == 0


def test_every_world_collides_where_its_derivation_says():
    # the collisions the derivations rely on, each confirmed by brute force
    # over all triangle pairs
    sc, _, _ = K.world_env_before_self()
    w, r = oracle.world_from_scene(sc)
    feas, env, slf, fp = w.feasible_batch(r, np.zeros((1, 4)), brute=True, nthreads=1)
    assert env[0] == 1 and slf[0] == 1 and tuple(fp[0]) == (4, 0)
    for make in K.WORLDS:
        sc, ov, exp = make()
        w, r = oracle.world_from_scene(sc)
        for a, b, en in ov:
            w.set_collision_enabled(a, b, en)
        fp_b = w.feasible_batch(r, np.zeros((1, sc.robot.n)), brute=True, nthreads=1)[3]
        assert tuple(int(x) for x in fp_b[0]) == exp, make.__name__


@pytest.mark.parametrize("mut", [oracle.MUT_STABLE_ERASE, oracle.MUT_LINKS_FIRST, oracle.MUT_NO_REJECT,
                                 oracle.MUT_ENV_MAJOR],
                         ids=["stable_erase", "links_reduced_first", "no_quick_reject", "env_outermost"])
def test_known_answers_detect_planted_policy_mistake(mut):
    """Each planted mistake changes at least one world's first pair."""
    with oracle.mutation(mut):
        wrong = [m.__name__ for m in K.WORLDS if _oracle_first_pair(m)[0] != m()[2]]
    assert wrong, "no known-answer world detects this mistake"
    # and the oracle is back to Klampt's policy afterwards
    assert all(_oracle_first_pair(m)[0] == m()[2] for m in K.WORLDS)
