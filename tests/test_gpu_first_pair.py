"""The HIP path reports the hand-worked first colliding pair of each
known-answer world (tests/first_pair_worlds.py), through the C ABI."""
import numpy as np
import pytest

from klampt_amd import engine
from tests import first_pair_worlds as K

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("make", K.WORLDS, ids=[f.__name__ for f in K.WORLDS])
def test_gpu_first_pair_matches_hand_worked_answer(ctx, make):
    sc, overrides, expected = make()
    print("world", make.__name__, flush=True)
    w, r = engine.world_from_scene(ctx, sc)
    for a, b, en in overrides:
        w.set_collision_enabled(a, b, en)
    Q = np.zeros((3, sc.robot.n), np.float32)     # three identical configurations
    feas, env, slf, fp = w.feasible_batch(Q, flags=True, first_pair=True)
    assert np.all(feas == 0)
    for c in range(3):
        assert tuple(int(x) for x in fp[c]) == expected
