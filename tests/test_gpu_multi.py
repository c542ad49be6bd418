"""Heterogeneous batching: many robots (6..30 DOF, branching above 12,
200..5000-triangle links) with uneven configuration counts in one launch,
equal to per-robot launches and to the FP64 oracle."""
import numpy as np
import pytest

import oracle
from klampt_amd import engine, synthetic as S
from tests.conftest import engine_params

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hetero(ctx):
    soup = S.clutter_soup(0, target_tris=60_000)
    robots, counts = S.hetero_robots(seed=3, nrobots=8, cfg_range=(20, 300))
    terrain = engine.Geom.trimesh(ctx, soup.verts, soup.tris)
    worlds = []
    for spec in robots:
        w = engine.World(ctx)
        w.add_terrain(terrain)
        w.add_robot(engine.Robot.from_spec(ctx, spec))
        worlds.append(w)
    Qs = [S.uniform_configs(spec, c, seed=10 + i) for i, (spec, c) in enumerate(zip(robots, counts))]
    return soup, robots, counts, worlds, Qs


def test_multi_equals_single_and_oracle(hetero):
    soup, robots, counts, worlds, Qs = hetero
    multi = engine.feasible_multi(worlds, [0] * len(worlds), counts, Qs)
    for b, (spec, w, Q) in enumerate(zip(robots, worlds, Qs)):
        single = w.feasible_batch(Q)[0]
        assert np.array_equal(multi[b], single), f"robot {b}: multi != single"
        sc = S.Scene(spec, terrains=[soup])
        ow, orb = oracle.world_from_scene(sc)
        fo = ow.feasible_batch(orb, Q)[0]
        assert np.array_equal(multi[b], fo), f"robot {b} ({spec.n} DOF): GPU != oracle"
    print("hetero feasible fractions:", [round(float(m.mean()), 2) for m in multi])


def test_autotune_keeps_results(hetero):
    import torch
    soup, robots, counts, worlds, Qs = hetero
    dQ = torch.from_numpy(np.concatenate([q.ravel() for q in Qs])).cuda()
    before = engine.feasible_multi_dev(worlds, [0] * len(worlds), counts, dQ).cpu().numpy()
    choice = engine.autotune(worlds, [0] * len(worlds), counts, dQ)
    print("autotune choice (cpw, leaf_min, scap, waves/CU, ms):", choice)
    after = engine.feasible_multi_dev(worlds, [0] * len(worlds), counts, dQ).cpu().numpy()
    assert np.array_equal(before, after)
    engine.set_tuning(worlds[0].ctx)  # back to defaults


def test_multi_fp64_pass_matches_oracle(ctx, hetero):
    """Heterogeneous batch with every configuration re-checked in FP64 (frames
    formed from q through the per-configuration Q offsets)."""
    soup, robots, counts, worlds, Qs = hetero
    with engine_params(ctx, force_recheck=1):
        multi = engine.feasible_multi(worlds, [0] * len(worlds), counts, Qs)
    for b, (spec, Q) in enumerate(zip(robots, Qs)):
        ow, orb = oracle.world_from_scene(S.Scene(spec, terrains=[soup]))
        assert np.array_equal(multi[b], ow.feasible_batch(orb, Q)[0]), f"robot {b} ({spec.n} DOF)"
