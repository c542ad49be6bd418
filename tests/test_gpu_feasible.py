"""GPU parity of FK, collide and batched feasibility against the FP64 oracle,
through the C ABI.

Bar: booleans and first-colliding pairs bit-exact; a mismatch is tolerated
only for a configuration inside the 1e-6 m contact band (some candidate pair
whose FP64 signed separation is within 1e-6 m of contact), and such
configurations are counted.
"""
import math

import numpy as np
import pytest

import oracle
from klampt_amd import engine, synthetic as S
from tests.conftest import engine_params

pytestmark = pytest.mark.gpu

BAND = 1e-6


@pytest.fixture(scope="module")
def c1(ctx):
    sc = S.boxes_scene(0)
    w, r = engine.world_from_scene(ctx, sc)
    ow, orb = oracle.world_from_scene(sc)
    return sc, w, r, ow, orb


@pytest.fixture(scope="module")
def c2(ctx):
    sc = S.soup_scene(0)
    w, r = engine.world_from_scene(ctx, sc)
    ow, orb = oracle.world_from_scene(sc)
    return sc, w, r, ow, orb


def _in_band(ow, orb, Q):
    sep = ow.distance_batch(orb, Q, signed=True)
    return np.abs(sep).min(axis=1) < BAND


def _check_exact(ow, orb, Q, got, want, what):
    bad = np.nonzero(np.any(np.asarray(got).reshape(len(Q), -1) != np.asarray(want).reshape(len(Q), -1), axis=1))[0]
    if len(bad):
        band = _in_band(ow, orb, Q[bad])
        assert band.all(), f"{what}: {int((~band).sum())} mismatches outside the contact band, e.g. config {bad[~band][:5]}"
    return len(bad)


def test_fk_fp32_pose_error(ctx):
    spec = S.arm7()
    r = engine.Robot.from_spec(ctx, spec)
    Q = S.uniform_configs(spec, 4096, seed=3)
    T = r.fk_batch(Q)
    ow = oracle.OracleWorld()
    orb = ow.add_robot(spec)
    err_t = err_r = 0.0
    for c in range(0, 4096, 16):
        T64 = ow.fk(orb, Q[c].astype(np.float64))
        err_t = max(err_t, np.abs(T[c][:, 9:] - T64[:, 9:]).max())
        err_r = max(err_r, np.abs(T[c][:, :9] - T64[:, :9]).max())
    print(f"FP32 FK max error: translation {err_t:.3e} m, rotation entry {err_r:.3e}")
    assert err_t < 2e-6 and err_r < 2e-6


def test_c1_feasibility_bit_exact(c1):
    sc, w, r, ow, orb = c1
    Q = S.uniform_configs(sc.robot, 4096, seed=1)
    feas, env, slf, fp = w.feasible_batch(Q, flags=True, first_pair=True)
    f2, e2, s2, fp2 = ow.feasible_batch(orb, Q)
    nb = _check_exact(ow, orb, Q, np.c_[feas, env, slf, fp], np.c_[f2, e2, s2, fp2], "C1")
    print(f"C1: {len(Q)} configs, feasible {feas.mean():.3f}, in-band mismatches {nb}, "
          f"fp64 rechecks {w.last_stats()['fp64_rechecks']}")
    assert 0.1 < feas.mean() < 0.9 and env.any() and slf.any()


def test_modes_agree(c1):
    sc, w, r, ow, orb = c1
    Q = S.uniform_configs(sc.robot, 2048, seed=9)
    f_any = w.feasible_batch(Q)[0]
    f_grp, env, slf, _ = w.feasible_batch(Q, flags=True)
    f_ord = w.feasible_batch(Q, flags=True, first_pair=True)[0]
    assert np.array_equal(f_any, f_grp) and np.array_equal(f_any, f_ord)
    jl = np.all((Q >= sc.robot.qmin) & (Q <= sc.robot.qmax), axis=1)
    assert np.array_equal(f_grp.astype(bool), jl & ~env.astype(bool) & ~slf.astype(bool))


def test_joint_limits_closed_interval(c1):
    sc, w, r, ow, orb = c1
    def f32_le(x):
        f = np.float32(x)
        return f if float(f) <= x else np.nextafter(f, np.float32(-np.inf))

    def f32_ge(x):
        f = np.float32(x)
        return f if float(f) >= x else np.nextafter(f, np.float32(np.inf))
    q = np.zeros((3, 7), np.float32)
    q[0, 3] = f32_le(sc.robot.qmax[3])                                  # on the closed bound
    q[1, 3] = np.nextafter(f32_le(sc.robot.qmax[3]), np.float32(10))    # just above qmax
    q[2, 3] = np.nextafter(f32_ge(sc.robot.qmin[3]), np.float32(-10))   # just below qmin
    f = w.feasible_batch(q)[0]
    f2 = ow.feasible_batch(orb, q)[0]
    assert np.array_equal(f, f2)
    assert f[1] == 0 and f[2] == 0


def test_c2_soup_subsample_bit_exact(c2):
    sc, w, r, ow, orb = c2
    Q = S.uniform_configs(sc.robot, 4096, seed=2)
    feas, env, slf, fp = w.feasible_batch(Q, flags=True, first_pair=True)
    f2, e2, s2, fp2 = ow.feasible_batch(orb, Q)
    nb = _check_exact(ow, orb, Q, np.c_[feas, env, slf, fp], np.c_[f2, e2, s2, fp2], "C2")
    print(f"C2 subsample: feasible {feas.mean():.3f} env {env.mean():.3f} self {slf.mean():.3f}; "
          f"in-band {nb}; fp64 rechecks {w.last_stats()['fp64_rechecks']}")


def test_c2_full_batch_properties(c2):
    """65,536 configurations: deterministic, consistent across modes, and a
    random subset equal to the oracle."""
    sc, w, r, ow, orb = c2
    Q = S.uniform_configs(sc.robot, 65536, seed=0)
    f1 = w.feasible_batch(Q)[0]
    f2 = w.feasible_batch(Q)[0]
    assert np.array_equal(f1, f2)
    fg, env, slf, _ = w.feasible_batch(Q, flags=True)
    assert np.array_equal(f1, fg)
    idx = np.random.default_rng(0).choice(65536, 1024, replace=False)
    fo = ow.feasible_batch(orb, Q[idx])[0]
    _check_exact(ow, orb, Q[idx], f1[idx], fo, "C2 full")


def test_geometry_collides_pairs(ctx):
    rng = np.random.default_rng(8)
    cap = S.capsule_mesh(0.05, 0.3, axis=2)
    box = S.box_mesh([0.1, 0.05, 0.08], edge=0.02)
    ga, gb = engine.Geom.trimesh(ctx, cap.verts, cap.tris), engine.Geom.trimesh(ctx, box.verts, box.tris)
    ow = oracle.OracleWorld()
    oa, ob = ow.add_mesh(cap), ow.add_mesh(box)
    hits = 0
    for _ in range(64):
        Ra, ta = S.random_rotation(rng), rng.uniform(-0.2, 0.2, 3)
        Rb, tb = S.random_rotation(rng), rng.uniform(-0.2, 0.2, 3)
        ga.set_transform(Ra.T.ravel(), ta)
        gb.set_transform(Rb.T.ravel(), tb)
        got = ga.collides(gb)
        want = ow.mesh_collide(oa, np.r_[Ra.T.ravel(), ta], ob, np.r_[Rb.T.ravel(), tb])
        if got != want:
            sep = ow.mesh_separation(oa, np.r_[Ra.T.ravel(), ta], ob, np.r_[Rb.T.ravel(), tb])
            assert abs(sep) < BAND
        hits += want
    assert 5 < hits < 60
    empty = engine.Geom.trimesh(ctx, np.zeros((0, 3)), np.zeros((0, 3), np.int32))
    assert not ga.collides(empty)


def test_self_collides_single_config(ctx):
    spec = S.arm7()
    r = engine.Robot.from_spec(ctx, spec)
    ow = oracle.OracleWorld()
    orb = ow.add_robot(spec)
    ow.init_default()
    Q = S.uniform_configs(spec, 32, seed=5)
    f2, e2, s2, _ = ow.feasible_batch(orb, Q)
    for c in range(32):
        r.set_config(Q[c].astype(np.float64))
        assert r.self_collides() == bool(s2[c])


def test_empty_batch_and_no_geometry(ctx, c1):
    sc, w, r, ow, orb = c1
    f = w.feasible_batch(np.zeros((0, 7), np.float32))[0]
    assert f.shape == (0,)
    spec = S.arm7()
    spec.link_meshes = [None] * 7
    bare = engine.Robot.from_spec(ctx, spec)
    wb = engine.World(ctx)
    wb.add_robot(bare)
    Q = S.uniform_configs(spec, 100, seed=1)
    assert wb.feasible_batch(Q)[0].all()


def test_fk64_matches_oracle(ctx):
    """kg_fk_batch64: the FP64 FK of the broad phase equals the oracle's."""
    spec = S.arm7()
    r = engine.Robot.from_spec(ctx, spec)
    Q = S.uniform_configs(spec, 256, seed=5).astype(np.float64)
    T = r.fk_batch64(Q)
    ow = oracle.OracleWorld()
    orb = ow.add_robot(spec)
    err = max(float(np.abs(T[c] - ow.fk(orb, Q[c])).max()) for c in range(0, 256, 4))
    assert err < 1e-12


def test_set_self_pairs_replaces_defaults(ctx):
    """kg_robot_set_self_pairs: only the listed pairs are self-checked."""
    import dataclasses
    sc = S.boxes_scene(0)
    pairs = np.array([[0, 3], [1, 5], [6, 2]], np.int32)
    w, r = engine.world_from_scene(ctx, sc)
    r.set_self_pairs(pairs)
    assert r.self_collision_enabled(2, 6) and not r.self_collision_enabled(0, 2)
    spec2 = dataclasses.replace(sc.robot, self_pairs=np.sort(pairs, axis=1))
    ow, orb = oracle.world_from_scene(S.Scene(spec2, sc.terrains, sc.objects))
    Q = S.uniform_configs(sc.robot, 2048, seed=9)
    feas, env, slf, fp = w.feasible_batch(Q, flags=True, first_pair=True)
    f64, e64, s64, p64 = ow.feasible_batch(orb, Q)
    assert np.array_equal(w.pair_list(0), ow.pair_list(orb))
    n = _check_exact(ow, orb, Q, np.c_[feas, env, slf, fp], np.c_[f64, e64, s64, p64], "custom self pairs")
    assert n <= 2


def test_timing_levels(ctx, c1):
    """Timing level 0 records no events, 1 the traversal kernel's, 2 also the
    broad phase and FP64 pass; results do not depend on it."""
    sc, w, r, ow, orb = c1
    Q = S.uniform_configs(sc.robot, 512, seed=41)
    want = w.feasible_batch(Q)[0]
    try:
        for level in (0, 1, 2):
            ctx.set_timing(level)
            ctx.kernel_times(reset=True)
            got = w.feasible_batch(Q)[0]
            st = w.last_stats()
            _, count = ctx.kernel_times(reset=True)
            assert np.array_equal(got, want)
            assert count == (0 if level == 0 else 1)
            assert (st["t_trav32_ns"] > 0) == (level >= 1)
            assert (st["t_prep_ns"] > 0) == (level == 2)
    finally:
        ctx.set_timing(1)
    with pytest.raises(RuntimeError):
        ctx.set_timing(3)


def test_fp64_pass_alone_matches_oracle(ctx, c1, c2):
    """Every configuration sent through the exact FP64 pass (which forms its
    link frames from q): the oracle's answers in all three result modes."""
    with engine_params(ctx, force_recheck=1):
        _fp64_pass_alone(c1, c2)


def _fp64_pass_alone(c1, c2):
    for name, (sc, w, r, ow, orb), seed in (("C1", c1, 51), ("C2", c2, 52)):
        Q = S.uniform_configs(sc.robot, 1024, seed=seed)
        f_any = w.feasible_batch(Q)[0]
        assert w.last_stats()["fp64_rechecks"] == 1024
        f_grp, e_grp, s_grp, _ = w.feasible_batch(Q, flags=True)
        feas, env, slf, fp = w.feasible_batch(Q, flags=True, first_pair=True)
        f2, e2, s2, fp2 = ow.feasible_batch(orb, Q)
        _check_exact(ow, orb, Q, f_any, f2, name + " any")
        _check_exact(ow, orb, Q, np.c_[f_grp, e_grp, s_grp], np.c_[f2, e2, s2], name + " groups")
        _check_exact(ow, orb, Q, np.c_[feas, env, slf, fp], np.c_[f2, e2, s2, fp2], name + " ordered")


def test_double_configuration_on_a_double_limit(ctx):
    """A float64 configuration exactly on a float64 joint limit is inside it
    (closed interval, RobotCSpace.cpp:610-630): the f64 entry point checks
    the doubles as given, where a float32 cast would round pi/2 up past the
    limit.  The planner layer uses it, so a start on a limit is accepted."""
    from klampt_amd import plan, robotsim
    spec = S.arm7()
    spec.qmin = np.full(7, -math.pi / 2)
    spec.qmax = np.full(7, math.pi / 2)
    sc = S.Scene(spec)
    w, r = engine.world_from_scene(ctx, sc)
    ow, orb = oracle.world_from_scene(sc)
    q = np.zeros((2, 7))
    q[0, 3] = math.pi / 2
    q[1, 1] = -math.pi / 2
    assert float(np.float32(math.pi / 2)) > math.pi / 2
    f64 = w.feasible_batch(q)[0]
    assert np.array_equal(f64, ow.feasible_batch(orb, q)[0]) and f64.all()
    assert not w.feasible_batch(q.astype(np.float32))[0][0]     # the float32 cast leaves the limit
    import torch
    assert w.feasible_batch_dev(torch.from_numpy(q).cuda()).cpu().numpy().all()
    W = robotsim.WorldModel()
    R = robotsim.RobotModel.from_spec(spec)
    W.addRobot(R)
    space = plan.RobotCSpace(R, W, eps=0.05)
    assert space.feasible(list(q[0])) and space.feasible(list(q[1]))
    mp = plan.MotionPlan(space, type="prm", knn=5)
    mp.setEndpoints(list(q[0]), list(q[1]))
    assert space.visible_batch(q[:1], q[1:])[0] in (True, False)
