"""Overflow of the traversal's spill region never yields a wrong answer.

Two dense concentric spheres about 0.1 mm apart (no contact) keep nearly
every node pair of their trees overlapping, so the traversal pools far more
tasks than a wave's LDS pool holds.  With the spill capacity set to 0 the
waves must drop pooled work: the feasibility result is then reported as
infeasible (never as a false "feasible"), the distance as a certified lower
bound, each counted in the status word that kg_ctx_status_dev copies on the
device, and the host-pointer calls fail with KG_E_LIMIT.  With the default
capacity the same queries are exact.
"""
import numpy as np
import pytest

import oracle
from klampt_amd import _lib, engine, synthetic as S
from tests.conftest import engine_params

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def spheres(ctx):
    a = S.ellipsoid_mesh([0.1, 0.1, 0.1], edge=0.005)
    b = S.ellipsoid_mesh([0.1001, 0.1001, 0.1001], edge=0.005)
    assert a.tris.shape[0] > 10000 and a.tris.shape == b.tris.shape
    spec = S.RobotSpec(name="ball", parents=np.array([-1], np.int32),
                       T0=np.r_[np.eye(3).ravel(), [0.0, 0.0, 0.0]].reshape(1, 12),
                       axes=np.array([[0.0, 0.0, 1.0]]), jtype=np.zeros(1, np.int32),
                       qmin=np.array([-np.pi]), qmax=np.array([np.pi]), link_meshes=[a])
    sc = S.Scene(spec, [], [(b, np.r_[np.eye(3).ravel(), [0.0, 0.0, 0.0]])])
    w, r = engine.world_from_scene(ctx, sc)
    ow, orb = oracle.world_from_scene(sc)
    return sc, w, ow, orb


def test_exact_with_default_capacity(ctx, spheres):
    sc, w, ow, orb = spheres
    Q = np.array([[0.0], [0.3], [1.1]], np.float32)
    feas, env, slf, fp = w.feasible_batch(Q, flags=True, first_pair=True)
    assert np.array_equal(feas, ow.feasible_batch(orb, Q)[0])
    assert feas.all()
    st = w.last_stats()
    print("default capacity:", st)
    assert st["uncertain"] == 0
    d = w.distance_batch(Q)
    d64 = ow.distance_batch(orb, Q)
    print("distances", d.ravel(), d64.ravel())
    assert np.abs(d - d64).max() < 1e-5 and d64.min() > 5e-5
    assert w.last_stats()["uncertain"] == 0


def test_dropped_work_is_reported_not_wrong(ctx, spheres):
    import torch
    sc, w, ow, orb = spheres
    Q = np.array([[0.0], [0.3], [1.1]], np.float32)
    d64 = ow.distance_batch(orb, Q)
    dQ = torch.from_numpy(Q).cuda()
    with engine_params(ctx, spill_cap=0):
        print("feasibility with spill_cap=0", flush=True)
        for flags in (False, True):
            feas = torch.empty(3, dtype=torch.uint8, device="cuda")
            env = torch.empty(3, dtype=torch.uint8, device="cuda") if flags else None
            slf = torch.empty(3, dtype=torch.uint8, device="cuda") if flags else None
            fp = torch.empty((3, 2), dtype=torch.int32, device="cuda") if flags else None
            w.feasible_batch_dev(dQ, feasible=feas, env=env, self_hit=slf, first_pair=fp)
            status = ctx.status_dev().cpu().numpy()
            f = feas.cpu().numpy()
            print("status", status, "feasible", f)
            assert status[0] == 1                       # the spill region filled
            assert status[1] >= 1                       # some configuration left open
            # every configuration is truly feasible: exactly the open ones are
            # reported infeasible (never the other way round)
            assert int((f == 0).sum()) == status[1]
            if flags:
                unc = fp.cpu().numpy()[:, 0] == -2
                assert int(unc.sum()) == status[1]
                assert np.all(f[unc] == 0) and np.all(f[~unc] == 1)
                assert np.array_equal(env.cpu().numpy() == 1, unc) and not slf.cpu().numpy().any()
                assert np.all(fp.cpu().numpy()[~unc] == -1)
            assert w.last_stats()["uncertain"] == status[1]
        with pytest.raises(_lib.KlamptGPUError) as ei:
            w.feasible_batch(Q)
        assert ei.value.code == -5
        print("distances with spill_cap=0", flush=True)
        dd = w.distance_batch_dev(dQ)
        status = ctx.status_dev().cpu().numpy()
        d = dd.cpu().numpy()
        print("status", status, "lower bounds", d.ravel(), "exact", d64.ravel())
        assert status[1] > 0
        # a certified lower bound of the FP64 distance (FP32 rounding aside)
        assert np.all(d <= d64 + 1e-6) and np.all(d >= 0)
        with pytest.raises(_lib.KlamptGPUError):
            w.distance_batch(Q)
    # parameters restored: exact again
    feas = w.feasible_batch(Q)[0]
    assert feas.all() and w.last_stats()["uncertain"] == 0


def test_params_roundtrip_and_validation(ctx):
    p = ctx.params()
    assert p["scap"] == 256 and p["spill_cap"] == 16384 and p["force_recheck"] == 0
    with pytest.raises(_lib.KlamptGPUError):
        ctx.set_params(scap=1024)
    with pytest.raises(_lib.KlamptGPUError):
        ctx.set_params(sdf_lanes=12)
    assert ctx.params() == p
    old = ctx.set_params(cpw=8, static_pct=0)
    assert ctx.params()["cpw"] == 8 and old == p
    ctx.set_params(**old)
    assert ctx.params() == p
