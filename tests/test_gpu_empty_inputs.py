"""Empty and degenerate inputs of the device-pointer entry points: nothing is
launched or written for an empty batch, and the next call is unaffected;
zero-length segments (a == b) are free; a world with no obstacles gives no
sweep terms; a serving pipeline of empty batches runs."""
import numpy as np
import pytest

import oracle
from klampt_amd import engine, synthetic as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def boxes(ctx):
    sc = S.boxes_scene(0)
    w, r = engine.world_from_scene(ctx, sc)
    ow, orb = oracle.world_from_scene(sc)
    return sc, w, ow, orb


def test_empty_device_batches(ctx, boxes):
    import torch
    sc, w, ow, orb = boxes
    n = sc.robot.n
    e32 = torch.empty((0, n), dtype=torch.float32, device="cuda")
    e64 = torch.empty((0, n), dtype=torch.float64, device="cuda")
    assert w.feasible_batch_dev(e32).numel() == 0
    assert w.feasible_batch_dev(e64).numel() == 0
    assert w.edge_batch_dev(e32, e32, eps=0.01).numel() == 0
    d, info, coef = w.sweep_batch_dev(e64, e64, 0.2)
    assert d.numel() == 0
    d, info, coef = w.sweep_batch(np.zeros((0, n)), np.zeros((0, n)), 0.2)
    assert d.shape == (0, n, len(sc.objects))
    # the next batch is unaffected
    Q = S.uniform_configs(sc.robot, 512, seed=60)
    got = w.feasible_batch_dev(torch.from_numpy(Q).cuda()).cpu().numpy()
    assert np.array_equal(got, ow.feasible_batch(orb, Q)[0])


def test_zero_length_and_single_step_segments(ctx, boxes):
    """a == b has no interior sample (n = 0): free; a segment of one step has
    none either (n = 1); both, as Klampt's checker, whatever the endpoints."""
    import torch
    sc, w, ow, orb = boxes
    Q = S.uniform_configs(sc.robot, 64, seed=61)
    B1 = Q.copy()
    B1[:, 0] += np.float32(0.004)          # one step at eps 0.01 (L-infinity)
    A = np.concatenate([Q, Q]).astype(np.float32)
    B = np.concatenate([Q, B1]).astype(np.float32)
    ns = torch.empty(128, dtype=torch.int32, device="cuda")
    fu = w.edge_batch_dev(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), eps=0.01, norm_inf=True,
                          nsamples=ns).cpu().numpy()
    assert np.all(ns.cpu().numpy()[:64] == 0) and np.all(ns.cpu().numpy()[64:] == 1)
    assert np.all(fu == -1.0)
    fo, _, _ = ow.edge_batch(orb, A, B, eps=0.01, norm_inf=True)
    assert np.all(fo == -1.0)


def test_world_without_obstacles_has_no_sweep_terms(ctx):
    sc = S.Scene(S.arm7())
    w, r = engine.world_from_scene(ctx, sc)
    assert w.num_obstacles() == 0
    Q0 = S.uniform_configs(sc.robot, 8, seed=62).astype(np.float64)
    d, info, coef = w.sweep_batch(Q0, Q0 + 0.1, 0.5)
    assert d.shape == (8, 7, 0)


def test_serving_pipeline_of_empty_batches(ctx, boxes):
    import torch
    sc, w, ow, orb = boxes
    sv = engine.Serve(ctx, feas=[(w, 0, 0)], edges=[(w, 0, 0)], eps=0.01)
    sv.run()
    torch.cuda.synchronize()
    st = ctx.status_dev().cpu().numpy()
    assert st[0] == 0 and st[1] == 0
    sv.close()
    with pytest.raises(RuntimeError):
        engine.Serve(ctx, feas=[], edges=[])
