"""Feasibility batches of one context issued on several streams run
concurrently on the context's workspace slots.  Each batch must still equal
the same batch run alone; plain calls between them (distances, host-pointer
batches) keep their order; a change to the world between the waves of
batches (a plan rebuild) is seen by the batches after it and not by those
before it."""
import numpy as np
import pytest

import oracle
from klampt_amd import engine, synthetic as S

pytestmark = pytest.mark.gpu


def test_batches_on_three_streams_equal_serial(ctx):
    import torch
    sc = S.soup_scene(0, target_tris=60_000)
    w, r = engine.world_from_scene(ctx, sc)
    sets = [S.uniform_configs(sc.robot, 8192 + 1000 * k, seed=40 + k) for k in range(6)]
    serial = [w.feasible_batch_dev(torch.from_numpy(Q).cuda()).cpu().numpy() for Q in sets]
    dQ = [torch.from_numpy(Q).cuda() for Q in sets]
    outs = [torch.full((len(Q),), 7, dtype=torch.uint8, device="cuda") for Q in sets]
    streams = [torch.cuda.Stream() for _ in range(3)]
    torch.cuda.synchronize()
    print("6 batches on 3 streams, a distance call between them", flush=True)
    for k in range(6):
        with torch.cuda.stream(streams[k % 3]):
            w.feasible_batch_dev(dQ[k], feasible=outs[k])
        if k == 2:
            # a plain call in the middle (its own workspaces, context order)
            d = w.distance_batch(sets[0][:64])
    torch.cuda.synchronize()
    for k in range(6):
        assert np.array_equal(outs[k].cpu().numpy(), serial[k]), f"batch {k}"
    ow, orb = oracle.world_from_scene(sc)
    assert np.abs(d.astype(np.float64) - ow.distance_batch(orb, sets[0][:64])).max() < 1e-5
    st = ctx.status_dev().cpu().numpy()
    assert st[0] == 0 and st[1] == 0


def test_world_change_between_concurrent_batches(ctx):
    """Batches in flight on two streams, then an obstacle is added (the next
    batch rebuilds the tables): the earlier batches answer for the world
    without it, the later ones with it, as the oracle says."""
    import torch
    sc = S.boxes_scene(2)
    w, r = engine.world_from_scene(ctx, sc)
    Qa = S.uniform_configs(sc.robot, 4096, seed=50)
    Qb = S.uniform_configs(sc.robot, 4096, seed=51)
    ow, orb = oracle.world_from_scene(sc)
    before = [ow.feasible_batch(orb, Q)[0] for Q in (Qa, Qb)]
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    da, db = torch.from_numpy(Qa).cuda(), torch.from_numpy(Qb).cuda()
    oa = torch.empty(4096, dtype=torch.uint8, device="cuda")
    ob = torch.empty(4096, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    with torch.cuda.stream(s1):
        w.feasible_batch_dev(da, feasible=oa)
    with torch.cuda.stream(s2):
        w.feasible_batch_dev(db, feasible=ob)
    # a large box through the middle of the workspace
    box = S.box_mesh(np.array([0.3, 0.3, 0.3]))
    xf = np.r_[np.eye(3).ravel(), 0.2, 0.2, 0.4]
    g = engine.Geom.trimesh(ctx, box.verts, box.tris)
    w.add_rigid_object(g, xf)
    sc.objects.append((box, xf))
    oa2 = torch.empty(4096, dtype=torch.uint8, device="cuda")
    with torch.cuda.stream(s1):
        w.feasible_batch_dev(da, feasible=oa2)
    torch.cuda.synchronize()
    assert np.array_equal(oa.cpu().numpy(), before[0]) and np.array_equal(ob.cpu().numpy(), before[1])
    ow2, orb2 = oracle.world_from_scene(sc)
    after = ow2.feasible_batch(orb2, Qa)[0]
    assert not np.array_equal(after, before[0])
    assert np.array_equal(oa2.cpu().numpy(), after)
