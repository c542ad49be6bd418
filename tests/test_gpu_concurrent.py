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


def test_edge_batches_on_two_streams_equal_serial(ctx):
    """Device-pointer edge batches take the context's edge workspace slots:
    batches issued alternately on two streams run concurrently and each
    equals the same batch run alone (and the oracle); a feasibility batch
    and a host-pointer edge call issued between them keep their order."""
    import torch
    sc = S.soup_scene(0, target_tris=60_000)
    w, r = engine.world_from_scene(ctx, sc)
    sets = [S.edge_segments(sc.robot, 1024 + 256 * k, eps=0.02, smin=5, smax=150, seed=80 + k) for k in range(4)]
    dev = [(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()) for A, B in sets]
    serial = [w.edge_batch_dev(dA, dB, eps=0.02).cpu().numpy() for dA, dB in dev]
    outs = [torch.full((len(A),), 7.0, dtype=torch.float32, device="cuda") for A, _ in sets]
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    Qf = S.uniform_configs(sc.robot, 4096, seed=85)
    dQ = torch.from_numpy(Qf).cuda()
    fq = torch.empty(4096, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    print("4 edge batches on 2 streams, a feasibility batch and a host edge call between them", flush=True)
    for k in range(4):
        with torch.cuda.stream(streams[k % 2]):
            w.edge_batch_dev(dev[k][0], dev[k][1], eps=0.02, out=outs[k])
            if k == 1:
                w.feasible_batch_dev(dQ, feasible=fq)
        if k == 2:
            host = w.edge_batch(sets[0][0], sets[0][1], eps=0.02)
    torch.cuda.synchronize()
    for k in range(4):
        assert np.array_equal(outs[k].cpu().numpy(), serial[k]), f"edge batch {k}"
    assert np.array_equal(host, serial[0])
    ow, orb = oracle.world_from_scene(sc)
    assert np.array_equal(fq.cpu().numpy(), ow.feasible_batch(orb, Qf)[0])
    fo, _, _ = ow.edge_batch(orb, sets[1][0], sets[1][1], eps=0.02, norm_inf=False)
    assert np.array_equal(serial[1], fo.astype(np.float32))
    st = ctx.status_dev().cpu().numpy()
    assert st[0] == 0 and st[1] == 0


def test_feasibility_batch_captured_in_a_graph(ctx):
    """kg_feasible_batch_dev captured in a HIP graph after one warm-up call
    per workspace slot: replays with new inputs copied into the captured
    buffer, and eager work between them, equal the eager results."""
    import torch
    sc = S.boxes_scene(3)
    w, r = engine.world_from_scene(ctx, sc)
    sets = [S.uniform_configs(sc.robot, 4096, seed=90 + k) for k in range(3)]
    eager = [w.feasible_batch(Q)[0] for Q in sets]
    dQ = torch.from_numpy(sets[0]).cuda()
    out = torch.empty(4096, dtype=torch.uint8, device="cuda")
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(3):
            w.feasible_batch_dev(dQ, feasible=out)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    print("capturing a feasibility batch", flush=True)
    with torch.cuda.graph(g, stream=side):
        w.feasible_batch_dev(dQ, feasible=out)
    torch.cuda.synchronize()
    for rep in range(2):
        for k, Q in enumerate(sets):
            dQ.copy_(torch.from_numpy(Q))
            out.fill_(7)
            w.distance_batch(sets[0][:16])
            g.replay()
            torch.cuda.synchronize()
            assert np.array_equal(out.cpu().numpy(), eager[k]), f"replay {rep}, set {k}"
