"""Serving pipeline (kg_serve_*): a heterogeneous feasibility batch and the
edge batches of several robots captured by the library into one HIP graph.
Every replay, with new inputs written into the fixed buffers and eager work
of the program and synchronizations between replays, equals the eager calls
it replaces (kg_feasible_multi_dev, kg_edge_batch_dev), which the other tests
compare with the oracle; the graph is recaptured when the engine parameters
change and when a workspace address moves, and replays equal eager calls
again afterwards."""
import numpy as np
import pytest

from klampt_amd import engine, synthetic as S
from tests.conftest import engine_params

pytestmark = pytest.mark.gpu

EPS = 0.02


@pytest.fixture(scope="module")
def pipeline(ctx):
    soup = S.clutter_soup(1, target_tris=60_000)
    robots, _ = S.hetero_robots(seed=5, nrobots=4, cfg_range=(20, 300))
    terrain = engine.Geom.trimesh(ctx, soup.verts, soup.tris)
    specs = [S.arm7()] + list(robots)
    worlds = []
    for spec in specs:
        w = engine.World(ctx)
        w.add_terrain(terrain)
        w.add_robot(engine.Robot.from_spec(ctx, spec))
        worlds.append(w)
    counts = [400, 150, 90, 260, 40]
    # edges of the arm and of the longest robot
    longest = int(np.argmax([s.n for s in specs]))
    egroups = [(0, 700), (longest, 120)]
    return specs, worlds, counts, egroups


def _inputs(specs, counts, egroups, seed):
    Qs = [S.uniform_configs(s, c, seed=seed * 10 + b) for b, (s, c) in enumerate(zip(specs, counts))]
    segs = [S.edge_segments(specs[r], m, eps=EPS, smin=5, smax=150, seed=seed * 10 + 7 + g)
            for g, (r, m) in enumerate(egroups)]
    return Qs, segs


def _eager(specs, worlds, counts, egroups, Qs, segs):
    import torch
    dQ = torch.from_numpy(np.concatenate([q.ravel() for q in Qs])).cuda()
    feas = engine.feasible_multi_dev(worlds, [0] * len(worlds), counts, dQ).cpu().numpy()
    us = []
    for (r, m), (A, B) in zip(egroups, segs):
        u = worlds[r].edge_batch_dev(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), eps=EPS,
                                     norm_inf=False).cpu().numpy()
        us.append(u)
    return feas, np.concatenate(us)


def _load(sv, Qs, segs):
    import torch
    sv.Q.copy_(torch.from_numpy(np.concatenate([q.ravel() for q in Qs])))
    sv.A.copy_(torch.from_numpy(np.concatenate([a.ravel() for a, _ in segs])))
    sv.B.copy_(torch.from_numpy(np.concatenate([b.ravel() for _, b in segs])))
    sv.feasible.fill_(7)
    sv.first_u.fill_(7.0)


def test_serve_graph_replays_equal_eager(ctx, pipeline):
    import torch
    specs, worlds, counts, egroups = pipeline
    sets = [_inputs(specs, counts, egroups, seed) for seed in range(3)]
    eager = [_eager(specs, worlds, counts, egroups, Qs, segs) for Qs, segs in sets]
    print("building the serving pipeline", flush=True)
    sv = engine.Serve(ctx, feas=[(w, 0, c) for w, c in zip(worlds, counts)],
                      edges=[(worlds[r], 0, m) for r, m in egroups], eps=EPS, norm_inf=False)
    assert sv.captures() == 1
    for rep in range(2):
        for k, (Qs, segs) in enumerate(sets):
            _load(sv, Qs, segs)
            # eager work of the program between replays
            worlds[0].feasible_batch(S.uniform_configs(specs[0], 64, seed=100 + k))
            sv.run()
            torch.cuda.synchronize()
            f, u = sv.feasible.cpu().numpy(), sv.first_u.cpu().numpy()
            assert np.array_equal(f, eager[k][0]), f"replay {rep}, set {k}: feasibility"
            assert np.array_equal(u, eager[k][1]), f"replay {rep}, set {k}: edges"
            st = ctx.status_dev().cpu().numpy()
            assert st[0] == 0 and st[1] == 0, st
    assert sv.captures() == 1, "the graph was recaptured without a reason"
    fr = [round(float(x.mean()), 2) for x in np.split(eager[0][0], np.cumsum(counts)[:-1])]
    print(f"replays equal eager: feasible fractions {fr}, free edges {(eager[0][1] < 0).mean():.2f}")

    # new engine parameters: recaptured, still equal
    with engine_params(ctx, cpw=4, edge_w0=4):
        Qs, segs = sets[1]
        _load(sv, Qs, segs)
        sv.run()
        torch.cuda.synchronize()
        assert sv.captures() == 2
        ref = _eager(specs, worlds, counts, egroups, Qs, segs)
        assert np.array_equal(sv.feasible.cpu().numpy(), ref[0])
        assert np.array_equal(sv.first_u.cpu().numpy(), ref[1])

    # a larger eager batch grows the shared workspaces: the next run recaptures
    n0 = sv.captures()
    worlds[0].feasible_batch(S.uniform_configs(specs[0], 200_000, seed=9))
    Qs, segs = sets[2]
    _load(sv, Qs, segs)
    sv.run()
    torch.cuda.synchronize()
    assert sv.captures() == n0 + 1
    assert np.array_equal(sv.feasible.cpu().numpy(), eager[2][0])
    assert np.array_equal(sv.first_u.cpu().numpy(), eager[2][1])
    sv.close()


def test_serve_feasibility_only_and_edges_only(ctx, pipeline):
    import torch
    specs, worlds, counts, egroups = pipeline
    (Qs, segs) = _inputs(specs, counts, egroups, 5)
    ref = _eager(specs, worlds, counts, egroups, Qs, segs)
    a = engine.Serve(ctx, feas=[(w, 0, c) for w, c in zip(worlds, counts)])
    a.Q.copy_(torch.from_numpy(np.concatenate([q.ravel() for q in Qs])))
    a.run()
    b = engine.Serve(ctx, edges=[(worlds[r], 0, m) for r, m in egroups], eps=EPS)
    b.A.copy_(torch.from_numpy(np.concatenate([x.ravel() for x, _ in segs])))
    b.B.copy_(torch.from_numpy(np.concatenate([y.ravel() for _, y in segs])))
    b.run()
    torch.cuda.synchronize()
    assert np.array_equal(a.feasible.cpu().numpy(), ref[0])
    assert np.array_equal(b.first_u.cpu().numpy(), ref[1])
    a.close()
    b.close()
