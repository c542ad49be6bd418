"""GPU parity at the workloads' own sizes.

* Heterogeneous batching: the 64 robots of 6-30 links (branched above 12,
  200-5000-triangle links, 100-10000 configurations each, 104,371 in all)
  against the 200,538-triangle clutter soup, in one launch; the first 64
  configurations of every robot are compared with the FP64 oracle (booleans
  bit-exact outside the 1e-6 m contact band).
* Pair distances on the 200k-triangle soup: 1,024 configurations x 22 pairs
  within 1e-5 m of the FP64 oracle.
* SDF penalty on the 256^3 grid at 1 cm built from the 1M-point cloud, over
  all 16,384 x 32 waypoints in one call: loss and gradient of a 2,048-waypoint
  subsample against the FP64 Klampt-rule gradient; and on a planar field of
  the same 256^3 / 1 cm grid, gradients against FP64 central differences
  (h = 1e-4) within 1e-4 relative.
"""
import numpy as np
import pytest

import oracle
from klampt_amd import engine, synthetic as S

pytestmark = pytest.mark.gpu

BAND = 1e-6


@pytest.fixture(scope="module")
def soup():
    return S.soup_scene(0)


def test_hetero_64_robots_full_launch_vs_oracle(ctx, soup):
    import torch
    terrain = engine.Geom.trimesh(ctx, soup.terrains[0].verts, soup.terrains[0].tris)
    robots, counts = S.hetero_robots(seed=0, nrobots=64)
    assert sum(counts) == 104371
    worlds = []
    for spec in robots:
        w = engine.World(ctx)
        w.add_terrain(terrain)
        w.add_robot(engine.Robot.from_spec(ctx, spec))
        worlds.append(w)
    Qs = [S.uniform_configs(sp, c, seed=20 + i) for i, (sp, c) in enumerate(zip(robots, counts))]
    dQ = torch.from_numpy(np.concatenate([q.ravel() for q in Qs])).cuda()
    print("launching 64 robots, %d configurations" % sum(counts), flush=True)
    out = engine.feasible_multi_dev(worlds, [0] * 64, counts, dQ).cpu().numpy()
    parts = np.split(out, np.cumsum(counts)[:-1])
    mism = 0
    for b, (spec, Q) in enumerate(zip(robots, Qs)):
        ow, orb = oracle.world_from_scene(S.Scene(spec, terrains=soup.terrains))
        q = Q[:64]
        fo = ow.feasible_batch(orb, q)[0]
        bad = np.nonzero(parts[b][:64] != fo)[0]
        if len(bad):
            sep = ow.distance_batch(orb, q[bad], signed=True)
            assert np.all(np.abs(sep).min(axis=1) < BAND), f"robot {b} ({spec.n} links): mismatch outside the band"
            mism += len(bad)
    print(f"64 robots: {64 * 64} configurations compared, {mism} in-band mismatches, "
          f"feasible fraction {out.mean():.3f}")


def test_soup_pair_distances_1024(ctx, soup):
    w, r = engine.world_from_scene(ctx, soup)
    ow, orb = oracle.world_from_scene(soup)
    Q = S.uniform_configs(soup.robot, 1024, seed=3)
    d = w.distance_batch(Q)
    d64 = ow.distance_batch(orb, Q)
    err = np.abs(d.astype(np.float64) - d64)
    print(f"soup distances: {d.shape}, max |err| {err.max():.2e} m, zeros {int((d64 == 0).sum())}")
    # one column per kg_world_pair_list pair, the oracle's in the same order
    assert d.shape == d64.shape == (1024, len(w.pair_list()))
    assert err.max() < 1e-5
    assert np.array_equal(d == 0, d64 == 0)


@pytest.fixture(scope="module")
def arm_points():
    spec = S.arm7()
    pts = [S.surface_samples(m, 64, seed=i) for i, m in enumerate(spec.link_meshes)]
    return spec, np.concatenate(pts), np.full(7, 64, np.int32)


def _near_face(ow, orb, g, Q, pts, ppl, dsafe, tol_cells=2e-4):
    """Per configuration: does an active point (d < dsafe) lie within
    tol_cells of a cell face of the grid of cell centres (where FP32 and FP64
    points may fall in different cells and the trilinear gradient jumps)?"""
    h = (g.bmax - g.bmin) / np.array(g.values.shape)
    off = np.concatenate([[0], np.cumsum(ppl)])
    g64 = S.SDFGrid(g.bmin, g.bmax, g.values.astype(np.float64))
    out = np.zeros(len(Q), bool)
    for c, q in enumerate(Q):
        T = ow.fk(orb, q.astype(np.float64))
        for k in range(len(ppl)):
            R, t = T[k][:9].reshape(3, 3).T, T[k][9:]
            X = pts[off[k]:off[k + 1]] @ R.T + t
            d, _ = oracle.sdf_eval(g64, X)
            u = (X - g.bmin) / h - 0.5
            frac = np.abs(u - np.round(u)).min(axis=1)
            if np.any((d < dsafe) & (frac < tol_cells)):
                out[c] = True
                break
    return out


def test_sdf_penalty_256_cloud_full_batch(ctx, arm_points):
    import torch
    spec, pts, ppl = arm_points
    P, inside = S.noisy_cloud(0, npoints=1_000_000)
    g = S.sdf_from_cloud(P, inside, dims=256, voxel=0.01)
    del P
    sdf = engine.Geom.sdf(ctx, g.bmin, g.bmax, g.values)
    r = engine.Robot.from_spec(ctx, spec)
    Qall = S.trajectories(spec, 16384, 32, seed=4)
    assert Qall.shape == (524288, 7)
    dQ = torch.from_numpy(Qall).cuda()
    dpts = torch.from_numpy(pts.astype(np.float32)).cuda()
    print("penalty over 524,288 waypoints", flush=True)
    loss = r.sdf_penalty_dev(sdf, dQ, dpts, ppl, 0.05).cpu().numpy()
    grad = r.sdf_penalty_grad_dev(sdf, dQ, dpts, ppl, 0.05,
                                  dloss=torch.ones(len(Qall), device="cuda")).cpu().numpy()
    idx = np.random.default_rng(1).choice(len(Qall), 2048, replace=False)
    Q = Qall[idx]
    ow = oracle.OracleWorld()
    orb = ow.add_robot(spec)
    g64 = S.SDFGrid(g.bmin, g.bmax, g.values.astype(np.float64))
    l64, gk = ow.sdf_penalty(orb, g64, Q.astype(np.float64), pts, ppl, dsafe=0.05, grad=True)
    active = l64 > 1e-6
    assert active.sum() > 200
    assert np.all(np.abs(loss[idx] - l64) <= 1e-4 * np.maximum(1e-2, l64))
    scale = np.abs(gk).max(axis=1, keepdims=True) + 1e-30
    rel = (np.abs(grad[idx] - gk) / scale).max(axis=1)
    bad = active & (rel > 1e-4)
    near = _near_face(ow, orb, g, Q[bad], pts, ppl, dsafe=0.05) if bad.any() else np.zeros(0, bool)
    print(f"256^3 cloud: {active.sum()} active of 2048, {bad.sum()} above 1e-4 (near a cell face: "
          f"{int(near.sum())}), 99% quantile {np.quantile(rel[active], 0.99):.2e}")
    assert near.all()
    assert np.quantile(rel[active], 0.99) < 1e-4


def test_sdf_penalty_256_plane_vs_central_differences(ctx, arm_points):
    spec, pts, ppl = arm_points
    dims, h = 256, 0.01
    bmin = np.array([-1.28, -1.28, -0.3])
    bmax = bmin + dims * h
    c = (np.arange(dims) + 0.5) * h
    n = np.array([0.3, -0.2, 1.0])
    n /= np.linalg.norm(n)
    # planar field n.x - 0.9 over the grid (float32 values; the FP64 checks
    # use the same float32 values, so the field is the same on both sides)
    vals = (((bmin[0] + c)[:, None, None] * n[0] + (bmin[1] + c)[None, :, None] * n[1])
            + (bmin[2] + c)[None, None, :] * n[2] - 0.9).astype(np.float32)
    g = S.SDFGrid(bmin, bmax, vals)
    sdf = engine.Geom.sdf(ctx, bmin, bmax, vals)
    r = engine.Robot.from_spec(ctx, spec)
    Q = (S.trajectories(spec, 32, 32, seed=5) * 0.3).astype(np.float32)
    ow = oracle.OracleWorld()
    orb = ow.add_robot(spec)
    # every sample point inside the box of cell centres (no clamping)
    lo, hi = bmin + h / 2, bmax - h / 2
    off = np.concatenate([[0], np.cumsum(ppl)])
    keep = []
    for i, q in enumerate(Q):
        T = ow.fk(orb, q.astype(np.float64))
        X = np.concatenate([pts[off[k]:off[k + 1]] @ T[k][:9].reshape(3, 3) + T[k][9:] for k in range(7)])
        keep.append(bool(np.all((X > lo + 1e-3) & (X < hi - 1e-3))))
    Q = Q[np.array(keep)]
    assert len(Q) > 500
    loss = r.sdf_penalty(sdf, Q, pts, ppl, dsafe=0.5)
    dq = r.sdf_penalty_grad(sdf, Q, pts, ppl, dsafe=0.5)
    g64 = S.SDFGrid(bmin, bmax, vals.astype(np.float64))
    l64, _ = ow.sdf_penalty(orb, g64, Q.astype(np.float64), pts, ppl, dsafe=0.5, grad=False)
    _, gfd = ow.sdf_penalty(orb, g64, Q.astype(np.float64), pts, ppl, dsafe=0.5, grad=True, fd_h=1e-4)
    active = l64 > 1e-3
    assert active.sum() > 100
    assert np.all(np.abs(loss - l64) <= 1e-4 * np.maximum(1.0, l64))
    scale = np.abs(gfd).max(axis=1, keepdims=True)
    rel = (np.abs(dq - gfd) / scale)[active]
    print(f"256^3 plane: {active.sum()} active, max rel grad err vs FD {rel.max():.2e}")
    assert rel.max() < 1e-4
