"""Device-side edge checks (kg_edge_batch_dev / _f64_dev, and the host-pointer
calls that run the same device rounds) against the FP64 oracle.

Bar: first_u bit-exact (float32 of the oracle's FP64 u) outside the 1e-6 m
contact band; per-segment step counts equal to numpy's ceil(d / eps); the
round schedule's limits reported (-2, status word [1]) and never turned into
a wrong answer; a captured HIP graph of the call replays to the eager result.
"""
import numpy as np
import pytest

import oracle
from klampt_amd import engine, synthetic as S
from tests.conftest import engine_params

pytestmark = pytest.mark.gpu

BAND = 1e-6


def _steps(A, B, eps, norm_inf):
    d = A.astype(np.float64) - B.astype(np.float64)
    dist = np.abs(d).max(axis=1) if norm_inf else np.sqrt((d * d).sum(axis=1))
    return np.ceil(dist / eps).astype(np.int64)


def _check_band(ow, orb, A, B, eps, ns, bad):
    """Segments whose first_u differs must have a bisection sample within
    the contact band."""
    for s in bad:
        n = int(ns[s])
        a, b = A[s].astype(np.float64), B[s].astype(np.float64)
        Q = np.array([a * (1 - i / n) + b * (i / n) for i in oracle.bisection_order(n)])
        sep = ow.distance_batch(orb, Q, signed=True)
        assert (np.abs(sep).min(axis=1) < BAND).any(), f"segment {s} differs outside the contact band"


@pytest.fixture(scope="module")
def soup(ctx):
    sc = S.soup_scene(0)
    w, r = engine.world_from_scene(ctx, sc)
    ow, orb = oracle.world_from_scene(sc)
    return sc, w, ow, orb


def test_c4_full_8192_segments_dev_vs_oracle(ctx, soup):
    """The edge workload at its own size: 8192 segments of the 7-DOF arm in
    the 200k-triangle soup, L-infinity step 0.01 rad, 10..500 samples."""
    import torch
    sc, w, ow, orb = soup
    A, B = S.edge_segments(sc.robot, 8192, eps=0.01, smin=10, smax=500, seed=8)
    dA, dB = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
    ns_dev = torch.empty(8192, dtype=torch.int32, device="cuda")
    print("C4 8192 segments, device rounds", flush=True)
    fu = w.edge_batch_dev(dA, dB, eps=0.01, norm_inf=True, nsamples=ns_dev).cpu().numpy()
    status = ctx.status_dev().cpu().numpy()
    fo, ns, nchk = ow.edge_batch(orb, A, B, eps=0.01, norm_inf=True)
    assert np.array_equal(ns_dev.cpu().numpy(), ns)
    assert np.array_equal(ns, _steps(A, B, 0.01, True))
    bad = np.nonzero(fu != fo.astype(np.float32))[0]
    print(f"C4: free {(fo < 0).mean():.3f}, samples checked by the oracle {int(nchk.sum())}, "
          f"mismatches {len(bad)}, status {status}")
    assert status[1] == 0
    _check_band(ow, orb, A, B, 0.01, ns, bad)
    # the host-pointer call runs the same rounds
    fh = w.edge_batch(A, B, eps=0.01, norm_inf=True)
    assert np.array_equal(fh, fu)


@pytest.mark.parametrize("norm_inf", [False, True])
def test_f64_endpoints_dev_and_host(ctx, soup, norm_inf):
    import torch
    sc, w, ow, orb = soup
    A, B = S.edge_segments(sc.robot, 512, eps=0.02, smin=5, smax=300, seed=11)
    A64 = A.astype(np.float64) + 1e-9 * np.random.default_rng(1).standard_normal(A.shape)
    B64 = B.astype(np.float64)
    fo, ns, _ = ow.edge_batch(orb, A64, B64, eps=0.02, norm_inf=norm_inf)
    fd = w.edge_batch_dev(torch.from_numpy(A64).cuda(), torch.from_numpy(B64).cuda(), eps=0.02,
                          norm_inf=norm_inf).cpu().numpy()
    fh = w.edge_batch(A64, B64, eps=0.02, norm_inf=norm_inf)
    assert np.array_equal(fd, fh)
    bad = np.nonzero(fd != fo.astype(np.float32))[0]
    print(f"f64 endpoints (Linf={norm_inf}): mismatches {len(bad)}")
    _check_band(ow, orb, A64, B64, 0.02, ns, bad)


def test_small_budget_rounds_and_chunks(ctx):
    """A 16 MiB round budget on the 16-box scene (127 items per sample)
    holds about 1,700 samples: 3,000 segments run as two chunks of
    budget-limited rounds, with the same answers."""
    sc = S.boxes_scene(0)
    w, r = engine.world_from_scene(ctx, sc)
    ow, orb = oracle.world_from_scene(sc)
    A, B = S.edge_segments(sc.robot, 3000, eps=0.01, smin=10, smax=400, seed=12)
    fo, ns, _ = ow.edge_batch(orb, A, B, eps=0.01, norm_inf=False)
    with engine_params(ctx, edge_budget_mb=16, edge_w0=4):
        print("3000 segments, 16 MiB budget", flush=True)
        fu = w.edge_batch(A, B, eps=0.01, norm_inf=False)
    bad = np.nonzero(fu != fo.astype(np.float32))[0]
    print(f"small budget: mismatches {len(bad)}")
    _check_band(ow, orb, A, B, 0.01, ns, bad)
    fd = w.edge_batch(A, B, eps=0.01, norm_inf=False)
    assert np.array_equal(fd, fu)


def test_max_steps_limit_reported_not_wrong(ctx, soup):
    """Segments longer than max_steps come out -2 and are counted; a hit the
    covered positions hold is still the exact first one."""
    import torch
    sc, w, ow, orb = soup
    A, B = S.edge_segments(sc.robot, 1024, eps=0.01, smin=10, smax=400, seed=13)
    fo, ns, nchk = ow.edge_batch(orb, A, B, eps=0.01, norm_inf=True)
    max_steps = 60
    fu = w.edge_batch_dev(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), eps=0.01, norm_inf=True,
                          max_steps=max_steps).cpu().numpy()
    status = ctx.status_dev().cpu().numpy()
    covered = max_steps - 1
    short = ns - 1 <= covered
    hit_early = (fo >= 0) & (nchk - 1 < covered)
    exact = short | hit_early
    und = ~exact
    print(f"max_steps {max_steps}: exact {exact.sum()}, undecided {und.sum()}, status {status}")
    assert und.sum() > 0 and exact.sum() > 0
    assert np.all(fu[und] == -2) and status[1] == und.sum()
    bad = np.nonzero(exact & (fu != fo.astype(np.float32)))[0]
    _check_band(ow, orb, A, B, 0.01, ns, bad)
    # the joint-range bound (max_steps = 0) covers every segment
    fz = w.edge_batch_dev(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), eps=0.01,
                          norm_inf=True).cpu().numpy()
    assert ctx.status_dev().cpu().numpy()[1] == 0
    assert np.all(fz >= -1)


def test_edge_dev_graph_replay(ctx, soup):
    """kg_edge_batch_dev captured in a HIP graph (after a warm-up call of the
    same size) replays to the eager result, with new inputs copied into the
    captured buffers and eager work and synchronization between replays."""
    import torch
    sc, w, ow, orb = soup
    sets = [S.edge_segments(sc.robot, 1024, eps=0.01, smin=10, smax=300, seed=20 + k) for k in range(3)]
    eager = [w.edge_batch(A, B, eps=0.01, norm_inf=True) for A, B in sets]
    dA = torch.from_numpy(sets[0][0]).cuda()
    dB = torch.from_numpy(sets[0][1]).cuda()
    out = torch.empty(1024, dtype=torch.float32, device="cuda")
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        # one warm-up call per edge workspace slot: the capture allocates nothing
        for _ in range(2):
            w.edge_batch_dev(dA, dB, eps=0.01, norm_inf=True, out=out)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    print("capturing the edge rounds", flush=True)
    with torch.cuda.graph(g, stream=side):
        w.edge_batch_dev(dA, dB, eps=0.01, norm_inf=True, out=out)
    torch.cuda.synchronize()
    for rep in range(2):
        for k, (A, B) in enumerate(sets):
            dA.copy_(torch.from_numpy(A))
            dB.copy_(torch.from_numpy(B))
            out.fill_(7.0)
            # eager work of the program between replays
            w.feasible_batch(S.uniform_configs(sc.robot, 64, seed=rep * 10 + k))
            g.replay()
            torch.cuda.synchronize()
            got = out.cpu().numpy()
            assert np.array_equal(got, eager[k]), f"replay {rep}, set {k}"
    print("graph replays equal the eager results")


def test_long_robot_edges_vs_oracle(ctx, soup):
    """Edge rounds for robots of more than 8 links (a device-count broad
    phase in groups of 16 and 32 lanes, world frames by pointer jumping): an
    11-link chain and a 20-link branched robot in the soup, and the same
    segments with the one-lane-per-configuration broad phase."""
    sc, _, _, _ = soup
    robots, _ = S.hetero_robots(seed=0, nrobots=4)
    for spec in (robots[3], robots[2]):
        assert spec.n > 8
        scene = S.Scene(spec, terrains=sc.terrains)
        w, r = engine.world_from_scene(ctx, scene)
        ow, orb = oracle.world_from_scene(scene)
        A, B = S.edge_segments(spec, 256, eps=0.02, smin=5, smax=200, seed=17)
        print(f"{spec.n}-link robot, 256 segments", flush=True)
        fu = w.edge_batch(A, B, eps=0.02, norm_inf=True)
        fo, ns, _ = ow.edge_batch(orb, A, B, eps=0.02, norm_inf=True)
        bad = np.nonzero(fu != fo.astype(np.float32))[0]
        print(f"  free {(fo < 0).mean():.2f}, mismatches {len(bad)}")
        _check_band(ow, orb, A, B, 0.02, ns, bad)
        with engine_params(ctx, prep_group=0):
            fs = w.edge_batch(A, B, eps=0.02, norm_inf=True)
        assert np.array_equal(fs, fu)
